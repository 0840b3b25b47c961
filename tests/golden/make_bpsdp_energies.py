"""Writes bpsdp_energies.json: converged BPSDP energies of the CPU oracle for
a small random active space under D, DQG, DQGT1, DQGT2 and DQGT (T1 + T2)
positivity, plus the brute-force FCI energy of the same Hamiltonian.

The GPU tests compare the engine's converged energies with these values; the
CPU tests check the bound chain E(D) <= E(DQG) <= E(DQGT*) <= E(FCI).
Run from the repository root:  python tests/golden/make_bpsdp_energies.py
(about five minutes on 8 cores).
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(HERE, "..", "..", "oracle"), os.path.join(HERE, "..")]

import fci_tools as ft  # noqa: E402
import oracle_py as orc  # noqa: E402

CASES = [dict(amopi=[4], nalpha=2, nbeta=1, seed=0)]
POSITIVITY = ["D", "DQG", "DQGT1", "DQGT2", "DQGT"]


def main():
    out = []
    for cs in CASES:
        n = sum(cs["amopi"])
        h1, eri = ft.random_hamiltonian(n, cs["seed"], None)
        F = ft.Fock(n)
        efci, _ = F.ground_state(h1, eri, cs["nalpha"], cs["nbeta"],
                                 s_target=(cs["nalpha"] - cs["nbeta"]) / 2)
        rec = dict(cs, fci=efci, r_convergence=1e-6, e_convergence=1e-6, energies={})
        for pos in POSITIVITY:
            L = orc.Layout(cs["amopi"], cs["nalpha"], cs["nbeta"], None, pos)
            c, b = L.build_c(h1, eri), L.build_b()
            x, z, y = L.guess_random()
            r = L.bpsdp(c, b, x, z, y, r_conv=1e-6, e_conv=1e-6, maxiter=40000)
            assert r["converged"], pos
            rec["energies"][pos] = dict(energy=r["energy_primal"], oiter=r["oiter"],
                                        dimx=L.dimx, ncon=L.ncon)
            print(pos, rec["energies"][pos], flush=True)
        out.append(rec)
    with open(os.path.join(HERE, "bpsdp_energies.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
