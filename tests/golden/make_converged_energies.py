"""Writes converged_energies.json: converged BPSDP energies of the CPU oracle
(the reference's plain CG and a LAPACK eigen split in every outer iteration)
at R/E_CONVERGENCE 1e-6 for the layouts the GPU loop is checked on:

  h8, h10, h12   linear H_n STO-3G chains, spacing 1.8 bohr, D2h, (n e, n o),
                 DQG + spin (the bond-scan and H50 family)
  n2_layout      the N2 (10e,8o) D2h active space, amopi [2,0,1,1,0,2,1,1],
                 5 alpha / 5 beta, with a seeded random D2h Hamiltonian
  c1_9           C1, 9 orbitals, 4 alpha / 4 beta, seeded random Hamiltonian
                 (blocks of 81 and 72: every block takes the block-Jacobi path)

Each record keeps the energy (enuc included for the chains), the outer
iteration count, dimx / ncon and a checksum of the integrals, so a test can
tell a changed Hamiltonian from a changed solver.

Run from the repository root:
    python tests/golden/make_converged_energies.py [case ...]
(h12 takes the longest, about 15 minutes on 8 cores; cases given on the
command line are recomputed and merged into the existing file).
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.join(HERE, "..", "..")
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(HERE, ".."), ROOT]

OUT = os.path.join(HERE, "converged_energies.json")


def case_inputs(name):
    """(amopi, nalpha, nbeta, h1, eri, enuc) of a named case; the GPU tests
    call this too, so both sides build identical integrals."""
    import fci_tools as ft
    if name.startswith("h"):
        from v2rdm_casscf_amd.hamiltonians import hydrogen_chain
        n = int(name[1:])
        d = hydrogen_chain(n, 1.8)
        return d["amopi"], n // 2, n // 2, d["h1"], d["eri"], d["enuc"]
    if name == "n2_layout":
        amopi = [2, 0, 1, 1, 0, 2, 1, 1]
        irreps = [h for h, m in enumerate(amopi) for _ in range(m)]
        h1, eri = ft.random_hamiltonian(8, 3, irreps)
        return amopi, 5, 5, h1, eri, 0.0
    if name == "c1_9":
        h1, eri = ft.random_hamiltonian(9, 4, None)
        return [9], 4, 4, h1, eri, 0.0
    raise ValueError(name)


def checksum(h1, eri):
    return float(np.sum(h1 * h1) + np.sum(eri * eri))


CASES = ["n2_layout", "c1_9", "h8", "h10", "h12"]


def main(names):
    import oracle_py as orc
    old = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in names:
        amopi, na, nb, h1, eri, enuc = case_inputs(name)
        L = orc.Layout(amopi, na, nb, None, "DQG")
        c, b = L.build_c(h1, eri), L.build_b()
        x, z, y = L.guess_random()
        t0 = time.time()
        r = L.bpsdp(c, b, x, z, y, r_conv=1e-6, e_conv=1e-6, maxiter=200000)
        assert r["converged"], name
        old[name] = dict(amopi=amopi, nalpha=na, nbeta=nb, positivity="DQG",
                         r_convergence=1e-6, e_convergence=1e-6,
                         energy=r["energy_primal"] + enuc, oiter=r["oiter"], dimx=L.dimx,
                         ncon=L.ncon, checksum=checksum(h1, eri),
                         oracle_seconds=round(time.time() - t0, 1))
        print(name, old[name], flush=True)
        with open(OUT, "w") as f:
            json.dump(old, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or CASES)
