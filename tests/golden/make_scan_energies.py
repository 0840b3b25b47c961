"""Writes scan_energies.json: converged BPSDP energies of the CPU oracle (plain
CG and a LAPACK split every outer iteration, R/E_CONVERGENCE 1e-6) for the
eight-point bond scan make_hchain_scan(8, seed=0): linear H_n / STO-3G chains,
n in {8, ..., 16}, spacings uniform in [0.7, 3.0] Angstrom, D2h, DQG + spin.
The batched GPU solver is checked point by point against them.

Run from the repository root (points given on the command line are
recomputed and merged):  python tests/golden/make_scan_energies.py [index ...]
Stretched chains converge slowly: points 2, 3, 5 and 6 took 52 000 - 134 000
outer iterations and 10 - 30 minutes each on 8 cores; points 0 and 1 (H16 at
2.57 A, H14 at 2.80 A) need hours each and are not in the file."""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.join(HERE, "..", "..")
sys.path[:0] = [os.path.join(ROOT, "oracle"), ROOT]

OUT = os.path.join(HERE, "scan_energies.json")


def main(which):
    import oracle_py as orc
    from v2rdm_casscf_amd.bondscan import make_hchain_scan
    from v2rdm_casscf_amd.hamiltonians import ANGSTROM_TO_BOHR, hydrogen_chain
    pts = make_hchain_scan(8, seed=0)
    for i in which:
        n, r = pts[i]
        d = hydrogen_chain(n, r * ANGSTROM_TO_BOHR)
        L = orc.Layout(d["amopi"], n // 2, n // 2, None, "DQG")
        c, b = L.build_c(d["h1"], d["eri"]), L.build_b()
        x, z, y = L.guess_random()
        t0 = time.time()
        res = L.bpsdp(c, b, x, z, y, r_conv=1e-6, e_conv=1e-6, maxiter=200000)
        rec = dict(index=i, n=n, spacing_angstrom=r, energy=res["energy_primal"] + d["enuc"],
                   oiter=res["oiter"], converged=res["converged"], dimx=L.dimx, ncon=L.ncon,
                   oracle_seconds=round(time.time() - t0, 1))
        print(rec, flush=True)
        old = json.load(open(OUT)) if os.path.exists(OUT) else {}
        old[str(i)] = rec
        with open(OUT, "w") as f:
            json.dump(old, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main([int(a) for a in sys.argv[1:]] or list(range(8)))
