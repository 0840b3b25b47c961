"""Writes oracle_loops.json: the converged oracle solves of the small cases in
tests/oracle_loops.py (plain CG, LAPACK split every outer iteration,
R/E_CONVERGENCE 1e-6).  Run from the repository root:
    python tests/golden/make_oracle_loops.py"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.join(HERE, "..", "..")
sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle"), ROOT]


def main():
    import oracle_loops as ol
    out = {}
    for key in ol.CASES:
        out[key] = ol.solve(key)
        print(key, out[key], flush=True)
    with open(ol.PATH, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
