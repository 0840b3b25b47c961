"""Writes layout_sizes.json: the variable / constraint counts that the
reference's own test outputs print ("Total number of variables/constraints"),
with the active space each run used.  Reads the reference checkout (given as
argv[1], default /root/reference) only when regenerating the fixture.
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"

# (test dir, occurrence index, amopi, nalpha_active, nbeta_active, multiplicity,
#  positivity, year of the output file; 2016/2017 builds predate the two
#  maximal-spin row blocks added later (2 * gems_ab[0] rows), and the 2016
#  DQGT2 output predates today's T2 block layout)
CASES = [
    ("v2rdm1", 0, [1, 0, 1, 1, 0, 1, 1, 1], 3, 3, 1, "DQG", 2016),
    ("v2rdm2", 0, [1, 0, 1, 1, 0, 1, 1, 1], 3, 3, 1, "DQG", 2016),
    ("v2rdm4", 0, [0, 1, 1, 3, 1, 3, 0, 1], 5, 5, 1, "DQG", 2016),
    ("v2rdm4", 1, [0, 1, 1, 3, 1, 3, 0, 1], 6, 4, 3, "DQG", 2016),
    ("v2rdm6", 0, [1, 0, 1, 1, 0, 1, 1, 1], 3, 3, 1, "DQG", 2017),
    ("v2rdm7", 0, [6], 3, 3, 1, "DQ", 2018),
    ("v2rdm7", 1, [0, 0, 1, 2, 1, 2, 0, 0], 3, 3, 1, "DQ", 2018),
    ("v2rdm8", 0, [2], 1, 1, 1, "D", 2020),
]

out = []
for d, occ, amopi, na, nb, mult, pos, year in CASES:
    path = os.path.join(REF, "tests", d, "output.ref")
    txt = open(path).read().splitlines()
    nv = [(i + 1, int(re.search(r"(\d+)\s*$", l).group(1))) for i, l in enumerate(txt)
          if "Total number of variables" in l]
    nc = [(i + 1, int(re.search(r"(\d+)\s*$", l).group(1))) for i, l in enumerate(txt)
          if "Total number of constraints" in l]
    out.append(dict(source=f"tests/{d}/output.ref:{nv[occ][0]}-{nc[occ][0]}",
                    amopi=amopi, nalpha=na, nbeta=nb, multiplicity=mult,
                    positivity=pos, year=year, dimx=nv[occ][1], ncon=nc[occ][1]))
json.dump(out, open(os.path.join(os.path.dirname(__file__), "layout_sizes.json"), "w"),
          indent=1)
print(json.dumps(out, indent=1))
