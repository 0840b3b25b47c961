"""Variational bound chain of the oracle's converged energies (tests/golden):
E(D) <= E(DQG) <= E(DQGT1), E(DQGT2) <= E(DQGT) <= E(FCI).

Each added positivity condition is satisfied by the exact RDMs and shrinks the
feasible set, so the lower bound can only rise and must stay below FCI (the
rows themselves are pinned by exact-RDM feasibility in test_oracle_pinning).
"""
import json
import os

import oracle_py as orc

HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 2e-6   # R/E_CONVERGENCE 1e-6 of each run


def _cases():
    return json.load(open(os.path.join(HERE, "golden", "bpsdp_energies.json")))


def test_bound_chain():
    for cs in _cases():
        e = {k: v["energy"] for k, v in cs["energies"].items()}
        assert e["D"] <= e["DQG"] + TOL
        assert e["DQG"] <= e["DQGT1"] + TOL
        assert e["DQG"] <= e["DQGT2"] + TOL
        if "DQGT" in e:
            assert max(e["DQGT1"], e["DQGT2"]) <= e["DQGT"] + TOL
        for k, v in e.items():
            assert v <= cs["fci"] + TOL, k


def test_golden_sizes_match_oracle_layout():
    for cs in _cases():
        for pos, v in cs["energies"].items():
            L = orc.Layout(cs["amopi"], cs["nalpha"], cs["nbeta"], None, pos)
            assert (L.dimx, L.ncon) == (v["dimx"], v["ncon"]), pos
