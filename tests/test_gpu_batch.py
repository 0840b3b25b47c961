"""The batched solver (v2g_batch_*): many problems in every launch.

* maps of a mixed batch (D2h chains, the N2 (10e,8o) active space, C1 with
  blocks above 64) against the oracle per problem, 1e-12;
* the device-built b, c and RANDOM guess of each problem equal the
  single-problem engine's bit for bit;
* converged energies of every problem against the oracle's
  (tests/golden/converged_energies.json, scan_energies.json), 1e-6 Eh at
  R/E_CONVERGENCE 1e-6, problems leaving the batch as they converge.
"""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle_py as orc  # noqa: E402

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _case(name):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_converged_energies import case_inputs
    return case_inputs(name)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_batch_maps_match_oracle_per_problem():
    from v2rdm_casscf_amd.batch import BatchSolver
    probs = [dict(amopi=[4, 0, 0, 0, 0, 4, 0, 0], nalpha=4, nbeta=4),
             dict(amopi=[2, 0, 1, 1, 0, 2, 1, 1], nalpha=5, nbeta=5),
             dict(amopi=[9], nalpha=4, nbeta=4),
             dict(amopi=[5], nalpha=3, nbeta=2, positivity="DQGT2"),
             dict(amopi=[6, 0, 0, 0, 0, 6, 0, 0], nalpha=6, nbeta=6)]
    bs = BatchSolver(probs)
    Ls = [orc.Layout(p["amopi"], p["nalpha"], p["nbeta"], None, p.get("positivity", "DQG"))
          for p in probs]
    assert [bs.xoff[q + 1] - bs.xoff[q] for q in range(len(probs))] == [L.dimx for L in Ls]
    assert [bs.yoff[q + 1] - bs.yoff[q] for q in range(len(probs))] == [L.ncon for L in Ls]
    rng = np.random.default_rng(50)
    x = rng.normal(size=bs.dimx)
    y = rng.normal(size=bs.ncon)
    ax = bs.Au(torch.from_numpy(x).cuda()).cpu().numpy()
    aty = bs.ATu(torch.from_numpy(y).cuda()).cpu().numpy()
    for q, L in enumerate(Ls):
        xs, ys = slice(bs.xoff[q], bs.xoff[q + 1]), slice(bs.yoff[q], bs.yoff[q + 1])
        assert _rel(ax[ys], L.Au(x[xs])) < 1e-12, q
        assert _rel(aty[xs], L.ATu(y[ys])) < 1e-12, q


def test_batch_inputs_equal_single_engine():
    from v2rdm_casscf_amd import V2RDMSolver
    from v2rdm_casscf_amd.batch import BatchSolver
    names = ["h8", "n2_layout", "c1_9", "h10"]
    cases = [_case(nm) for nm in names]
    bs = BatchSolver([dict(amopi=a, nalpha=na, nbeta=nb) for a, na, nb, _, _, _ in cases])
    bs.set_integrals([c[3] for c in cases], [c[4] for c in cases])
    for q, (a, na, nb, h1, eri, _) in enumerate(cases):
        s = V2RDMSolver(a, na, nb, 1, "DQG")
        s.GetIntegrals(h1, eri)
        s.BuildConstraints()
        s.Guess()
        assert torch.equal(bs.x_of(q, bs.c), s.c), names[q]
        assert torch.equal(bs.y_of(q, bs.b), s.b), names[q]
        assert torch.equal(bs.x_of(q), s.x) and torch.equal(bs.x_of(q, bs.z), s.z), names[q]
        assert torch.equal(bs.y_of(q), s.y), names[q]


def test_batch_energies_match_oracle():
    """Five problems of different sizes and symmetries in one batch."""
    from v2rdm_casscf_amd.batch import BatchSolver
    gold = json.load(open(os.path.join(HERE, "golden", "converged_energies.json")))
    names = ["h12", "c1_9", "h10", "n2_layout", "h8"]
    cases = [_case(nm) for nm in names]
    bs = BatchSolver([dict(amopi=a, nalpha=na, nbeta=nb) for a, na, nb, _, _, _ in cases])
    bs.set_integrals([c[3] for c in cases], [c[4] for c in cases])
    out = bs.compute_energy(enucs=[c[5] for c in cases])
    for nm, r in zip(names, out):
        print(f"[batch {nm}] E {r['energy']:.10f} oracle {gold[nm]['energy']:.10f} "
              f"outer {r['oiter']} (oracle {gold[nm]['oiter']}) exact re-splits {r['exact_splits']}",
              flush=True)
    for nm, r in zip(names, out):
        assert r["converged"], nm
        assert abs(r["energy"] - gold[nm]["energy"]) < 1e-6, nm
        assert r["ep"] < 1e-6 and r["ed"] < 1e-6, nm


def test_batch_scan_matches_oracle():
    """Points of the eight-point bond scan make_hchain_scan(8, seed=0) (H8 ...
    H16, 0.7-3.0 Angstrom) solved as one batch, every point against the
    oracle's converged energy.  Stretched chains take 10^5 outer iterations or
    more; the test takes the points of tests/golden/scan_energies.json whose
    oracle solve converged within 40 000."""
    from v2rdm_casscf_amd.batch import BatchSolver
    from v2rdm_casscf_amd.bondscan import make_hchain_scan
    from v2rdm_casscf_amd.hamiltonians import ANGSTROM_TO_BOHR, hydrogen_chain
    path = os.path.join(HERE, "golden", "scan_energies.json")
    gold = json.load(open(path)) if os.path.exists(path) else {}
    pts = make_hchain_scan(8, seed=0)
    which = sorted(int(k) for k, g in gold.items() if g["converged"] and g["oiter"] <= 40000)
    if len(which) < 2:
        pytest.skip("scan_energies.json holds fewer than two short points; "
                    "run tests/golden/make_scan_energies.py")
    for i in which:
        assert tuple(pts[i]) == (gold[str(i)]["n"], gold[str(i)]["spacing_angstrom"]), i
    hams = [hydrogen_chain(pts[i][0], pts[i][1] * ANGSTROM_TO_BOHR) for i in which]
    bs = BatchSolver([dict(amopi=d["amopi"], nalpha=d["nalpha"], nbeta=d["nbeta"]) for d in hams])
    bs.set_integrals([d["h1"] for d in hams], [d["eri"] for d in hams])
    out = bs.compute_energy(enucs=[d["enuc"] for d in hams])
    for i, r in zip(which, out):
        g = gold[str(i)]
        print(f"[scan {i}] H{g['n']} r={g['spacing_angstrom']:.3f} E {r['energy']:.10f} "
              f"oracle {g['energy']:.10f} outer {r['oiter']} (oracle {g['oiter']})", flush=True)
    for i, r in zip(which, out):
        assert r["converged"], i
        assert abs(r["energy"] - gold[str(i)]["energy"]) < 1e-6, i
