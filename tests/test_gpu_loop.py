"""The BPSDP loop as it runs by default (tracked eigen split: one block-Jacobi
sweep with one cyclic sweep per pair sub-problem per outer iteration,
extrapolated CG start) against the oracle's converged energies
(tests/golden/converged_energies.json, written by
tests/golden/make_converged_energies.py: plain CG and a LAPACK split every
iteration), 1e-6 Eh at R/E_CONVERGENCE 1e-6.

Also: every launch shape of the eigen split gives the same split bit for bit,
the autotuner settles on one of them without changing results, and a solve
continued by a second call (resume) reaches the same energy.
"""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "converged_energies.json")))


def _inputs(name):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_converged_energies import case_inputs, checksum
    if name not in GOLD:
        pytest.skip(f"no oracle energy for {name} in tests/golden/converged_energies.json")
    amopi, na, nb, h1, eri, enuc = case_inputs(name)
    g = GOLD[name]
    assert abs(checksum(h1, eri) - g["checksum"]) <= 1e-10 * g["checksum"], "integrals changed"
    return amopi, na, nb, h1, eri, enuc, g


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("name,shape", [
    ("n2_layout", None),
    ("c1_9", None),
    ("c1_9", (2, 1024)),    # the launch shape of the large problems (H50)
    ("h8", None),
    ("h10", None),
    ("h12", None),
    ("h12", (2, 1024)),
])
def test_loop_energy_matches_oracle(name, shape):
    from v2rdm_casscf_amd import V2RDMSolver
    amopi, na, nb, h1, eri, enuc, g = _inputs(name)
    print(f"[loop {name} shape {shape}]", flush=True)
    s = V2RDMSolver(amopi, na, nb, 1, "DQG", options=dict(MAXITER=200000))
    assert (s.dimx, s.nconstraints) == (g["dimx"], g["ncon"])
    if shape:
        s.set_eig_shape(*shape)
        assert s.eig_shape() == shape
    out = s.compute_energy(h1, eri, enuc=enuc)
    print(f"  E {out['energy']:.10f} (oracle {g['energy']:.10f}), {out['oiter']} outer "
          f"(oracle {g['oiter']}), exact re-splits {out['exact_splits']}", flush=True)
    assert out["converged"]
    assert abs(out["energy"] - g["energy"]) < 1e-6
    assert out["ep"] < 1e-6 and out["ed"] < 1e-6
    if max(s.dimensions) > 64:
        # blocks beyond 64 are tracked with one sweep per iteration: the
        # accepted iteration's split was redone to convergence
        assert out["exact_splits"] >= 1
    else:
        # every block is solved whole (to convergence) in every iteration
        assert out["exact_splits"] == 0


SHAPES = [(1, 256), (1, 1024), (2, 256), (2, 1024)]


@pytest.mark.parametrize("name", ["c1_9", "h16"])
def test_eig_shapes_bit_identical(name):
    """Cold then warm Update_xz under every launch shape: identical bits."""
    import oracle_py as orc
    from v2rdm_casscf_amd import V2RDMSolver
    amopi = [9] if name == "c1_9" else [8, 0, 0, 0, 0, 8, 0, 0]
    na = 4 if name == "c1_9" else 8
    L = orc.Layout(amopi, na, na, 1, "DQG")
    rng = np.random.default_rng(31)
    y, c, x0 = rng.normal(size=L.ncon), rng.normal(size=L.dimx), rng.normal(size=L.dimx)
    y2 = y + 1e-3 * rng.normal(size=L.ncon)
    res = {}
    for shape in SHAPES:
        s = V2RDMSolver(amopi, na, na, 1, "DQG")
        s.set_eig_shape(*shape)
        x, z = _dev(x0), _dev(np.zeros(L.dimx))
        s.update_xz(0.9, _dev(y), _dev(c), x, z)
        s.update_xz(0.9, _dev(y2), _dev(c), x, z)
        res[shape] = (x.cpu().numpy(), z.cpu().numpy())
        del s
    xr, zr = L.update_xz(0.9, y2, c, L.update_xz(0.9, y, c, x0)[0])
    for shape in SHAPES:
        assert np.array_equal(res[shape][0], res[SHAPES[0]][0]), shape
        assert np.array_equal(res[shape][1], res[SHAPES[0]][1]), shape
    assert np.linalg.norm(res[SHAPES[0]][0] - xr) < 1e-10 * np.linalg.norm(xr)


def test_autotune_settles_without_changing_results():
    """The autotuner times every shape on the first warm splits (two each)
    and keeps one; the splits it makes meanwhile equal a fixed shape's."""
    from v2rdm_casscf_amd import V2RDMSolver
    amopi, na = [8, 0, 0, 0, 0, 8, 0, 0], 8
    rng = np.random.default_rng(32)
    tuned = V2RDMSolver(amopi, na, na, 1, "DQG")
    fixed = V2RDMSolver(amopi, na, na, 1, "DQG")
    fixed.set_eig_shape(1, 256)
    n, m = tuned.nconstraints, tuned.dimx
    c, x0 = rng.normal(size=m), rng.normal(size=m)
    xs = {id(tuned): _dev(x0), id(fixed): _dev(x0)}
    zs = {id(tuned): _dev(np.zeros(m)), id(fixed): _dev(np.zeros(m))}
    y = rng.normal(size=n)
    tuned.autotune(True)
    for k in range(10):
        y = y + 1e-4 * rng.normal(size=n)
        for s in (tuned, fixed):
            s.update_xz(1.0, _dev(y), _dev(c), xs[id(s)], zs[id(s)])
        assert torch.equal(xs[id(tuned)], xs[id(fixed)]), k
        assert torch.equal(zs[id(tuned)], zs[id(fixed)]), k
    assert tuned.eig_shape() in SHAPES


def test_resumed_solve_reaches_the_same_energy():
    """compute_energy(resume=True) continues mu, the iteration count and the
    errors of the previous call; split into two calls, H8 converges to the
    oracle's energy in about as many outer iterations as one call."""
    from v2rdm_casscf_amd import V2RDMSolver
    amopi, na, nb, h1, eri, enuc, g = _inputs("h8")
    s = V2RDMSolver(amopi, na, nb, 1, "DQG", options=dict(MAXITER=700))
    first = s.compute_energy(h1, eri, enuc=enuc)
    assert first["oiter"] == 700 and not first["converged"]
    s.options["MAXITER"] = 200000
    out = s.compute_energy(enuc=enuc, resume=True)
    assert out["converged"]
    assert abs(out["energy"] - g["energy"]) < 1e-6
    one = V2RDMSolver(amopi, na, nb, 1, "DQG", options=dict(MAXITER=200000))
    ref = one.compute_energy(h1, eri, enuc=enuc)
    assert abs(out["oiter"] - ref["oiter"]) <= 0.05 * ref["oiter"] + 10
