"""v2g_cg_solve_noalloc (the CG solve that allocates nothing: its vectors live
in a static 64 MiB device array) against v2g_cg_solve, called through the C
ABI: the same kernels on the same inputs give the same bits.  Two calls in a
row, the second on a smaller problem (H12 then H8 layouts), then a problem
beyond the scratch limit, which must be rejected with a clear error."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle_py as orc  # noqa: E402

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _solve(s, fn, rhs, y0, precond):
    y = y0.clone()
    it, res = ctypes.c_int(), ctypes.c_double()
    torch.cuda.synchronize()
    rc = fn(s.h, _p(rhs), _p(y), ctypes.c_double(1e-9), 500, precond, 0, ctypes.byref(it),
            ctypes.byref(res))
    return rc, y, it.value, res.value


@pytest.mark.parametrize("precond", [0, 2])
def test_noalloc_cg_matches_cg(precond):
    from v2rdm_casscf_amd import V2RDMSolver
    from v2rdm_casscf_amd._lib import load
    lib = load()
    for n in (12, 8):                      # the second call on a smaller problem
        amopi = [n // 2, 0, 0, 0, 0, n // 2, 0, 0]
        print(f"[noalloc H{n} precond {precond}]", flush=True)
        s = V2RDMSolver(amopi, n // 2, n // 2, 1, "DQG")
        L = orc.Layout(amopi, n // 2, n // 2, 1, "DQG")
        rng = np.random.default_rng(40 + n)
        rhs = _dev(L.Au(rng.normal(size=L.dimx)))
        y0 = _dev(rng.normal(size=L.ncon))
        if precond == 2:
            s.set_precond_diag(_dev(rng.uniform(0.5, 2.0, size=L.ncon)))
        rc1, y1, it1, r1 = _solve(s, lib.v2g_cg_solve, rhs, y0, precond)
        rc2, y2, it2, r2 = _solve(s, lib.v2g_cg_solve_noalloc, rhs, y0, precond)
        assert rc1 == 0 and rc2 == 0, lib.v2g_last_error()
        assert it1 == it2 and r1 == r2
        assert torch.equal(y1, y2)
        assert r2 < 1e-9


def test_noalloc_cg_rejects_problems_beyond_the_scratch():
    from v2rdm_casscf_amd import V2RDMSolver
    from v2rdm_casscf_amd._lib import load
    lib = load()
    s = V2RDMSolver([24], 12, 12, 1, "DQG")   # 3 ncon + dimx = 12.4 M doubles > 8.4 M
    rhs = torch.zeros(s.nconstraints, dtype=torch.float64, device="cuda")
    y0 = torch.zeros_like(rhs)
    rc, _, _, _ = _solve(s, lib.v2g_cg_solve_noalloc, rhs, y0, 0)
    assert rc == -1
    assert b"allocation-free" in lib.v2g_last_error()
