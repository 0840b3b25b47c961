"""The algebra of the fused CG iteration, checked on the oracle's maps (CPU).

The GPU iteration (csrc/cg.hip with csrc/maps.hip) takes p.(A A^T p) as
||A^T p||^2 from the A^T launch and updates the residual row by row inside the
A.x launch instead of storing A p.  In exact arithmetic this is the recurrence
of CGSolver::solve (cg_solver.cc:125-177); here both forms run on the same
operator (the oracle's bpsdp_Au / bpsdp_ATu on real layouts) and must give the
same iterates to rounding, and the oracle's own CG must agree with both.  A
planted mistake (p.Ap taken as p.p, the operator dropped) must break the
agreement.
"""
import numpy as np
import pytest

import oracle_py as orc


def _cg(L, rhs, y0, tol, maxiter, fused, wrong=False):
    """Plain CG on A A^T y = rhs: break after the residual test, count after
    the p update (the reference's order)."""
    y = y0.copy()
    r = rhs - L.Au(L.ATu(y))
    p = r.copy()
    rz = r @ r
    it = 0
    while it < maxiter:
        t = L.ATu(p)
        if fused:
            pap = t @ t if not wrong else p @ p   # wrong: the operator dropped
            alpha = rz / pap
            r = r - alpha * L.Au(t)
        else:
            ap = L.Au(t)
            alpha = rz / (p @ ap)
            r = r - alpha * ap
        y = y + alpha * p
        rz_new = r @ r
        if np.sqrt(rz_new) < tol:
            break
        p = r + (rz_new / rz) * p
        rz = rz_new
        it += 1
    return y, it


@pytest.mark.parametrize("amopi,na,nb", [([4], 2, 2), ([2, 0, 0, 0, 0, 2, 0, 0], 2, 2),
                                         ([3, 1, 1, 0, 0, 1, 1, 1], 3, 3)])
def test_fused_form_matches_reference_form(amopi, na, nb):
    L = orc.Layout(amopi, na, nb, 1, "DQG")
    rng = np.random.default_rng(11)
    rhs = L.Au(rng.normal(size=L.dimx))        # consistent: in the range of A A^T
    y0 = rng.normal(size=L.ncon)
    ya, ita = _cg(L, rhs, y0, 1e-9, 400, fused=False)
    yb, itb = _cg(L, rhs, y0, 1e-9, 400, fused=True)
    yo, ito = L.cg_solve(rhs, y0, 1e-9, 400)
    assert abs(ita - itb) <= 1 and abs(ita - ito) <= max(2, ita // 10)
    # A^T y is what the outer loop uses; y is unique only up to null(A^T)
    ref = L.ATu(yo)
    for y in (ya, yb):
        assert np.linalg.norm(L.ATu(y) - ref) <= 1e-7 * np.linalg.norm(ref)
    # the iterates of the two forms agree to rounding, not just at the end
    y1, _ = _cg(L, rhs, y0, 0.0, 5, fused=False)
    y2, _ = _cg(L, rhs, y0, 0.0, 5, fused=True)
    assert np.linalg.norm(y1 - y2) <= 1e-12 * np.linalg.norm(y1)


def test_a_wrong_fused_denominator_is_caught():
    L = orc.Layout([4], 2, 2, 1, "DQG")
    rng = np.random.default_rng(12)
    rhs = L.Au(rng.normal(size=L.dimx))
    y0 = rng.normal(size=L.ncon)
    y1, _ = _cg(L, rhs, y0, 0.0, 5, fused=False)
    y2, _ = _cg(L, rhs, y0, 0.0, 5, fused=True, wrong=True)
    assert np.linalg.norm(y1 - y2) > 1e-6 * np.linalg.norm(y1)
