"""Pins the CPU oracle (oracle/v2rdm_oracle.c) to the reference's algorithm.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign or index, a transposed operand, a wrong block size) fails it:
  * layout sizes against the counts the reference's own outputs print;
  * A.x = b for exact RDMs of brute-force FCI states, every block built as a
    Gram matrix of Fock-space vectors (never from the constraint formulas);
  * c.x = E_FCI (cost vector, tei.cc:144-313);
  * adjointness of A^T against A;
  * CG against a dense least-squares solve;
  * Update_xz against its closed form;
  * the full loop: N=2 and 2-hole DQG energies equal FCI (N-representability is
    exact there), any N gives a lower bound.
"""
import json
import os

import numpy as np
import pytest

import fci_tools as ft
import oracle_py as orc

HERE = os.path.dirname(os.path.abspath(__file__))

LAYOUTS = [
    # amopi, nalpha, nbeta, multiplicity, positivity, constrain_sz
    ([4], 2, 2, 1, "DQG", True),
    ([4], 2, 1, 2, "DQG", True),
    ([5], 3, 1, 3, "DQG", True),
    ([2, 1, 1, 0], 2, 2, 1, "DQG", True),
    ([1, 1, 1, 1, 0, 1, 0, 0], 2, 2, 1, "DQG", True),
    ([4], 2, 2, 1, "DQG", False),
    ([4], 2, 2, 1, "DQ", True),
    ([4], 2, 2, 1, "DG", True),
    ([4], 2, 2, 1, "DQGT2", True),
    ([5], 3, 2, 2, "DQGT2", True),
    ([2, 1, 1, 1], 2, 2, 1, "DQGT2", True),
    ([4], 2, 2, 1, "DQGT1", True),
    ([5], 3, 2, 2, "DQGT1", True),
    ([3, 2], 3, 2, 2, "DQGT1", True),
    ([2, 1, 1, 1], 3, 3, 1, "DQGT", True),
]


def _layout(amopi, na, nb, mult, pos, sz=True):
    return orc.Layout(amopi, na, nb, mult, pos, constrain_sz=sz)


def _ham(L, seed=0):
    n = L.amo
    irreps = [L.irrep(p) for p in range(n)]
    return ft.random_hamiltonian(n, seed, irreps if L.nirrep > 1 else None)


def test_layout_sizes_match_reference_outputs():
    cases = json.load(open(os.path.join(HERE, "golden", "layout_sizes.json")))
    for cs in cases:
        L = orc.Layout(cs["amopi"], cs["nalpha"], cs["nbeta"], cs["multiplicity"],
                       cs["positivity"])
        assert L.dimx == cs["dimx"], cs["source"]
        extra = 0
        if cs["year"] < 2018 and "G" in cs["positivity"]:
            extra = 2 * L.gems_ab(0)   # maximal-spin rows added after 2017
        assert L.ncon == cs["ncon"] + extra, cs["source"]


@pytest.mark.parametrize("amopi,na,nb,mult,pos,sz", LAYOUTS)
def test_exact_rdms_are_feasible(amopi, na, nb, mult, pos, sz):
    L = _layout(amopi, na, nb, mult, pos, sz)
    h, eri = _ham(L)
    F = ft.Fock(L.amo)
    E, psi = F.ground_state(h, eri, na, nb, s_target=(na - nb) / 2)
    x = ft.exact_x(F, psi, L)
    b = L.build_b()
    r = L.Au(x) - b
    for name, st, cnt in L.row_families():
        assert np.abs(r[st:st + cnt]).max() < 1e-11, name
    c = L.build_c(h, eri)
    assert abs(c @ x - E) < 1e-10


def test_feasibility_detects_a_planted_sign_error():
    """The feasibility check must fail when one family's coefficient is wrong:
    flip the sign of the D2aa block of x (equivalent to a sign slip in every
    row that reads it) and expect residuals."""
    L = _layout([4], 2, 2, 1, "DQG")
    h, eri = _ham(L)
    F = ft.Fock(L.amo)
    E, psi = F.ground_state(h, eri, 2, 2, s_target=0)
    x = ft.exact_x(F, psi, L)
    o = L.offset("D2AA", 0)
    g = L.gems_aa(0)
    x[o:o + g * g] *= -1.0
    assert np.abs(L.Au(x) - L.build_b()).max() > 1e-3


def test_t1_feasibility_detects_planted_sign_errors():
    """T1 rows read the mixed-spin G2 block (aab rows) and the G2 super block
    (aaa rows): a sign slip in either shows up as T1 residuals."""
    L = _layout([5], 3, 2, 2, "DQGT1")
    h, eri = _ham(L)
    F = ft.Fock(L.amo)
    E, psi = F.ground_state(h, eri, 3, 2, s_target=0.5)
    fam = {name: (st, cnt) for name, st, cnt in L.row_families()}
    for blk, rows in (("G2BA", "t1aab"), ("G2AA", "t1aaa"), ("Q2BB", "t1bba")):
        x = ft.exact_x(F, psi, L)
        st, cnt = fam[rows]
        assert np.abs((L.Au(x) - L.build_b())[st:st + cnt]).max() < 1e-11
        o = L.offset(blk, 0)
        g = {"G2BA": L.gems_ab(0), "G2AA": 2 * L.gems_ab(0), "Q2BB": L.gems_aa(0)}[blk]
        x[o:o + g * g] *= -1.0
        assert np.abs((L.Au(x) - L.build_b())[st:st + cnt]).max() > 1e-3, blk


@pytest.mark.parametrize("amopi,na,nb,mult,pos,sz", LAYOUTS)
def test_adjointness(amopi, na, nb, mult, pos, sz):
    L = _layout(amopi, na, nb, mult, pos, sz)
    rng = np.random.default_rng(1)
    x = rng.normal(size=L.dimx)
    y = rng.normal(size=L.ncon)
    lhs = L.Au(x) @ y
    rhs = x @ L.ATu(y)
    assert abs(lhs - rhs) < 1e-12 * (abs(lhs) + np.linalg.norm(x) * np.linalg.norm(y))


def test_cg_matches_dense_solve():
    L = _layout([3], 1, 1, 1, "DQG")
    rng = np.random.default_rng(2)
    A = np.stack([L.Au(e) for e in np.eye(L.dimx)], axis=1)   # ncon x dimx
    rhs = A @ rng.normal(size=L.dimx)                          # consistent rhs
    y, it = L.cg_solve(rhs, np.zeros(L.ncon), 1e-11, 5000)
    AAT = A @ A.T
    assert np.linalg.norm(AAT @ y - rhs) < 1e-9
    # A^T y is unique even though A A^T is singular (dependent rows)
    ylsq = np.linalg.lstsq(AAT, rhs, rcond=None)[0]
    assert np.allclose(A.T @ y, A.T @ ylsq, atol=1e-8)


def test_update_xz_closed_form():
    L = _layout([2, 1, 1, 0], 2, 2, 1, "DQG")
    rng = np.random.default_rng(3)
    y = rng.normal(size=L.ncon)
    c = rng.normal(size=L.dimx)
    x0 = rng.normal(size=L.dimx)
    mu = 0.7
    x, z = L.update_xz(mu, y, c, x0)
    M = L.ATu(y) - c + mu * x0
    off = 0
    for n in L.block_dims:
        if n == 0:
            continue
        blk = M[off:off + n * n].reshape(n, n)
        S = 0.5 * (blk + blk.T)
        xb = x[off:off + n * n].reshape(n, n)
        zb = z[off:off + n * n].reshape(n, n)
        w, V = np.linalg.eigh(S)
        assert np.allclose(mu * xb - zb, S, atol=1e-12)
        assert np.allclose(xb, (V * np.maximum(w, 0)) @ V.T / mu, atol=1e-12)
        assert np.linalg.eigvalsh(0.5 * (xb + xb.T)).min() > -1e-12
        assert np.linalg.eigvalsh(0.5 * (zb + zb.T)).min() > -1e-12
        assert np.abs(xb @ zb).max() < 1e-10
        off += n * n


def test_random_guess_is_glibc_rand_sequence():
    """srand(0); rand() -- glibc treats seed 0 as 1; first outputs are the
    published glibc TYPE_3 values 1804289383, 846930886, 1681692777."""
    L = _layout([2], 1, 1, 1, "D")
    x, z, y = L.guess_random()
    RAND_MAX = 2147483647
    first = [1804289383, 846930886, 1681692777]
    assert np.isclose(x[0], (first[0] / RAND_MAX - 1.0) * 2.0)
    assert np.isclose(z[0], (first[1] / RAND_MAX - 1.0) * 2.0)
    assert np.isclose(x[1], (first[2] / RAND_MAX - 1.0) * 2.0)
    assert x.min() >= -2.0 and x.max() <= 0.0


def _loop_energy(amopi, na, nb, pos, seed=0, conv=1e-5, **kw):
    L = orc.Layout(amopi, na, nb, None, pos, **kw)
    h, eri = _ham(L, seed)
    F = ft.Fock(L.amo)
    E, _ = F.ground_state(h, eri, na, nb, s_target=(na - nb) / 2)
    c = L.build_c(h, eri)
    b = L.build_b()
    x, z, y = L.guess_random()
    st = L.bpsdp(c, b, x, z, y, r_conv=conv, e_conv=conv, maxiter=20000)
    return E, st


def test_two_electrons_dqg_is_exact():
    E, st = _loop_energy([4], 1, 1, "DQG")
    assert st["converged"]
    assert abs(st["energy_primal"] - E) < 1e-4


def test_two_electrons_d_is_exact():
    E, st = _loop_energy([4], 1, 1, "D")
    assert st["converged"]
    assert abs(st["energy_primal"] - E) < 1e-4


def test_two_holes_dqg_is_exact():
    E, st = _loop_energy([4], 3, 3, "DQG")
    assert st["converged"]
    assert abs(st["energy_primal"] - E) < 1e-4


def test_resumed_loop_is_the_same_iteration():
    """A solve split into two calls (the second resuming mu, the iteration
    count and the errors of the first) walks the same trajectory as one call:
    same outer iteration count (to a few iterations in 7 000: the OpenMP
    reductions do not fix their summation order, so two runs drift apart in
    the last bits) and energy."""
    L = orc.Layout([4], 2, 2, None, "DQG")
    h, eri = _ham(L, 0)
    c, b = L.build_c(h, eri), L.build_b()
    x, z, y = L.guess_random()
    one = L.bpsdp(c, b, x.copy(), z.copy(), y.copy(), maxiter=20000)
    first = L.bpsdp(c, b, x, z, y, maxiter=150)
    assert first["oiter"] == 150 and not first["converged"]
    second = L.bpsdp(c, b, x, z, y, maxiter=20000, resume=first)
    assert one["converged"] and second["converged"]
    assert abs(second["oiter"] - one["oiter"]) <= max(2, one["oiter"] // 500)
    assert abs(second["energy_primal"] - one["energy_primal"]) < 1e-8


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    n, nin, nh = 13, 6, 8
    f = rng.normal(size=(n, nin)).astype(np.float32)
    t = rng.normal(size=n).astype(np.float32)
    wt = rng.uniform(0.5, 1.5, n).astype(np.float32)
    W1 = rng.normal(0, 0.5, (nin, nh)).astype(np.float32)
    b1 = rng.normal(0, 0.1, nh).astype(np.float32)
    W2 = rng.normal(0, 0.5, (nh, nh)).astype(np.float32)
    b2 = rng.normal(0, 0.1, nh).astype(np.float32)
    w3 = rng.normal(0, 0.5, nh).astype(np.float32)
    b3 = np.float32(0.1)
    loss, g = orc.mlp_loss_grad(f, t, wt, W1, b1, W2, b2, w3, b3)
    out = orc.mlp_forward(f, W1, b1, W2, b2, w3, b3)
    assert abs(loss - np.mean(wt * (out - t) ** 2)) < 1e-5
    params = [W1, b1, W2, b2, w3]
    for pi, P in enumerate(params):
        flat = P.reshape(-1)
        for k in range(0, flat.size, max(1, flat.size // 5)):
            eps = 1e-2
            old = flat[k]
            flat[k] = old + eps
            lp = orc.mlp_loss_grad(f, t, wt, W1, b1, W2, b2, w3, b3)[0]
            flat[k] = old - eps
            lm = orc.mlp_loss_grad(f, t, wt, W1, b1, W2, b2, w3, b3)[0]
            flat[k] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[pi][k]) < 2e-3 * (1 + abs(fd)), (pi, k)


# ---------------------------------------------------------------------------
# spin-adapted Q2 / G2 (SPIN_ADAPT_Q2 / _G2, q2.cc:260-670, g2.cc:235-856)
# ---------------------------------------------------------------------------

SA_LAYOUTS = [
    # amopi, nalpha, nbeta, positivity, constrain_spin, sa_q2, sa_g2, energy order
    ([4], 2, 2, "DQG", True, True, False, None),
    ([4], 2, 2, "DQG", True, True, False, [2, 0, 3, 1]),
    ([4], 2, 2, "DQG", False, True, True, [1, 3, 0, 2]),
    ([2, 1, 1, 0], 2, 2, "DQG", False, True, True, [0, 2, 1, 3]),
    ([1, 1, 1, 1, 0, 1, 0, 0], 2, 2, "DQ", True, True, False, [4, 0, 1, 2, 3]),
    ([5], 3, 3, "DG", False, False, True, None),
]


def _sa_layout(amopi, na, nb, pos, spin, saq, sag, eo):
    return orc.Layout(amopi, na, nb, 1, pos, constrain_spin=spin, spin_adapt_q2=saq,
                      spin_adapt_g2=sag, energy_order=eo)


@pytest.mark.parametrize("amopi,na,nb,pos,spin,saq,sag,eo", SA_LAYOUTS)
def test_exact_rdms_are_feasible_spin_adapted(amopi, na, nb, pos, spin, saq, sag, eo):
    """The spin-adapted blocks of an FCI singlet, each built as the Gram
    matrix of its singlet / triplet pair operators, satisfy every row."""
    L = _sa_layout(amopi, na, nb, pos, spin, saq, sag, eo)
    h, eri = _ham(L)
    F = ft.Fock(L.amo)
    E, psi = F.ground_state(h, eri, na, nb, s_target=0)
    x = ft.exact_x(F, psi, L)
    r = L.Au(x) - L.build_b()
    for name, st, cnt in L.row_families():
        assert np.abs(r[st:st + cnt]).max() < 1e-11, name
    assert abs(L.build_c(h, eri) @ x - E) < 1e-10


def test_spin_adapted_feasibility_detects_planted_errors():
    """Each planted mistake breaks its family's rows: the triplet sign of Q2t,
    singlet and triplet particle-hole blocks swapped, and the weight of the
    i = j singlet geminals (the pair operator of i = j is sqrt2 a+_ib a+_ia)."""
    L = _sa_layout([4], 2, 2, "DQG", False, True, True, [1, 3, 0, 2])
    h, eri = _ham(L)
    F = ft.Fock(L.amo)
    E, psi = F.ground_state(h, eri, 2, 2, s_target=0)
    fam = {name: (st, cnt) for name, st, cnt in L.row_families()}

    def resid(x, rows):
        st, cnt = fam[rows]
        return np.abs((L.Au(x) - L.build_b())[st:st + cnt]).max()

    x0 = ft.exact_x(F, psi, L)
    assert resid(x0, "q2t") < 1e-11 and resid(x0, "g2s") < 1e-11 and resid(x0, "q2s") < 1e-11
    x = x0.copy()
    o, g = L.offset("Q2T", 0), L.gems_aa(0)
    x[o:o + g * g] *= -1.0
    assert resid(x, "q2t") > 1e-3
    x = x0.copy()
    g = L.gems_ab(0)
    os_, ot = L.offset("G2S", 0), L.offset("G2T", 0)
    x[os_:os_ + g * g], x[ot:ot + g * g] = x0[ot:ot + g * g], x0[os_:os_ + g * g]
    assert resid(x, "g2s") > 1e-3
    x = x0.copy()
    g0 = L.gems_00(0)
    o = L.offset("Q2S", 0)
    diag = [a for a, (i, j) in enumerate(L.bas_00(0)) if i == j]
    for a in diag:
        for b_ in range(g0):
            x[o + a * g0 + b_] /= np.sqrt(2.0)
            x[o + b_ * g0 + a] /= np.sqrt(2.0)
    assert resid(x, "q2s") > 1e-3


def test_singlet_geminals_follow_the_energy_order():
    """Convention (basis.cc:626-651, any order is mathematically valid): the
    singlet geminals run over orbital-energy ranks, higher rank outer, pair
    (orbital of the higher rank, orbital of the lower); None = Pitzer order."""
    eo = [2, 0, 3, 1]
    L = _sa_layout([2, 2], 2, 2, "DQG", True, True, False, eo)
    sym = [L.irrep(p) for p in range(4)]
    for h in range(2):
        want = [(eo[a], eo[b]) for a in range(4) for b in range(a + 1) if sym[eo[a]] ^ sym[eo[b]] == h]
        assert L.bas_00(h) == want
    P = _sa_layout([3], 2, 2, "DQG", True, True, False, None)
    assert P.bas_00(0) == [(a, b) for a in range(3) for b in range(a + 1)]
    assert P.gems_00(0) == 6


@pytest.mark.parametrize("amopi,na,nb,pos,spin,saq,sag,eo", SA_LAYOUTS)
def test_adjointness_spin_adapted(amopi, na, nb, pos, spin, saq, sag, eo):
    L = _sa_layout(amopi, na, nb, pos, spin, saq, sag, eo)
    rng = np.random.default_rng(2)
    x = rng.normal(size=L.dimx)
    y = rng.normal(size=L.ncon)
    lhs = L.Au(x) @ y
    rhs = x @ L.ATu(y)
    assert abs(lhs - rhs) < 1e-12 * (abs(lhs) + np.linalg.norm(x) * np.linalg.norm(y))


def test_spin_adapted_sizes():
    """dimx / ncon of the spin-adapted layouts from their block counts
    (v2rdm_solver.cc:597-635, 976-1015): Q2 trades gab^2 + 2 gaa^2 for
    g00^2 + 3 gaa^2, G2 trades 6 gab^2 for 4 gab^2."""
    P = orc.Layout([3, 1, 2, 0], 3, 3, 1, "DQG", constrain_spin=False)
    Q = orc.Layout([3, 1, 2, 0], 3, 3, 1, "DQG", constrain_spin=False, spin_adapt_q2=True,
                   spin_adapt_g2=True)
    dq = sum(Q.gems_00(h) ** 2 + Q.gems_aa(h) ** 2 - Q.gems_ab(h) ** 2 for h in range(4))
    dg = sum(-2 * Q.gems_ab(h) ** 2 for h in range(4))
    assert Q.dimx - P.dimx == dq + dg and Q.ncon - P.ncon == dq + dg
    assert sum(Q.gems_00(h) for h in range(4)) == 6 * 7 // 2


def test_two_holes_dqg_is_exact_with_spin_adapted_q2():
    E, st = _loop_energy([4], 3, 3, "DQG", spin_adapt_q2=True, energy_order=[3, 1, 0, 2])
    assert st["converged"]
    assert abs(st["energy_primal"] - E) < 1e-4


def test_spin_adapted_energies_bracket():
    """Spin-adapted Q2 is the plain Q2 condition in a rotated basis (same
    energy); spin-adapted G2 keeps only the singlet and triplet projections of
    the like-spin particle-hole block (a relaxation: energy not above the
    plain one, both above the D-only bound and below FCI)."""
    kw = dict(seed=3, conv=1e-5)
    E, plain = _loop_energy([4], 2, 2, "DQG", **kw)
    _, saq = _loop_energy([4], 2, 2, "DQG", spin_adapt_q2=True, **kw)
    assert plain["converged"] and saq["converged"]
    assert abs(saq["energy_primal"] - plain["energy_primal"]) < 2e-5
    _, nospin = _loop_energy([4], 2, 2, "DQG", constrain_spin=False, **kw)
    _, sag = _loop_energy([4], 2, 2, "DQG", constrain_spin=False, spin_adapt_g2=True, **kw)
    _, donly = _loop_energy([4], 2, 2, "D", constrain_spin=False, **kw)
    assert sag["converged"] and nospin["converged"] and donly["converged"]
    assert sag["energy_primal"] <= nospin["energy_primal"] + 2e-5
    assert donly["energy_primal"] <= sag["energy_primal"] + 2e-5
    assert sag["energy_primal"] <= E + 2e-5
