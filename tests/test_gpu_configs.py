"""Parity of the HIP engine with the CPU oracle at the benchmark
configurations' real sizes (not toy layouts).

* per-call A.x, A^T.y and A(A^T u) at 1e-12 normwise relative on the full
  layouts of: the N2 (10e,8o) D2h active space, synthetic n=24 C1 DQG, H_8 /
  H_12 / H_16 chains (the bond-scan sizes), H_50 (50e,50o) D2h DQG, and
  synthetic n=16 C1 DQGT2 (T2 blocks of 6016);
* Update_xz, cold and then warm from the kept eigenvectors, at 1e-10 on
  H_50's blocks (up to 2500, padded 2560: two lanes, 1024-thread pairs) and on
  the 6016-wide T2aaa block of n=16 DQGT2 (checked against numpy's eigh of that
  block's M).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle_py as orc  # noqa: E402

pytestmark = pytest.mark.gpu

CONFIGS = {
    # name: (amopi, nalpha, nbeta, positivity)
    "n2_10e8o_d2h": ([2, 0, 1, 1, 0, 2, 1, 1], 5, 5, "DQG"),
    "syn24_c1": ([24], 12, 12, "DQG"),
    "h8": ([4, 0, 0, 0, 0, 4, 0, 0], 4, 4, "DQG"),
    "h12": ([6, 0, 0, 0, 0, 6, 0, 0], 6, 6, "DQG"),
    "h16": ([8, 0, 0, 0, 0, 8, 0, 0], 8, 8, "DQG"),
    "h50": ([25, 0, 0, 0, 0, 25, 0, 0], 25, 25, "DQG"),
    "syn16_dqgt2": ([16], 8, 8, "DQGT2"),
}


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _pair(name):
    from v2rdm_casscf_amd import V2RDMSolver
    amopi, na, nb, pos = CONFIGS[name]
    print(f"[config {name}] amopi={amopi} na={na} nb={nb} {pos}", flush=True)
    s = V2RDMSolver(amopi, na, nb, 1, pos)
    L = orc.Layout(amopi, na, nb, 1, pos)
    assert (s.dimx, s.nconstraints) == (L.dimx, L.ncon)
    assert s.dimensions == L.block_dims
    return s, L


@pytest.mark.parametrize("name", list(CONFIGS))
def test_maps_at_config_size(name):
    s, L = _pair(name)
    rng = np.random.default_rng(5)
    x = rng.normal(size=L.dimx)
    y = rng.normal(size=L.ncon)
    ax = s.bpsdp_Au(torch.empty(L.ncon, dtype=torch.float64, device="cuda"), _dev(x)).cpu().numpy()
    ref = L.Au(x)
    fam = {nm: _rel(ax[st:st + cnt], ref[st:st + cnt]) for nm, st, cnt in L.row_families()}
    assert _rel(ax, ref) < 1e-12, fam
    del ax, ref
    aty = s.bpsdp_ATu(torch.empty(L.dimx, dtype=torch.float64, device="cuda"), _dev(y)).cpu().numpy()
    assert _rel(aty, L.ATu(y)) < 1e-12
    del aty
    out = s.cg_Ax(L.ncon, torch.empty(L.ncon, dtype=torch.float64, device="cuda"), _dev(y))
    assert _rel(out.cpu().numpy(), L.cg_Ax(y)) < 1e-12


def test_update_xz_h50_cold_and_warm():
    """All 28 blocks of H_50 (2500, ten of 1250, ...): the default two-lane,
    1024-thread shape, cold to convergence, then warm from the kept
    eigenvectors after a small change of y."""
    s, L = _pair("h50")
    assert s.eig_shape() == (2, 1024)
    rng = np.random.default_rng(11)
    y = rng.normal(size=L.ncon)
    c = rng.normal(size=L.dimx)
    x0 = rng.normal(size=L.dimx)
    xr, zr = L.update_xz(0.8, y, c, x0)
    x, z = _dev(x0), _dev(np.zeros(L.dimx))
    s.update_xz(0.8, _dev(y), _dev(c), x, z)
    assert _rel(x.cpu().numpy(), xr) < 1e-10
    assert _rel(z.cpu().numpy(), zr) < 1e-10
    y2 = y + 1e-3 * rng.normal(size=L.ncon)
    xr2, zr2 = L.update_xz(0.8, y2, c, xr)
    s.update_xz(0.8, _dev(y2), _dev(c), x, z)
    assert _rel(x.cpu().numpy(), xr2) < 1e-10
    assert _rel(z.cpu().numpy(), zr2) < 1e-10


def _block_split(M, mu):
    w, V = np.linalg.eigh(0.5 * (M + M.T))
    xb = (V * (np.maximum(w, 0.0) / mu)) @ V.T
    return xb, mu * xb - 0.5 * (M + M.T)


def test_update_xz_t2_block_cold_and_warm():
    """n=16 C1 DQGT2: the 6016-wide T2aaa block (and every other block on the
    GPU) split cold, then warm; the T2aaa block is held against numpy's eigh
    of that block's M = sym(A^T y - c + mu x): x to 1e-10, z = mu x - M to
    3e-10 (a block 100 times wider than the 1e-10 cases; the backward errors
    of both eigensolvers grow with n, and z, the smaller part of M, carries
    x's absolute error times mu)."""
    s, L = _pair("syn16_dqgt2")
    dims = L.block_dims
    offs = np.concatenate([[0], np.cumsum([d * d for d in dims])]).astype(np.int64)
    bi = dims.index(max(dims))
    n, o = dims[bi], int(offs[bi])
    assert n == 6016
    mu = 0.8
    rng = np.random.default_rng(12)
    y = rng.normal(size=L.ncon)
    c = rng.normal(size=L.dimx)
    x0 = rng.normal(size=L.dimx)

    def m_block(yv, xv):
        aty = L.ATu(yv)
        return (aty[o:o + n * n] - c[o:o + n * n] + mu * xv[o:o + n * n]).reshape(n, n)

    xb, zb = _block_split(m_block(y, x0), mu)
    x, z = _dev(x0), _dev(np.zeros(L.dimx))
    s.update_xz(mu, _dev(y), _dev(c), x, z)
    xg = x[o:o + n * n].cpu().numpy().reshape(n, n)
    zg = z[o:o + n * n].cpu().numpy().reshape(n, n)
    assert _rel(xg, xb) < 1e-10
    assert _rel(zg, zb) < 3e-10
    y2 = y + 1e-3 * rng.normal(size=L.ncon)
    # the reference for the warm call starts from the x the GPU holds
    xfull = np.zeros(L.dimx)
    xfull[o:o + n * n] = xg.ravel()
    xb2, zb2 = _block_split(m_block(y2, xfull), mu)
    s.update_xz(mu, _dev(y2), _dev(c), x, z)
    assert _rel(x[o:o + n * n].cpu().numpy().reshape(n, n), xb2) < 1e-10
    assert _rel(z[o:o + n * n].cpu().numpy().reshape(n, n), zb2) < 3e-10
