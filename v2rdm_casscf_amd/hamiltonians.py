"""Active-space Hamiltonians for the BPSDP engine (no Psi4 needed).

synthetic_cholesky(n, seed): the seeded synthetic model of the benchmark
configurations -- h symmetric with diagonal -linspace(2.0, 0.2, n) and
off-diagonal N(0, 0.05); (pq|rs) = sum_L B^L_pq B^L_rs with L = 3n symmetric
B^L whose upper triangle (diagonal included) is N(0, 0.1).  numpy
default_rng(seed) draws h's upper triangle first, then B^0, B^1, ...

All integrals are returned in chemists' notation, full arrays (n x n and
n^4), Pitzer order, which is what V2RDMSolver.GetIntegrals consumes.
"""
import numpy as np


def synthetic_cholesky(n, seed=0, nL=None, sigma_h=0.05, sigma_b=0.1):
    rng = np.random.default_rng(seed)
    h = np.diag(-np.linspace(2.0, 0.2, n))
    iu = np.triu_indices(n, 1)
    h[iu] = rng.normal(0.0, sigma_h, len(iu[0]))
    h = np.triu(h) + np.triu(h, 1).T
    nL = 3 * n if nL is None else nL
    B = np.zeros((nL, n, n))
    iu0 = np.triu_indices(n)
    for L in range(nL):
        B[L][iu0] = rng.normal(0.0, sigma_b, len(iu0[0]))
        B[L] = np.triu(B[L]) + np.triu(B[L], 1).T
    Bm = B.reshape(nL, n * n)
    eri = (Bm.T @ Bm).reshape(n, n, n, n)
    return h, eri


# ---------------------------------------------------------------------------
# Linear H_n chains in STO-3G (closed-form s-Gaussian integrals)
# ---------------------------------------------------------------------------
# STO-3G hydrogen 1s (zeta = 1.24): exponents and contraction coefficients of
# the published basis set.
_STO3G_H_ALPHA = np.array([3.42525091, 0.62391373, 0.16885540])
_STO3G_H_COEF = np.array([0.15432897, 0.53532814, 0.44463454])
ANGSTROM_TO_BOHR = 1.8897261246257702


def _boys0(t):
    from scipy.special import erf
    t = np.asarray(t, dtype=np.float64)
    out = np.ones_like(t)
    m = t > 1e-12
    st = np.sqrt(t[m])
    out[m] = 0.5 * np.sqrt(np.pi) * erf(st) / st
    small = ~m
    out[small] = 1.0 - t[small] / 3.0
    return out


def sto3g_chain_ao(n, spacing_bohr):
    """AO integrals of H_n on the z axis, centred at the origin.

    Returns S, T+V (core Hamiltonian), ERI (mu nu|la si) full array, E_nuc."""
    z = (np.arange(n) - (n - 1) / 2.0) * spacing_bohr
    a = _STO3G_H_ALPHA
    c = _STO3G_H_COEF * (2.0 * a / np.pi) ** 0.75
    # primitive data: index (atom, prim)
    pz = np.repeat(z, 3)
    pa = np.tile(a, n)
    pc = np.tile(c, n)
    own = np.repeat(np.arange(n), 3)
    A, B = np.meshgrid(np.arange(3 * n), np.arange(3 * n), indexing="ij")
    p = pa[A] + pa[B]
    mu = pa[A] * pa[B] / p
    dAB2 = (pz[A] - pz[B]) ** 2
    K = np.exp(-mu * dAB2)
    Pz = (pa[A] * pz[A] + pa[B] * pz[B]) / p
    cc = pc[A] * pc[B]
    Sp = cc * (np.pi / p) ** 1.5 * K
    Tp = cc * mu * (3.0 - 2.0 * mu * dAB2) * (np.pi / p) ** 1.5 * K
    Vp = np.zeros_like(Sp)
    for C in range(n):
        Vp -= cc * 2.0 * np.pi / p * K * _boys0(p * (Pz - z[C]) ** 2)

    def contract2(M):
        return np.add.reduceat(np.add.reduceat(M, np.arange(0, 3 * n, 3), axis=0),
                               np.arange(0, 3 * n, 3), axis=1)

    S = contract2(Sp)
    H = contract2(Tp + Vp)
    # ERI over unique contracted pairs (mu <= nu)
    iu = np.triu_indices(n)
    npair = len(iu[0])
    # primitive pair blocks for each contracted pair: 9 primitives
    prim_idx = np.array([[3 * m + i for i in range(3)] for m in range(n)])
    PA = prim_idx[iu[0]][:, :, None].repeat(3, axis=2).reshape(npair, 9)
    PB = prim_idx[iu[1]][:, None, :].repeat(3, axis=1).reshape(npair, 9)
    pp = p[PA, PB]
    pK = (cc * K)[PA, PB]
    pP = Pz[PA, PB]
    eri_pair = np.zeros((npair, npair))
    chunk = max(1, 4_000_000 // (npair * 81))
    pref = 2.0 * np.pi ** 2.5
    for s in range(0, npair, chunk):
        e = slice(s, min(npair, s + chunk))
        p1 = pp[e][:, None, :, None]
        q1 = pp[None, :, None, :]
        ptot = p1 + q1
        t = p1 * q1 / ptot * (pP[e][:, None, :, None] - pP[None, :, None, :]) ** 2
        val = pref / (p1 * q1 * np.sqrt(ptot)) * pK[e][:, None, :, None] * pK[None, :, None, :] \
            * _boys0(t)
        eri_pair[e] = val.sum(axis=(2, 3))
    eri = np.zeros((n, n, n, n))
    P0, P1 = iu
    full = np.zeros((n, n, npair))
    full[P0, P1] = np.eye(npair)
    full[P1, P0] = np.eye(npair)
    tmp = full.reshape(n * n, npair) @ eri_pair          # (mn, pair)
    eri = (tmp @ full.reshape(n * n, npair).T).reshape(n, n, n, n)
    enuc = 0.0
    for i in range(n):
        for j in range(i):
            enuc += 1.0 / abs(z[i] - z[j])
    return S, H, eri, enuc


def _rhf(S, H, eri, nocc, maxiter=500, tol=1e-10):
    """Closed-shell RHF with DIIS; returns orbital energies and coefficients."""
    w, U = np.linalg.eigh(S)
    X = U @ np.diag(w ** -0.5) @ U.T
    e, Cp = np.linalg.eigh(X.T @ H @ X)
    C = X @ Cp
    D = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    focks, errs = [], []
    for it in range(maxiter):
        J = np.einsum("pqrs,rs->pq", eri, D)
        Kx = np.einsum("prqs,rs->pq", eri, D)
        F = H + J - 0.5 * Kx
        err = F @ D @ S - S @ D @ F
        focks.append(F)
        errs.append(err)
        if len(focks) > 8:
            focks.pop(0)
            errs.pop(0)
        if np.abs(err).max() < tol:
            break
        nb = len(focks)
        Bm = -np.ones((nb + 1, nb + 1))
        Bm[-1, -1] = 0.0
        for i in range(nb):
            for j in range(nb):
                Bm[i, j] = np.vdot(errs[i], errs[j])
        rhs = np.zeros(nb + 1)
        rhs[-1] = -1.0
        try:
            cf = np.linalg.solve(Bm, rhs)[:nb]
            F = sum(ci * Fi for ci, Fi in zip(cf, focks))
        except np.linalg.LinAlgError:
            pass
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        C = X @ Cp
        D = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    return e, C


def hydrogen_chain(n, spacing_bohr):
    """Full-valence (n e, n o) active space of linear H_n / STO-3G in D2h.

    The chain lies on z (as Psi4 orients linear molecules), so the 1s
    functions span Ag (symmetric under inversion) and B1u (antisymmetric).
    RHF orbitals are computed per irrep from inversion-adapted combinations.
    Returns dict(amopi, h1, eri, enuc, nalpha, nbeta, e_rhf); orbitals in
    Pitzer order (all Ag, then all B1u), integrals in chemists' notation."""
    if n % 2:
        raise ValueError("even chains only (closed shell, no centre atom)")
    S, H, eri, enuc = sto3g_chain_ao(n, spacing_bohr)
    m = n // 2
    # inversion-adapted combinations: g_k = (phi_k + phi_{n-1-k}), u_k = (phi_k - phi_{n-1-k})
    G = np.zeros((n, m))
    U = np.zeros((n, m))
    for k in range(m):
        G[k, k] = G[n - 1 - k, k] = 1.0
        U[k, k] = 1.0
        U[n - 1 - k, k] = -1.0
    e_all, C_ao = _rhf(S, H, eri, n // 2)
    # per-irrep RHF: project the converged Fock operator onto each irrep
    nocc = n // 2
    D = 2.0 * C_ao[:, :nocc] @ C_ao[:, :nocc].T
    J = np.einsum("pqrs,rs->pq", eri, D)
    Kx = np.einsum("prqs,rs->pq", eri, D)
    F = H + J - 0.5 * Kx
    cols, eps = [], []
    for B in (G, U):
        Sb = B.T @ S @ B
        Fb = B.T @ F @ B
        w, V = np.linalg.eigh(Sb)
        Xb = V @ np.diag(w ** -0.5) @ V.T
        e, Cp = np.linalg.eigh(Xb.T @ Fb @ Xb)
        cols.append(B @ Xb @ Cp)
        eps.append(e)
    C = np.hstack(cols)                     # Pitzer order: Ag then B1u
    h1 = C.T @ H @ C
    g = np.einsum("pqrs,pi->iqrs", eri, C)
    g = np.einsum("iqrs,qj->ijrs", g, C)
    g = np.einsum("ijrs,rk->ijks", g, C)
    g = np.einsum("ijks,sl->ijkl", g, C)
    e_rhf = 0.0
    occ = np.argsort(np.concatenate(eps))[:nocc]
    for i in occ:
        e_rhf += 2.0 * h1[i, i]
        for j in occ:
            e_rhf += 2.0 * g[i, i, j, j] - g[i, j, i, j]
    amopi = [m, 0, 0, 0, 0, m, 0, 0]
    return dict(amopi=amopi, h1=h1, eri=g, enuc=enuc, nalpha=n // 2, nbeta=n // 2,
                e_rhf=e_rhf + enuc, orbital_energies=np.concatenate(eps))
