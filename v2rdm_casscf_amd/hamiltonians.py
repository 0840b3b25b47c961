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
