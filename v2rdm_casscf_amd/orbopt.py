"""Orbital optimization around the GPU BPSDP solve: the v2RDM-CASSCF cycle.

The reference alternates its BPSDP iteration with orbital steps when
OPTIMIZE_ORBITALS is on (v2rdm_solver.cc:1896-1911): once the iteration has
converged, RotateOrbitals packs the spin-free active densities
(PackSpatialDensity, v2rdm_solver.cc:3664-3773), hands them with the
full-space integrals to its Fortran orbital optimizer (OrbOpt /
focas_interface, fortran.h:41-58), and RepackIntegrals rebuilds the cost
vector c from the rotated integrals (tei.cc:144-228); the BPSDP iteration
then continues from its x, y, z and mu until both the BPSDP errors and the
orbital step have converged.

Here the densities come off the device through v2g_pack_spatial_density (the
reference's packed layout), the orbital step is this module's own host code
(the Fortran optimizer is not part of this engine), c is rebuilt on the device
from the rotated active integrals with the core folded in (frozen_core,
tei.cc:229-313), and the solve resumes (V2RDMSolver.compute_energy(...,
resume=True)).

The orbital step is a quasi-Newton step on the non-redundant rotations of one
irrep (core-active, core-virtual, active-virtual; active-active when asked,
ORBOPT_ACTIVE_ACTIVE_ROTATIONS): gradient from the generalized Fock matrix,
diagonal Hessian model, backtracking on the exact energy with the densities
held fixed.  Orbitals rotate as C <- C expm(-kappa).
"""
import numpy as np
import scipy.linalg

from .fcidump import frozen_core


def unpack_spatial_density(d1, d2, amopi, bas00):
    """Dense spin-free active densities from the packed PackSpatialDensity
    layout: D1[p, q] (D1a + D1b) and G[p, q, r, s] in chemists' order with
    E = sum h_pq D1_pq + 1/2 sum (pq|rs) G_pqrs, symmetrized over the eight
    index permutations the integrals share.

    bas00[h]: the singlet geminals (i, j) of irrep h in their order.  A packed
    entry (ik, jl) holds half the sum of G over the distinct permutations of
    (ik|jl), so that the two-electron energy is sum_entries d2 (ik|jl)."""
    amo = sum(amopi)
    D1 = np.zeros((amo, amo))
    o, off = 0, 0
    for n in amopi:
        for i in range(n):
            for j in range(i, n):
                v = d1[o + j * (j + 1) // 2 + i]
                D1[off + i, off + j] = D1[off + j, off + i] = v
        o += n * (n + 1) // 2
        off += n
    G = np.zeros((amo,) * 4)
    base = 0
    for pairs in bas00:
        g0 = len(pairs)
        for a in range(g0):
            for b in range(a + 1):
                v = d2[base + a * (a + 1) // 2 + b]
                (i, k), (j, l) = pairs[a], pairs[b]
                perms = {(i, k, j, l), (k, i, j, l), (i, k, l, j), (k, i, l, j),
                         (j, l, i, k), (l, j, i, k), (j, l, k, i), (l, j, k, i)}
                # v = 1/2 sum over the distinct permutations; each permutation
                # of the symmetrized density carries 2 v / (their number)
                w = 2.0 * v / len(perms)
                for p in perms:
                    G[p] = w
        base += g0 * (g0 + 1) // 2
    return D1, G


def transform(h, eri, U):
    """h, (pq|rs) in the orbitals C U"""
    h2 = U.T @ h @ U
    g = np.einsum("pqrs,pi->iqrs", eri, U, optimize=True)
    g = np.einsum("iqrs,qj->ijrs", g, U, optimize=True)
    g = np.einsum("ijrs,rk->ijks", g, U, optimize=True)
    g = np.einsum("ijks,sl->ijkl", g, U, optimize=True)
    return h2, g


class Spaces:
    """Core / active / virtual orbitals per irrep (Pitzer order within the
    full orbital list, which runs irrep by irrep): the reference's
    RESTRICTED_DOCC, ACTIVE and RESTRICTED_UOCC."""

    def __init__(self, nmopi, docc, active):
        self.nmopi = list(nmopi)
        self.docc = list(docc)
        self.active = list(active)
        self.core, self.act, self.virt, self.sym = [], [], [], []
        o = 0
        for h, n in enumerate(self.nmopi):
            c, a = self.docc[h], self.active[h]
            if c + a > n:
                raise ValueError("restricted_docc + active exceeds the orbitals of an irrep")
            self.core += list(range(o, o + c))
            self.act += list(range(o + c, o + c + a))
            self.virt += list(range(o + c + a, o + n))
            self.sym += [h] * n
            o += n
        self.n = o

    def pairs(self, active_active=False):
        """non-redundant rotations (m, n), m > n, same irrep"""
        kind = {}
        for p in self.core:
            kind[p] = 0
        for p in self.act:
            kind[p] = 1
        for p in self.virt:
            kind[p] = 2
        out = []
        for m in range(self.n):
            for n in range(m):
                if self.sym[m] != self.sym[n]:
                    continue
                if kind[m] == kind[n] and not (active_active and kind[m] == 1):
                    continue
                out.append((m, n))
        return out


def energy(h, eri, sp, D1, G):
    """core + active energy of densities D1, G (active) in the orbitals of h, eri"""
    h_eff, g_act, e_core = frozen_core(h, eri, sp.core, sp.act)
    return e_core + float(np.einsum("pq,pq", h_eff, D1)) + 0.5 * float(np.einsum("pqrs,pqrs", g_act, G))


def gradient_hessian(h, eri, sp, D1, G):
    """orbital gradient 2 (F_mn - F_nm) on the full orbital space (for
    C <- C expm(-kappa)) and a diagonal Hessian model, from the inactive /
    active Fock matrices and the generalized Fock matrix F"""
    n = sp.n
    core, act = sp.core, sp.act
    FI = h.copy()
    if core:
        FI += 2.0 * np.einsum("mncc->mn", eri[:, :, core][:, :, :, core]) \
            - np.einsum("mccn->mn", eri[:, core][:, :, core])
    ga = eri[:, :, act][:, :, :, act]
    FA = np.einsum("vw,mnvw->mn", D1, ga) - 0.5 * np.einsum("vw,mvwn->mn", D1, eri[:, act][:, :, act])
    F = np.zeros((n, n))
    for i in core:
        F[i, :] = 2.0 * (FI[:, i] + FA[:, i])
    gm = eri[:, act][:, :, act][:, :, :, act]          # (n w | x y), w, x, y active
    Fa = D1 @ FI[act, :] + np.einsum("vwxy,nwxy->vn", G, gm)
    for t, p in enumerate(act):
        F[p, :] = Fa[t, :]
    grad = 2.0 * (F - F.T)
    # diagonal Hessian model (inactive / active Fock and generalized Fock
    # diagonals, the usual two-step CASSCF approximation)
    Ft = np.diag(FI + FA)
    Fd = np.diag(F)
    dact = np.zeros(n)
    for t, p in enumerate(act):
        dact[p] = D1[t, t]
    kind = np.zeros(n, dtype=int)
    kind[sp.act] = 1
    kind[sp.virt] = 2
    hdiag = np.zeros((n, n))
    for m in range(n):
        for q in range(m):
            lo, hi = (m, q) if kind[m] < kind[q] else (q, m)   # lower / higher space
            pair = (kind[lo], kind[hi])
            if pair == (0, 2):                 # core-virtual
                v = 4.0 * (Ft[hi] - Ft[lo])
            elif pair == (1, 2):               # active-virtual
                v = 2.0 * dact[lo] * Ft[hi] - 2.0 * Fd[lo]
            elif pair == (0, 1):               # core-active
                v = 4.0 * Ft[hi] + 2.0 * dact[hi] * Ft[lo] - 4.0 * Ft[lo] - 2.0 * Fd[hi]
            elif pair == (1, 1):               # active-active
                v = 2.0 * (dact[m] * Ft[q] + dact[q] * Ft[m]) - 2.0 * (Fd[m] + Fd[q])
            else:
                v = 0.0
            hdiag[m, q] = hdiag[q, m] = v
    return grad, hdiag


def orbital_step(h, eri, sp, D1, G, maxiter=20, gtol=1e-6, active_active=False, max_step=0.5):
    """Orbital optimization with the densities held fixed (one RotateOrbitals
    call): returns (U, energy before, energy after, gradient norm at the
    start, gradient norm at the end, iterations).  U is the accumulated
    rotation; the integrals of the rotated orbitals are transform(h, eri, U)."""
    pairs = sp.pairs(active_active)
    U = np.eye(sp.n)
    hc, gc = h, eri
    e0 = e = energy(hc, gc, sp, D1, G)
    gnorm = gnorm0 = 0.0
    it = 0
    for it in range(1, maxiter + 1):
        grad, hdiag = gradient_hessian(hc, gc, sp, D1, G)
        gv = np.array([grad[m, q] for m, q in pairs])
        gnorm = float(np.linalg.norm(gv))
        if it == 1:
            gnorm0 = gnorm
        if gnorm < gtol or not pairs:
            it -= 1
            break
        hv = np.array([hdiag[m, q] for m, q in pairs])
        step = -gv / np.maximum(hv, 0.05)
        s = float(np.max(np.abs(step)))
        if s > max_step:
            step *= max_step / s
        accepted = False
        for _ in range(12):
            K = np.zeros((sp.n, sp.n))
            for (m, q), k in zip(pairs, step):
                K[m, q], K[q, m] = k, -k
            R = scipy.linalg.expm(-K)
            h2, g2 = transform(hc, gc, R)
            e2 = energy(h2, g2, sp, D1, G)
            if e2 <= e + 1e-14:
                accepted = True
                break
            step *= 0.5
        if not accepted:
            break
        U = U @ R
        hc, gc, e = h2, g2, e2
    return U, e0, e, gnorm0, gnorm, it


def v2rdm_casscf(h, eri, enuc, spaces, nalpha, nbeta, positivity="DQG", options=None,
                 solver_kwargs=None, maxiter=50, orbopt_maxiter=20,
                 gtol=1e-6, etol=1e-8, active_active=False, progress=None):
    """v2RDM-CASSCF: BPSDP in the active space of `spaces` on the GPU,
    alternated with orbital steps until both converge (the OPTIMIZE_ORBITALS
    branch of v2RDMSolver::compute_energy).

    h, eri: one- and two-electron integrals over all orbitals of `spaces`
    (chemists' notation), nalpha / nbeta: active electrons.  Returns a dict
    with the total energy, the macro-iteration count, the orbital rotation
    accumulated and the final solver."""
    from .solver import V2RDMSolver
    sp = spaces
    amopi = sp.active
    s = V2RDMSolver(amopi, nalpha, nbeta, 1, positivity, options=options, **(solver_kwargs or {}))
    hc, gc = np.array(h, dtype=np.float64), np.array(eri, dtype=np.float64)
    Utot = np.eye(sp.n)
    d1dim, d2dim = s.spatial_density_dims()
    bas00 = s.singlet_geminals()
    h_eff, g_act, e_core = frozen_core(hc, gc, sp.core, sp.act)
    out = s.compute_energy(h_eff, g_act)
    macro, converged = 0, False
    history = []
    for macro in range(1, maxiter + 1):
        d1, d2 = s.pack_spatial_density()
        D1, G = unpack_spatial_density(d1, d2, amopi, bas00)
        U, e0, e1, gnorm0, gnorm, micro = orbital_step(hc, gc, sp, D1, G, orbopt_maxiter, gtol,
                                                       active_active)
        history.append(dict(macro=macro, energy=out["energy_primal"] + e_core + enuc,
                            orbital_gradient=gnorm0, orbital_energy_change=e1 - e0,
                            micro=micro, bpsdp_oiter=out["oiter"]))
        if progress:
            progress(history[-1])
        # stopping rule of the reference (v2rdm_solver.cc:1896-1911 with
        # RotateOrbitals' test): the orbital step changes the energy by less
        # than ORBOPT_ENERGY_CONVERGENCE with a gradient below
        # ORBOPT_GRADIENT_CONVERGENCE, on a converged BPSDP iteration
        if abs(e1 - e0) < etol and gnorm0 < gtol and out["converged"]:
            converged = True
            break
        hc, gc = transform(hc, gc, U)
        Utot = Utot @ U
        h_eff, g_act, e_core = frozen_core(hc, gc, sp.core, sp.act)
        out = s.compute_energy(h_eff, g_act, resume=True)
    return dict(energy=out["energy_primal"] + e_core + enuc, converged=int(converged),
                macro=macro, rotation=Utot, history=history, solver=s, stats=out,
                h=hc, eri=gc)
