"""Integral and RDM files in the layout of the reference's FCIDUMP writer
(v2RDMSolver::FCIDUMP, fcidump.cc:67-282), and the frozen-core reduction
that turns full-space integrals into the active-space h1 / (pq|rs) the engine
takes (RepackIntegrals / FrozenCoreEnergy, tei.cc:144-313).

Both files are flat binary records of one double and four 4-byte ints,
(value, p, q, r, s), orbitals numbered from 1:

  <name>.int   (pq|rs) for p <= q, r <= s, pq <= rs (triangular pair index),
               symmetry-allowed only (fcidump.cc:118-144); then h_pq for every
               p, q of each irrep with r = s = 0 (:146-169); then one record
               (enuc, 0, 0, 0, 0) (:170-174).  Orbitals are ordered core
               (by irrep), then active (by irrep), then virtual (:87-110).
  <name>.rdm   the spin-summed 2-RDM element for every pair of ab geminals of
               every irrep (value, i, j, k, l) = <a+_i a+_j a_l a_k> summed
               over spins, then the 1-RDM (value, i, j, 0, 0) (:176-268);
               active orbitals only, numbered from 1 (written by the engine:
               V2RDMSolver.write_fcidump_rdm).
"""
import numpy as np

RECORD = np.dtype([("v", "<f8"), ("p", "<i4"), ("q", "<i4"), ("r", "<i4"), ("s", "<i4")])


def _tri(p, q):
    return p * (p + 1) // 2 + q if p >= q else q * (q + 1) // 2 + p


def write_integrals(path, h1, eri, enuc, orbsym):
    """Write <name>.int for orbitals already in file order (core, active,
    virtual, each by irrep); orbsym[p] is the D2h irrep of orbital p (XOR
    products), h1 symmetric, eri (pq|rs) with the 8-fold symmetry."""
    n = h1.shape[0]
    orbsym = np.asarray(orbsym)
    recs = []
    for p in range(n):
        for q in range(p, n):
            hpq = orbsym[p] ^ orbsym[q]
            pq = _tri(p, q)
            for r in range(n):
                for s in range(r, n):
                    if orbsym[r] ^ orbsym[s] != hpq or pq > _tri(r, s):
                        continue
                    recs.append((eri[p, q, r, s], p + 1, q + 1, r + 1, s + 1))
    for p in range(n):
        for q in range(n):
            if orbsym[p] == orbsym[q]:
                recs.append((h1[p, q], p + 1, q + 1, 0, 0))
    recs.append((enuc, 0, 0, 0, 0))
    np.array(recs, dtype=RECORD).tofile(path)


def read_integrals(path, norb):
    """Read <name>.int: returns (h1, eri, enuc) over all `norb` orbitals of
    the file, eri filled with its 8-fold symmetry; integrals the file does not
    list (symmetry-forbidden) are zero."""
    recs = np.fromfile(path, dtype=RECORD)
    p, q, r, s, v = (recs[k] for k in ("p", "q", "r", "s", "v"))
    if np.any((p < 0) | (p > norb) | (q > norb) | (r > norb) | (s > norb)):
        raise ValueError(f"{path}: orbital index outside 1..{norb}")
    h1 = np.zeros((norb, norb))
    eri = np.zeros((norb,) * 4)
    two = (r > 0) & (s > 0)
    one = (p > 0) & (q > 0) & (r == 0) & (s == 0)
    core = (p == 0) & (q == 0) & (r == 0) & (s == 0)
    if core.sum() != 1:
        raise ValueError(f"{path}: expected exactly one nuclear-repulsion record")
    i, j, k, l, w = p[two] - 1, q[two] - 1, r[two] - 1, s[two] - 1, v[two]
    for a, b, c, d in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                       (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)):
        eri[a, b, c, d] = w
    h1[p[one] - 1, q[one] - 1] = v[one]
    return h1, eri, float(v[core][0])


def read_rdm(path, norb):
    """Read <name>.rdm: returns (d1, d2) with d1[i, j] the spin-summed 1-RDM
    and d2[i, j, k, l] = <a+_i a+_j a_l a_k> summed over spins."""
    recs = np.fromfile(path, dtype=RECORD)
    p, q, r, s, v = (recs[k] for k in ("p", "q", "r", "s", "v"))
    d1 = np.zeros((norb, norb))
    d2 = np.zeros((norb,) * 4)
    two = (r > 0) & (s > 0)
    d2[p[two] - 1, q[two] - 1, r[two] - 1, s[two] - 1] = v[two]
    one = ~two
    d1[p[one] - 1, q[one] - 1] = v[one]
    return d1, d2


def rdm_energy(d1, d2, h1, eri):
    """E = sum h_ij d1_ij + 1/2 sum (ik|jl) d2_ijkl (active part, no enuc)."""
    return float(np.einsum("ij,ij->", h1, d1) + 0.5 * np.einsum("ijkl,ikjl->", d2, eri))


def frozen_core(h1, eri, core, active):
    """Active-space integrals with the `core` orbitals doubly occupied
    (tei.cc:229-313): h_eff = h + sum_c [2 (pq|cc) - (pc|cq)] on the active
    orbitals, (pq|rs) restricted to them, and the core energy
    E_core = sum_c 2 h_cc + sum_cd [2 (cc|dd) - (cd|dc)]."""
    core = list(core)
    active = list(active)
    J = np.einsum("pqcc->pq", eri[:, :, core, :][:, :, :, core]) if core else 0.0
    K = np.einsum("pccq->pq", eri[:, core, :, :][:, :, core, :]) if core else 0.0
    fock_core = 2.0 * J - K if core else np.zeros_like(h1)
    ix = np.ix_(active, active)
    h_eff = h1[ix] + fock_core[ix]
    e_core = 0.0
    if core:
        cc = np.ix_(core, core)
        e_core = float(2.0 * np.trace(h1[cc]) + np.trace(fock_core[cc]))
    eri_act = eri[np.ix_(active, active, active, active)]
    return h_eff, eri_act, e_core
