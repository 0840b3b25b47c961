"""Brute-force FCI and exact reduced density matrices for tiny active spaces.

Used by the oracle pinning tests.  Every block of the v2RDM primal vector is
built here from its physical definition as a Gram matrix <O_a^dag O_b> of
Fock-space vectors O_b |psi>, never from the reference's constraint formulas,
so that A . x_exact = b is a genuine check of the oracle's rows.

Mode ordering: alpha spin-orbital p -> mode p, beta -> mode n + p.
"""
import itertools

import numpy as np
import scipy.sparse as sp


class Fock:
    def __init__(self, n):
        self.n = n
        self.nm = 2 * n
        self.dim = 1 << self.nm
        states = np.arange(self.dim)
        self._ann = []
        for m in range(self.nm):
            occ = (states >> m) & 1
            src = states[occ == 1]
            dst = src ^ (1 << m)
            below = src & ((1 << m) - 1)
            sign = np.where(np.array([bin(v).count("1") for v in below]) % 2 == 0, 1.0, -1.0)
            A = sp.csr_matrix((sign, (dst, src)), shape=(self.dim, self.dim))
            self._ann.append(A)
        self.popa = np.array([bin(s & ((1 << n) - 1)).count("1") for s in states])
        self.popb = np.array([bin(s >> n).count("1") for s in states])

    def a(self, p, spin):
        return self._ann[p + (self.n if spin else 0)]

    def c(self, p, spin):
        return self._ann[p + (self.n if spin else 0)].T.tocsr()

    def hamiltonian(self, h1, eri):
        """H = sum h_pq E_pq + 1/2 sum (pq|rs) (E_pq E_rs - delta_qr E_ps)."""
        n = self.n
        E = [[None] * n for _ in range(n)]
        for p in range(n):
            for q in range(n):
                E[p][q] = (self.c(p, 0) @ self.a(q, 0) + self.c(p, 1) @ self.a(q, 1)).tocsr()
        H = sp.csr_matrix((self.dim, self.dim))
        for p in range(n):
            for q in range(n):
                if h1[p, q] != 0.0:
                    H = H + h1[p, q] * E[p][q]
        for p in range(n):
            for q in range(n):
                F = sp.csr_matrix((self.dim, self.dim))
                for r in range(n):
                    for s in range(n):
                        v = eri[p, q, r, s]
                        if v != 0.0:
                            F = F + v * E[r][s]
                H = H + 0.5 * (E[p][q] @ F)
        for p in range(n):
            for s in range(n):
                v = sum(eri[p, q, q, s] for q in range(n))
                if v != 0.0:
                    H = H - 0.5 * v * E[p][s]
        return H.tocsr()

    def ground_state(self, h1, eri, na, nb, s_target=None, irrep_of=None, want_irrep=None):
        """Lowest eigenpair in the (na, nb) sector with S(S+1) = s_target(s_target+1)."""
        idx = np.where((self.popa == na) & (self.popb == nb))[0]
        H = self.hamiltonian(h1, eri)[idx][:, idx].toarray()
        w, V = np.linalg.eigh(H)
        S2 = self.s2_operator()[idx][:, idx].toarray()
        for k in range(len(w)):
            v = V[:, k]
            s2 = v @ S2 @ v
            if s_target is not None and abs(s2 - s_target * (s_target + 1)) > 1e-6:
                continue
            psi = np.zeros(self.dim)
            psi[idx] = v
            return w[k], psi
        raise RuntimeError("no state with the requested spin")

    def s2_operator(self):
        n = self.n
        Sp = sp.csr_matrix((self.dim, self.dim))
        Sz = sp.csr_matrix((self.dim, self.dim))
        for p in range(n):
            Sp = Sp + self.c(p, 0) @ self.a(p, 1)
            Sz = Sz + 0.5 * (self.c(p, 0) @ self.a(p, 0) - self.c(p, 1) @ self.a(p, 1))
        Sm = Sp.T.tocsr()
        return (Sm @ Sp + Sz @ Sz + Sz).tocsr()


def gram(vecs_a, vecs_b=None):
    """G[a, b] = <v_a | v_b>."""
    A = np.array(vecs_a)
    B = A if vecs_b is None else np.array(vecs_b)
    return A @ B.T


def exact_x(F, psi, L):
    """Primal vector x of the exact state psi in the oracle layout L.

    Blocks (each a Gram matrix of operator images of psi):
      D1s(i,k)   = <a+_i a_k>                 : vectors a_i psi
      Q1s(i,k)   = <a_i a+_k>                 : vectors a+_i psi
      D2ab(ij,kl)= <a+_ia a+_jb a_lb a_ka>    : vectors a_jb a_ia psi
      D2aa(ij,kl)= <a+_i a+_j a_l a_k>, i>j   : vectors a_j a_i psi
      D200       = singlet-pair Gram           : (a_jb a_ia + a_ib a_ja) psi / (sqrt2 d_ij)
      Q2ab(ij,kl)= <a_ia a_jb a+_lb a+_ka>    : vectors a+_jb a+_ia psi
      G2ab(ij,kl)= <a+_ia a_jb a+_lb a_ka>    : vectors a+_jb a_ia psi
      G2 super block: vectors [a+_ja a_ia ; a+_jb a_ib] psi
      spin-adapted Q2: Q2s over singlet geminals (i, j) (i = j included):
                   (a+_jb a+_ia + a+_ib a+_ja) psi / sqrt2; Q2t over aa
                   geminals: (a+_jb a+_ia - a+_ib a+_ja) psi / sqrt2;
                   Q2t(+1) / Q2t(-1) = Q2aa / Q2bb
      spin-adapted G2: G2s / G2t: (a+_ja a_ia +- a+_jb a_ib) psi / sqrt2;
                   G2t(+1) / G2t(-1) = G2ab / G2ba
      T1 blocks  = <B+ B> + <B B+>, B_ijk = a_k a_j a_i (aab: i<j alpha, k beta;
                   aaa: i<j<k), i.e. the D3 + Q3 matrix
      T2 blocks  = <B+ B> + <B B+> for B = the triplet operators of each block
    """
    n = L.amo
    x = np.zeros(L.dimx)
    sym = [L.irrep(p) for p in range(n)]
    orbs_of = {}
    for p in range(n):
        orbs_of.setdefault(sym[p], []).append(p)

    def ab_pairs(h):
        g = L.gems_ab(h)
        out = [None] * g
        for p in range(n):
            for q in range(n):
                if sym[p] ^ sym[q] == h:
                    out[L.ibas_ab(p, q)] = (p, q)
        return out

    def aa_pairs(h):
        g = L.gems_aa(h)
        out = [None] * g
        for p in range(n):
            for q in range(p):
                if sym[p] ^ sym[q] == h:
                    out[L.ibas_aa(p, q)] = (p, q)
        return out

    def put(name, h, M):
        o = L.offset(name, h)
        x[o:o + M.size] = M.ravel()

    a, c = F.a, F.c
    for h in range(L.nirrep):
        orbs = [p for p in range(n) if sym[p] == h]
        if not orbs:
            continue
        for s, (dn, qn) in enumerate((("D1A", "Q1A"), ("D1B", "Q1B"))):
            va = [a(p, s) @ psi for p in orbs]
            vc = [c(p, s) @ psi for p in orbs]
            put(dn, h, gram(va))
            put(qn, h, gram(vc))
    singlet = L.opts.nalpha == L.opts.nbeta
    for h in range(L.nirrep):
        P = ab_pairs(h)
        if not P:
            continue
        vd = [a(q, 1) @ (a(p, 0) @ psi) for p, q in P]
        put("D2AB", h, gram(vd))
        if L.opts.constrain_spin:
            d = lambda p, q: np.sqrt(2.0) if p == q else 1.0
            vs = [(a(q, 1) @ (a(p, 0) @ psi) + a(p, 1) @ (a(q, 0) @ psi)) / (np.sqrt(2.0) * d(p, q))
                  for p, q in P]
            if singlet:
                put("D200", h, gram(vs))
            else:
                vt = [(a(q, 1) @ (a(p, 0) @ psi) - a(p, 1) @ (a(q, 0) @ psi)) / np.sqrt(2.0)
                      for p, q in P]
                put("D200", h, gram(vs + vt))
        sa_q2 = getattr(L.opts, "spin_adapt_q2", 0)
        sa_g2 = getattr(L.opts, "spin_adapt_g2", 0)
        if L.opts.constrain_q2 and not sa_q2:
            vq = [c(q, 1) @ (c(p, 0) @ psi) for p, q in P]
            put("Q2AB", h, gram(vq))
        if L.opts.constrain_q2 and sa_q2:
            # singlet hole pairs (a+_jb a+_ia + a+_ib a+_ja) / sqrt2 over the
            # singlet geminals (i, j), i = j included
            P00 = L.bas_00(h)
            if P00:
                vs = [(c(j, 1) @ (c(i, 0) @ psi) + c(i, 1) @ (c(j, 0) @ psi)) / np.sqrt(2.0)
                      for i, j in P00]
                put("Q2S", h, gram(vs))
        if L.opts.constrain_g2 and not sa_g2:
            vg = [c(q, 1) @ (a(p, 0) @ psi) for p, q in P]      # a+_jb a_ia psi
            put("G2AB", h, gram(vg))
            vg2 = [c(q, 0) @ (a(p, 1) @ psi) for p, q in P]     # a+_ja a_ib psi
            put("G2BA", h, gram(vg2))
            vaa = [c(q, 0) @ (a(p, 0) @ psi) for p, q in P]
            vbb = [c(q, 1) @ (a(p, 1) @ psi) for p, q in P]
            put("G2AA", h, gram(vaa + vbb))
        if L.opts.constrain_g2 and sa_g2:
            # M_s = 0 singlet / triplet particle-hole pairs
            # (a+_ja a_ia +- a+_jb a_ib) / sqrt2, and the spin-flip pairs
            vaa = [c(q, 0) @ (a(p, 0) @ psi) for p, q in P]
            vbb = [c(q, 1) @ (a(p, 1) @ psi) for p, q in P]
            put("G2S", h, gram([(u + v) / np.sqrt(2.0) for u, v in zip(vaa, vbb)]))
            put("G2T", h, gram([(u - v) / np.sqrt(2.0) for u, v in zip(vaa, vbb)]))
            put("G2TP1", h, gram([c(q, 1) @ (a(p, 0) @ psi) for p, q in P]))
            put("G2TM1", h, gram([c(q, 0) @ (a(p, 1) @ psi) for p, q in P]))
    for h in range(L.nirrep):
        P = aa_pairs(h)
        if not P:
            continue
        sa_q2 = getattr(L.opts, "spin_adapt_q2", 0)
        for s, (dn, qn) in enumerate((("D2AA", "Q2AA"), ("D2BB", "Q2BB"))):
            vd = [a(q, s) @ (a(p, s) @ psi) for p, q in P]
            put(dn, h, gram(vd))
            if L.opts.constrain_q2:
                # like-spin hole pairs: Q2aa / Q2bb, or Q2t(+1) / Q2t(-1)
                vq = [c(q, s) @ (c(p, s) @ psi) for p, q in P]
                put(("Q2TM1" if s else "Q2TP1") if sa_q2 else qn, h, gram(vq))
        if L.opts.constrain_q2 and sa_q2:
            # M_s = 0 triplet hole pairs (a+_jb a+_ia - a+_ib a+_ja) / sqrt2
            vt = [(c(q, 1) @ (c(p, 0) @ psi) - c(p, 1) @ (c(q, 0) @ psi)) / np.sqrt(2.0)
                  for p, q in P]
            put("Q2T", h, gram(vt))
    def anti(ops):
        """<B+ B> + <B B+> as Gram of B psi and B+ psi."""
        v1 = [op @ psi for op in ops]
        v2 = [op.T @ psi for op in ops]
        return gram(v1) + gram(v2)

    if getattr(L.opts, "constrain_t1", 0):
        for h in range(L.nirrep):
            trip = [(i, j, k) for i in range(n) for j in range(n) for k in range(n)
                    if sym[i] ^ sym[j] ^ sym[k] == h]
            aab = [t for t in trip if t[0] < t[1]]
            aaa = [t for t in trip if t[0] < t[1] < t[2]]
            for name, s1, s2 in (("T1AAB", 0, 1), ("T1BBA", 1, 0)):
                ops = [a(k, s2) @ a(j, s1) @ a(i, s1) for i, j, k in aab]
                if ops:
                    put(name, h, anti(ops))
            for name, s1 in (("T1AAA", 0), ("T1BBB", 1)):
                ops = [a(k, s1) @ a(j, s1) @ a(i, s1) for i, j, k in aaa]
                if ops:
                    put(name, h, anti(ops))
    if L.opts.constrain_t2:
        for h in range(L.nirrep):
            aab, aba = [], []
            for i in range(n):
                for j in range(n):
                    for k in range(n):
                        if sym[i] ^ sym[j] ^ sym[k] != h:
                            continue
                        aba.append((i, j, k))
                        if i < j:
                            aab.append((i, j, k))

            # T2aab: B_ijk = a+_kb a_ja a_ia ; T2bba: spins swapped
            for name, s1, s2 in (("T2AAB", 0, 1), ("T2BBA", 1, 0)):
                ops = [c(k, s2) @ a(j, s1) @ a(i, s1) for i, j, k in aab]
                if ops:
                    put(name, h, anti(ops))
            # big blocks: [a+_ka a_ja a_ia (aab) ; a+_kb a_jb a_ia (aba)] and spin flip
            for name, s1, s2 in (("T2AAA", 0, 1), ("T2BBB", 1, 0)):
                ops = [c(k, s1) @ a(j, s1) @ a(i, s1) for i, j, k in aab]
                ops += [c(k, s2) @ a(j, s2) @ a(i, s1) for i, j, k in aba]
                if ops:
                    put(name, h, anti(ops))
    return x


def random_hamiltonian(n, seed, irreps=None, nL=None, scale=0.1):
    """Symmetric h and Cholesky-form (pq|rs) = sum_L B^L_pq B^L_rs.

    With irreps given, every integral not totally symmetric under the
    orbital irreps (XOR product) is zeroed, so the Hamiltonian has that
    point-group symmetry."""
    rng = np.random.default_rng(seed)
    h = rng.normal(0, 0.3, (n, n))
    h = 0.5 * (h + h.T) - np.diag(np.linspace(2.0, 0.2, n))
    nL = nL or 2 * n
    B = rng.normal(0, scale, (nL, n, n))
    B = 0.5 * (B + B.transpose(0, 2, 1))
    if irreps is not None:
        ir = np.array(irreps)
        mask = (ir[:, None] ^ ir[None, :]) == 0
        h = h * mask
        # split each B^L into one irrep component so that (pq|rs) stays symmetric
        comp = rng.integers(0, max(ir) + 1, nL)
        for L_ in range(nL):
            m = (ir[:, None] ^ ir[None, :]) == comp[L_]
            B[L_] *= m
    eri = np.einsum("lpq,lrs->pqrs", B, B)
    return h, eri
