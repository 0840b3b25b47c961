// rowterms.hpp -- the constraint rows of A, as code that runs on the device
// (maps.hip) and, unchanged, on the host (the bounds-checked emulator that
// validates every index against the CPU oracle before a kernel runs).
//
// Row formulas follow d2.cc:296-547, spin.cc:274-493, q2.cc:801-950,
// g2.cc:1061-1260, t1.cc:433-809 and t2.cc:1251-1861, and the spin-adapted
// q2.cc:260-467 and g2.cc:235-546; the A^T gathers invert the first six.
#pragma once
#include <cmath>
#include <cstdint>

#include "devlayout.hpp"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define V2G_HD __host__ __device__
#else
#define V2G_HD
#endif

namespace v2g {
namespace rows {

V2G_HD inline int P0(int packed) { return packed & 0xffff; }
V2G_HD inline int P1(int packed) { return packed >> 16; }
V2G_HD inline int T0(int packed) { return packed & 0xff; }
V2G_HD inline int T1(int packed) { return (packed >> 8) & 0xff; }
V2G_HD inline int T2i(int packed) { return (packed >> 16) & 0xff; }

// e = q d + r for an element index inside one block: 32-bit unsigned
// division whenever d^2 fits (every block below 65536 x 65536), 64-bit
// otherwise.  The test is uniform across a workgroup.
V2G_HD inline void split_index(int64_t e, int64_t d, int &q, int &r)
{
    if (d <= 65535) {
        const uint32_t ue = (uint32_t)e, ud = (uint32_t)d;
        const uint32_t qq = ue / ud;
        q = (int)qq;
        r = (int)(ue - qq * ud);
    } else {
        q = (int)(e / d);
        r = (int)(e - (int64_t)q * d);
    }
}

// ----------------------------------------------------------------------------
// row terms of the elementwise families (A.x)
// ----------------------------------------------------------------------------

template <class F>
V2G_HD inline void t2_terms(const DevLayout &L, int kind, int s, int h, int64_t e,
                                         F &f, bool self)
{
    const int amo = L.amo;
    const int *gab = L.gab, *gaa = L.gaa, *sym = L.sym;
    auto GAB = [&](int p, int q) -> int64_t { return gab[p * amo + q]; };
    auto GAA = [&](int p, int q) -> int64_t { return gaa[p * amo + q]; };
    auto D1 = [&](int w, int k, int n, double c) {
        const int h3 = sym[k];
        const int na3 = L.amopi[h3];
        f(L.xoff[w][h3] + (int64_t)(n - L.poff[h3]) * na3 + (k - L.poff[h3]), c);
    };
    if (kind == RF_T2AAB) {
        // T2aab / T2bba rows, t2.cc:1256-1390
        const int t = L.trip_aab[h];
        const int *bt = L.baab + L.baab_off[h];
        int ijk, lmn;
        split_index(e, t, ijk, lmn);
        const int a = bt[ijk], b = bt[lmn];
        const int i = T0(a), j = T1(a), k = T2i(a), l = T0(b), m = T1(b), n = T2i(b);
        if (self) f(L.xoff[s ? XB_T2BBA : XB_T2AAB][h] + e, -1.0);
        const int w2 = s ? XB_D2BB : XB_D2AA, wd = s ? XB_D1A : XB_D1B;
        if (k == n) {
            const int hij = sym[i] ^ sym[j];
            f(L.xoff[w2][hij] + GAA(i, j) * L.gems_aa[hij] + GAA(l, m), 1.0);
        }
        if (j == m && i == l) D1(wd, k, n, 1.0);
        if (j == l && i == m) D1(wd, k, n, -1.0);
        auto AB = [&](int p, int q, int r, int v, double c) {
            const int hh = sym[p] ^ sym[q];
            const int64_t a1 = s ? GAB(q, p) : GAB(p, q), a2 = s ? GAB(v, r) : GAB(r, v);
            f(L.xoff[XB_D2AB][hh] + a1 * L.gems_ab[hh] + a2, c);
        };
        if (i == l) AB(j, n, m, k, -1.0);
        if (j == l) AB(i, n, m, k, 1.0);
        if (i == m) AB(j, n, l, k, 1.0);
        if (j == m) AB(i, n, l, k, -1.0);
        return;
    }
    // big blocks T2aaa(+abb) / T2bbb(+baa), t2.cc:1392-1858
    const int ta = L.trip_aab[h], tb = L.trip_aba[h];
    const int64_t T = (int64_t)ta + tb;
    int R, C;
        split_index(e, T, R, C);
    const bool rA = R < ta, cA = C < ta;
    const int pr = rA ? L.baab[L.baab_off[h] + R] : L.baba[L.baba_off[h] + (R - ta)];
    const int pc = cA ? L.baab[L.baab_off[h] + C] : L.baba[L.baba_off[h] + (C - ta)];
    const int i = T0(pr), j = T1(pr), k = T2i(pr), l = T0(pc), m = T1(pc), n = T2i(pc);
    if (self) f(L.xoff[s ? XB_T2BBB : XB_T2AAA][h] + e, -1.0);
    const int w2 = s ? XB_D2BB : XB_D2AA, w2o = s ? XB_D2AA : XB_D2BB;
    const int wd = s ? XB_D1B : XB_D1A, wdo = s ? XB_D1A : XB_D1B;
    auto AA = [&](int w, int p, int q, int r, int v, double c) {
        if (p == q || r == v) return;
        const int hh = sym[p] ^ sym[q];
        double sg = 1.0;
        if (p > q) sg = -sg;
        if (r > v) sg = -sg;
        f(L.xoff[w][hh] + GAA(p, q) * L.gems_aa[hh] + GAA(r, v), c * sg);
    };
    auto AB = [&](int p1, int q1, int p2, int q2, double c) {
        // D2ab entry (p1 q1 | p2 q2) for the alpha-first orientation
        const int hh = sym[p1] ^ sym[q1];
        f(L.xoff[XB_D2AB][hh] + GAB(p1, q1) * L.gems_ab[hh] + GAB(p2, q2), c);
    };
    if (rA && cA) {
        if (k == n) {
            const int hij = sym[i] ^ sym[j];
            f(L.xoff[w2][hij] + GAA(i, j) * L.gems_aa[hij] + GAA(l, m), 1.0);
        }
        if (j == m && i == l) D1(wd, k, n, 1.0);
        if (j == l && i == m) D1(wd, k, n, -1.0);
        if (i == l) AA(w2, n, j, k, m, -1.0);
        if (j == l) AA(w2, n, i, k, m, 1.0);
        if (i == m) AA(w2, n, j, k, l, 1.0);
        if (j == m) AA(w2, n, i, k, l, -1.0);
    } else if (rA) {
        if (i == l) { if (s) AB(n, j, m, k, 1.0); else AB(j, n, k, m, 1.0); }
        if (j == l) { if (s) AB(n, i, m, k, -1.0); else AB(i, n, k, m, -1.0); }
    } else if (cA) {
        if (i == l) { if (s) AB(j, n, k, m, 1.0); else AB(n, j, m, k, 1.0); }
        if (i == m) { if (s) AB(j, n, k, l, -1.0); else AB(n, j, l, k, -1.0); }
    } else {
        if (k == n) {
            if (s) AB(j, i, m, l, 1.0); else AB(i, j, l, m, 1.0);
        }
        if (j == m && i == l) D1(wdo, k, n, 1.0);
        if (i == l) AA(w2o, n, j, k, m, -1.0);
        if (j == m) { if (s) AB(n, i, k, l, -1.0); else AB(i, n, l, k, -1.0); }
    }
}

// T1 rows, t1.cc:433-809.  kind RF_T1AAB: T1aab (s = 0) / T1bba (s = 1) over
// i<j triplets; RF_T1AAA: T1aaa / T1bbb over i<j<k triplets.  At most one
// term of each delta group fires, so a row never repeats a column.
template <class F>
V2G_HD inline void t1_terms(const DevLayout &L, int kind, int s, int h, int64_t e, F &f,
                            bool self)
{
    const int amo = L.amo;
    const int *gab = L.gab, *gaa = L.gaa, *sym = L.sym;
    auto GAB = [&](int p, int q) -> int64_t { return gab[p * amo + q]; };
    auto GAA = [&](int p, int q) -> int64_t { return gaa[p * amo + q]; };
    const int wq = s ? XB_Q2BB : XB_Q2AA;
    if (kind == RF_T1AAB) {
        const int t = L.trip_aab[h];
        const int *bt = L.baab + L.baab_off[h];
        int ijk, lmn;
        split_index(e, t, ijk, lmn);
        const int a = bt[ijk], b = bt[lmn];
        const int i = T0(a), j = T1(a), k = T2i(a), l = T0(b), m = T1(b), n = T2i(b);
        if (self) f(L.xoff[s ? XB_T1BBA : XB_T1AAB][h] + e, -1.0);
        if (k == n) {
            const int hij = sym[i] ^ sym[j];
            f(L.xoff[wq][hij] + GAA(i, j) * L.gems_aa[hij] + GAA(l, m), 1.0);
        }
        // D2ab (mn|ik): T1aab reads alpha-first pairs, T1bba the swapped ones
        auto AB = [&](int p1, int q1, int p2, int q2, double c) {
            const int hh = sym[p1] ^ sym[q1];
            const int64_t a1 = s ? GAB(q1, p1) : GAB(p1, q1), a2 = s ? GAB(q2, p2) : GAB(p2, q2);
            f(L.xoff[XB_D2AB][hh] + a1 * L.gems_ab[hh] + a2, c);
        };
        if (j == l) AB(m, n, i, k, -1.0);
        if (l == i) AB(m, n, j, k, 1.0);
        // mixed-spin G2: G2ba for T1aab, G2ab for T1bba
        const int wg = s ? XB_G2AB : XB_G2BA;
        if (j == m) {
            const int hh = sym[n] ^ sym[i];
            f(L.xoff[wg][hh] + GAB(n, i) * L.gems_ab[hh] + GAB(k, l), -1.0);
        }
        if (i == m) {
            const int hh = sym[n] ^ sym[j];
            f(L.xoff[wg][hh] + GAB(n, j) * L.gems_ab[hh] + GAB(k, l), 1.0);
        }
        return;
    }
    const int t = L.trip_aaa[h];
    const int *bt = L.baaa + L.baaa_off[h];
    int ijk, lmn;
        split_index(e, t, ijk, lmn);
    const int a = bt[ijk], b = bt[lmn];
    const int i = T0(a), j = T1(a), k = T2i(a), l = T0(b), m = T1(b), n = T2i(b);
    if (self) f(L.xoff[s ? XB_T1BBB : XB_T1AAA][h] + e, -1.0);
    const int w2 = s ? XB_D2BB : XB_D2AA, wd = s ? XB_D1B : XB_D1A;
    auto Q = [&](int p, int q, double c) {   // Q2 (pq|lm)
        const int hh = sym[p] ^ sym[q];
        f(L.xoff[wq][hh] + GAA(p, q) * L.gems_aa[hh] + GAA(l, m), c);
    };
    auto D = [&](int p, int q, double c) {   // D2 (nm|pq)
        const int hh = sym[p] ^ sym[q];
        f(L.xoff[w2][hh] + GAA(n, m) * L.gems_aa[hh] + GAA(p, q), c);
    };
    auto D1 = [&](int p, double c) {         // D1 (n, p)
        const int h3 = sym[n];
        f(L.xoff[wd][h3] + (int64_t)(n - L.poff[h3]) * L.amopi[h3] + (p - L.poff[h3]), c);
    };
    auto G = [&](int p, int q, int r, double c) {   // G2 super block (np|qr), spin quadrant s
        const int hh = sym[n] ^ sym[p];
        const int64_t g = L.gems_ab[hh], sh = s ? g : 0;
        f(L.xoff[XB_G2AA][hh] + (GAB(n, p) + sh) * (2 * g) + (GAB(q, r) + sh), c);
    };
    if (k == n) Q(i, j, 1.0);
    if (j == n) Q(i, k, -1.0);
    if (i == n) Q(j, k, 1.0);
    if (l == k) D(j, i, 1.0);
    if (j == l) D(k, i, -1.0);
    if (l == i) D(k, j, 1.0);
    if (k == m) {
        if (j == l) D1(i, -1.0);
        G(i, j, l, 1.0);
    }
    if (j == m) {
        if (k == l) D1(i, 1.0);
        G(i, k, l, -1.0);
    }
    if (i == m) {
        if (k == l) D1(j, -1.0);
        G(j, k, l, 1.0);
    }
}

// spin-adapted Q2 (kind RF_Q2SA, q2.cc:260-467) and G2 (RF_G2SA,
// g2.cc:235-546) rows: row e of sub-block s of irrep h.  Constants of the
// plain rows are carried by Q1 / D1 terms, so b = 0 on all of them.
template <class F>
V2G_HD inline void sa_terms(const DevLayout &L, int kind, int s, int h, int64_t e, F &f, bool self)
{
    const int amo = L.amo;
    const int *gab = L.gab, *gaa = L.gaa, *sym = L.sym;
    auto GAB = [&](int p, int q) -> int64_t { return gab[p * amo + q]; };
    auto GAA = [&](int p, int q) -> int64_t { return gaa[p * amo + q]; };
    // one-particle entry (p, q) of block w, p and q of one irrep
    auto D1 = [&](int w, int p, int q, double c) {
        const int h3 = sym[p], n = L.amopi[h3];
        f(L.xoff[w][h3] + (int64_t)(p - L.poff[h3]) * n + (q - L.poff[h3]), c);
    };
    if (kind == RF_Q2SA) {
        if (s == 0) {
            // Q2s over singlet geminals (i, j), both orders with weight 1/2,
            // q2.cc:264-319
            const int g0 = L.gems_00[h];
            int IJ, KL;
            split_index(e, g0, IJ, KL);
            const int pij = L.b00[L.b00_off[h] + IJ], pkl = L.b00[L.b00_off[h] + KL];
            const int i = P0(pij), j = P1(pij), k = P0(pkl), l = P1(pkl);
            if (self) f(L.xoff[XB_Q2S][h] + e, -1.0);
            const int64_t g = L.gems_ab[h], b = L.xoff[XB_D2AB][h];
            const int64_t ijd = GAB(i, j), jid = GAB(j, i), kld = GAB(k, l), lkd = GAB(l, k);
            f(b + kld * g + ijd, 0.5);
            f(b + lkd * g + ijd, 0.5);
            f(b + kld * g + jid, 0.5);
            f(b + lkd * g + jid, 0.5);
            if (j == l) { D1(XB_Q1A, i, k, 0.5); D1(XB_D1B, i, k, -0.5); }
            if (i == k) { D1(XB_Q1A, l, j, 0.5); D1(XB_D1B, l, j, -0.5); }
            if (j == k) { D1(XB_Q1A, l, i, 0.5); D1(XB_D1B, l, i, -0.5); }
            if (i == l) { D1(XB_Q1A, k, j, 0.5); D1(XB_D1B, k, j, -0.5); }
            return;
        }
        const int ga = L.gems_aa[h];
        int IJ, KL;
        split_index(e, ga, IJ, KL);
        const int pij = L.baa[L.baa_off[h] + IJ], pkl = L.baa[L.baa_off[h] + KL];
        const int i = P0(pij), j = P1(pij), k = P0(pkl), l = P1(pkl);
        if (s == 1) {
            // Q2t (M_s = 0 triplet), antisymmetric D2ab combination, q2.cc:321-377
            if (self) f(L.xoff[XB_Q2T][h] + e, -1.0);
            const int64_t g = L.gems_ab[h], b = L.xoff[XB_D2AB][h];
            const int64_t ijd = GAB(i, j), jid = GAB(j, i), kld = GAB(k, l), lkd = GAB(l, k);
            f(b + kld * g + ijd, 0.5);
            f(b + lkd * g + ijd, -0.5);
            f(b + kld * g + jid, -0.5);
            f(b + lkd * g + jid, 0.5);
            if (j == l) { D1(XB_Q1A, i, k, 0.5); D1(XB_D1B, i, k, -0.5); }
            if (i == k) { D1(XB_Q1A, l, j, 0.5); D1(XB_D1B, l, j, -0.5); }
            if (j == k) { D1(XB_Q1A, l, i, -0.5); D1(XB_D1B, l, i, 0.5); }
            if (i == l) { D1(XB_Q1A, k, j, -0.5); D1(XB_D1B, k, j, 0.5); }
            return;
        }
        // Q2t(+1) from D2aa, Q2t(-1) from D2bb, q2.cc:378-465
        const bool bb = s == 3;
        const int wq = bb ? XB_Q1B : XB_Q1A, wd = bb ? XB_D1B : XB_D1A;
        if (self) f(L.xoff[bb ? XB_Q2TM1 : XB_Q2TP1][h] + e, -1.0);
        f(L.xoff[bb ? XB_D2BB : XB_D2AA][h] + (int64_t)KL * ga + IJ, 1.0);
        if (j == l) D1(wq, i, k, 1.0);
        if (j == k) D1(wd, l, i, 1.0);
        if (i == l) D1(wq, j, k, -1.0);
        if (i == k) D1(wd, l, j, -1.0);
        return;
    }
    // spin-adapted G2 over ab pairs (i, j), (k, l)
    const int g = L.gems_ab[h];
    int IJ, KL;
    split_index(e, g, IJ, KL);
    const int pij = L.bab[L.bab_off[h] + IJ], pkl = L.bab[L.bab_off[h] + KL];
    const int i = P0(pij), j = P1(pij), k = P0(pkl), l = P1(pkl);
    const int h2 = sym[i] ^ sym[l];
    const int64_t g2 = L.gems_ab[h2], b2 = L.xoff[XB_D2AB][h2];
    if (s < 2) {
        // G2s (s = 0) and G2t (s = 1): the like-spin blocks and the mixed
        // ones with weight 1/2, the latter with sign +1 / -1; g2.cc:240-376
        const double sg = s ? -0.5 : 0.5;
        if (self) f(L.xoff[s ? XB_G2T : XB_G2S][h] + e, -1.0);
        if (j == l) { D1(XB_D1A, i, k, 0.5); D1(XB_D1B, i, k, 0.5); }
        if (i != l && k != j) {
            const double c = (i < l) == (k < j) ? -0.5 : 0.5;
            const int64_t col = GAA(i, l) * L.gems_aa[h2] + GAA(k, j);
            f(L.xoff[XB_D2AA][h2] + col, c);
            f(L.xoff[XB_D2BB][h2] + col, c);
        }
        f(b2 + GAB(i, l) * g2 + GAB(j, k), sg);
        f(b2 + GAB(l, i) * g2 + GAB(k, j), sg);
        return;
    }
    // G2t(+1) and G2t(-1): the rows of G2ab / G2ba, g2.cc:378-469
    const bool m1 = s == 3;
    if (self) f(L.xoff[m1 ? XB_G2TM1 : XB_G2TP1][h] + e, -1.0);
    if (j == l) D1(m1 ? XB_D1B : XB_D1A, i, k, 1.0);
    if (m1) f(b2 + GAB(l, i) * g2 + GAB(j, k), -1.0);
    else f(b2 + GAB(i, l) * g2 + GAB(k, j), -1.0);
}

// the rows walked by the second pass of A^T.y (T1, T2, spin-adapted Q2 / G2):
// the caller writes the self term, these scatter the rest
template <class F>
V2G_HD inline void scatter_terms(const DevLayout &L, int kind, int s, int h, int64_t e, F &f,
                                 bool self)
{
    if (kind == RF_Q2SA || kind == RF_G2SA) sa_terms(L, kind, s, h, e, f, self);
    else if (kind == RF_T1AAB || kind == RF_T1AAA) t1_terms(L, kind, s, h, e, f, self);
    else t2_terms(L, kind, s, h, e, f, self);
}

template <class F>
V2G_HD inline void elem_terms(const DevLayout &L, const Segment &sg, int64_t e, F &f)
{
    const int amo = L.amo, h = sg.h, s = sg.sub;
    const int *gab = L.gab, *gaa = L.gaa, *sym = L.sym;
    auto GAB = [&](int p, int q) -> int64_t { return gab[p * amo + q]; };
    auto GAA = [&](int p, int q) -> int64_t { return gaa[p * amo + q]; };
    switch (sg.kind) {
    case RF_D1Q1: {  // d2.cc:370-388
        const int n = sg.dim;
        int i, j;
        split_index(e, n, i, j);
        f(L.xoff[s ? XB_D1B : XB_D1A][h] + j * n + i, 1.0);
        f(L.xoff[s ? XB_Q1B : XB_Q1A][h] + e, 1.0);
        return;
    }
    case RF_SPIN_D1:  // spin.cc:324-328
        f(L.xoff[XB_D1A][h] + e, 1.0);
        f(L.xoff[XB_D1B][h] + e, -1.0);
        return;
    case RF_SPIN_D2:  // spin.cc:330-334
        f(L.xoff[XB_D2AA][h] + e, 1.0);
        f(L.xoff[XB_D2BB][h] + e, -1.0);
        return;
    case RF_SPIN_ANTI: {  // spin.cc:336-373
        const int ga = sg.dim;
        int IJ, KL;
        split_index(e, ga, IJ, KL);
        const int pij = L.baa[L.baa_off[h] + IJ], pkl = L.baa[L.baa_off[h] + KL];
        const int i = P0(pij), j = P1(pij), k = P0(pkl), l = P1(pkl);
        const int64_t g = L.gems_ab[h], b = L.xoff[XB_D2AB][h];
        const int64_t ijb = GAB(i, j), jib = GAB(j, i), klb = GAB(k, l), lkb = GAB(l, k);
        f(L.xoff[s ? XB_D2BB : XB_D2AA][h] + e, 1.0);
        f(b + ijb * g + klb, -0.5);
        f(b + jib * g + klb, 0.5);
        f(b + ijb * g + lkb, 0.5);
        f(b + jib * g + lkb, -0.5);
        return;
    }
    case RF_SPIN_D200: {  // spin.cc:375-395
        const int g = sg.dim;
        int IJ, KL;
        split_index(e, g, IJ, KL);
        const int pij = L.bab[L.bab_off[h] + IJ], pkl = L.bab[L.bab_off[h] + KL];
        const int i = P0(pij), j = P1(pij), k = P0(pkl), l = P1(pkl);
        const int64_t ji = GAB(j, i), lk = GAB(l, k), b = L.xoff[XB_D2AB][h];
        const double dij = (i == j) ? 1.4142135623730951 : 1.0;
        const double dkl = (k == l) ? 1.4142135623730951 : 1.0;
        const double c = -0.5 / (dij * dkl);
        f(L.xoff[XB_D200][h] + e, 1.0);
        f(b + (int64_t)IJ * g + KL, c);
        f(b + ji * g + KL, c);
        f(b + (int64_t)IJ * g + lk, c);
        f(b + ji * g + lk, c);
        return;
    }
    case RF_SPIN_D200_4: {  // spin.cc:396-469
        const int g = sg.dim;
        const int64_t G = 2 * (int64_t)g;
        int R, C;
        split_index(e, G, R, C);
        const int a = R >= g, bb = C >= g;
        const int IJ = R - a * g, KL = C - bb * g;
        const int pij = L.bab[L.bab_off[h] + IJ], pkl = L.bab[L.bab_off[h] + KL];
        const int i = P0(pij), j = P1(pij), k = P0(pkl), l = P1(pkl);
        const int64_t ji = GAB(j, i), lk = GAB(l, k), b = L.xoff[XB_D2AB][h];
        const double dij = (i == j) ? 1.4142135623730951 : 1.0;
        const double dkl = (k == l) ? 1.4142135623730951 : 1.0;
        double c1, c2, c3, c4;
        if (!a && !bb) { const double c = -0.5 / (dij * dkl); c1 = c2 = c3 = c4 = c; }
        else if (!a) { const double c = 0.5 / dij; c1 = -c; c2 = -c; c3 = c; c4 = c; }
        else if (!bb) { const double c = 0.5 / dkl; c1 = -c; c2 = c; c3 = -c; c4 = c; }
        else { c1 = -0.5; c2 = 0.5; c3 = 0.5; c4 = -0.5; }
        f(L.xoff[XB_D200][h] + e, 1.0);
        f(b + (int64_t)IJ * g + KL, c1);
        f(b + ji * g + KL, c2);
        f(b + (int64_t)IJ * g + lk, c3);
        f(b + ji * g + lk, c4);
        return;
    }
    case RF_Q2AB: {  // q2.cc:813-849
        const int g = sg.dim;
        int IJ, KL;
        split_index(e, g, IJ, KL);
        const int pij = L.bab[L.bab_off[h] + IJ], pkl = L.bab[L.bab_off[h] + KL];
        const int i = P0(pij), j = P1(pij), k = P0(pkl), l = P1(pkl);
        f(L.xoff[XB_D2AB][h] + e, 1.0);
        f(L.xoff[XB_Q2AB][h] + e, -1.0);
        if (j == l) {
            const int h3 = sym[i], n = L.amopi[h3];
            f(L.xoff[XB_D1A][h3] + (k - L.poff[h3]) * n + (i - L.poff[h3]), -1.0);
        }
        if (i == k) {
            const int h3 = sym[j], n = L.amopi[h3];
            f(L.xoff[XB_D1B][h3] + (l - L.poff[h3]) * n + (j - L.poff[h3]), -1.0);
        }
        return;
    }
    case RF_Q2AA: {  // q2.cc:851-948
        const int ga = sg.dim;
        int IJ, KL;
        split_index(e, ga, IJ, KL);
        const int pij = L.baa[L.baa_off[h] + IJ], pkl = L.baa[L.baa_off[h] + KL];
        const int i = P0(pij), j = P1(pij), k = P0(pkl), l = P1(pkl);
        const int wd = s ? XB_D1B : XB_D1A;
        f(L.xoff[s ? XB_D2BB : XB_D2AA][h] + e, 1.0);
        f(L.xoff[s ? XB_Q2BB : XB_Q2AA][h] + e, -1.0);
        if (j == l) {
            const int h3 = sym[i], n = L.amopi[h3];
            f(L.xoff[wd][h3] + (k - L.poff[h3]) * n + (i - L.poff[h3]), -1.0);
        }
        if (j == k) {
            const int h3 = sym[i], n = L.amopi[h3];
            f(L.xoff[wd][h3] + (l - L.poff[h3]) * n + (i - L.poff[h3]), 1.0);
        }
        if (i == l) {
            const int h3 = sym[j], n = L.amopi[h3];
            f(L.xoff[wd][h3] + (k - L.poff[h3]) * n + (j - L.poff[h3]), 1.0);
        }
        if (i == k) {
            const int h3 = sym[j], n = L.amopi[h3];
            f(L.xoff[wd][h3] + (l - L.poff[h3]) * n + (j - L.poff[h3]), -1.0);
        }
        return;
    }
    case RF_G2AB: {  // g2.cc:1067-1127
        const int g = sg.dim;
        int IJ, KL;
        split_index(e, g, IJ, KL);
        const int pij = L.bab[L.bab_off[h] + IJ], pkl = L.bab[L.bab_off[h] + KL];
        const int i = P0(pij), j = P1(pij), k = P0(pkl), l = P1(pkl);
        f(L.xoff[s ? XB_G2BA : XB_G2AB][h] + e, -1.0);
        if (j == l) {
            const int h3 = sym[i], n = L.amopi[h3];
            f(L.xoff[s ? XB_D1B : XB_D1A][h3] + (i - L.poff[h3]) * n + (k - L.poff[h3]), 1.0);
        }
        const int h2 = sym[i] ^ sym[l];
        const int64_t r2 = s ? GAB(l, i) : GAB(i, l), c2 = s ? GAB(j, k) : GAB(k, j);
        f(L.xoff[XB_D2AB][h2] + r2 * L.gems_ab[h2] + c2, -1.0);
        return;
    }
    case RF_G2AA: {  // g2.cc:1128-1257
        const int g = sg.dim;
        const int64_t G = 2 * (int64_t)g;
        int R, C;
        split_index(e, G, R, C);
        const int ra = R >= g, cb = C >= g;
        const int IJ = R - ra * g, KL = C - cb * g;
        const int pij = L.bab[L.bab_off[h] + IJ], pkl = L.bab[L.bab_off[h] + KL];
        const int i = P0(pij), j = P1(pij), k = P0(pkl), l = P1(pkl);
        const int h2 = sym[i] ^ sym[l];
        f(L.xoff[XB_G2AA][h] + e, -1.0);
        if (ra == cb) {
            if (j == l) {
                const int h3 = sym[i], n = L.amopi[h3];
                f(L.xoff[ra ? XB_D1B : XB_D1A][h3] + (i - L.poff[h3]) * n + (k - L.poff[h3]), 1.0);
            }
            if (i != l && k != j) {
                const double sgn = ((i < l) == (k < j)) ? 1.0 : -1.0;
                f(L.xoff[ra ? XB_D2BB : XB_D2AA][h2] + GAA(i, l) * L.gems_aa[h2] + GAA(k, j), -sgn);
            }
        } else if (!ra) {
            f(L.xoff[XB_D2AB][h2] + GAB(i, l) * L.gems_ab[h2] + GAB(j, k), 1.0);
        } else {
            f(L.xoff[XB_D2AB][h2] + GAB(l, i) * L.gems_ab[h2] + GAB(k, j), 1.0);
        }
        return;
    }
    case RF_T2AAB:
    case RF_T2BIG:
        t2_terms(L, sg.kind, s, h, e, f, true);
        return;
    case RF_T1AAB:
    case RF_T1AAA:
        t1_terms(L, sg.kind, s, h, e, f, true);
        return;
    case RF_Q2SA:
    case RF_G2SA:
        sa_terms(L, sg.kind, s, h, e, f, true);
        return;
    default:
        return;
    }
}

// one wavefront per row (traces, S^2, contractions, maximal-spin rows): the
// lane-strided share of the row's terms.  Columns are distinct across the
// terms of one row except in the trace rows (handled in the diagonal pass).
template <class F>
V2G_HD inline void wave_terms(const DevLayout &L, const Segment &sg, int64_t item,
                                           int lane, F &f)
{
    const int amo = L.amo;
    const int *gab = L.gab, *gaa = L.gaa, *sym = L.sym;
    switch (sg.kind) {
    case RF_TRACE: {  // d2.cc:303-366
        for (int pq = lane; pq < amo * amo; pq += 64) {
            const int i = pq / amo, j = pq % amo, h = sym[i] ^ sym[j];
            const int64_t gb = L.gems_ab[h], ga = L.gems_aa[h];
            if (sg.sub == 0) {
                const int64_t ij = gab[i * amo + j];
                f(L.xoff[XB_D2AB][h] + ij * gb + ij, 1.0);
            } else if (sg.sub == 1 || sg.sub == 2) {
                if (i == j) continue;
                const int64_t ij = gaa[i * amo + j];
                f(L.xoff[sg.sub == 1 ? XB_D2AA : XB_D2BB][h] + ij * ga + ij, 1.0);
            } else {
                int64_t ij = gab[i * amo + j];
                f(L.xoff[XB_D2AB][h] + ij * gb + ij, 2.0);
                if (i == j) continue;
                ij = gaa[i * amo + j];
                f(L.xoff[XB_D2AA][h] + ij * ga + ij, 1.0);
                f(L.xoff[XB_D2BB][h] + ij * ga + ij, 1.0);
            }
        }
        return;
    }
    case RF_S2: {  // spin.cc:281-292
        for (int pq = lane; pq < amo * amo; pq += 64) {
            const int i = pq / amo, j = pq % amo, h = sym[i] ^ sym[j];
            f(L.xoff[XB_D2AB][h] + (int64_t)gab[i * amo + j] * L.gems_ab[h] + gab[j * amo + i], 1.0);
        }
        return;
    }
    case RF_CONTR: {  // d2.cc:398-545
        const int h = sg.h, n = sg.dim;
        const int i = (int)(item / n), j = (int)(item % n);
        const int ii = i + L.poff[h], jj = j + L.poff[h];
        const int sub = sg.sub;
        const bool beta = (sub == 1 || sub == 3 || sub == 5);
        const bool use_ab = (sub == 0 || sub == 1 || sub >= 4);
        const bool use_aa = (sub >= 2);
        if (lane == 0) {
            double coef;
            switch (sub) {
            case 0: coef = L.nb; break;
            case 1: coef = L.na; break;
            case 2: coef = L.na - 1.0; break;
            case 3: coef = L.nb - 1.0; break;
            default: coef = L.na + L.nb - 1.0; break;
            }
            f(L.xoff[beta ? XB_D1B : XB_D1A][h] + i * n + j, coef);
        }
        for (int k = lane; k < amo; k += 64) {
            const int h2 = sym[ii] ^ sym[k];
            if (use_ab) {
                const int64_t ik = beta ? gab[k * amo + ii] : gab[ii * amo + k];
                const int64_t jk = beta ? gab[k * amo + jj] : gab[jj * amo + k];
                f(L.xoff[XB_D2AB][h2] + ik * L.gems_ab[h2] + jk, -1.0);
            }
            if (use_aa && ii != k && jj != k) {
                const int64_t ik = gaa[ii * amo + k], jk = gaa[jj * amo + k];
                const double sg2 = ((ii < k) == (jj < k)) ? 1.0 : -1.0;
                f(L.xoff[beta ? XB_D2BB : XB_D2AA][h2] + ik * L.gems_aa[h2] + jk, -sg2);
            }
        }
        return;
    }
    case RF_MAXSPIN: {  // spin.cc:471-491
        const int64_t g0 = L.gems_ab[0], kl = item;
        for (int i = lane; i < amo; i += 64) {
            const int64_t ii = gab[i * amo + i];
            f(L.xoff[XB_G2BA][0] + (sg.sub == 0 ? kl * g0 + ii : ii * g0 + kl), 1.0);
        }
        return;
    }
    default:
        return;
    }
}

// squared row norm of an elementwise row: merge repeated columns first
struct Collect {
    int64_t col[16];
    double coef[16];
    int n = 0;
    V2G_HD inline void operator()(int64_t c, double v)
    {
        for (int k = 0; k < n; k++)
            if (col[k] == c) { coef[k] += v; return; }
        col[n] = c;
        coef[n] = v;
        n++;
    }
    V2G_HD inline double norm2() const
    {
        double s = 0.0;
        for (int k = 0; k < n; k++) s += coef[k] * coef[k];
        return s;
    }
};

struct SumSq {
    double acc = 0.0;
    V2G_HD inline void operator()(int64_t, double v) { acc += v * v; }
};

// ----------------------------------------------------------------------------
// processing order.  Rows (or primal entries) of a block are written where
// they belong, but the kernels may visit them in another order.  In the G2ba
// and G2 super-block rows (and in the D2ab entries of A^T), the D2ab / G2ba
// operand's column index runs over the second orbital of a pair, so one row
// pair touches single doubles spread over many cache lines; the row pair with
// the transposed orbitals touches the neighbouring doubles of the same lines.
// Visiting the row pairs in transposed order puts those neighbours in
// consecutive workgroups, so the lines are reused from L2.
// ----------------------------------------------------------------------------

// ab pair index of the transposed pair of the pair at position pos of irrep h
V2G_HD inline int ab_transposed(const DevLayout &L, int h, int pos)
{
    const int pr = L.bab[L.bab_off[h] + pos];
    return L.gab[P1(pr) * L.amo + P0(pr)];
}

// row offset (within its segment) visited at position pos of an A.x segment
V2G_HD inline int64_t row_of(const DevLayout &L, const Segment &sg, int64_t pos)
{
    if (sg.kind == RF_G2AB && sg.sub == 1) {
        int IJ, KL;
        split_index(pos, sg.dim, IJ, KL);
        return (int64_t)ab_transposed(L, sg.h, IJ) * sg.dim + KL;
    }
    if (sg.kind == RF_G2AA) {
        const int g = sg.dim;
        int R, C;
        split_index(pos, 2 * (int64_t)g, R, C);
        const int half = R >= g;
        return ((int64_t)half * g + ab_transposed(L, sg.h, R - half * g)) * (2 * (int64_t)g) + C;
    }
    return pos;
}

// primal entry (within its segment) visited at position pos of an A^T.y segment
V2G_HD inline int64_t entry_of(const DevLayout &L, const Segment &sg, int64_t pos)
{
    if (sg.kind == XS_D2AB) {
        int PQ, RS;
        split_index(pos, sg.dim, PQ, RS);
        return (int64_t)ab_transposed(L, sg.h, PQ) * sg.dim + RS;
    }
    return pos;
}

// ----------------------------------------------------------------------------
// A^T.y gathers
// ----------------------------------------------------------------------------

V2G_HD inline double atu_d2ab(const DevLayout &L, const Segment &sg, int64_t e,
                           const double *__restrict__ y)
{
    const int amo = L.amo, h = sg.h, g = sg.dim;
    const int *gab = L.gab, *gaa = L.gaa, *sym = L.sym;
    auto GAB = [&](int p, int q) -> int64_t { return gab[p * amo + q]; };
    int IJ, KL;
        split_index(e, g, IJ, KL);
    const int pij = L.bab[L.bab_off[h] + IJ], pkl = L.bab[L.bab_off[h] + KL];
    const int p = P0(pij), q = P1(pij), r = P0(pkl), s = P1(pkl);
    double v = 0.0;
    // trace rows
    if (IJ == KL) v += L.sz ? y[L.rowoff[RF_TRACE][0][0]] : 2.0 * y[L.rowoff[RF_TRACE][3][0]];
    // D2ab -> D1a : row (p, r) of irrep sym p when q == s
    if (q == s) {
        const int hp = sym[p], n = L.amopi[hp];
        v -= y[L.rowoff[RF_CONTR][L.sz ? 0 : 4][hp] + (p - L.poff[hp]) * n + (r - L.poff[hp])];
    }
    // D2ab -> D1b : row (q, s) when p == r
    if (p == r) {
        const int hq = sym[q], n = L.amopi[hq];
        v -= y[L.rowoff[RF_CONTR][L.sz ? 1 : 5][hq] + (q - L.poff[hq]) * n + (s - L.poff[hq])];
    }
    if (L.spin) {
        if (p == s && q == r) v += y[L.rowoff[RF_S2][0][0]];
        const int64_t qp = GAB(q, p), sr = GAB(s, r);
        const double dpq = (p == q) ? 1.4142135623730951 : 1.0;
        const double drs = (r == s) ? 1.4142135623730951 : 1.0;
        if (L.singlet) {
            if (p != q && r != s) {
                const int64_t ga = L.gems_aa[h];
                const int64_t ia = gaa[p * amo + q], ka = gaa[r * amo + s];
                const double sg2 = ((p > q) == (r > s)) ? 1.0 : -1.0;
                v += -0.5 * sg2 *
                     (y[L.rowoff[RF_SPIN_ANTI][0][h] + ia * ga + ka] + y[L.rowoff[RF_SPIN_ANTI][1][h] + ia * ga + ka]);
            }
            const int64_t R0 = L.rowoff[RF_SPIN_D200][0][h];
            const double c = -0.5 / (dpq * drs);
            v += c * (y[R0 + (int64_t)IJ * g + KL] + y[R0 + qp * g + KL] + y[R0 + (int64_t)IJ * g + sr] +
                      y[R0 + qp * g + sr]);
        } else {
            const int64_t G = 2 * (int64_t)g;
            const int64_t R4 = L.rowoff[RF_SPIN_D200_4][0][h];
            for (int blk = 0; blk < 4; blk++) {
                const int a = blk >> 1, b = blk & 1;
                double c1, c2, c3, c4;
                if (!a && !b) { const double c = -0.5 / (dpq * drs); c1 = c2 = c3 = c4 = c; }
                else if (!a) { const double c = 0.5 / dpq; c1 = -c; c2 = -c; c3 = c; c4 = c; }
                else if (!b) { const double c = 0.5 / drs; c1 = -c; c2 = c; c3 = -c; c4 = c; }
                else { c1 = -0.5; c2 = 0.5; c3 = 0.5; c4 = -0.5; }
                const int64_t base = R4 + (int64_t)a * g * G + (int64_t)b * g;
                v += c1 * y[base + IJ * G + KL] + c2 * y[base + qp * G + KL] +
                     c3 * y[base + IJ * G + sr] + c4 * y[base + qp * G + sr];
            }
        }
    }
    if (L.q2) v += y[L.rowoff[RF_Q2AB][0][h] + e];
    if (L.g2) {
        {   // G2ab row ((p,s),(r,q))
            const int h1 = sym[p] ^ sym[s];
            v -= y[L.rowoff[RF_G2AB][0][h1] + GAB(p, s) * L.gems_ab[h1] + GAB(r, q)];
        }
        {   // G2ba row ((q,r),(s,p))
            const int h2 = sym[q] ^ sym[r];
            v -= y[L.rowoff[RF_G2AB][1][h2] + GAB(q, r) * L.gems_ab[h2] + GAB(s, p)];
        }
        {   // G2aabb row ((p,r),(s,q))
            const int h3 = sym[p] ^ sym[r];
            const int64_t g3 = L.gems_ab[h3];
            v += y[L.rowoff[RF_G2AA][0][h3] + GAB(p, r) * 2 * g3 + g3 + GAB(s, q)];
        }
        {   // G2bbaa row ((q,s),(r,p))
            const int h4 = sym[q] ^ sym[s];
            const int64_t g4 = L.gems_ab[h4];
            v += y[L.rowoff[RF_G2AA][0][h4] + (g4 + GAB(q, s)) * 2 * g4 + GAB(r, p)];
        }
    }
    return v;
}

V2G_HD inline double atu_d2aa(const DevLayout &L, const Segment &sg, int64_t e,
                           const double *__restrict__ y)
{
    const int amo = L.amo, h = sg.h, ga = sg.dim, sp = sg.sub;
    const int *gab = L.gab, *sym = L.sym;
    int IJ, KL;
        split_index(e, ga, IJ, KL);
    const int pij = L.baa[L.baa_off[h] + IJ], pkl = L.baa[L.baa_off[h] + KL];
    const int p1 = P0(pij), p2 = P1(pij), r1 = P0(pkl), r2 = P1(pkl);
    double v = 0.0;
    if (IJ == KL) v += 2.0 * y[L.rowoff[RF_TRACE][L.sz ? 1 + sp : 3][0]];
    // contraction rows: common orbital k of the two pairs
    const int csub = L.sz ? 2 + sp : 4 + sp;
    for (int a = 0; a < 2; a++) {
        const int k = a ? p2 : p1, ii = a ? p1 : p2;
        int jj = -1;
        if (r1 == k) jj = r2;
        else if (r2 == k) jj = r1;
        if (jj < 0) continue;
        const int hi = sym[ii], n = L.amopi[hi];
        const double sg2 = ((ii < k) == (jj < k)) ? 1.0 : -1.0;
        v -= sg2 * y[L.rowoff[RF_CONTR][csub][hi] + (ii - L.poff[hi]) * n + (jj - L.poff[hi])];
    }
    if (L.spin && L.singlet) {
        v += (sp ? -1.0 : 1.0) * y[L.rowoff[RF_SPIN_D2][0][h] + e];
        v += y[L.rowoff[RF_SPIN_ANTI][sp][h] + e];
    }
    if (L.q2) v += y[L.rowoff[RF_Q2AA][sp][h] + e];
    if (L.g2) {
        // G2aaaa / G2bbbb rows ((i,j),(k,l)) with {i,l} = pair 1, {k,j} = pair 2
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                const int i = a ? p2 : p1, l = a ? p1 : p2;
                const int k = b ? r2 : r1, j = b ? r1 : r2;
                const int hh = sym[i] ^ sym[j];
                const int64_t gg = L.gems_ab[hh];
                const int64_t IJb = gab[i * amo + j], KLb = gab[k * amo + l];
                const int64_t pos = sp ? (gg + IJb) * 2 * gg + gg + KLb : IJb * 2 * gg + KLb;
                const double sg2 = ((i < l) == (k < j)) ? 1.0 : -1.0;
                v -= sg2 * y[L.rowoff[RF_G2AA][0][hh] + pos];
            }
    }
    return v;
}

V2G_HD inline double atu_d1_wave(const DevLayout &L, const Segment &sg, int64_t item, int lane,
                              const double *__restrict__ y)
{
    const int amo = L.amo, h = sg.h, n = sg.dim, sp = sg.sub;
    const int *gab = L.gab, *gaa = L.gaa, *sym = L.sym;
    const int a = (int)(item / n), b = (int)(item % n);
    const int A = a + L.poff[h], B = b + L.poff[h];
    double v = 0.0;
    if (lane == 0) {
        v += y[L.rowoff[RF_D1Q1][sp][h] + b * n + a];
        if (L.sz) {
            v += (sp ? L.na : L.nb) * y[L.rowoff[RF_CONTR][sp][h] + item];
            v += (sp ? L.nb - 1.0 : L.na - 1.0) * y[L.rowoff[RF_CONTR][2 + sp][h] + item];
        } else {
            v += (L.na + L.nb - 1.0) * y[L.rowoff[RF_CONTR][4 + sp][h] + item];
        }
        if (L.spin && L.singlet) v += (sp ? -1.0 : 1.0) * y[L.rowoff[RF_SPIN_D1][0][h] + item];
    }
    for (int t = lane; t < amo; t += 64) {
        const int ht = sym[B] ^ sym[t];   // = sym[A] ^ sym[t]
        const int64_t g = L.gems_ab[ht];
        if (L.q2) {
            if (!sp) v -= y[L.rowoff[RF_Q2AB][0][ht] + (int64_t)gab[B * amo + t] * g + gab[A * amo + t]];
            else v -= y[L.rowoff[RF_Q2AB][0][ht] + (int64_t)gab[t * amo + B] * g + gab[t * amo + A]];
            if (t != A && t != B) {
                const int64_t ga = L.gems_aa[ht];
                const double sg2 = ((B > t) == (A > t)) ? 1.0 : -1.0;
                v -= sg2 * y[L.rowoff[RF_Q2AA][sp][ht] + (int64_t)gaa[B * amo + t] * ga + gaa[A * amo + t]];
            }
        }
        if (L.g2) {
            const int64_t ia = gab[A * amo + t], ib = gab[B * amo + t];
            v += y[L.rowoff[RF_G2AB][sp][ht] + ia * g + ib];
            const int64_t pos = sp ? (g + ia) * 2 * g + g + ib : ia * 2 * g + ib;
            v += y[L.rowoff[RF_G2AA][0][ht] + pos];
        }
    }
    return v;  // this lane's share; the caller sums the wavefront
}

V2G_HD inline double atu_self(const DevLayout &L, const Segment &sg, int64_t e,
                           const double *__restrict__ y)
{
    double v = sg.coef * y[sg.aux + e];
    if (sg.sub == 1) {  // G2ba of irrep 0 also appears in the maximal-spin rows
        const int g = sg.dim;
        int IJ, KL;
        split_index(e, g, IJ, KL);
        const int pij = L.bab[L.bab_off[0] + IJ], pkl = L.bab[L.bab_off[0] + KL];
        if (P0(pkl) == P1(pkl)) v += y[L.rowoff[RF_MAXSPIN][0][0] + IJ];
        if (P0(pij) == P1(pij)) v += y[L.rowoff[RF_MAXSPIN][1][0] + KL];
    }
    return v;
}


}  // namespace rows
}  // namespace v2g
