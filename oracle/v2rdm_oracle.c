/*
 * v2rdm_oracle.c -- CPU restatement of the v2RDM-CASSCF BPSDP hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity oracle for the HIP path and the timed
 * CPU baseline.  Loaded by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg; the product never links it.
 *
 * Parity pinning of each part (checks live in tests/test_oracle_*.py):
 *   layout sizes      dimx / ncon against the counts printed in the
 *                     reference's tests/v2rdm{1..8}/output.ref (golden).
 *   Au rows + b       exact-RDM feasibility: A.x(FCI RDMs) = b to 1e-12 for
 *                     brute-force FCI states (C1 and D2h, singlet and triplet);
 *                     T1/T2 blocks built physically as <B+B> + <BB+>.
 *   ATu               adjointness <A x, y> = <x, A^T y> against Au.
 *   c, energy         c.x(FCI RDMs) = E_FCI.
 *   CG                against a dense least-squares solve of A A^T y = rhs.
 *   Update_xz         closed form: mu x - z = sym(M), x,z PSD, x z = 0.
 *   BPSDP loop        N=2 (D exact) and 2-hole (Q exact) energies = FCI;
 *                     D2h and C1 runs of one Hamiltonian agree.
 *   Conventions that no mathematical check can see (the transposition of a
 *   D1 index inside a block whose exact value is symmetric) follow the
 *   reference source line by line; the HIP path is held to them by
 *   random-input per-call parity.
 */
#define _GNU_SOURCE
#include "v2rdm_oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXAMO 128

struct orc_layout {
    orc_options o;
    int nirrep, amo, na, nb;
    int sym[ORC_MAXAMO];
    int poff[8];
    int *gab, *gaa;                   /* amo*amo */
    int gems_ab[8], gems_aa[8];
    int *bab[8], *baa[8];             /* pairs (first, second) */
    int gems_00[8];                   /* singlet geminals (SPIN_ADAPT_Q2) */
    int *b00[8];                      /* their pairs (i, j), energy order */
    int trip_aab[8], trip_aba[8], trip_aaa[8];
    int *baab[8], *baba[8], *baaa[8]; /* triplets (i, j, k) */
    long off[ORC_OFF_COUNT][8];
    long dimx, ncon;
    int nblocks;
    int *bdim;
    long *boff;
    int nfam;
    char famname[32][24];
    long famstart[32], famcount[32];
};

static double wall(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* layout: basis.cc:57-548, 655-757; v2rdm_solver.cc:560-1270               */
/* ------------------------------------------------------------------------ */

static void add_family(orc_layout *L, const char *name, long *cursor, long count)
{
    if (count <= 0) return;
    int f = L->nfam++;
    snprintf(L->famname[f], 24, "%s", name);
    L->famstart[f] = *cursor;
    L->famcount[f] = count;
    *cursor += count;
}

orc_layout *orc_layout_create(const orc_options *opt)
{
    orc_layout *L = (orc_layout *)calloc(1, sizeof(orc_layout));
    L->o = *opt;
    L->nirrep = opt->nirrep;
    int amo = 0;
    for (int h = 0; h < L->nirrep; h++) {
        L->poff[h] = amo;
        for (int i = 0; i < opt->amopi[h]; i++) L->sym[amo++] = h;
    }
    if (amo > ORC_MAXAMO) { free(L); return NULL; }
    /* spin adaptation is refused where the reference refuses it
     * (v2rdm_solver.cc:538-561: with T1 / T2, and for M_s > 0) and where its
     * rows would address a block it never lays out (the maximal-spin rows
     * of CONSTRAIN_SPIN read G2ba, spin.cc:250-268, which SPIN_ADAPT_G2
     * replaces) */
    if ((opt->spin_adapt_q2 || opt->spin_adapt_g2) &&
        (opt->constrain_t1 || opt->constrain_t2 || opt->multiplicity > 1)) {
        free(L);
        return NULL;
    }
    if (opt->spin_adapt_g2 && opt->constrain_g2 && opt->constrain_spin) { free(L); return NULL; }
    L->amo = amo;
    L->na = opt->nalpha;
    L->nb = opt->nbeta;
    L->gab = (int *)malloc(sizeof(int) * amo * amo);
    L->gaa = (int *)malloc(sizeof(int) * amo * amo);
    for (int e = 0; e < amo * amo; e++) { L->gab[e] = -999; L->gaa[e] = -999; }
    /* geminals: outer loop i, inner j, pair stored (j, i); basis.cc:439-452,528-545 */
    for (int h = 0; h < L->nirrep; h++) {
        L->bab[h] = (int *)malloc(sizeof(int) * 2 * (amo * amo + 1));
        L->baa[h] = (int *)malloc(sizeof(int) * 2 * (amo * amo + 1));
        int nab = 0, naa = 0;
        for (int i = 0; i < amo; i++)
            for (int j = 0; j < amo; j++) {
                if ((L->sym[i] ^ L->sym[j]) != h) continue;
                int first = j, second = i;
                L->gab[first * amo + second] = nab;
                L->bab[h][2 * nab] = first;
                L->bab[h][2 * nab + 1] = second;
                nab++;
                if (first <= second) continue;
                L->gaa[first * amo + second] = naa;
                L->gaa[second * amo + first] = naa;
                L->baa[h][2 * naa] = first;
                L->baa[h][2 * naa + 1] = second;
                naa++;
            }
        L->gems_ab[h] = nab;
        L->gems_aa[h] = naa;
    }
    /* singlet geminals: energy-ordered pairs (ieo outer, jeo <= ieo inner),
     * stored (i(ieo), j(jeo)); basis.cc:626-651 */
    for (int h = 0; h < L->nirrep; h++) {
        L->b00[h] = (int *)malloc(sizeof(int) * (amo * (amo + 1) + 2));
        L->gems_00[h] = 0;
    }
    for (int ie = 0; ie < amo; ie++) {
        int i = opt->energy_order ? opt->energy_order[ie] : ie;
        for (int je = 0; je <= ie; je++) {
            int j = opt->energy_order ? opt->energy_order[je] : je;
            int h = L->sym[i] ^ L->sym[j];
            int n00 = L->gems_00[h]++;
            L->b00[h][2 * n00] = i;
            L->b00[h][2 * n00 + 1] = j;
        }
    }
    /* triplets: basis.cc:655-757 (T1 and T2); aba: all (i,j,k), aab: i < j,
     * aaa: i < j < k, each in (i outer, k inner) order */
    if (opt->constrain_t2 || opt->constrain_t1) {
        for (int h = 0; h < L->nirrep; h++) {
            L->baab[h] = (int *)malloc(sizeof(int) * 3 * (amo * amo * amo + 1));
            L->baba[h] = (int *)malloc(sizeof(int) * 3 * (amo * amo * amo + 1));
            L->baaa[h] = (int *)malloc(sizeof(int) * 3 * (amo * amo * amo + 1));
            int caba = 0, caab = 0, caaa = 0;
            for (int i = 0; i < amo; i++)
                for (int j = 0; j < amo; j++)
                    for (int k = 0; k < amo; k++) {
                        if ((L->sym[i] ^ L->sym[j] ^ L->sym[k]) != h) continue;
                        L->baba[h][3 * caba] = i; L->baba[h][3 * caba + 1] = j;
                        L->baba[h][3 * caba + 2] = k;
                        caba++;
                        if (i >= j) continue;
                        L->baab[h][3 * caab] = i; L->baab[h][3 * caab + 1] = j;
                        L->baab[h][3 * caab + 2] = k;
                        caab++;
                        if (j >= k) continue;
                        L->baaa[h][3 * caaa] = i; L->baaa[h][3 * caaa + 1] = j;
                        L->baaa[h][3 * caaa + 2] = k;
                        caaa++;
                    }
            L->trip_aba[h] = caba;
            L->trip_aab[h] = caab;
            L->trip_aaa[h] = caaa;
        }
    }
    const int *g = L->gems_ab, *ga = L->gems_aa, *n = opt->amopi;
    int singlet = opt->nalpha == opt->nbeta;
    /* x offsets in dimensions_ order, v2rdm_solver.cc:698-844 */
    long o = 0;
    int nb = 0;
    int *bdim = (int *)malloc(sizeof(int) * 8 * ORC_OFF_COUNT);
#define BLOCKS(w, dimexpr) for (int h = 0; h < L->nirrep; h++) { long d_ = (dimexpr); \
        L->off[w][h] = o; o += d_ * d_; bdim[nb++] = (int)d_; }
    BLOCKS(ORC_OFF_D2AB, g[h])
    BLOCKS(ORC_OFF_D2AA, ga[h])
    BLOCKS(ORC_OFF_D2BB, ga[h])
    if (opt->constrain_spin) {
        if (singlet) { BLOCKS(ORC_OFF_D200, g[h]) }
        else { BLOCKS(ORC_OFF_D200, 2 * g[h]) }
    }
    BLOCKS(ORC_OFF_D1A, n[h])
    BLOCKS(ORC_OFF_D1B, n[h])
    BLOCKS(ORC_OFF_Q1A, n[h])
    BLOCKS(ORC_OFF_Q1B, n[h])
    if (opt->constrain_q2 && !opt->spin_adapt_q2) {
        BLOCKS(ORC_OFF_Q2AB, g[h])
        BLOCKS(ORC_OFF_Q2AA, ga[h])
        BLOCKS(ORC_OFF_Q2BB, ga[h])
    } else if (opt->constrain_q2) {      /* v2rdm_solver.cc:597-609 */
        BLOCKS(ORC_OFF_Q2S, L->gems_00[h])
        BLOCKS(ORC_OFF_Q2T, ga[h])
        BLOCKS(ORC_OFF_Q2TP1, ga[h])
        BLOCKS(ORC_OFF_Q2TM1, ga[h])
    }
    if (opt->constrain_g2 && !opt->spin_adapt_g2) {
        BLOCKS(ORC_OFF_G2AB, g[h])
        BLOCKS(ORC_OFF_G2BA, g[h])
        BLOCKS(ORC_OFF_G2AA, 2 * g[h])
    } else if (opt->constrain_g2) {      /* v2rdm_solver.cc:623-635 */
        BLOCKS(ORC_OFF_G2S, g[h])
        BLOCKS(ORC_OFF_G2T, g[h])
        BLOCKS(ORC_OFF_G2TP1, g[h])
        BLOCKS(ORC_OFF_G2TM1, g[h])
    }
    if (opt->constrain_t1) {             /* v2rdm_solver.cc:809-825 */
        BLOCKS(ORC_OFF_T1AAA, L->trip_aaa[h])
        BLOCKS(ORC_OFF_T1BBB, L->trip_aaa[h])
        BLOCKS(ORC_OFF_T1AAB, L->trip_aab[h])
        BLOCKS(ORC_OFF_T1BBA, L->trip_aab[h])
    }
    if (opt->constrain_t2) {
        BLOCKS(ORC_OFF_T2AAA, L->trip_aab[h] + L->trip_aba[h])
        BLOCKS(ORC_OFF_T2BBB, L->trip_aab[h] + L->trip_aba[h])
        BLOCKS(ORC_OFF_T2AAB, L->trip_aab[h])
        BLOCKS(ORC_OFF_T2BBA, L->trip_aab[h])
    }
#undef BLOCKS
    L->dimx = o;
    L->nblocks = nb;
    L->bdim = bdim;
    L->boff = (long *)malloc(sizeof(long) * nb);
    long bo = 0;
    for (int i = 0; i < nb; i++) { L->boff[i] = bo; bo += (long)bdim[i] * bdim[i]; }

    /* constraint rows, in bpsdp_Au order (v2rdm_solver.cc:887-1127) */
    long r = 0;
    long sa = 0, sab = 0, saa = 0;
    for (int h = 0; h < L->nirrep; h++) {
        sa += (long)n[h] * n[h];
        sab += (long)g[h] * g[h];
        saa += (long)ga[h] * ga[h];
    }
    add_family(L, "d2_trace", &r, opt->constrain_sz ? 3 : 1);
    add_family(L, "d1q1", &r, 2 * sa);
    add_family(L, "d2_contract", &r, (opt->constrain_sz ? 4 : 2) * sa);
    if (opt->constrain_spin) {
        add_family(L, "spin_s2", &r, 1);
        if (singlet) {
            add_family(L, "spin_d1", &r, sa);
            add_family(L, "spin_d2aa_bb", &r, saa);
            add_family(L, "spin_d2aa_ab", &r, saa);
            add_family(L, "spin_d2bb_ab", &r, saa);
            add_family(L, "spin_d200", &r, sab);
        } else {
            add_family(L, "spin_d200_4", &r, 4 * sab);
        }
        if (opt->constrain_g2) add_family(L, "spin_maxspin", &r, 2L * g[0]);
    }
    if (opt->constrain_q2 && !opt->spin_adapt_q2) {
        add_family(L, "q2ab", &r, sab);
        add_family(L, "q2aa", &r, saa);
        add_family(L, "q2bb", &r, saa);
    } else if (opt->constrain_q2) {      /* v2rdm_solver.cc:976-989 */
        long s00 = 0;
        for (int h = 0; h < L->nirrep; h++) s00 += (long)L->gems_00[h] * L->gems_00[h];
        add_family(L, "q2s", &r, s00);
        add_family(L, "q2t", &r, saa);
        add_family(L, "q2t_p1", &r, saa);
        add_family(L, "q2t_m1", &r, saa);
    }
    if (opt->constrain_g2 && !opt->spin_adapt_g2) {
        add_family(L, "g2ab", &r, sab);
        add_family(L, "g2ba", &r, sab);
        add_family(L, "g2aa", &r, 4 * sab);
    } else if (opt->constrain_g2) {      /* v2rdm_solver.cc:1003-1015 */
        add_family(L, "g2s", &r, sab);
        add_family(L, "g2t", &r, sab);
        add_family(L, "g2t_p1", &r, sab);
        add_family(L, "g2t_m1", &r, sab);
    }
    if (opt->constrain_t1) {             /* t1.cc:433-809 row order: aab, bba, aaa, bbb */
        long s1 = 0, s3 = 0;
        for (int h = 0; h < L->nirrep; h++) {
            s1 += (long)L->trip_aab[h] * L->trip_aab[h];
            s3 += (long)L->trip_aaa[h] * L->trip_aaa[h];
        }
        add_family(L, "t1aab", &r, s1);
        add_family(L, "t1bba", &r, s1);
        add_family(L, "t1aaa", &r, s3);
        add_family(L, "t1bbb", &r, s3);
    }
    if (opt->constrain_t2) {
        long s1 = 0, s2 = 0;
        for (int h = 0; h < L->nirrep; h++) {
            long t = L->trip_aab[h], T = (long)L->trip_aab[h] + L->trip_aba[h];
            s1 += t * t; s2 += T * T;
        }
        add_family(L, "t2aab", &r, s1);
        add_family(L, "t2bba", &r, s1);
        add_family(L, "t2aaa", &r, s2);
        add_family(L, "t2bbb", &r, s2);
    }
    L->ncon = r;
    return L;
}

void orc_layout_destroy(orc_layout *L)
{
    if (!L) return;
    for (int h = 0; h < L->nirrep; h++) {
        free(L->bab[h]); free(L->baa[h]);
        free(L->baab[h]); free(L->baba[h]); free(L->baaa[h]);
        free(L->b00[h]);
    }
    free(L->gab); free(L->gaa); free(L->bdim); free(L->boff);
    free(L);
}

long orc_dimx(const orc_layout *L) { return L->dimx; }
long orc_ncon(const orc_layout *L) { return L->ncon; }
int orc_nblocks(const orc_layout *L) { return L->nblocks; }
void orc_block_dims(const orc_layout *L, int *dims) { memcpy(dims, L->bdim, sizeof(int) * L->nblocks); }
long orc_offset(const orc_layout *L, int which, int h) { return L->off[which][h]; }
int orc_gems_ab(const orc_layout *L, int h) { return L->gems_ab[h]; }
int orc_gems_aa(const orc_layout *L, int h) { return L->gems_aa[h]; }
int orc_trip_aab(const orc_layout *L, int h) { return L->trip_aab[h]; }
int orc_gems_00(const orc_layout *L, int h) { return L->gems_00[h]; }

/* ------------------------------------------------------------------------ */
/* PackSpatialDensity, v2rdm_solver.cc:3664-3773; INDEX is psi4's packed     */
/* lower-triangle index (libciomr), max (max + 1) / 2 + min                   */
/* ------------------------------------------------------------------------ */

static long packed(long a, long b) { return a > b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a; }

long orc_spatial_d1_dim(const orc_layout *L)
{
    long d = 0;
    for (int h = 0; h < L->nirrep; h++) d += (long)L->o.amopi[h] * (L->o.amopi[h] + 1) / 2;
    return d;
}

long orc_spatial_d2_dim(const orc_layout *L)
{
    long d = 0;
    for (int h = 0; h < L->nirrep; h++) d += (long)L->gems_00[h] * (L->gems_00[h] + 1) / 2;
    return d;
}

/* index of the singlet geminal (p, q) of irrep sym p ^ sym q */
static int sgem(const orc_layout *L, int p, int q)
{
    int h = L->sym[p] ^ L->sym[q];
    for (int n = 0; n < L->gems_00[h]; n++) {
        int a = L->b00[h][2 * n], b = L->b00[h][2 * n + 1];
        if ((a == p && b == q) || (a == q && b == p)) return n;
    }
    return -1;
}

void orc_pack_spatial_density(const orc_layout *L, const double *x, double *d1, double *d2)
{
    const int amo = L->amo;
    long base[9] = {0};
    for (int h = 0; h < L->nirrep; h++)
        base[h + 1] = base[h] + (long)L->gems_00[h] * (L->gems_00[h] + 1) / 2;
    memset(d2, 0, sizeof(double) * base[L->nirrep]);
    int *s00 = (int *)malloc(sizeof(int) * amo * amo);
    for (int p = 0; p < amo; p++)
        for (int q = 0; q < amo; q++) s00[p * amo + q] = sgem(L, p, q);
#define AB(h, r, c) x[L->off[ORC_OFF_D2AB][h] + (long)(r) * L->gems_ab[h] + (c)]
#define AA(w, h, r, c) x[L->off[w][h] + (long)(r) * L->gems_aa[h] + (c)]
    for (int h = 0; h < L->nirrep; h++) {
        for (int ij = 0; ij < L->gems_ab[h]; ij++) {
            int i = L->bab[h][2 * ij], j = L->bab[h][2 * ij + 1];
            int ji = L->gab[j * amo + i];
            for (int kl = 0; kl < L->gems_ab[h]; kl++) {
                int k = L->bab[h][2 * kl], l = L->bab[h][2 * kl + 1];
                int lk = L->gab[l * amo + k];
                int hik = L->sym[i] ^ L->sym[k], hkj = L->sym[k] ^ L->sym[j];
                int ik = s00[i * amo + k], jl = s00[j * amo + l];
                int kj = L->gab[k * amo + j], il = L->gab[i * amo + l];
                int jk = L->gab[j * amo + k], li = L->gab[l * amo + i];
                double dlj = (l == j) ? 0.0 : 1.0, dik = (i == k) ? 0.0 : 1.0;
                double val = 0.0;
                /* alpha-beta and beta-alpha, d2ab (ij,kl) and its i<->k, j<->l images */
                val += 0.5 * AB(h, ij, kl);
                val += 0.5 * AB(hkj, kj, il) * dlj;
                val += 0.5 * AB(hkj, il, kj) * dik;
                val += 0.5 * AB(h, kl, ij) * dlj * dik;
                val += 0.5 * AB(h, ji, lk);
                val += 0.5 * AB(hkj, jk, li) * dlj;
                val += 0.5 * AB(hkj, li, jk) * dik;
                val += 0.5 * AB(h, lk, ji) * dlj * dik;
                if (i != j && k != l) {             /* alpha-alpha, beta-beta */
                    double sg = (double)((i < j ? 1 : -1) * (k < l ? 1 : -1));
                    int ija = L->gaa[i * amo + j], kla = L->gaa[k * amo + l];
                    val += 0.5 * sg * AA(ORC_OFF_D2AA, h, ija, kla);
                    val += 0.5 * sg * AA(ORC_OFF_D2AA, h, kla, ija) * dlj * dik;
                    val += 0.5 * sg * AA(ORC_OFF_D2BB, h, ija, kla);
                    val += 0.5 * sg * AA(ORC_OFF_D2BB, h, kla, ija) * dlj * dik;
                }
                if (k != j && i != l) {
                    double sg = (double)((k < j ? 1 : -1) * (i < l ? 1 : -1));
                    int kja = L->gaa[k * amo + j], ila = L->gaa[i * amo + l];
                    val += 0.5 * sg * AA(ORC_OFF_D2AA, hkj, kja, ila) * dlj;
                    val += 0.5 * sg * AA(ORC_OFF_D2AA, hkj, ila, kja) * dik;
                    val += 0.5 * sg * AA(ORC_OFF_D2BB, hkj, kja, ila) * dlj;
                    val += 0.5 * sg * AA(ORC_OFF_D2BB, hkj, ila, kja) * dik;
                }
                if (ik != jl) val *= 2.0;          /* off-diagonal elements count twice */
                d2[base[hik] + packed(ik, jl)] = val;
            }
        }
    }
#undef AB
#undef AA
    free(s00);
    long o = 0;
    for (int h = 0; h < L->nirrep; h++) {
        int n = L->o.amopi[h];
        for (int i = 0; i < n; i++)
            for (int j = i; j < n; j++)
                d1[o + packed(i, j)] = x[L->off[ORC_OFF_D1A][h] + (long)i * n + j] +
                                       x[L->off[ORC_OFF_D1B][h] + (long)i * n + j];
        o += (long)n * (n + 1) / 2;
    }
}
void orc_bas_00(const orc_layout *L, int h, int *pairs)
{
    memcpy(pairs, L->b00[h], sizeof(int) * 2 * L->gems_00[h]);
}
int orc_trip_aba(const orc_layout *L, int h) { return L->trip_aba[h]; }
int orc_trip_aaa(const orc_layout *L, int h) { return L->trip_aaa[h]; }
int orc_ibas_ab(const orc_layout *L, int p, int q) { return L->gab[p * L->amo + q]; }
int orc_ibas_aa(const orc_layout *L, int p, int q) { return p == q ? -1 : L->gaa[p * L->amo + q]; }
int orc_orbital_irrep(const orc_layout *L, int p) { return L->sym[p]; }

int orc_row_families(const orc_layout *L, int max, char names[][24], long *start, long *count)
{
    int n = L->nfam < max ? L->nfam : max;
    for (int f = 0; f < n; f++) {
        memcpy(names[f], L->famname[f], 24);
        start[f] = L->famstart[f];
        count[f] = L->famcount[f];
    }
    return L->nfam;
}

/* ------------------------------------------------------------------------ */
/* maps: one body, two expansions                                           */
/* ------------------------------------------------------------------------ */

#define ROW_BEGIN(r) { long r_ = (r); double acc_ = 0.0;
#define TERM(col, coef) acc_ += (coef) * u[(col)];
#define ROW_END A[r_] = acc_; }
#define PARFOR _Pragma("omp parallel for schedule(static)")
#define MAPS_FN orc_Au_impl
#include "maps_body.inc"
#undef ROW_BEGIN
#undef TERM
#undef ROW_END
#undef PARFOR
#undef MAPS_FN

#define ROW_BEGIN(r) { long r_ = (r); double y_ = u[r_];
#define TERM(col, coef) A[(col)] += (coef) * y_;
#define ROW_END }
#define PARFOR
#define MAPS_FN orc_ATu_impl
#include "maps_body.inc"
#undef ROW_BEGIN
#undef TERM
#undef ROW_END
#undef PARFOR
#undef MAPS_FN

void orc_Au(const orc_layout *L, const double *x, double *Ax)
{
    memset(Ax, 0, sizeof(double) * L->ncon);          /* v2rdm_solver.cc:3023 */
    orc_Au_impl(L, x, Ax);
}

void orc_ATu(const orc_layout *L, const double *y, double *ATy)
{
    memset(ATy, 0, sizeof(double) * L->dimx);         /* v2rdm_solver.cc:3117 */
    orc_ATu_impl(L, y, ATy);
}

void orc_cg_Ax(const orc_layout *L, const double *u, double *out, double *scratch)
{
    orc_ATu(L, u, scratch);                           /* v2rdm_solver.cc:3207-3213 */
    orc_Au(L, scratch, out);
}

/* ------------------------------------------------------------------------ */
/* BLAS-1 helpers                                                           */
/* ------------------------------------------------------------------------ */

static double ddot(long n, const double *a, const double *b)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}
static void daxpy(long n, double a, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) y[i] += a * x[i];
}

/* CGSolver::solve, cg_solver.cc:125-177 */
int orc_cg_solve(const orc_layout *L, const double *rhs, double *y, double tol,
                 int maxiter, double *Ap_out)
{
    long n = L->ncon;
    double *p = (double *)malloc(sizeof(double) * n);
    double *r = (double *)malloc(sizeof(double) * n);
    double *Ap = Ap_out ? Ap_out : (double *)malloc(sizeof(double) * n);
    double *scr = (double *)malloc(sizeof(double) * L->dimx);
    orc_cg_Ax(L, y, Ap, scr);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) { r[i] = rhs[i] - Ap[i]; p[i] = r[i]; }
    int iter = 0;
    do {
        orc_cg_Ax(L, p, Ap, scr);
        double rr = ddot(n, r, r);
        double pap = ddot(n, p, Ap);
        double alpha = rr / pap;
        daxpy(n, alpha, p, y);
        daxpy(n, -alpha, Ap, r);
        double rrnew = ddot(n, r, r);
        double nrm = sqrt(rrnew);
        double beta = rrnew / rr;
        if (nrm < tol) break;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) p[i] = beta * p[i] + r[i];
        iter++;
    } while (iter < maxiter);
    free(p); free(r); free(scr);
    if (!Ap_out) free(Ap);
    return iter;
}

/* ------------------------------------------------------------------------ */
/* LAPACK (psi4 Matrix::diagonalize -> dsyev) through MKL, else Jacobi      */
/* ------------------------------------------------------------------------ */

typedef void (*dsyev_t)(const char *, const char *, const int *, double *, const int *,
                        double *, double *, const int *, int *);
typedef void (*dgemm_t)(const char *, const char *, const int *, const int *, const int *,
                        const double *, const double *, const int *, const double *,
                        const int *, const double *, double *, const int *);
static dsyev_t p_dsyev = NULL;
static dgemm_t p_dgemm = NULL;
static int lapack_tried = 0;

static void lapack_init(void)
{
    if (lapack_tried) return;
    lapack_tried = 1;
    if (getenv("ORC_NO_MKL")) return;
    const char *names[] = {"libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so.1", "libmkl_rt.so",
                           "/opt/conda/lib/libmkl_rt.so"};
    void *hnd = NULL;
    for (int i = 0; i < 4 && !hnd; i++) hnd = dlopen(names[i], RTLD_NOW | RTLD_GLOBAL);
    if (!hnd) return;
    /* use the GNU OpenMP runtime the oracle itself runs on */
    int (*setlayer)(int) = (int (*)(int))dlsym(hnd, "MKL_Set_Threading_Layer");
    if (setlayer) setlayer(1); /* MKL_THREADING_GNU */
    p_dsyev = (dsyev_t)dlsym(hnd, "dsyev_");
    p_dgemm = (dgemm_t)dlsym(hnd, "dgemm_");
    if (!p_dsyev || !p_dgemm) { p_dsyev = NULL; p_dgemm = NULL; }
}

int orc_lapack_backend(void)
{
    lapack_init();
    return p_dsyev ? 1 : 0;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* cyclic Jacobi (fallback when no LAPACK is loadable); a: n x n row-major,
 * on exit w ascending, v row-major with eigenvector j in column j */
static void jacobi_eig(int n, double *a, double *w, double *v)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) v[(long)i * n + j] = (i == j);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double e = a[(long)i * n + j] * a[(long)i * n + j];
                tot += e;
                if (i != j) off += e;
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = a[(long)p * n + q];
                if (fabs(apq) < 1e-300) continue;
                double app = a[(long)p * n + p], aqq = a[(long)q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = a[(long)k * n + p], akq = a[(long)k * n + q];
                    a[(long)k * n + p] = c * akp - s * akq;
                    a[(long)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = a[(long)p * n + k], aqk = a[(long)q * n + k];
                    a[(long)p * n + k] = c * apk - s * aqk;
                    a[(long)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = v[(long)k * n + p], vkq = v[(long)k * n + q];
                    v[(long)k * n + p] = c * vkp - s * vkq;
                    v[(long)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++) w[i] = a[(long)i * n + i];
    /* sort ascending (selection sort on columns) */
    for (int i = 0; i < n - 1; i++) {
        int m = i;
        for (int j = i + 1; j < n; j++) if (w[j] < w[m]) m = j;
        if (m != i) {
            double t = w[i]; w[i] = w[m]; w[m] = t;
            for (int k = 0; k < n; k++) {
                double tv = v[(long)k * n + i];
                v[(long)k * n + i] = v[(long)k * n + m];
                v[(long)k * n + m] = tv;
            }
        }
    }
}

/* one block of Update_xz, v2rdm_solver.cc:3237-3286.  M is the (already
 * symmetrized) n x n block; xblk = V+ diag(l+/mu) V+^T, zblk = -V- diag(l-) V-^T */
void orc_psd_split(int n, const double *M, double mu, double *xblk, double *zblk)
{
    lapack_init();
    long nn = (long)n * n;
    double *a = (double *)malloc(sizeof(double) * nn);
    double *w = (double *)malloc(sizeof(double) * n);
    double *vp = (double *)malloc(sizeof(double) * nn);   /* scaled vectors */
    double *vq = (double *)malloc(sizeof(double) * nn);   /* plain vectors  */
    memcpy(a, M, sizeof(double) * nn);
    /* after the call, eigenvector j is the contiguous run a[j*n .. j*n+n) */
    if (p_dsyev) {
        int lwork = -1, info = 0;
        double wq;
        p_dsyev("V", "U", &n, a, &n, w, &wq, &lwork, &info);
        lwork = (int)wq;
        double *work = (double *)malloc(sizeof(double) * (lwork > 1 ? lwork : 1));
        p_dsyev("V", "U", &n, a, &n, w, work, &lwork, &info);
        free(work);
    } else {
        double *v = (double *)malloc(sizeof(double) * nn);
        jacobi_eig(n, a, w, v);
        for (int j = 0; j < n; j++)
            for (int q = 0; q < n; q++) a[(long)j * n + q] = v[(long)q * n + j];
        free(v);
    }
    for (int part = 0; part < 2; part++) {
        double *out = part ? zblk : xblk;
        int k = 0;
        for (int j = 0; j < n; j++) {
            int take = part ? (w[j] < 0.0) : (w[j] > 0.0);
            if (!take) continue;
            double sc = part ? -w[j] : w[j] / mu;
            for (int q = 0; q < n; q++) {
                vp[(long)k * n + q] = a[(long)j * n + q] * sc;
                vq[(long)k * n + q] = a[(long)j * n + q];
            }
            k++;
        }
        if (k == 0) { memset(out, 0, sizeof(double) * nn); continue; }
        if (p_dgemm) {
            /* column-major view: VP, VQ are n x k; out = VP * VQ^T */
            double one = 1.0, zero = 0.0;
            p_dgemm("N", "T", &n, &n, &k, &one, vp, &n, vq, &n, &zero, out, &n);
        } else {
            for (int p = 0; p < n; p++)
                for (int q = 0; q < n; q++) {
                    double s = 0.0;
                    for (int m = 0; m < k; m++) s += vp[(long)m * n + q] * vq[(long)m * n + p];
                    out[(long)p * n + q] = s;
                }
        }
    }
    free(a); free(w); free(vp); free(vq);
}

/* Update_xz, v2rdm_solver.cc:3216-3289 */
void orc_update_xz(const orc_layout *L, double mu, const double *y, const double *c,
                   double *x, double *z)
{
    long dimx = L->dimx;
    double *aty = (double *)malloc(sizeof(double) * dimx);
    orc_ATu(L, y, aty);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < dimx; i++) aty[i] = aty[i] - c[i] + mu * x[i];
    for (int b = 0; b < L->nblocks; b++) {
        int n = L->bdim[b];
        if (n == 0) continue;
        long o = L->boff[b], nn = (long)n * n;
        double *m = (double *)malloc(sizeof(double) * nn);
        for (int p = 0; p < n; p++)
            for (int q = p; q < n; q++) {
                double d = 0.5 * (aty[o + (long)p * n + q] + aty[o + (long)q * n + p]);
                m[(long)p * n + q] = m[(long)q * n + p] = d;
            }
        orc_psd_split(n, m, mu, x + o, z + o);
        free(m);
    }
    free(aty);
}

/* ------------------------------------------------------------------------ */
/* path inputs                                                              */
/* ------------------------------------------------------------------------ */

/* BuildConstraints, v2rdm_solver.cc:2518-3017 */
void orc_build_b(const orc_layout *L, double *b)
{
    memset(b, 0, sizeof(double) * L->ncon);
    double na = L->na, nb = L->nb, n = na + nb;
    long off = 0;
    if (L->o.constrain_sz) {
        b[off++] = na * nb;
        b[off++] = na * (na - 1.0);
        b[off++] = nb * (nb - 1.0);
    } else {
        b[off++] = n * (n - 1.0);
    }
    for (int s = 0; s < 2; s++)
        for (int h = 0; h < L->nirrep; h++) {
            int m = L->o.amopi[h];
            for (int i = 0; i < m; i++)
                for (int j = 0; j < m; j++) b[off + i * m + j] = (double)(i == j);
            off += (long)m * m;
        }
    long sa = 0;
    for (int h = 0; h < L->nirrep; h++) sa += (long)L->o.amopi[h] * L->o.amopi[h];
    off += (L->o.constrain_sz ? 4 : 2) * sa;
    if (L->o.constrain_spin) {
        double ms = (L->o.multiplicity - 1.0) / 2.0;
        b[off++] = 0.5 * (na + nb) + ms * ms - ms * (ms + 1.0);
        if (L->o.nalpha == L->o.nbeta) {
            for (int h = 0; h < L->nirrep; h++) {
                off += (long)L->o.amopi[h] * L->o.amopi[h];
                off += 3L * L->gems_aa[h] * L->gems_aa[h];
                off += (long)L->gems_ab[h] * L->gems_ab[h];
            }
        } else {
            for (int h = 0; h < L->nirrep; h++) off += 4L * L->gems_ab[h] * L->gems_ab[h];
        }
        if (L->o.constrain_g2) off += 2L * L->gems_ab[0];
    }
    /* the spin-adapted Q2 rows carry no constant: b = 0 there
     * (v2rdm_solver.cc:2709-2745) */
    if (L->o.constrain_q2 && !L->o.spin_adapt_q2) {
        for (int h = 0; h < L->nirrep; h++) {
            int g = L->gems_ab[h];
            for (int ij = 0; ij < g; ij++) {
                int i = L->bab[h][2 * ij], j = L->bab[h][2 * ij + 1];
                for (int kl = 0; kl < g; kl++) {
                    int k = L->bab[h][2 * kl], l = L->bab[h][2 * kl + 1];
                    b[off + (long)ij * g + kl] = -(double)((i == k) * (j == l));
                }
            }
            off += (long)g * g;
        }
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < L->nirrep; h++) {
                int g = L->gems_aa[h];
                for (int ij = 0; ij < g; ij++) {
                    int i = L->baa[h][2 * ij], j = L->baa[h][2 * ij + 1];
                    for (int kl = 0; kl < g; kl++) {
                        int k = L->baa[h][2 * kl], l = L->baa[h][2 * kl + 1];
                        b[off + (long)ij * g + kl] = -(double)((i == k) * (j == l)) + (double)((i == l) * (j == k));
                    }
                }
                off += (long)g * g;
            }
    }
    /* G2, T1 and T2 rows have b = 0 (v2rdm_solver.cc:2821-2860) */
}

/* RepackIntegrals + FrozenCoreEnergy (no core), tei.cc:144-228, 229-313:
 * D1 blocks get h, D2ab (ik|jl), D2aa/bb (ik|jl)-(il|jk); zero elsewhere */
void orc_build_c(const orc_layout *L, const double *h1, const double *eri, double *c)
{
    memset(c, 0, sizeof(double) * L->dimx);
    long a = L->amo;
#define ERI(p, q, r, s) eri[(((long)(p) * a + (q)) * a + (r)) * a + (s)]
    for (int h = 0; h < L->nirrep; h++) {
        int g = L->gems_ab[h];
        for (int ij = 0; ij < g; ij++) {
            int i = L->bab[h][2 * ij], j = L->bab[h][2 * ij + 1];
            for (int kl = 0; kl < g; kl++) {
                int k = L->bab[h][2 * kl], l = L->bab[h][2 * kl + 1];
                c[L->off[ORC_OFF_D2AB][h] + (long)ij * g + kl] = ERI(i, k, j, l);
            }
        }
    }
    for (int h = 0; h < L->nirrep; h++) {
        int g = L->gems_aa[h];
        for (int ij = 0; ij < g; ij++) {
            int i = L->baa[h][2 * ij], j = L->baa[h][2 * ij + 1];
            for (int kl = 0; kl < g; kl++) {
                int k = L->baa[h][2 * kl], l = L->baa[h][2 * kl + 1];
                double v = ERI(i, k, j, l) - ERI(i, l, j, k);
                c[L->off[ORC_OFF_D2AA][h] + (long)ij * g + kl] = v;
                c[L->off[ORC_OFF_D2BB][h] + (long)ij * g + kl] = v;
            }
        }
    }
#undef ERI
    for (int h = 0; h < L->nirrep; h++) {
        int m = L->o.amopi[h];
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) {
                double v = h1[(long)(i + L->poff[h]) * a + (j + L->poff[h])];
                c[L->off[ORC_OFF_D1A][h] + i * m + j] = v;
                c[L->off[ORC_OFF_D1B][h] + i * m + j] = v;
            }
    }
}

/* RANDOM guess, v2rdm_solver.cc:2479-2490 */
void orc_guess_random(const orc_layout *L, double *x, double *z, double *y)
{
    srand(0);
    for (long i = 0; i < L->dimx; i++) {
        x[i] = ((double)rand() / RAND_MAX - 1.0) * 2.0;
        z[i] = ((double)rand() / RAND_MAX - 1.0) * 2.0;
    }
    for (long i = 0; i < L->ncon; i++) y[i] = ((double)rand() / RAND_MAX - 1.0) * 2.0;
}

/* ------------------------------------------------------------------------ */
/* BPSDP outer loop, v2rdm_solver.cc:1685-1912 (OPTIMIZE_ORBITALS false)    */
/* ------------------------------------------------------------------------ */
int orc_bpsdp(const orc_layout *L, const orc_params *p, const double *c, const double *b,
              double *x, double *z, double *y, double max_seconds, orc_stats *st)
{
    long dimx = L->dimx, ncon = L->ncon;
    double *Ax = (double *)malloc(sizeof(double) * ncon);
    double *B = (double *)malloc(sizeof(double) * ncon);
    double *ATy = (double *)malloc(sizeof(double) * dimx);
    double tau = p->tau, mu = p->mu0 > 0.0 ? p->mu0 : 1.0;
    double ep = p->ep0, ed = p->ed0, egap = 0.0, eprimal = 0.0, edual = 0.0;
    int oiter = p->oiter0 > 0 ? p->oiter0 : 0, converged = 0;
    const int oiter_first = oiter;
    long cgtot = 0;
    double tcg = 0.0, txz = 0.0, t0 = wall();
    do {
        double t1 = wall();
        orc_Au(L, x, Ax);                                     /* :1784-1786 */
#pragma omp parallel for schedule(static)
        for (long i = 0; i < ncon; i++) Ax[i] = -tau * mu * (Ax[i] - b[i]);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < dimx; i++) z[i] = c[i] - z[i];   /* :1789-1790 */
        orc_Au(L, z, B);                                      /* :1791 */
#pragma omp parallel for schedule(static)
        for (long i = 0; i < ncon; i++) B[i] += Ax[i];        /* :1794 */
        double cgtol;                                         /* :1797-1805 */
        if (oiter == 0 || (oiter == oiter_first && !(ep > 0.0 && ed > 0.0))) cgtol = 0.01;
        else cgtol = (ep > ed) ? 0.01 * ed : 0.01 * ep;
        if (cgtol < p->cg_convergence) cgtol = p->cg_convergence;
        int iiter = orc_cg_solve(L, B, y, cgtol, p->cg_maxiter, Ax);  /* :1808 */
        cgtot += iiter;
        double t2 = wall();
        tcg += t2 - t1;
        orc_update_xz(L, mu, y, c, x, z);                     /* :1819 */
        txz += wall() - t2;
        orc_ATu(L, y, ATy);                                   /* :1829-1832 */
#pragma omp parallel for schedule(static)
        for (long i = 0; i < dimx; i++) ATy[i] = ATy[i] + z[i] - c[i];
        ed = sqrt(ddot(dimx, ATy, ATy));
        orc_Au(L, x, Ax);                                     /* :1835-1837 */
#pragma omp parallel for schedule(static)
        for (long i = 0; i < ncon; i++) Ax[i] -= b[i];
        ep = sqrt(ddot(ncon, Ax, Ax));
        if (oiter % p->mu_update_frequency == 0 && oiter > 0) mu = mu * ep / ed;  /* :1840-1848 */
        double ecur = ddot(dimx, c, x);                       /* :1851-1852 */
        edual = ddot(ncon, b, y);
        if (p->verbose)
            fprintf(stderr, "      %5i %5i %14.8f %14.8f %11.4e %7.3f %10.5e %10.5e\n", oiter,
                    iiter, ecur, edual, fabs(ecur - edual), mu, ep, ed);
        oiter++;
        egap = fabs(ecur - edual);
        eprimal = ecur;
        if (oiter >= p->maxiter) break;
        if (max_seconds > 0 && wall() - t0 > max_seconds) break;
    } while (ep > p->r_convergence || ed > p->r_convergence || egap > p->e_convergence);
    converged = !(ep > p->r_convergence || ed > p->r_convergence || egap > p->e_convergence);
    if (st) {
        st->oiter = oiter; st->cg_iters = cgtot; st->converged = converged;
        st->energy_primal = eprimal; st->energy_dual = edual; st->ep = ep; st->ed = ed;
        st->mu = mu; st->t_cg = tcg; st->t_xz = txz; st->t_total = wall() - t0;
    }
    free(Ax); free(B); free(ATy);
    return converged ? 0 : 1;
}

/* ------------------------------------------------------------------------ */
/* MLP preconditioner reference (no reference counterpart; checks the      */
/* product's forward / backward kernels).  s = softplus-free form:          */
/*   h1 = tanh(W1^T f + b1), h2 = tanh(W2^T h1 + b2), out = w3.h2 + b3      */
/* ------------------------------------------------------------------------ */
void orc_mlp_forward(int n, int nin, int nh, const float *feat, const float *W1,
                     const float *b1, const float *W2, const float *b2, const float *w3,
                     float b3, float *out)
{
    double h1[256], h2[256];
    for (int e = 0; e < n; e++) {
        const float *f = feat + (long)e * nin;
        for (int j = 0; j < nh; j++) {
            double s = b1[j];
            for (int i = 0; i < nin; i++) s += (double)W1[i * nh + j] * f[i];
            h1[j] = tanh(s);
        }
        for (int j = 0; j < nh; j++) {
            double s = b2[j];
            for (int i = 0; i < nh; i++) s += (double)W2[i * nh + j] * h1[i];
            h2[j] = tanh(s);
        }
        double o = b3;
        for (int j = 0; j < nh; j++) o += (double)w3[j] * h2[j];
        out[e] = (float)o;
    }
}

double orc_mlp_loss_grad(int n, int nin, int nh, const float *feat, const float *target,
                         const float *wt, const float *W1, const float *b1, const float *W2,
                         const float *b2, const float *w3, float b3, double *gW1, double *gb1,
                         double *gW2, double *gb2, double *gw3, double *gb3)
{
    double h1[256], h2[256], d2[256], d1[256];
    memset(gW1, 0, sizeof(double) * nin * nh);
    memset(gb1, 0, sizeof(double) * nh);
    memset(gW2, 0, sizeof(double) * nh * nh);
    memset(gb2, 0, sizeof(double) * nh);
    memset(gw3, 0, sizeof(double) * nh);
    *gb3 = 0.0;
    double loss = 0.0;
    for (int e = 0; e < n; e++) {
        const float *f = feat + (long)e * nin;
        for (int j = 0; j < nh; j++) {
            double s = b1[j];
            for (int i = 0; i < nin; i++) s += (double)W1[i * nh + j] * f[i];
            h1[j] = tanh(s);
        }
        for (int j = 0; j < nh; j++) {
            double s = b2[j];
            for (int i = 0; i < nh; i++) s += (double)W2[i * nh + j] * h1[i];
            h2[j] = tanh(s);
        }
        double o = b3;
        for (int j = 0; j < nh; j++) o += (double)w3[j] * h2[j];
        double go;
        if (target) {
            double r = o - target[e];
            double w = wt ? wt[e] : 1.0;
            loss += w * r * r;
            go = 2.0 * w * r / n;
        } else {
            go = wt[e];   /* upstream gradient d(loss)/d(out) given per sample */
        }
        *gb3 += go;
        for (int j = 0; j < nh; j++) {
            gw3[j] += go * h2[j];
            d2[j] = go * w3[j] * (1.0 - h2[j] * h2[j]);
            gb2[j] += d2[j];
        }
        for (int i = 0; i < nh; i++) {
            double s = 0.0;
            for (int j = 0; j < nh; j++) {
                gW2[i * nh + j] += h1[i] * d2[j];
                s += W2[i * nh + j] * d2[j];
            }
            d1[i] = s * (1.0 - h1[i] * h1[i]);
            gb1[i] += d1[i];
        }
        for (int i = 0; i < nin; i++)
            for (int j = 0; j < nh; j++) gW1[i * nh + j] += f[i] * d1[j];
    }
    return loss / n;
}
