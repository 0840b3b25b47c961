/*
 * v2rdm_oracle.h -- CPU restatement of the v2RDM-CASSCF boundary-point SDP
 * iteration (test infrastructure only).
 *
 * This library is the parity oracle and the timed CPU baseline.  It is only
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg.  Nothing in the product (v2rdm_casscf_amd/) links or calls it.
 *
 * Every function restates a reference routine (cited file:line, paths relative
 * to the edeprince3/v2rdm_casscf source tree):
 *   orc_layout_create   basis.cc:57-548, 655-757 (geminal / triplet tables),
 *                       v2rdm_solver.cc:488-1270 (flags, dimx_, offsets,
 *                       nconstraints_, dimensions_)
 *   orc_Au / orc_ATu    v2rdm_solver.cc:3020-3065 / 3114-3158 dispatch over
 *                       d2.cc:296-547 / 58-293, spin.cc:274-493 / 58-271,
 *                       q2.cc:801-950 / 953-1074, g2.cc:1061-1260 / 1263-1457,
 *                       spin-adapted q2.cc:260-467 / 470-670,
 *                       g2.cc:235-546 / 549-856,
 *                       t1.cc:433-809 / 812-1173,
 *                       t2.cc:1251-1861 / 3399-3993 (the *_slow maps)
 *   orc_cg_Ax           v2rdm_solver.cc:3207-3213
 *   orc_cg_solve        cg_solver.cc:125-177
 *   orc_update_xz       v2rdm_solver.cc:3216-3289 (LAPACK dsyev + dgemm)
 *   orc_build_b         v2rdm_solver.cc:2518-3017
 *   orc_build_c         tei.cc:144-313 (no frozen / restricted core)
 *   orc_guess_random    v2rdm_solver.cc:2479-2490 (glibc rand after srand(0))
 *   orc_bpsdp           v2rdm_solver.cc:1685-1912 (no orbital optimization)
 *   orc_pack_spatial_density  v2rdm_solver.cc:3664-3773 (PackSpatialDensity)
 *
 * Parity status of each part is recorded in the header of v2rdm_oracle.c.
 */
#ifndef V2RDM_ORACLE_H
#define V2RDM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_options {
    int nirrep;          /* 1, 2, 4 or 8 (D2h subgroups, Cotton order)        */
    int amopi[8];        /* active orbitals per irrep                         */
    int nalpha, nbeta;   /* active electrons                                  */
    int multiplicity;    /* 2S+1                                              */
    int constrain_q2, constrain_g2, constrain_t2;
    int constrain_spin, constrain_sz;
    int constrain_t1;    /* POSITIVITY DQGT1 / DQGT1T2                        */
    int spin_adapt_q2;   /* SPIN_ADAPT_Q2 (plugin.cc:110): Q2s, Q2t, Q2t(+1),  */
                         /* Q2t(-1) blocks instead of Q2ab, Q2aa, Q2bb       */
    int spin_adapt_g2;   /* SPIN_ADAPT_G2 (plugin.cc:108): G2s, G2t, G2t(+1),  */
                         /* G2t(-1) blocks instead of G2ab, G2ba, G2aa       */
    /* active orbitals (Pitzer index) from lowest to highest orbital energy:
     * the order of the singlet geminals of SPIN_ADAPT_Q2 (basis.cc:626-651);
     * NULL = Pitzer order */
    const int *energy_order;
} orc_options;

typedef struct orc_layout orc_layout;

orc_layout *orc_layout_create(const orc_options *opt);
void orc_layout_destroy(orc_layout *L);

long orc_dimx(const orc_layout *L);
long orc_ncon(const orc_layout *L);
int  orc_nblocks(const orc_layout *L);
/* dims[i] = dimension of block i of x, in dimensions_ order */
void orc_block_dims(const orc_layout *L, int *dims);
/* named offsets into x: which = one of the ORC_OFF_* codes, h = irrep */
long orc_offset(const orc_layout *L, int which, int h);
int  orc_gems_ab(const orc_layout *L, int h);
int  orc_gems_aa(const orc_layout *L, int h);
int  orc_trip_aab(const orc_layout *L, int h);
int  orc_trip_aba(const orc_layout *L, int h);
int  orc_trip_aaa(const orc_layout *L, int h);
/* singlet geminals of SPIN_ADAPT_Q2: count, and the pairs (i, j) of irrep h
 * in their order (pairs has 2 * orc_gems_00(L, h) entries) */
int  orc_gems_00(const orc_layout *L, int h);
void orc_bas_00(const orc_layout *L, int h, int *pairs);
/* geminal tables: ab index of (p,q); aa index of (p,q) (-1 if p==q) */
int  orc_ibas_ab(const orc_layout *L, int p, int q);
int  orc_ibas_aa(const orc_layout *L, int p, int q);
int  orc_orbital_irrep(const orc_layout *L, int p);
/* row ranges of each constraint family in reference order:
 * returns number of families; names/starts/counts filled up to max */
int  orc_row_families(const orc_layout *L, int max, char names[][24],
                      long *start, long *count);

enum {
    ORC_OFF_D2AB = 0, ORC_OFF_D2AA, ORC_OFF_D2BB, ORC_OFF_D200,
    ORC_OFF_D1A, ORC_OFF_D1B, ORC_OFF_Q1A, ORC_OFF_Q1B,
    ORC_OFF_Q2AB, ORC_OFF_Q2AA, ORC_OFF_Q2BB,
    ORC_OFF_G2AB, ORC_OFF_G2BA, ORC_OFF_G2AA,
    ORC_OFF_T2AAA, ORC_OFF_T2BBB, ORC_OFF_T2AAB, ORC_OFF_T2BBA,
    ORC_OFF_T1AAA, ORC_OFF_T1BBB, ORC_OFF_T1AAB, ORC_OFF_T1BBA,
    ORC_OFF_Q2S, ORC_OFF_Q2T, ORC_OFF_Q2TP1, ORC_OFF_Q2TM1,
    ORC_OFF_G2S, ORC_OFF_G2T, ORC_OFF_G2TP1, ORC_OFF_G2TM1,
    ORC_OFF_COUNT
};

/* PackSpatialDensity (v2rdm_solver.cc:3664-3773): spin-free active-space
 * D1 (per irrep, packed lower triangle of D1a + D1b) and D2 (per irrep of the
 * singlet geminals, packed lower triangle over singlet-geminal pairs) */
long orc_spatial_d1_dim(const orc_layout *L);
long orc_spatial_d2_dim(const orc_layout *L);
void orc_pack_spatial_density(const orc_layout *L, const double *x, double *d1, double *d2);

/* constraint maps; callee overwrites the output (bpsdp_Au / bpsdp_ATu) */
void orc_Au(const orc_layout *L, const double *x, double *Ax);
void orc_ATu(const orc_layout *L, const double *y, double *ATy);
/* A (A^T u); scratch has dimx entries */
void orc_cg_Ax(const orc_layout *L, const double *u, double *out, double *scratch);

/* plain CG on A A^T, warm start from y; returns iterations (iter_) */
int orc_cg_solve(const orc_layout *L, const double *rhs, double *y,
                 double tol, int maxiter, double *Ap_out);

/* Update_xz: ATy = A^T y - c + mu x ; per block eig split into x, z */
void orc_update_xz(const orc_layout *L, double mu, const double *y,
                   const double *c, double *x, double *z);
/* eigen-split of one symmetric block (row-major n x n): the ± parts */
void orc_psd_split(int n, const double *M, double mu, double *xblk, double *zblk);

void orc_build_b(const orc_layout *L, double *b);
/* c from active-space integrals: h1 (amo x amo, Pitzer order, full square)
 * and eri in chemists' notation (pq|rs), amo^4 full array */
void orc_build_c(const orc_layout *L, const double *h1, const double *eri, double *c);
void orc_guess_random(const orc_layout *L, double *x, double *z, double *y);

typedef struct orc_params {
    double tau, r_convergence, e_convergence, cg_convergence;
    int maxiter, cg_maxiter, mu_update_frequency;
    int verbose;
    /* continuation of an earlier run (all zero: a fresh solve from mu = 1):
     * mu, outer iterations done and the last primal / dual errors */
    double mu0;
    int oiter0;
    double ep0, ed0;
} orc_params;

typedef struct orc_stats {
    int oiter;            /* outer iterations performed                      */
    long cg_iters;        /* total CG iterations                             */
    int converged;
    double energy_primal, energy_dual, ep, ed, mu;
    double t_cg, t_xz, t_total;  /* seconds: CG phase, Update_xz phase, loop  */
} orc_stats;

/* full BPSDP loop; x, z, y in/out (guess on entry); c, b inputs.
 * max_seconds > 0 stops early after that much wall time (for timing). */
int orc_bpsdp(const orc_layout *L, const orc_params *p, const double *c,
              const double *b, double *x, double *z, double *y,
              double max_seconds, orc_stats *st);

/* LAPACK backend in use: 1 = MKL dsyev, 0 = built-in Jacobi */
int orc_lapack_backend(void);
int orc_num_threads(void);

/* per-entry MLP preconditioner oracle (2 hidden layers, tanh), fp32 weights
 * laid out as in the product: W1[in][h], b1[h], W2[h][h], b2[h], w3[h], b3 */
void orc_mlp_forward(int n, int nin, int nh, const float *feat, const float *W1,
                     const float *b1, const float *W2, const float *b2,
                     const float *w3, float b3, float *out);
/* loss = sum_i wt_i (t_i - s_i)^2 / n with s_i = out_i; gradients of weights;
 * target = NULL: wt_i is the upstream gradient d(loss)/d(out_i) itself (the
 * form the online training feeds the backward pass) and the loss is 0 */
double orc_mlp_loss_grad(int n, int nin, int nh, const float *feat, const float *target,
                         const float *wt, const float *W1, const float *b1,
                         const float *W2, const float *b2, const float *w3, float b3,
                         double *gW1, double *gb1, double *gW2, double *gb2,
                         double *gw3, double *gb3);

#ifdef __cplusplus
}
#endif
#endif
