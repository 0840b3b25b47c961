/*
 * v2g.h -- C ABI of the MI355X boundary-point SDP (BPSDP) engine of
 * v2RDM-CASSCF.  Drop-in for the hot path of v2RDMSolver (v2rdm_solver.h):
 * the constraint maps, the CG solve of A A^T y = rhs, the per-block eigen
 * split and the outer iteration.
 *
 * Conventions: plain C, int status (0 = success, negative = error, message
 * from v2g_last_error()), no exceptions cross the boundary.  Vector arguments
 * named *_dev are device pointers (hipMalloc / torch CUDA tensors) holding
 * FP64 data in the reference's layout (x: dimx entries in dimensions_ block
 * order; y: ncon entries in bpsdp_Au row order).  Calls run on the handle's
 * stream and return after it has drained, unless stated otherwise.
 */
#ifndef V2G_H
#define V2G_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct v2g_solver *v2g_handle;

/* Problem description; mirrors what v2RDMSolver::common_init derives from
 * Psi4 options (v2rdm_solver.cc:488-545 POSITIVITY -> constrain_*,
 * CONSTRAIN_SPIN, CONSTRAIN_SZ; plugin.cc:102-114). */
typedef struct v2g_layout_desc {
    int nirrep;            /* 1, 2, 4 or 8 (D2h subgroup, Cotton order)      */
    int amopi[8];          /* active orbitals per irrep                      */
    int nalpha, nbeta;     /* active electrons                               */
    int multiplicity;      /* 2S+1                                           */
    int constrain_q2, constrain_g2, constrain_t2;
    int constrain_spin, constrain_sz;
    int constrain_t1;      /* POSITIVITY DQGT1 / DQGT1T2 / DQGT              */
    int spin_adapt_q2;     /* SPIN_ADAPT_Q2 (plugin.cc:110): Q2s, Q2t, Q2t(+1),*/
                           /* Q2t(-1) blocks (q2.cc:260-670)                 */
    int spin_adapt_g2;     /* SPIN_ADAPT_G2 (plugin.cc:108): G2s, G2t, G2t(+1),*/
                           /* G2t(-1) blocks (g2.cc:235-856); needs           */
                           /* constrain_spin = 0, whose maximal-spin rows     */
                           /* address G2ba (spin.cc:250-268)                  */
    const int *energy_order; /* active orbitals from lowest to highest       */
                           /* orbital energy: the order of the singlet       */
                           /* geminals of spin-adapted Q2 (basis.cc:626-651);*/
                           /* NULL = Pitzer order                            */
} v2g_layout_desc;

/* Replaces v2RDMSolver construction + common_init sizes (v2rdm_solver.cc:
 * 129-1675): builds the geminal tables (basis.cc:439-757), block offsets and
 * the constraint-row plan on the device. */
int v2g_create(const v2g_layout_desc *desc, v2g_handle *out);
int v2g_destroy(v2g_handle h);
/* dimx_, nconstraints_, dimensions_.size() (v2rdm_solver.h:161-340) */
int v2g_dims(v2g_handle h, long long *dimx, long long *ncon, int *nblocks);
/* dimensions_ (v2rdm_solver.cc:1129-1270) */
int v2g_block_dims(v2g_handle h, int *dims);
/* run on this HIP stream instead of the handle's own (NULL = own) */
int v2g_set_stream(v2g_handle h, void *hip_stream);

/* v2RDMSolver::bpsdp_Au(A, u) (v2rdm_solver.cc:3020-3065): Ax = A x */
int v2g_Au(v2g_handle h, const double *x_dev, double *Ax_dev);
/* v2RDMSolver::bpsdp_ATu(A, u) (v2rdm_solver.cc:3114-3158): ATy = A^T y */
int v2g_ATu(v2g_handle h, const double *y_dev, double *ATy_dev);
/* v2RDMSolver::cg_Ax (v2rdm_solver.cc:3207-3213): out = A (A^T u) */
int v2g_cg_Ax(v2g_handle h, const double *u_dev, double *out_dev);
/* diag(A A^T): squared norm of every constraint row (no reference
 * counterpart; the Jacobi preconditioner) */
int v2g_row_norm2(v2g_handle h, double *d_dev);

/* CGSolver::solve (cg_solver.cc:125-177) with evaluate_Ap = cg_Ax:
 * warm start from y, stop when ||r||_2 < tol or after maxiter iterations.
 * precond: 0 none (the reference), 1 Jacobi 1/diag(A A^T), 2 the diagonal
 * given to v2g_set_precond_diag (learned); flexible != 0 uses the
 * Polak-Ribiere beta.  iters receives CGSolver::total_iterations(). */
int v2g_cg_solve(v2g_handle h, const double *rhs_dev, double *y_dev, double tol, int maxiter,
                 int precond, int flexible, int *iters, double *resnorm);
/* The same solve without any allocation (no reference counterpart; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The CG vectors live in a
 * static device array of 64 MiB per GPU instead of handle buffers: problems
 * with 3 ncon + dimx + the row-tile count (+ about 8 k) doubles beyond
 * 8 388 608 are rejected (-1).  precond 0 or 2 only.  Calls on one GPU share
 * that array, so calls that overlap in time on one GPU race. */
int v2g_cg_solve_noalloc(v2g_handle h, const double *rhs_dev, double *y_dev, double tol,
                         int maxiter, int precond, int flexible, int *iters, double *resnorm);
/* inverse-diagonal used by precond = 2 (ncon entries, device) */
int v2g_set_precond_diag(v2g_handle h, const double *minv_dev);

/* v2RDMSolver::Update_xz (v2rdm_solver.cc:3216-3289): M = A^T y - c + mu x,
 * per block x = P+(M)/mu, z = -P-(M); x is read and overwritten.  The
 * eigenvectors are kept for the next call (warm start). */
int v2g_update_xz(v2g_handle h, double mu, const double *y_dev, const double *c_dev,
                  double *x_dev, double *z_dev);
/* forget the kept eigenvectors (next update_xz starts cold) */
int v2g_reset_eig(v2g_handle h);

/* ---- launch shape of the eigen split (no reference counterpart) ----
 * lanes: 1 or 2 groups of blocks advancing their block-Jacobi rounds on their
 * own streams; pair_threads: 256 or 1024 threads per 64 x 64 pair
 * sub-problem.  0 keeps the current value.  Every shape gives the same split
 * bit for bit; the default follows the largest padded block (two lanes and
 * 1024 threads from 512 on). */
int v2g_set_eig_shape(v2g_handle h, int lanes, int pair_threads);
int v2g_get_eig_shape(v2g_handle h, int *lanes, int *pair_threads);
/* on != 0: the next warm splits time every shape (two calls each, block-Jacobi
 * rounds only, HIP events) and keep the fastest per sweep; the choice is read
 * back with v2g_get_eig_shape.  on = 0 stops a tuning in progress. */
int v2g_set_autotune(v2g_handle h, int on);

/* ---- learned preconditioner (no reference counterpart) ----
 * A per-row MLP (14 inputs: one-hot constraint family x11, log diag(A A^T),
 * asinh(rhs/rms), asinh(y_prev/rms); tanh hidden layers 64, 64; one output g)
 * gives M^-1_ii = exp(clamp(g_i, -3, 3)), divided by diag(A A^T)_ii with the
 * Jacobi base (v2g_params.mlp_base = 1).  Inside v2g_bpsdp (precond 3) it is
 * trained online during the first train_iters outer iterations, whose CG
 * solves use the base method, on the spectral objective log rho(r_hi) -
 * log rho(r_lo) (rho the Rayleigh quotient of the preconditioned operator;
 * r_lo the base solve's final residual, r_hi from power iterations), then
 * used in flexible CG, guarded by the true residual.  Weights are FP32,
 * packed W1[14][64], b1[64], W2[64][64], b2[64], w3[64], b3 (5185 floats). */
int v2g_mlp_num_params(void);
/* seeded init; the last layer starts at zero, i.e. exactly Jacobi */
int v2g_mlp_init(v2g_handle h, unsigned long long seed);
int v2g_mlp_get_weights(v2g_handle h, float *w_host);
int v2g_mlp_set_weights(v2g_handle h, const float *w_host);
/* loss = sum_b wt_b (out_b - tgt_b)^2 / nb and its gradient (FP64 out,
 * device pointers; feat is nb x 14 FP32) -- the training kernels; with
 * tgt_dev = NULL, wt_dev holds the upstream gradient d(loss)/d(out_b) per
 * sample (the form online training uses) and loss is 0 */
int v2g_mlp_loss_grad(v2g_handle h, int nb, const float *feat_dev, const float *tgt_dev,
                      const float *wt_dev, double *grad_dev, double *loss_host);
/* the training signal of the spectral objective for every row: dout_dev[i]
 * (FP32) = d L / d g_i, L = log rho(r_hi) - log rho(r_lo) with S the current
 * learned diagonal from (rhs, y); objective receives L */
int v2g_mlp_spectral_grad(v2g_handle h, const double *rhs_dev, const double *y_dev,
                          const double *rhi_dev, const double *rlo_dev, float *dout_dev,
                          double *objective);
/* one Adam step (lr) on the current weights with grad_dev */
int v2g_mlp_adam_step(v2g_handle h, const double *grad_dev, double lr);
/* inference: minv for every constraint row from rhs and y (device) */
int v2g_mlp_minv(v2g_handle h, const double *rhs_dev, const double *y_dev, double *minv_dev);

typedef struct v2g_params {
    double tau;                  /* TAU_PARAMETER (plugin.cc:97)          */
    double r_convergence;        /* R_CONVERGENCE                         */
    double e_convergence;        /* E_CONVERGENCE                         */
    double cg_convergence;       /* CG_CONVERGENCE                        */
    int maxiter;                 /* MAXITER                               */
    int cg_maxiter;              /* CG_MAXITER                            */
    int mu_update_frequency;     /* MU_UPDATE_FREQUENCY                   */
    int precond;                 /* as v2g_cg_solve, or 3: learned online  */
    int flexible;
    int verbose;                 /* 1: one line per outer iteration;      */
                                 /* 2 (batches): a progress line on       */
                                 /* stderr every 20 s                     */
    double max_seconds;          /* > 0: stop after this wall time        */
    int train_iters;             /* precond 3: outer iterations of data   */
    int train_batch;             /* precond 3: rows sampled per iteration */
    int train_steps;             /* precond 3: Adam steps per iteration   */
    double train_lr;             /* precond 3: Adam learning rate         */
    int eig_warm_sweeps;         /* > 0: block-Jacobi sweeps per warm      */
                                 /* Update_xz inside the loop (0: converge) */
    int eig_inner_sweeps;        /* > 0: cyclic-Jacobi sweeps per 64 x 64  */
                                 /* pair sub-problem of a warm call (0: 12) */
    int cg_extrapolate;          /* 1: start each CG from 2 y_k - y_(k-1)  */
                                 /* (linear extrapolation of the last two */
                                 /* solutions); 0: from y_k (reference)   */
    /* resume (a solve continued by a second call, or from a checkpoint):
     * the state of the last iteration done, as the previous call's stats
     * report it; all zero for a fresh solve (mu starts at 1, as in
     * v2rdm_solver.cc:1735) */
    double mu0;                  /* > 0: mu to continue with               */
    int oiter0;                  /* outer iterations already done          */
    double ep0, ed0;             /* their last primal / dual errors (CG    */
                                 /* tolerance of the first iteration)      */
    int mlp_base;                /* precond 3: 0 the learned diagonal      */
                                 /* scales the identity (plain CG), 1 the  */
                                 /* Jacobi diagonal                        */
    double train_penalty;        /* precond 3: weight of (1/2) mean(g^2),  */
                                 /* which holds the learned diagonal near  */
                                 /* its base                               */
    int eig_pair_order;          /* capped warm splits: 0 every round's    */
                                 /* 64 x 64 pairs in cyclic order; 1 cyclic */
                                 /* in the first round of a sweep, then    */
                                 /* only the pairs across the two column   */
                                 /* blocks (32 instead of 63 steps)        */
} v2g_params;

typedef struct v2g_stats {
    int oiter;                   /* outer iterations done                 */
    long long cg_iters;          /* total CG iterations                   */
    int converged;
    double energy_primal, energy_dual, ep, ed, mu;
    double t_cg, t_xz, t_total;  /* seconds: CG phase, Update_xz, loop    */
    int eig_sweeps;              /* block-Jacobi sweeps, last Update_xz   */
    int precond_active;          /* preconditioner of the last CG solve   */
    long long cg_iters_train;    /* CG iterations during the training phase */
    int train_outer;             /* outer iterations in the training phase */
    double mlp_loss;             /* last training loss                     */
    int fallback;                /* 1 if the learned PC was abandoned      */
    int exact_splits;            /* iterations whose tracked split was     */
                                 /* redone to convergence before the       */
                                 /* stopping rule was applied              */
} v2g_stats;

/* v2RDMSolver::compute_energy outer loop (v2rdm_solver.cc:1778-1912, orbital
 * optimization off): x, z, y in/out (the guess on entry), c and b inputs,
 * all device-resident for the whole solve.  Energies exclude enuc/efzc. */
int v2g_bpsdp(v2g_handle h, const v2g_params *p, const double *c_dev, const double *b_dev,
              double *x_dev, double *z_dev, double *y_dev, v2g_stats *st);

/* callback invoked after each outer iteration inside v2g_bpsdp when set
 * (hook for the learned preconditioner's online training); return nonzero
 * to stop the loop */
typedef int (*v2g_iter_callback)(void *user, int oiter, const v2g_stats *st);
int v2g_set_iter_callback(v2g_handle h, v2g_iter_callback cb, void *user);

/* Path inputs, built on the host (one-time work):
 * v2RDMSolver::BuildConstraints (v2rdm_solver.cc:2518-3017): b, ncon entries */
int v2g_build_b(v2g_handle h, double *b_host);
/* RepackIntegrals + FrozenCoreEnergy without core orbitals (tei.cc:144-313):
 * h1 = amo x amo one-electron integrals, eri = amo^4 (pq|rs) (chemists'
 * notation), both in Pitzer order, row-major; c has dimx entries */
int v2g_build_c(v2g_handle h, const double *h1_host, const double *eri_host, double *c_host);
/* TPDM_GUESS RANDOM (v2rdm_solver.cc:2479-2490): srand(0), then rand() for
 * x, z (interleaved) and y, each in [-2, 0] */
int v2g_guess_random(v2g_handle h, double *x_host, double *z_host, double *y_host);

/* The same three inputs built on the device, for batched runs (no reference
 * counterpart beyond the functions above; results equal theirs bit for bit):
 * b (ncon) and c (dimx) from device h1 (amo x amo) and eri (amo^4); the guess
 * from glibc's rand() sequence after srand(0), restated on the device. */
int v2g_build_b_device(v2g_handle h, double *b_dev);
int v2g_build_c_device(v2g_handle h, const double *h1_dev, const double *eri_dev, double *c_dev);
int v2g_guess_random_device(v2g_handle h, double *x_dev, double *z_dev, double *y_dev);

/* ---- RDM outputs (x on the device, results in host memory) ----
 * v2RDMSolver::get_opdm (python_api.cc:376-396): per irrep h an amopi[h] x
 * amopi[h] block, blocks concatenated in irrep order; the spin-summed
 * D1a + D1b (the reference adds the D1a block twice, which is the same for
 * the spin-constrained singlets it is used with). */
int v2g_get_opdm(v2g_handle h, const double *x_dev, double *opdm_host);
/* v2RDMSolver::get_tpdm (python_api.cc:398-458): amo^2 x amo^2, element
 * [i amo + k][j amo + l]; C1 only (-1 with the reference's message otherwise) */
int v2g_get_tpdm(v2g_handle h, const double *x_dev, double *tpdm_host);
/* Orbital-optimization hand-off: v2RDMSolver::PackSpatialDensity
 * (v2rdm_solver.cc:3664-3773) from x on the device.  d1 (d1dim doubles):
 * per irrep the packed lower triangle of D1a + D1b; d2 (d2dim doubles): per
 * irrep of the singlet geminals (basis.cc:626-651, in the layout's orbital-
 * energy order) the packed lower triangle over singlet-geminal pairs of the
 * spin-summed 2-RDM, as the reference's orbital optimizer (OrbOpt,
 * fortran.h:41-58) reads them. */
int v2g_spatial_density_dims(v2g_handle h, long long *d1dim, long long *d2dim);
/* the singlet geminals the packed D2 runs over: counts[h] per irrep and the
 * pairs (i, j) (active orbitals, Pitzer order), irrep after irrep
 * (ibas_00_sym / bas_00_sym of basis.cc:626-651) */
int v2g_singlet_geminals(v2g_handle h, int *counts, int *pairs);
int v2g_pack_spatial_density(v2g_handle h, const double *x_dev, double *d1_host,
                             double *d2_host);
/* the RDM file of v2RDMSolver::FCIDUMP (fcidump.cc:176-268): binary records
 * (double value, int i, j, k, l) of the spin-summed 2-RDM, then the 1-RDM
 * with k = l = 0; active orbitals numbered from 1 in Pitzer order (no core) */
int v2g_write_fcidump_rdm(v2g_handle h, const double *x_dev, const char *path);

/* ---- batches of independent problems (bond scans; no reference
 * counterpart: the reference solves one problem per process) ----
 * All problems' vectors are concatenated: problem q's x / z / c entries start
 * at xoff[q], its y / b rows at yoff[q] (v2g_batch_offsets fills nprob + 1
 * entries, the last one the totals).  Every kernel launch of the iteration
 * covers the whole batch; each problem keeps its own mu, errors, CG scalars
 * and stopping rule, and leaves the batch when it converges. */
typedef struct v2g_batch *v2g_batch_handle;
int v2g_batch_create(const v2g_layout_desc *descs, int nprob, v2g_batch_handle *out);
int v2g_batch_destroy(v2g_batch_handle h);
int v2g_batch_offsets(v2g_batch_handle h, long long *xoff, long long *yoff);
/* bpsdp_Au / bpsdp_ATu of every problem at once (concatenated vectors) */
int v2g_batch_Au(v2g_batch_handle h, const double *x_dev, double *Ax_dev);
int v2g_batch_ATu(v2g_batch_handle h, const double *y_dev, double *ATy_dev);
/* b, c (from per-problem device h1 / eri arrays) and the RANDOM guess of
 * every problem, on the device */
int v2g_batch_inputs(v2g_batch_handle h, const double *const *h1_dev, const double *const *eri_dev,
                     double *c_dev, double *b_dev, double *x_dev, double *z_dev, double *y_dev);
/* the BPSDP loop of every problem (precond 0, the reference's plain CG);
 * stats receives nprob records */
int v2g_batch_bpsdp(v2g_batch_handle h, const v2g_params *p, const double *c_dev,
                    const double *b_dev, double *x_dev, double *z_dev, double *y_dev,
                    v2g_stats *stats);

const char *v2g_last_error(void);
const char *v2g_version(void);

#ifdef __cplusplus
}
#endif
#endif
