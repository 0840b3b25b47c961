// kernels.hpp -- launch wrappers of the HIP kernels (maps, CG, eigen split,
// learned preconditioner).  Every wrapper enqueues on the given stream and
// returns without synchronizing.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "devlayout.hpp"

namespace v2g {

struct MapPlan {
    const DevLayout *dlay = nullptr;
    const Segment *rowsegs = nullptr;
    const int64_t *row_tile_start = nullptr;
    int nrowsegs = 0;
    int64_t row_tiles = 0;
    const Segment *xsegs = nullptr;
    const int64_t *x_tile_start = nullptr;
    int nxsegs = 0;
    int64_t x_tiles = 0;
    const Segment *t2segs = nullptr;
    const int64_t *t2_tile_start = nullptr;
    int nt2segs = 0;
    int64_t t2_tiles = 0;
    // precomputed (D2 column << 1 | sign, D1 column) of the G2 rows, indexed
    // by each G2 row segment's aux + visiting position (null: not built)
    const int2 *g2idx = nullptr;
    // dlay holds one layout per problem (Segment::prob picks it); the done flag
    // of problem p is done[p * done_stride] (0: one flag for the whole plan)
    int done_stride = 0;
};

int64_t tiles_for(const Segment &s);
// fill the G2 column table for the row segments whose aux is >= 0; D1 columns
// of x are [d1lo, d1hi)
void build_g2_index(const MapPlan &p, const std::vector<Segment> &rowsegs, int64_t d1lo,
                    int64_t d1hi, int2 *idx, hipStream_t st);

// The residual update of a CG iteration fused into A.x (r is set): A.x is not
// stored; r -= alpha (A.x) with alpha = scal[CG_ALPHA], and part[k ntiles +
// tile] (k = 0, 1, 2) receives the tile's sums of r.z, r.r and r_old.z of the
// new residual, z = minv r (minv null: z = r); ntiles = the plan's row_tiles
struct CgFuse {
    double *r = nullptr;
    const double *minv = nullptr;
    const double *scal = nullptr;
    double *part = nullptr;
    int64_t ntiles = 0;
};

// dotv / rowpart (optional): rowpart[tile] = sum over the tile's rows of dotv . Ax
void launch_au(const MapPlan &p, const double *x, double *Ax, hipStream_t st,
               const double *done = nullptr, const double *dotv = nullptr,
               double *rowpart = nullptr, CgFuse fz = {});
// sqpart (optional, layouts without second-pass rows only): sqpart[tile] =
// the tile's sum of squares of A^T.y (x_tiles entries)
void launch_atu(const MapPlan &p, const double *y, double *ATy, hipStream_t st,
                const double *done = nullptr, double *sqpart = nullptr);
void launch_row_norm2(const MapPlan &p, double *d, hipStream_t st);

// ---- CG (cg.hip) ----
// device scalar slots of one CG solve; r.z and the done flag are ping-pong
// pairs indexed by the iteration parity
enum CgSlot : int {
    CG_RZ0 = 16, CG_RZ1 = 17, CG_DONE0 = 18, CG_DONE1 = 19, CG_NRM = 20, CG_ITER = 21,
    CG_ALPHA = 22, CG_BETA = 23, CG_PAP = 24, CG_TOL = 25, CG_MAXIT = 26,
    CG_IT0 = 27, CG_IT1 = 28,   // iteration counter, ping-pong like r.z
    CG_NRM0 = 29,               // ||r|| of the starting guess
    CG_NSLOTS = 32
};
constexpr int kRedBlocks = 2048;    // partial-sum slots per reduction (= grid of the vector kernels)
constexpr int kPartialArrays = 4;   // partial buffer holds 4 * kRedBlocks doubles

void cg_init(int64_t n, const double *b, const double *Ap, double *r, double *p, const double *minv,
             double *z, double *partial, double *scal, double tol, int maxiter, hipStream_t st);
// p.Ap from the per-tile partials of the fused A.x launch
void cg_dot_from_rows(int64_t ntiles, const double *rowpart, double *partial, const double *scal,
                      int parity, hipStream_t st);
// alpha = r.z / p.Ap of this parity with p.Ap = ||A^T p||^2 summed from the
// per-tile squares of the A^T.p launch (one workgroup, fixed order)
void cg_alpha_from_tiles(int64_t ntiles, const double *sqpart, double *scal, int parity,
                         hipStream_t st);
// the r.z, r.r and r_old.z partials of the fused A.x launch, per-tile sums
// gathered into the partial arrays k_finish reads
void cg_step_from_tiles(int64_t ntiles, const double *part, double *partial, const double *scal,
                        int parity, hipStream_t st);
// r -= alpha Ap and the partial dot products of the new residual
void cg_step(int64_t n, double *r, const double *Ap, const double *minv, double *partial,
             double *scal, int parity, hipStream_t st);
// y += alpha p, convergence test, p update.  The iteration number is read from
// the device counter of this parity, so a chunk of iterations can be captured
// once and replayed
void cg_finish_iter(int64_t n, double *y, double *r, double *p, const double *z, double *partial,
                    double *scal, int parity, int flexible, hipStream_t st);

// ---- vector helpers (cg.hip) ----
void vec_axpby(int64_t n, double a, const double *x, double b, const double *y, double *out,
               hipStream_t st);
void vec_fill(int64_t n, double v, double *out, hipStream_t st);
// out = a x + b y + cw w (w may be null; out may alias x) and kRedBlocks partial sums of
// out.out for vec_reduce
void vec_lincomb_sq(int64_t n, double a, const double *x, double b, const double *y, double cw,
                    const double *w, double *out, double *partial, hipStream_t st);
void vec_dot_partial(int64_t n, const double *a, const double *b, double *partial, hipStream_t st);
void vec_reduce(const double *partial, double *out, hipStream_t st);
// y <- 2 y - yh and yh <- old y (linear extrapolation of the CG starting guess)
void vec_extrapolate(int64_t n, double *y, double *yh, hipStream_t st);
void vec_mul(int64_t n, const double *a, const double *b, double *out, hipStream_t st);

}  // namespace v2g
