// eig.hpp -- launch wrappers of the FP64 eigen split (eig.hip) and the
// batched FP64 GEMM (gemm.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace v2g {

// cyclic sweeps of each 64 x 64 pair problem per round in a split that must
// converge (a direct Update_xz, a cold split, the exact re-split): the block
// sweeps repeat until none rotates, so one pass per pair problem and round
// reaches the same tolerance in the least time (H50 warm split with the
// cyclic order in every round: 175 ms with 1, 219 ms with 2, 349 ms with 12;
// 143 ms with 1 and the cross-pair order after each sweep's first round)
constexpr int kEigFullInner = 1;

struct EigBlock {
    int n;          // block dimension
    int N;          // padded dimension (multiple of 64) for the block-Jacobi path
    int64_t xoff;   // offset of the block in x / z / A^T y
    int64_t woff;   // offset in the N x N workspaces
    int64_t qoff;   // offset in the pair-rotation buffer (N/64 pairs of 64 x 64)
    int64_t poff;   // offset in the pair-active flags
    int prob;       // problem of a batch the block belongs to (0 alone)
    int pad;
};

// Batched splits: mu of each problem and which problems take part (null: one
// problem, the scalar mu, all blocks)
struct EigBatch {
    const double *mus = nullptr;
    const int *active = nullptr;
};

struct GemmDesc {
    const double *A;
    const double *B;
    double *C;
    int M, N, K;
    int lda, ldb, ldc;
    int ta, tb;     // 1: use the transpose of A / B
    int sym;        // 1: C is symmetric (M == N); only upper tiles are computed
};

void eig_small(const EigBlock *blocks, int nblocks, const double *aty, const double *c, double *x,
               double *z, double mu, hipStream_t st, EigBatch eb = {});
// max_N: the largest padded dimension among the blocks
void eig_build_m(const EigBlock *blocks, int nblocks, int max_N, const double *aty, const double *c,
                 const double *x, double mu, double *Mw, double *Aw, double *Vw, int cold,
                 hipStream_t st, EigBatch eb = {});
// tol[b] = 2e-14 ||M_b||_F (times 64 / N_b when tight); part holds 64 * nblocks doubles
void eig_tol(const EigBlock *blocks, int nblocks, const double *Mw, double *tol, double *part,
             int tight, hipStream_t st, EigBatch eb = {});
// one block-Jacobi round: the pair sub-problems (inner_sweeps caps the cyclic
// sweeps of each 64 x 64 pair; wide = 1: 1024-thread workgroups, for problems
// whose rounds are latency-bound; bip = 1: only the pairs across the two
// column blocks, in 32 steps), then the rotation of the A and V tiles
void eig_pairs(const EigBlock *blocks, int nblocks, int maxpairs, int round, const double *Aw,
               double *Qbuf, int *active, const double *tol, int *counter, int inner_sweeps,
               int wide, int bip, hipStream_t st, EigBatch eb = {});
void eig_apply(const EigBlock *blocks, int nblocks, int64_t maxtiles, int round, double *Aw,
               double *Vw, const double *Qbuf, const int *active, hipStream_t st, EigBatch eb = {});
void eig_scale(const EigBlock *blocks, int nblocks, const double *Aw, const double *Vw, double *Tw,
               double mu, hipStream_t st, EigBatch eb = {});
void eig_z(const EigBlock *blocks, int nblocks, const double *x, const double *Mw, double *z,
           double mu, hipStream_t st, EigBatch eb = {});

// C = op(A) op(B) for every descriptor (beta = 0); grid sized for the largest
void gemm_batched(const GemmDesc *descs, int ndesc, int max_m, int max_n, hipStream_t st);

}  // namespace v2g
