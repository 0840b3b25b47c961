// cg.hip -- device-resident conjugate gradient on A A^T y = rhs.
//
// Same recurrence and stopping rule as CGSolver::solve (cg_solver.cc:125-177):
// warm start from y, absolute ||r||_2 < tol tested after the update, iteration
// count incremented after the p update, at most maxiter iterations.  Options
// beyond the reference: a diagonal preconditioner M^-1 (Jacobi, or the
// learned scaling) and the flexible (Polak-Ribiere) beta.
//
// All scalars live in device memory.  Reductions are two-stage: a fixed grid
// of kRedBlocks blocks writes partial sums, and every block of the consuming
// kernel re-sums the same partials in the same order, so each block derives
// bit-identical alpha / beta without a grid barrier.  Iteration k uses the
// ping-pong slot k & 1 for r.z and the done flag, so no kernel reads a slot
// that a concurrent block of the same kernel writes.  Once the done flag is
// set, every later kernel of the solve returns at once; the host polls the
// flag every few iterations.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace v2g {

namespace {

constexpr int kThreads = 256;

enum : int { S_RZ0 = CG_RZ0, S_DONE0 = CG_DONE0, S_DONE1 = CG_DONE1, S_NRM = CG_NRM,
             S_ITER = CG_ITER, S_ALPHA = CG_ALPHA, S_BETA = CG_BETA, S_PAP = CG_PAP,
             S_TOL = CG_TOL, S_MAXIT = CG_MAXIT, S_IT0 = CG_IT0 };
// partial-sum arrays inside `partial` (kRedBlocks each)
enum : int { P_PAP = 0, P_RZ = 1, P_RR = 2, P_ROZ = 3, P_COUNT = 4 };

__device__ __forceinline__ double block_sum(double v, double *sh)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kThreads / 64; i++) s += sh[i];
    return s;  // valid in thread 0
}

// every block sums the kRedBlocks partials in the same fixed order
__device__ __forceinline__ double all_sum(const double *__restrict__ part, double *sh)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += kThreads) v += part[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < kThreads / 64; i++) s += sh[i];
    __syncthreads();
    return s;  // valid in every thread
}

__global__ void __launch_bounds__(kThreads)
k_init(int64_t n, const double *__restrict__ b, const double *__restrict__ Ap,
       double *__restrict__ r, double *__restrict__ p, const double *__restrict__ minv,
       double *__restrict__ z, double *__restrict__ partial)
{
    __shared__ double sh[kThreads / 64];
    double rz = 0.0, rr = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const double ri = b[i] - Ap[i];
        r[i] = ri;
        const double zi = minv ? minv[i] * ri : ri;
        if (z) z[i] = zi;
        p[i] = zi;
        rz += ri * zi;
        rr += ri * ri;
    }
    double s = block_sum(rz, sh);
    if (threadIdx.x == 0) partial[P_RZ * kRedBlocks + blockIdx.x] = s;
    s = block_sum(rr, sh);
    if (threadIdx.x == 0) partial[P_RR * kRedBlocks + blockIdx.x] = s;
}

__global__ void k_init_fin(double *__restrict__ partial, double *__restrict__ scal, double tol,
                           int maxiter)
{
    __shared__ double sh[kThreads / 64];
    const double rz = all_sum(partial + P_RZ * kRedBlocks, sh);
    const double rr = all_sum(partial + P_RR * kRedBlocks, sh);
    if (threadIdx.x == 0) {
        scal[CG_NRM0] = sqrt(rr);
        scal[S_RZ0] = rz;
        scal[S_DONE0] = (maxiter <= 0) ? 1.0 : 0.0;
        scal[S_DONE1] = 0.0;
        scal[S_ITER] = 0.0;
        scal[S_IT0] = 0.0;
        scal[S_IT0 + 1] = 0.0;
        scal[S_TOL] = tol;
        scal[S_MAXIT] = (double)maxiter;
        scal[S_NRM] = 0.0;
    }
}

// alpha = rz / pap; r -= alpha Ap; z = M^-1 r; partial r.z, r.r, r_old.z.
// y += alpha p is left to k_finish, which reads p anyway (one vector pass less)
__global__ void __launch_bounds__(kThreads)
k_step(int64_t n, double *__restrict__ r, const double *__restrict__ Ap,
       const double *__restrict__ minv, double *__restrict__ partial, double *__restrict__ scal,
       int parity)
{
    __shared__ double sh[kThreads / 64];
    if (scal[S_DONE0 + parity] != 0.0) return;
    const double pap = all_sum(partial + P_PAP * kRedBlocks, sh);
    const double alpha = scal[S_RZ0 + parity] / pap;
    double rz = 0.0, rr = 0.0, roz = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const double ro = r[i];
        const double rn = ro - alpha * Ap[i];
        r[i] = rn;
        const double zn = minv ? minv[i] * rn : rn;
        rz += rn * zn;
        rr += rn * rn;
        roz += ro * zn;
    }
    double t = block_sum(rz, sh);
    if (threadIdx.x == 0) partial[P_RZ * kRedBlocks + blockIdx.x] = t;
    t = block_sum(rr, sh);
    if (threadIdx.x == 0) partial[P_RR * kRedBlocks + blockIdx.x] = t;
    t = block_sum(roz, sh);
    if (threadIdx.x == 0) partial[P_ROZ * kRedBlocks + blockIdx.x] = t;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[S_ALPHA] = alpha;
        scal[S_PAP] = pap;
    }
}

// y += alpha p (alpha from k_step), convergence test and p = M^-1 r + beta p
// (z carries M^-1 here, not z)
__global__ void __launch_bounds__(kThreads)
k_finish(int64_t n, double *__restrict__ y, double *__restrict__ p, const double *__restrict__ r,
         const double *__restrict__ z, double *__restrict__ partial, double *__restrict__ scal,
         int parity, int flexible)
{
    __shared__ double sh[kThreads / 64];
    if (scal[S_DONE0 + parity] != 0.0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) scal[S_DONE0 + (parity ^ 1)] = 1.0;
        return;
    }
    const double alpha = scal[S_ALPHA];
    const double rz = all_sum(partial + P_RZ * kRedBlocks, sh);
    const double rr = all_sum(partial + P_RR * kRedBlocks, sh);
    const double roz = flexible ? all_sum(partial + P_ROZ * kRedBlocks, sh) : 0.0;
    const double nrm = sqrt(rr);
    const double rzold = scal[S_RZ0 + parity];
    const bool conv = nrm < scal[S_TOL];
    const int iter = (int)scal[S_IT0 + parity];
    const int iter_next = iter + 1;
    // block 0 writes only next-parity slots and values no block of this launch reads
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[S_NRM] = nrm;
        if (conv) {
            scal[S_DONE0 + (parity ^ 1)] = 1.0;
            scal[S_ITER] = (double)iter;
        } else {
            scal[S_ITER] = (double)iter_next;
            scal[S_DONE0 + (parity ^ 1)] = (iter_next >= (int)scal[S_MAXIT]) ? 1.0 : 0.0;
            scal[S_RZ0 + (parity ^ 1)] = rz;
        }
        scal[S_IT0 + (parity ^ 1)] = (double)iter_next;
    }
    if (conv) {
        for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * kThreads)
            y[i] += alpha * p[i];
        return;
    }
    const double beta = (rz - (flexible ? roz : 0.0)) / rzold;
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[S_BETA] = beta;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const double pi = p[i];
        y[i] += alpha * pi;
        p[i] = (z ? z[i] * r[i] : r[i]) + beta * pi;
    }
}

// partial[P_PAP][b] = sum of a contiguous chunk of the per-tile partials
__global__ void __launch_bounds__(kThreads)
k_rows_to_partials(int64_t ntiles, const double *__restrict__ rowpart, double *__restrict__ partial,
                   const double *__restrict__ scal, int parity)
{
    __shared__ double sh[kThreads / 64];
    if (scal[S_DONE0 + parity] != 0.0) return;
    const int64_t chunk = (ntiles + gridDim.x - 1) / gridDim.x;
    const int64_t lo = blockIdx.x * chunk, hi = lo + chunk < ntiles ? lo + chunk : ntiles;
    double s = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kThreads) s += rowpart[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) partial[P_PAP * kRedBlocks + blockIdx.x] = t;
}

// alpha = r.z / p.Ap with p.Ap = ||A^T p||^2 from the per-tile squares of the
// A^T.p launch: one workgroup of kAlphaThreads sums them in a fixed order
constexpr int kAlphaThreads = 1024;

__global__ void __launch_bounds__(kAlphaThreads)
k_alpha_tiles(int64_t ntiles, const double *__restrict__ sqpart, double *__restrict__ scal,
              int parity)
{
    __shared__ double sh[kAlphaThreads / 64];
    if (scal[S_DONE0 + parity] != 0.0) return;
    // four independent sums per thread keep four loads in flight
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    int64_t i = threadIdx.x;
    for (; i + 3 * kAlphaThreads < ntiles; i += 4 * kAlphaThreads) {
        v0 += sqpart[i];
        v1 += sqpart[i + kAlphaThreads];
        v2 += sqpart[i + 2 * kAlphaThreads];
        v3 += sqpart[i + 3 * kAlphaThreads];
    }
    for (; i < ntiles; i += kAlphaThreads) v0 += sqpart[i];
    double v = (v0 + v1) + (v2 + v3);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double pap = 0.0;
        for (int w = 0; w < kAlphaThreads / 64; w++) pap += sh[w];
        scal[S_ALPHA] = scal[S_RZ0 + parity] / pap;
        scal[S_PAP] = pap;
    }
}

// partial[P_RZ / P_RR / P_ROZ][b] = sums of contiguous chunks of the per-tile
// r.z, r.r, r_old.z of the fused A.x launch (part holds three arrays of ntiles)
__global__ void __launch_bounds__(kThreads)
k_step_tiles(int64_t ntiles, const double *__restrict__ part, double *__restrict__ partial,
             const double *__restrict__ scal, int parity)
{
    __shared__ double sh[kThreads / 64];
    if (scal[S_DONE0 + parity] != 0.0) return;
    const int64_t chunk = (ntiles + gridDim.x - 1) / gridDim.x;
    const int64_t lo = blockIdx.x * chunk, hi = lo + chunk < ntiles ? lo + chunk : ntiles;
    const int dst[3] = {P_RZ, P_RR, P_ROZ};
    for (int k = 0; k < 3; k++) {
        double s = 0.0;
        for (int64_t i = lo + threadIdx.x; i < hi; i += kThreads) s += part[k * ntiles + i];
        const double t = block_sum(s, sh);
        if (threadIdx.x == 0) partial[dst[k] * kRedBlocks + blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kThreads)
k_axpby(int64_t n, double a, const double *x, double b, const double *y, double *out)
{
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        out[i] = a * x[i] + (y ? b * y[i] : 0.0);
}

// out = a x + b y + cw w (w null: no third term) and the block partial sums
// of out.out (the grid and summation order of k_dotp)
__global__ void __launch_bounds__(kThreads)
k_lincomb_sq(int64_t n, double a, const double *x, double b, const double *y, double cw,
             const double *w, double *out, double *__restrict__ partial)
{
    __shared__ double sh[kThreads / 64];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        double v = a * x[i] + b * y[i];
        if (w) v = v + cw * w[i];
        out[i] = v;
        s += v * v;
    }
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kThreads)
k_fillv(int64_t n, double v, double *out)
{
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        out[i] = v;
}

__global__ void __launch_bounds__(kThreads)
k_dotp(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
       double *__restrict__ partial)
{
    __shared__ double sh[kThreads / 64];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        s += a[i] * b[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kThreads)
k_mul(int64_t n, const double *a, const double *b, double *out)
{
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads)
        out[i] = a[i] * b[i];
}

// CG starting guess from the last two solutions: y <- 2 y - yh, yh <- old y
__global__ void __launch_bounds__(kThreads)
k_extrapolate(int64_t n, double *__restrict__ y, double *__restrict__ yh)
{
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        const double yc = y[i];
        y[i] = 2.0 * yc - yh[i];
        yh[i] = yc;
    }
}

__global__ void k_reduce(const double *__restrict__ partial, double *__restrict__ out)
{
    __shared__ double sh[kThreads / 64];
    const double s = all_sum(partial, sh);
    if (threadIdx.x == 0) *out = s;
}

}  // namespace

// public wrappers (declared in kernels.hpp); `scal` must hold 32 doubles and
// `partial` P_COUNT * kRedBlocks doubles
void cg_init(int64_t n, const double *b, const double *Ap, double *r, double *p, const double *minv,
             double *z, double *partial, double *scal, double tol, int maxiter, hipStream_t st)
{
    hipLaunchKernelGGL(k_init, dim3(kRedBlocks), dim3(kThreads), 0, st, n, b, Ap, r, p, minv, z,
                       partial);
    hipLaunchKernelGGL(k_init_fin, dim3(1), dim3(kThreads), 0, st, partial, scal, tol, maxiter);
}

void cg_dot_from_rows(int64_t ntiles, const double *rowpart, double *partial, const double *scal,
                      int parity, hipStream_t st)
{
    hipLaunchKernelGGL(k_rows_to_partials, dim3(kRedBlocks), dim3(kThreads), 0, st, ntiles, rowpart,
                       partial, scal, parity);
}

void cg_alpha_from_tiles(int64_t ntiles, const double *sqpart, double *scal, int parity,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_alpha_tiles, dim3(1), dim3(kAlphaThreads), 0, st, ntiles, sqpart, scal,
                       parity);
}

void cg_step_from_tiles(int64_t ntiles, const double *part, double *partial, const double *scal,
                        int parity, hipStream_t st)
{
    hipLaunchKernelGGL(k_step_tiles, dim3(kRedBlocks), dim3(kThreads), 0, st, ntiles, part, partial,
                       scal, parity);
}

void cg_step(int64_t n, double *r, const double *Ap, const double *minv, double *partial,
             double *scal, int parity, hipStream_t st)
{
    hipLaunchKernelGGL(k_step, dim3(kRedBlocks), dim3(kThreads), 0, st, n, r, Ap, minv, partial,
                       scal, parity);
}

void cg_finish_iter(int64_t n, double *y, double *r, double *p, const double *z, double *partial,
                    double *scal, int parity, int flexible, hipStream_t st)
{
    hipLaunchKernelGGL(k_finish, dim3(kRedBlocks), dim3(kThreads), 0, st, n, y, p, r, z, partial,
                       scal, parity, flexible);
}

void vec_axpby(int64_t n, double a, const double *x, double b, const double *y, double *out,
               hipStream_t st)
{
    hipLaunchKernelGGL(k_axpby, dim3(kRedBlocks), dim3(kThreads), 0, st, n, a, x, b, y, out);
}

void vec_lincomb_sq(int64_t n, double a, const double *x, double b, const double *y, double cw,
                    const double *w, double *out, double *partial, hipStream_t st)
{
    hipLaunchKernelGGL(k_lincomb_sq, dim3(kRedBlocks), dim3(kThreads), 0, st, n, a, x, b, y, cw, w,
                       out, partial);
}

void vec_fill(int64_t n, double v, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_fillv, dim3(kRedBlocks), dim3(kThreads), 0, st, n, v, out);
}

void vec_dot_partial(int64_t n, const double *a, const double *b, double *partial, hipStream_t st)
{
    hipLaunchKernelGGL(k_dotp, dim3(kRedBlocks), dim3(kThreads), 0, st, n, a, b, partial);
}

void vec_mul(int64_t n, const double *a, const double *b, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_mul, dim3(kRedBlocks), dim3(kThreads), 0, st, n, a, b, out);
}

void vec_extrapolate(int64_t n, double *y, double *yh, hipStream_t st)
{
    hipLaunchKernelGGL(k_extrapolate, dim3(kRedBlocks), dim3(kThreads), 0, st, n, y, yh);
}

void vec_reduce(const double *partial, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(kThreads), 0, st, partial, out);
}

}  // namespace v2g
