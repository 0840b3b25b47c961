// eig.hip -- the per-block positive / negative split of Update_xz
// (v2rdm_solver.cc:3216-3289) on CDNA4, in FP64.
//
// For every block of the primal vector: M = sym(A^T y - c + mu x); M = V L V^T;
// x = V max(L,0)/mu V^T; z = mu x - M  (= -V min(L,0) V^T).
//
// Eigensolver: two-sided block Jacobi.  A block of dimension n is padded to
// N (a multiple of 64, zero rows / columns never couple) and cut into N/32
// column blocks.  Each round pairs the column blocks (circle method); a
// workgroup diagonalizes each 64x64 pair sub-matrix in LDS by parallel
// cyclic Jacobi, and a second kernel applies the pair rotations to A
// (tile-wise Q_P^T A_PR Q_R) and to V.  Sweeps repeat until a full sweep
// applies no rotation above tol = 1e-15 ||M||_F.
//
// Warm start: the eigenvectors of the previous Update_xz stay on the device;
// the next call starts from V^T M V, which is nearly diagonal once the BPSDP
// iteration settles, so most pairs need no rotation and are skipped.
// Blocks with n <= 64 are solved completely by one workgroup (no warm start).
#include <hip/hip_runtime.h>

#include "eig.hpp"

namespace v2g {

namespace {

constexpr int kT = 256;      // threads per workgroup
constexpr int kP = 64;       // pair sub-problem size
constexpr int kLd = 65;      // LDS row pitch (breaks bank conflicts on columns)

__device__ __forceinline__ void circle_pair(int m, int r, int k, int &a, int &b)
{
    // round r of the circle method on m players (m even): k = 0 pairs the
    // fixed player m-1 with r; k >= 1 pairs (r+k) and (r-k) mod (m-1)
    if (k == 0) { a = m - 1; b = r; return; }
    const int mm = m - 1;
    a = (r + k) % mm;
    b = (r - k + mm) % mm;
}

// parallel cyclic Jacobi on the kP x kP symmetric matrix S (pitch kLd) in LDS,
// accumulating rotations into Q.  Returns the number of rotations applied.
__device__ int lds_jacobi(double *S, double *Q, double tol, int maxsweeps, int *sh_cnt,
                          double *sh_c, double *sh_s, int *sh_pa, int *sh_pb)
{
    int total = 0;
    for (int sweep = 0; sweep < maxsweeps; sweep++) {
        if (threadIdx.x == 0) *sh_cnt = 0;
        __syncthreads();
        for (int r = 0; r < kP - 1; r++) {
            if (threadIdx.x < kP / 2) {
                int a, b;
                circle_pair(kP, r, threadIdx.x, a, b);
                const int p = a < b ? a : b, q = a < b ? b : a;
                const double apq = S[p * kLd + q];
                double c = 1.0, s = 0.0;
                if (fabs(apq) > tol) {
                    const double app = S[p * kLd + p], aqq = S[q * kLd + q];
                    const double th = (aqq - app) / (2.0 * apq);
                    const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                    c = 1.0 / sqrt(t * t + 1.0);
                    s = t * c;
                    atomicAdd(sh_cnt, 1);
                }
                sh_c[threadIdx.x] = c;
                sh_s[threadIdx.x] = s;
                sh_pa[threadIdx.x] = p;
                sh_pb[threadIdx.x] = q;
            }
            __syncthreads();
            // S <- R^T S R on every 2x2 (row pair, column pair) block
            for (int w = threadIdx.x; w < (kP / 2) * (kP / 2); w += kT) {
                const int P = w / (kP / 2), R = w % (kP / 2);
                const double cp = sh_c[P], sp = sh_s[P], cr = sh_c[R], sr = sh_s[R];
                if (sp == 0.0 && sr == 0.0) continue;
                const int p0 = sh_pa[P], p1 = sh_pb[P], r0 = sh_pa[R], r1 = sh_pb[R];
                const double b00 = S[p0 * kLd + r0], b01 = S[p0 * kLd + r1];
                const double b10 = S[p1 * kLd + r0], b11 = S[p1 * kLd + r1];
                // rows: R_P^T B
                const double u00 = cp * b00 - sp * b10, u01 = cp * b01 - sp * b11;
                const double u10 = sp * b00 + cp * b10, u11 = sp * b01 + cp * b11;
                // columns: (R_P^T B) R_R
                S[p0 * kLd + r0] = cr * u00 - sr * u01;
                S[p0 * kLd + r1] = sr * u00 + cr * u01;
                S[p1 * kLd + r0] = cr * u10 - sr * u11;
                S[p1 * kLd + r1] = sr * u10 + cr * u11;
            }
            // Q <- Q R on column pairs
            for (int w = threadIdx.x; w < kP * (kP / 2); w += kT) {
                const int row = w / (kP / 2), R = w % (kP / 2);
                const double cr = sh_c[R], sr = sh_s[R];
                if (sr == 0.0) continue;
                const int r0 = sh_pa[R], r1 = sh_pb[R];
                const double v0 = Q[row * kLd + r0], v1 = Q[row * kLd + r1];
                Q[row * kLd + r0] = cr * v0 - sr * v1;
                Q[row * kLd + r1] = sr * v0 + cr * v1;
            }
            __syncthreads();
        }
        const int cnt = *sh_cnt;
        __syncthreads();
        total += cnt;
        if (cnt == 0) break;
    }
    return total;
}

// ---- small blocks: whole problem in one workgroup --------------------------
__global__ void __launch_bounds__(kT)
k_small(const EigBlock *__restrict__ blocks, const double *__restrict__ aty,
        const double *__restrict__ c, double *__restrict__ x, double *__restrict__ z, double mu)
{
    __shared__ double S[kP * kLd];
    __shared__ double Q[kP * kLd];
    __shared__ double Mv[kP * kLd];
    __shared__ double sh_c[kP / 2], sh_s[kP / 2], lam[kP], red[kT / 64];
    __shared__ int sh_pa[kP / 2], sh_pb[kP / 2], sh_cnt;
    const EigBlock B = blocks[blockIdx.x];
    const int n = B.n;
    const int64_t o = B.xoff;
    double fro = 0.0;
    for (int w = threadIdx.x; w < kP * kP; w += kT) {
        const int p = w / kP, q = w % kP;
        double v = 0.0;
        if (p < n && q < n) {
            const int64_t ipq = o + (int64_t)p * n + q, iqp = o + (int64_t)q * n + p;
            v = 0.5 * ((aty[ipq] - c[ipq] + mu * x[ipq]) + (aty[iqp] - c[iqp] + mu * x[iqp]));
        }
        S[p * kLd + q] = v;
        Mv[p * kLd + q] = v;
        Q[p * kLd + q] = (p == q) ? 1.0 : 0.0;
        fro += v * v;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) fro += __shfl_xor(fro, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = fro;
    __syncthreads();
    double tot = 0.0;
    for (int i = 0; i < kT / 64; i++) tot += red[i];
    const double tol = 1e-15 * sqrt(tot) + 1e-300;
    lds_jacobi(S, Q, tol, 60, &sh_cnt, sh_c, sh_s, sh_pa, sh_pb);
    if (threadIdx.x < kP) {
        const double l = S[threadIdx.x * kLd + threadIdx.x];
        lam[threadIdx.x] = l > 0.0 ? l / mu : 0.0;
    }
    __syncthreads();
    for (int w = threadIdx.x; w < n * n; w += kT) {
        const int p = w / n, q = w % n;
        double s = 0.0;
        for (int j = 0; j < kP; j++) s += Q[p * kLd + j] * lam[j] * Q[q * kLd + j];
        x[o + w] = s;
        z[o + w] = mu * s - Mv[p * kLd + q];
    }
}

// ---- large blocks -----------------------------------------------------------
__global__ void __launch_bounds__(kT)
k_build_m(const EigBlock *__restrict__ blocks, const double *__restrict__ aty,
          const double *__restrict__ c, const double *__restrict__ x, double mu,
          double *__restrict__ Mw, double *__restrict__ Aw, double *__restrict__ Vw, int cold)
{
    const EigBlock B = blocks[blockIdx.y];
    const int64_t NN = (int64_t)B.N * B.N;
    const int n = B.n;
    for (int64_t w = blockIdx.x * (int64_t)kT + threadIdx.x; w < NN; w += (int64_t)gridDim.x * kT) {
        const int p = (int)(w / B.N), q = (int)(w % B.N);
        double v = 0.0;
        if (p < n && q < n) {
            const int64_t ipq = B.xoff + (int64_t)p * n + q, iqp = B.xoff + (int64_t)q * n + p;
            v = 0.5 * ((aty[ipq] - c[ipq] + mu * x[ipq]) + (aty[iqp] - c[iqp] + mu * x[iqp]));
        }
        Mw[B.woff + w] = v;
        if (cold) {
            Aw[B.woff + w] = v;
            Vw[B.woff + w] = (p == q) ? 1.0 : 0.0;
        }
    }
}

__global__ void __launch_bounds__(kT)
k_fro(const EigBlock *__restrict__ blocks, const double *__restrict__ Mw, double *__restrict__ tol)
{
    __shared__ double red[kT / 64];
    const EigBlock B = blocks[blockIdx.x];
    const int64_t NN = (int64_t)B.N * B.N;
    double s = 0.0;
    for (int64_t w = threadIdx.x; w < NN; w += kT) {
        const double v = Mw[B.woff + w];
        s += v * v;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kT / 64; i++) t += red[i];
        tol[blockIdx.x] = 2e-14 * sqrt(t) + 1e-300;
    }
}

// index of the 64 pair rows: first column block a (32 indices), then b
__device__ __forceinline__ int pair_index(int a, int b, int i)
{
    return i < 32 ? a * 32 + i : b * 32 + (i - 32);
}

__global__ void __launch_bounds__(kT)
k_pair_eig(const EigBlock *__restrict__ blocks, int round, const double *__restrict__ Aw,
           double *__restrict__ Qbuf, int *__restrict__ active, const double *__restrict__ tol,
           int *__restrict__ counter)
{
    __shared__ double S[kP * kLd];
    __shared__ double Q[kP * kLd];
    __shared__ double sh_c[kP / 2], sh_s[kP / 2];
    __shared__ int sh_pa[kP / 2], sh_pb[kP / 2], sh_cnt;
    const EigBlock B = blocks[blockIdx.y];
    const int nbk = B.N / 32;
    const int k = blockIdx.x;
    if (k >= nbk / 2 || round >= nbk - 1) return;
    int a, b;
    circle_pair(nbk, round, k, a, b);
    const double *A = Aw + B.woff;
    for (int w = threadIdx.x; w < kP * kP; w += kT) {
        const int i = w / kP, j = w % kP;
        S[i * kLd + j] = A[(int64_t)pair_index(a, b, i) * B.N + pair_index(a, b, j)];
        Q[i * kLd + j] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    const int rot = lds_jacobi(S, Q, tol[blockIdx.y], 40, &sh_cnt, sh_c, sh_s, sh_pa, sh_pb);
    double *Qo = Qbuf + B.qoff + (int64_t)k * kP * kP;
    for (int w = threadIdx.x; w < kP * kP; w += kT) Qo[w] = Q[(w / kP) * kLd + (w % kP)];
    if (threadIdx.x == 0) {
        active[B.poff + k] = rot > 0;
        if (rot > 0) atomicAdd(counter, 1);
    }
}

// A_PR <- Q_P^T A_PR Q_R for every pair tile; V[:, P] <- V[:, P] Q_P
__global__ void __launch_bounds__(kT)
k_apply(const EigBlock *__restrict__ blocks, int round, double *__restrict__ Aw,
        double *__restrict__ Vw, const double *__restrict__ Qbuf, const int *__restrict__ active)
{
    __shared__ double T[kP * kLd];
    __shared__ double Qs[kP * kLd];
    const EigBlock B = blocks[blockIdx.y];
    const int nbk = B.N / 32, np = nbk / 2;
    if (round >= nbk - 1) return;
    const int64_t ntA = (int64_t)np * np, ntV = (int64_t)(B.N / kP) * np;
    const int64_t t = blockIdx.x;
    if (t >= ntA + ntV) return;
    const bool isA = t < ntA;
    int P, R, rt = 0;
    if (isA) { P = (int)(t / np); R = (int)(t % np); }
    else { rt = (int)((t - ntA) / np); R = (int)((t - ntA) % np); P = -1; }
    const bool actP = isA && active[B.poff + P];
    const bool actR = active[B.poff + R];
    if (!actP && !actR) return;
    int pa = 0, pb = 0, ra, rb;
    if (isA) circle_pair(nbk, round, P, pa, pb);
    circle_pair(nbk, round, R, ra, rb);
    double *M = isA ? Aw + B.woff : Vw + B.woff;
    auto row_of = [&](int i) -> int64_t { return isA ? pair_index(pa, pb, i) : rt * kP + i; };
    for (int w = threadIdx.x; w < kP * kP; w += kT) {
        const int i = w / kP, j = w % kP;
        T[i * kLd + j] = M[row_of(i) * B.N + pair_index(ra, rb, j)];
    }
    const int ti = (threadIdx.x / 16) * 4, tj = (threadIdx.x % 16) * 4;  // 4x4 outputs
    double acc[4][4];
    if (actP) {
        const double *QP = Qbuf + B.qoff + (int64_t)P * kP * kP;
        for (int w = threadIdx.x; w < kP * kP; w += kT) Qs[(w / kP) * kLd + (w % kP)] = QP[w];
        __syncthreads();
        // acc = Q_P^T T
        for (int u = 0; u < 4; u++)
            for (int v = 0; v < 4; v++) acc[u][v] = 0.0;
        for (int kk = 0; kk < kP; kk++) {
            double qa[4], tb[4];
            for (int u = 0; u < 4; u++) qa[u] = Qs[kk * kLd + ti + u];
            for (int v = 0; v < 4; v++) tb[v] = T[kk * kLd + tj + v];
            for (int u = 0; u < 4; u++)
                for (int v = 0; v < 4; v++) acc[u][v] += qa[u] * tb[v];
        }
        __syncthreads();
        for (int u = 0; u < 4; u++)
            for (int v = 0; v < 4; v++) T[(ti + u) * kLd + tj + v] = acc[u][v];
    }
    if (actR) {
        const double *QR = Qbuf + B.qoff + (int64_t)R * kP * kP;
        __syncthreads();
        for (int w = threadIdx.x; w < kP * kP; w += kT) Qs[(w / kP) * kLd + (w % kP)] = QR[w];
        __syncthreads();
        for (int u = 0; u < 4; u++)
            for (int v = 0; v < 4; v++) acc[u][v] = 0.0;
        for (int kk = 0; kk < kP; kk++) {
            double ta[4], qb[4];
            for (int u = 0; u < 4; u++) ta[u] = T[(ti + u) * kLd + kk];
            for (int v = 0; v < 4; v++) qb[v] = Qs[kk * kLd + tj + v];
            for (int u = 0; u < 4; u++)
                for (int v = 0; v < 4; v++) acc[u][v] += ta[u] * qb[v];
        }
    } else {
        __syncthreads();
        for (int u = 0; u < 4; u++)
            for (int v = 0; v < 4; v++) acc[u][v] = T[(ti + u) * kLd + tj + v];
    }
    for (int u = 0; u < 4; u++)
        for (int v = 0; v < 4; v++)
            M[row_of(ti + u) * B.N + pair_index(ra, rb, tj + v)] = acc[u][v];
}

// T[:, j] = V[:, j] sqrt(max(lambda_j, 0) / mu), lambda_j = A[j][j]
__global__ void __launch_bounds__(kT)
k_scale_vecs(const EigBlock *__restrict__ blocks, const double *__restrict__ Aw,
             const double *__restrict__ Vw, double *__restrict__ Tw, double mu)
{
    const EigBlock B = blocks[blockIdx.y];
    const int64_t NN = (int64_t)B.N * B.N;
    for (int64_t w = blockIdx.x * (int64_t)kT + threadIdx.x; w < NN; w += (int64_t)gridDim.x * kT) {
        const int j = (int)(w % B.N);
        const double l = Aw[B.woff + (int64_t)j * B.N + j];
        Tw[B.woff + w] = l > 0.0 ? Vw[B.woff + w] * sqrt(l / mu) : 0.0;
    }
}

__global__ void __launch_bounds__(kT)
k_z(const EigBlock *__restrict__ blocks, const double *__restrict__ x,
    const double *__restrict__ Mw, double *__restrict__ z, double mu)
{
    const EigBlock B = blocks[blockIdx.y];
    const int n = B.n;
    const int64_t nn = (int64_t)n * n;
    for (int64_t w = blockIdx.x * (int64_t)kT + threadIdx.x; w < nn; w += (int64_t)gridDim.x * kT) {
        const int p = (int)(w / n), q = (int)(w % n);
        z[B.xoff + w] = mu * x[B.xoff + w] - Mw[B.woff + (int64_t)p * B.N + q];
    }
}

}  // namespace

void eig_small(const EigBlock *blocks, int nblocks, const double *aty, const double *c, double *x,
               double *z, double mu, hipStream_t st)
{
    if (nblocks > 0)
        hipLaunchKernelGGL(k_small, dim3(nblocks), dim3(kT), 0, st, blocks, aty, c, x, z, mu);
}

void eig_build_m(const EigBlock *blocks, int nblocks, const double *aty, const double *c,
                 const double *x, double mu, double *Mw, double *Aw, double *Vw, int cold,
                 hipStream_t st)
{
    hipLaunchKernelGGL(k_build_m, dim3(256, nblocks), dim3(kT), 0, st, blocks, aty, c, x, mu, Mw,
                       Aw, Vw, cold);
}

void eig_tol(const EigBlock *blocks, int nblocks, const double *Mw, double *tol, hipStream_t st)
{
    hipLaunchKernelGGL(k_fro, dim3(nblocks), dim3(kT), 0, st, blocks, Mw, tol);
}

void eig_round(const EigBlock *blocks, int nblocks, int maxpairs, int64_t maxtiles, int round,
               double *Aw, double *Vw, double *Qbuf, int *active, const double *tol, int *counter,
               hipStream_t st)
{
    hipLaunchKernelGGL(k_pair_eig, dim3(maxpairs, nblocks), dim3(kT), 0, st, blocks, round, Aw, Qbuf,
                       active, tol, counter);
    hipLaunchKernelGGL(k_apply, dim3((unsigned)maxtiles, nblocks), dim3(kT), 0, st, blocks, round, Aw,
                       Vw, Qbuf, active);
}

void eig_scale(const EigBlock *blocks, int nblocks, const double *Aw, const double *Vw, double *Tw,
               double mu, hipStream_t st)
{
    hipLaunchKernelGGL(k_scale_vecs, dim3(256, nblocks), dim3(kT), 0, st, blocks, Aw, Vw, Tw, mu);
}

void eig_z(const EigBlock *blocks, int nblocks, const double *x, const double *Mw, double *z,
           double mu, hipStream_t st)
{
    hipLaunchKernelGGL(k_z, dim3(256, nblocks), dim3(kT), 0, st, blocks, x, Mw, z, mu);
}

}  // namespace v2g
