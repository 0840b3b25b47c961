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
// (tile-wise Q_P^T A_PR Q_R on FP64 MFMA) and to V.  Sweeps repeat until a
// full sweep applies no rotation above tol = 2e-14 ||M||_F (64 / N times that
// when the split must converge; blocks n <= 64: 1e-15 ||M||_F).
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

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void circle_pair(int m, int r, int k, int &a, int &b)
{
    // round r of the circle method on m players (m even): k = 0 pairs the
    // fixed player m-1 with r; k >= 1 pairs (r+k) and (r-k) mod (m-1)
    if (k == 0) { a = m - 1; b = r; return; }
    const int mm = m - 1;
    a = (r + k) % mm;
    b = (r - k + mm) % mm;
}

// Jacobi rotation (c, s) that annihilates apq.  t = tan(theta), the smaller
// root: the textbook t = sgn(th) / (|th| + sqrt(th^2 + 1)), th = d / e with
// d = aqq - app, e = 2 apq, multiplied through by |e|, gives
// t = sgn(d) e / (|d| + sqrt(d^2 + e^2)): one division instead of three.
// c = 1 / sqrt(1 + t^2) (1 <= 1 + t^2 <= 2) from the hardware reciprocal square
// root and two Newton steps; s = t c.
__device__ __forceinline__ void jacobi_rotation(double app, double aqq, double apq, double &c,
                                                double &s)
{
    const double d = aqq - app, e = 2.0 * apq;
    const double r = sqrt(fma(d, d, e * e));
    double t;
    if (r > 0.0) {
        t = (d >= 0.0 ? e : -e) / (fabs(d) + r);
    } else {   // d^2 + e^2 underflows: the textbook form
        const double th = d / e;
        t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(fma(th, th, 1.0)));
    }
    const double x = fma(t, t, 1.0), hx = 0.5 * x;
    double y = __builtin_amdgcn_rsq(x);
    y = fma(y, fma(-hx * y, y, 0.5), y);
    y = fma(y, fma(-hx * y, y, 0.5), y);
    c = y;
    s = t * y;
}

// R_P^T B R_R for one 2 x 2 block B = [b00 b01; b10 b11] with R = [c s; -s c]
// (rows by (cp, sp), columns by (cr, sr)).  Every product and sum is spelled
// out with fma(), so the rounding does not depend on how the compiler
// schedules a particular kernel: each launch shape of the eigen split makes
// the same rotations bit for bit.
__device__ __forceinline__ void rot2x2(double cp, double sp, double cr, double sr, double b00,
                                       double b01, double b10, double b11, double (&o)[4])
{
    // rows: R_P^T B
    const double u00 = fma(cp, b00, -(sp * b10)), u01 = fma(cp, b01, -(sp * b11));
    const double u10 = fma(sp, b00, cp * b10), u11 = fma(sp, b01, cp * b11);
    // columns: (R_P^T B) R_R
    o[0] = fma(cr, u00, -(sr * u01));
    o[1] = fma(sr, u00, cr * u01);
    o[2] = fma(cr, u10, -(sr * u11));
    o[3] = fma(sr, u10, cr * u11);
}

// parallel cyclic Jacobi on the leading m x m part (m even, m <= kP) of the
// symmetric matrix S (pitch kLd) in LDS, accumulating rotations into Q; rows
// and columns beyond m are zero in S and the identity in Q, and stay so.
// Returns the number of rotations applied.
__device__ int lds_jacobi(double *S, double *Q, int m, double tol, int maxsweeps, int *sh_cnt,
                          double *sh_c, double *sh_s, int *sh_pa, int *sh_pb)
{
    int total = 0;
    const int h = m / 2;
    for (int sweep = 0; sweep < maxsweeps; sweep++) {
        if (threadIdx.x == 0) *sh_cnt = 0;
        __syncthreads();
        for (int r = 0; r < m - 1; r++) {
            if (threadIdx.x < h) {
                int a, b;
                circle_pair(m, r, threadIdx.x, a, b);
                const int p = a < b ? a : b, q = a < b ? b : a;
                const double apq = S[p * kLd + q];
                double c = 1.0, s = 0.0;
                if (fabs(apq) > tol) {
                    const double app = S[p * kLd + p], aqq = S[q * kLd + q];
                    jacobi_rotation(app, aqq, apq, c, s);
                    atomicAdd(sh_cnt, 1);
                }
                sh_c[threadIdx.x] = c;
                sh_s[threadIdx.x] = s;
                sh_pa[threadIdx.x] = p;
                sh_pb[threadIdx.x] = q;
            }
            __syncthreads();
            // S <- R^T S R on every 2x2 (row pair, column pair) block
            for (int w = threadIdx.x; w < h * h; w += kT) {
                const int P = w / h, R = w % h;
                const double cp = sh_c[P], sp = sh_s[P], cr = sh_c[R], sr = sh_s[R];
                if (sp == 0.0 && sr == 0.0) continue;
                const int p0 = sh_pa[P], p1 = sh_pb[P], r0 = sh_pa[R], r1 = sh_pb[R];
                const double b00 = S[p0 * kLd + r0], b01 = S[p0 * kLd + r1];
                const double b10 = S[p1 * kLd + r0], b11 = S[p1 * kLd + r1];
                double o[4];
                rot2x2(cp, sp, cr, sr, b00, b01, b10, b11, o);
                S[p0 * kLd + r0] = o[0];
                S[p0 * kLd + r1] = o[1];
                S[p1 * kLd + r0] = o[2];
                S[p1 * kLd + r1] = o[3];
            }
            // Q <- Q R on column pairs (rows beyond m are zero in these columns)
            for (int w = threadIdx.x; w < m * h; w += kT) {
                const int row = w / h, R = w % h;
                const double cr = sh_c[R], sr = sh_s[R];
                if (sr == 0.0) continue;
                const int r0 = sh_pa[R], r1 = sh_pb[R];
                const double v0 = Q[row * kLd + r0], v1 = Q[row * kLd + r1];
                Q[row * kLd + r0] = fma(cr, v0, -(sr * v1));
                Q[row * kLd + r1] = fma(sr, v0, cr * v1);
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

// The pair sub-problems of the large-block path use a leaner variant:
//  * S keeps only its upper triangle (element (i, j) lives at (min, max)); a
//    step updates the 528 pair blocks (P <= R) instead of all 1024;
//  * Q lives in registers: wave w, lane l holds column l, rows
//    kRowsPerWave w .. kRowsPerWave (w + 1) - 1.
//    A column rotation needs the partner column only, which sits in the
//    partner lane of the same wave, so Q moves through lane shuffles and never
//    touches LDS.  That leaves one 64 x 65 FP64 array of LDS per workgroup.
// In step r of the circle method, lane l is paired with 63 <-> r and
// l <-> (2r - l) mod 63 otherwise; circle_slot(r, l) is the pair's slot (the
// k of circle_pair(kP, r, k)).
// x mod 63 for 0 <= x < 189 without a division
__device__ __forceinline__ int mod63(int x)
{
    x = x >= 2 * (kP - 1) ? x - 2 * (kP - 1) : x;
    return x >= kP - 1 ? x - (kP - 1) : x;
}

__device__ __forceinline__ int circle_partner(int r, int l)
{
    if (l == kP - 1) return r;
    if (l == r) return kP - 1;
    return mod63(2 * r - l + (kP - 1));
}

__device__ __forceinline__ int circle_slot(int r, int l)
{
    if (l == kP - 1 || l == r) return 0;
    const int d = mod63(l - r + (kP - 1));
    return d <= kP / 2 - 1 ? d : mod63(r - l + (kP - 1));
}

constexpr int kUT = (kP / 2) * (kP / 2 + 1) / 2;   // 528 pair blocks with P <= R
// Workgroup size NT of the pair problems.  Large problems have a few hundred
// pair problems per round, so a step's latency sets the time: NT = 1024 (16
// waves, one S block and 4 rows of Q per thread) shortens it.  Many small
// problems at once (bond scans) are throughput-bound: NT = 256 does the same
// work in fewer instructions.
template <int NT> struct PairShape {
    static constexpr int kRowsPerWave = kP * 64 / NT;   // rows of Q a wave holds
    static constexpr int kUTper = (kUT + NT - 1) / NT;  // S blocks per thread
};

// Per step, wave 0 publishes each slot's rotation (sh_c, sh_s) and index pair
// (sh_pa < sh_pb) for the S update, and each lane's own signed rotation
// (sh_cl, sh_sl) for the Q update, so no other wave recomputes circle positions.
// Cross-pair order (bip != 0): step r pairs lane l < 32 (column block a)
// with lane 32 + ((l + r) mod 32) (column block b), so 32 steps rotate every
// pair with one index in each block once and no pair inside a block.  A
// tracking sweep uses it after its first round: that round's cyclic order has
// already rotated the pairs inside every column block (each block takes part
// in every round), so the sweep still visits every pair, in 32 + 31 fewer
// dependent steps per later round.
__device__ __forceinline__ int cross_partner(int r, int l)
{
    return l < kP / 2 ? kP / 2 + ((l + r) & (kP / 2 - 1)) : ((l - kP / 2 - r) & (kP / 2 - 1));
}

template <int NT>
__device__ int pair_jacobi(double *S, double (&q)[PairShape<NT>::kRowsPerWave], double tol,
                           int maxsweeps, const unsigned short *ut, int *sh_cnt, double *sh_c,
                           double *sh_s, int *sh_pa, int *sh_pb, double *sh_cl, double *sh_sl,
                           int bip)
{
    constexpr int kRowsPerWave = PairShape<NT>::kRowsPerWave, kUTper = PairShape<NT>::kUTper;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // this thread's upper-triangle pair blocks (fixed for the whole solve)
    int myP[kUTper], myR[kUTper];
#pragma unroll
    for (int u = 0; u < kUTper; u++) {
        const int w = threadIdx.x + u * NT;
        myP[u] = w < kUT ? (ut[w] & 0xff) : -1;
        myR[u] = w < kUT ? (ut[w] >> 8) : -1;
    }
    int total = 0;
    for (int sweep = 0; sweep < maxsweeps; sweep++) {
        if (threadIdx.x == 0) *sh_cnt = 0;
        __syncthreads();
        const int nsteps = bip ? kP / 2 : kP - 1;
        for (int r = 0; r < nsteps; r++) {
            const int pl = bip ? cross_partner(r, lane) : circle_partner(r, lane);
            if (wave == 0) {
                const int p = lane < pl ? lane : pl, qq = lane < pl ? pl : lane;
                bool rotated = false;
                if (lane == p) {
                    const double apq = S[p * kLd + qq];
                    double c = 1.0, sn = 0.0;
                    if (fabs(apq) > tol) {
                        const double app = S[p * kLd + p], aqq = S[qq * kLd + qq];
                        jacobi_rotation(app, aqq, apq, c, sn);
                        rotated = true;
                    }
                    const int k = bip ? p : circle_slot(r, lane);
                    sh_c[k] = c;
                    sh_s[k] = sn;
                    sh_pa[k] = p;
                    sh_pb[k] = qq;
                    // column p becomes c q_p - s q_qq, column qq c q_qq + s q_p
                    sh_cl[p] = c;
                    sh_sl[p] = -sn;
                    sh_cl[qq] = c;
                    sh_sl[qq] = sn;
                }
                const unsigned long long b = __ballot(rotated);
                if (lane == 0 && b) *sh_cnt += __popcll(b);
            }
            __syncthreads();
            // S <- R^T S R on the upper-triangle pair blocks.  A thread owns up
            // to kUTper blocks; all their LDS reads are issued before any
            // write (the blocks are disjoint), so their latencies overlap.
            {
                int addr[kUTper][4];
                double bv[kUTper][4], cf[kUTper][4];
                bool act[kUTper];
#pragma unroll
                for (int u = 0; u < kUTper; u++) {
                    act[u] = false;
                    if (myP[u] >= 0) {
                        const int P = myP[u], R = myR[u];
                        cf[u][0] = sh_c[P];
                        cf[u][1] = sh_s[P];
                        cf[u][2] = sh_c[R];
                        cf[u][3] = sh_s[R];
                        act[u] = !(cf[u][1] == 0.0 && cf[u][3] == 0.0);
                        const int p0 = sh_pa[P], p1 = sh_pb[P], r0 = sh_pa[R], r1 = sh_pb[R];
                        auto ad = [](int i, int j) { return i < j ? i * kLd + j : j * kLd + i; };
                        addr[u][0] = ad(p0, r0);
                        addr[u][1] = ad(p0, r1);
                        addr[u][2] = ad(p1, r0);
                        addr[u][3] = ad(p1, r1);
                    }
                }
#pragma unroll
                for (int u = 0; u < kUTper; u++)
                    if (act[u]) {
#pragma unroll
                        for (int e = 0; e < 4; e++) bv[u][e] = S[addr[u][e]];
                    }
#pragma unroll
                for (int u = 0; u < kUTper; u++)
                    if (act[u]) {
                        const double cp = cf[u][0], sp = cf[u][1], cr = cf[u][2], sr = cf[u][3];
                        const double b00 = bv[u][0], b01 = bv[u][1], b10 = bv[u][2], b11 = bv[u][3];
                        double o[4];
                        rot2x2(cp, sp, cr, sr, b00, b01, b10, b11, o);
#pragma unroll
                        for (int e = 0; e < 4; e++) S[addr[u][e]] = o[e];
                    }
            }
            // Q <- Q R in registers; every lane takes part in the shuffles
            {
                const double c = sh_cl[lane], sn = sh_sl[lane];
                // all shuffles first, then the arithmetic: one wait for the
                // whole batch instead of one per row
                double o[kRowsPerWave];
#pragma unroll
                for (int i = 0; i < kRowsPerWave; i++) o[i] = __shfl(q[i], pl, 64);
#pragma unroll
                for (int i = 0; i < kRowsPerWave; i++) q[i] = fma(c, q[i], sn * o[i]);
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
        const double *__restrict__ c, double *__restrict__ x, double *__restrict__ z, double mu0,
        EigBatch eb)
{
    __shared__ double S[kP * kLd];
    __shared__ double Q[kP * kLd];
    __shared__ double Mv[kP * kLd];
    __shared__ double sh_c[kP / 2], sh_s[kP / 2], lam[kP], red[kT / 64];
    __shared__ int sh_pa[kP / 2], sh_pb[kP / 2], sh_cnt;
    const EigBlock B = blocks[blockIdx.x];
    if (eb.active && !eb.active[B.prob]) return;
    const double mu = eb.mus ? eb.mus[B.prob] : mu0;
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
    // the circle runs over n rounded up to even indices, not all kP
    const int m = (n + 1) & ~1;
    lds_jacobi(S, Q, m, tol, 60, &sh_cnt, sh_c, sh_s, sh_pa, sh_pb);
    if (threadIdx.x < kP) {
        const double l = S[threadIdx.x * kLd + threadIdx.x];
        lam[threadIdx.x] = l > 0.0 ? l / mu : 0.0;
    }
    __syncthreads();
    for (int w = threadIdx.x; w < n * n; w += kT) {
        const int p = w / n, q = w % n;
        double s = 0.0;
        for (int j = 0; j < m; j++) s += Q[p * kLd + j] * lam[j] * Q[q * kLd + j];
        x[o + w] = s;
        z[o + w] = mu * s - Mv[p * kLd + q];
    }
}

// ---- large blocks -----------------------------------------------------------
// M = sym(A^T y - c + mu x) into the padded N x N workspace, one 64 x 64 tile
// pair (P, R), P <= R, per workgroup: both tiles are read with coalesced rows,
// the transpose goes through LDS, and the (R, P) tile is written as the
// transpose of the (P, R) one.  Cold calls also set A = M, V = I.
__global__ void __launch_bounds__(kT)
k_build_m(const EigBlock *__restrict__ blocks, const double *__restrict__ aty,
          const double *__restrict__ c, const double *__restrict__ x, double mu0,
          double *__restrict__ Mw, double *__restrict__ Aw, double *__restrict__ Vw, int cold,
          EigBatch eb)
{
    __shared__ double S[kP * kLd];
    const EigBlock B = blocks[blockIdx.y];
    if (eb.active && !eb.active[B.prob]) return;
    const double mu = eb.mus ? eb.mus[B.prob] : mu0;
    const int nt = B.N / kP;
    int64_t rem = blockIdx.x;
    if (rem >= (int64_t)nt * (nt + 1) / 2) return;
    int P = 0;
    while (rem >= nt - P) { rem -= nt - P; P++; }
    const int R = P + (int)rem;
    const int n = B.n;
    const int j = threadIdx.x & (kP - 1), i0 = threadIdx.x / kP;
    constexpr int kU = kP * kP / kT;
    auto g = [&](int p, int q) -> double {
        if (p >= n || q >= n) return 0.0;
        const int64_t i = B.xoff + (int64_t)p * n + q;
        return aty[i] - c[i] + mu * x[i];
    };
    // (R, P) tile into LDS, (P, R) tile into registers
    double v[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
        const int i = i0 + u * (kT / kP);
        S[i * kLd + j] = g(R * kP + i, P * kP + j);
        v[u] = g(P * kP + i, R * kP + j);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kU; u++) {
        const int i = i0 + u * (kT / kP);
        v[u] = 0.5 * (v[u] + S[j * kLd + i]);
    }
    auto put = [&](int p, int q, double val) {
        const int64_t w = B.woff + (int64_t)p * B.N + q;
        Mw[w] = val;
        if (cold) {
            Aw[w] = val;
            Vw[w] = (p == q) ? 1.0 : 0.0;
        }
    };
#pragma unroll
    for (int u = 0; u < kU; u++) put(P * kP + i0 + u * (kT / kP), R * kP + j, v[u]);
    if (P == R) return;
    __syncthreads();   // every thread has read S
#pragma unroll
    for (int u = 0; u < kU; u++) S[(i0 + u * (kT / kP)) * kLd + j] = v[u];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kU; u++) {
        const int i = i0 + u * (kT / kP);
        put(R * kP + i, P * kP + j, S[j * kLd + i]);
    }
}

constexpr int kFroParts = 64;

// partial squared Frobenius norms: block y, part x
__global__ void __launch_bounds__(kT)
k_fro_part(const EigBlock *__restrict__ blocks, const double *__restrict__ Mw,
           double *__restrict__ part, EigBatch eb)
{
    __shared__ double red[kT / 64];
    const EigBlock B = blocks[blockIdx.y];
    if (eb.active && !eb.active[B.prob]) return;
    const int64_t NN = (int64_t)B.N * B.N;
    double s = 0.0;
    for (int64_t w = blockIdx.x * (int64_t)kT + threadIdx.x; w < NN; w += (int64_t)kFroParts * kT) {
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
        part[blockIdx.y * kFroParts + blockIdx.x] = t;
    }
}

// rotation threshold of block b: 2e-14 ||M||_F, scaled by 64 / N for a split
// that must converge (its off-diagonal remainder then stays at the level of
// one 64-wide block's however large the block is)
__global__ void k_fro_fin(const EigBlock *__restrict__ blocks, int nblocks,
                          const double *__restrict__ part, double *__restrict__ tol, int tight)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    double t = 0.0;
    for (int i = 0; i < kFroParts; i++) t += part[b * kFroParts + i];
    const double scale = tight ? 64.0 / (double)blocks[b].N : 1.0;
    tol[b] = 2e-14 * scale * sqrt(t) + 1e-300;
}

// index of the 64 pair rows: first column block a (32 indices), then b
__device__ __forceinline__ int pair_index(int a, int b, int i)
{
    return i < 32 ? a * 32 + i : b * 32 + (i - 32);
}

// the column blocks of pair k in round r, lower block first: the rotated
// matrix A is kept as its upper triangle of 32 x 32 sub-blocks (sub-block
// (u, v) with u > v is stale; it is the transpose of (v, u)), and with a < b
// the off-diagonal sub-block of every pair problem is a stored one
__device__ __forceinline__ void ordered_pair(int m, int r, int k, int &a, int &b)
{
    circle_pair(m, r, k, a, b);
    if (a > b) { const int t = a; a = b; b = t; }
}

template <int NT>
__global__ void __launch_bounds__(NT)
k_pair_eig(const EigBlock *__restrict__ blocks, int round, const double *__restrict__ Aw,
           double *__restrict__ Qbuf, int *__restrict__ active, const double *__restrict__ tol,
           int *__restrict__ counter, int inner_sweeps, int bip, EigBatch eb)
{
    __shared__ double S[kP * kLd];
    __shared__ double sh_c[kP / 2], sh_s[kP / 2];
    __shared__ int sh_pa[kP / 2], sh_pb[kP / 2], sh_cnt;
    __shared__ double sh_cl[kP], sh_sl[kP];
    __shared__ unsigned short ut[kUT];
    const EigBlock B = blocks[blockIdx.y];
    const int nbk = B.N / 32;
    const int k = blockIdx.x;
    if (k >= nbk / 2 || round >= nbk - 1) return;
    if (eb.active && !eb.active[B.prob]) {   // a problem left out: no rotation
        if (threadIdx.x == 0) active[B.poff + k] = 0;
        return;
    }
    int a, b;
    ordered_pair(nbk, round, k, a, b);
    const double *A = Aw + B.woff;
    const double tl = tol[blockIdx.y];
    if (threadIdx.x == 0) sh_cnt = 0;
    for (int w = threadIdx.x; w < kUT; w += NT) {
        int P = 0, rem = w;
        while (rem >= kP / 2 - P) { rem -= kP / 2 - P; P++; }
        ut[w] = (unsigned short)(P | ((P + rem) << 8));
    }
    __syncthreads();
    int big = 0;
    {
        // thread (i0, j) loads rows i0, i0 + NT / kP, ... of column j, all
        // loads in flight at once
        // the upper triangle of the pair matrix is all pair_jacobi reads: its
        // (b, a) sub-block is stale in A and is not loaded
        const int j = threadIdx.x & (kP - 1), i0 = threadIdx.x / kP;
        const int64_t colj = pair_index(a, b, j);
        double v[kP * kP / NT];
#pragma unroll
        for (int u = 0; u < kP * kP / NT; u++) {
            const int i = i0 + u * (NT / kP);
            v[u] = (i >= 32 && j < 32) ? 0.0 : A[(int64_t)pair_index(a, b, i) * B.N + colj];
        }
#pragma unroll
        for (int u = 0; u < kP * kP / NT; u++) {
            const int i = i0 + u * (NT / kP);
            S[i * kLd + j] = v[u];
            big |= (i < j && fabs(v[u]) > tl);
        }
    }
    if (big) atomicOr(&sh_cnt, 1);
    __syncthreads();
    const int any_big = sh_cnt;  // uniform across the workgroup
    __syncthreads();             // every thread has read sh_cnt before it is reused
    if (any_big == 0) {          // already diagonal to tolerance: identity rotation
        if (threadIdx.x == 0) active[B.poff + k] = 0;
        return;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int kRowsPerWave = PairShape<NT>::kRowsPerWave;
    double q[kRowsPerWave];
#pragma unroll
    for (int i = 0; i < kRowsPerWave; i++) q[i] = (wave * kRowsPerWave + i == lane) ? 1.0 : 0.0;
    const int rot = pair_jacobi<NT>(S, q, tl, inner_sweeps, ut, &sh_cnt, sh_c, sh_s, sh_pa, sh_pb,
                                    sh_cl, sh_sl, bip);
    double *Qo = Qbuf + B.qoff + (int64_t)k * kP * kP;
#pragma unroll
    for (int i = 0; i < kRowsPerWave; i++) Qo[(wave * kRowsPerWave + i) * kP + lane] = q[i];
    if (threadIdx.x == 0) {
        active[B.poff + k] = rot > 0;
        if (rot > 0) atomicAdd(counter, 1);
    }
}

// A_PR <- Q_P^T A_PR Q_R for the pair tiles with P <= R (A is symmetric and
// kept as its upper triangle of 32 x 32 sub-blocks); V[:, R] <- V[:, R] Q_R.
// Four waves per SIMD: the LDS tile allows four workgroups per CU, and the
// bound keeps the registers within that (122 VGPRs).  FP64 MFMA
// v_mfma_f64_16x16x4f64: A[i][k] from lane i + 16k, B[k][j] from lane
// j + 16k, D[i][j] in lane j + 16 (i % 4), element i / 4.
__global__ void __launch_bounds__(kT, 4)
k_apply(const EigBlock *__restrict__ blocks, int round, double *__restrict__ Aw,
        double *__restrict__ Vw, const double *__restrict__ Qbuf, const int *__restrict__ active,
        EigBatch eb)
{
    // only the tile lives in LDS; the pair rotations Q_P, Q_R are read from
    // global memory (L2-resident: each one serves a row of tiles), which
    // halves the LDS footprint and doubles the workgroups per CU
    __shared__ double T[kP * kLd];
    const EigBlock B = blocks[blockIdx.y];
    const int nbk = B.N / 32, np = nbk / 2;
    if (round >= nbk - 1) return;
    if (eb.active && !eb.active[B.prob]) return;
    const int64_t ntA = (int64_t)np * (np + 1) / 2, ntV = (int64_t)(B.N / kP) * np;
    const int64_t t = blockIdx.x;
    if (t >= ntA + ntV) return;
    const bool isA = t < ntA;
    int P = -1, R, rt = 0;
    if (isA) {
        // t -> (P, R) with P <= R, row-major over the upper triangle
        int64_t rem = t;
        P = 0;
        while (rem >= np - P) { rem -= np - P; P++; }
        R = P + (int)rem;
    } else {
        rt = (int)((t - ntA) / np);
        R = (int)((t - ntA) % np);
    }
    const bool actP = isA && active[B.poff + P];
    const bool actR = active[B.poff + R];
    if (!actP && !actR) return;
    int pa = 0, pb = 0, ra, rb;
    if (isA) ordered_pair(nbk, round, P, pa, pb);
    ordered_pair(nbk, round, R, ra, rb);
    double *M = isA ? Aw + B.woff : Vw + B.woff;
    const int j = threadIdx.x & (kP - 1), i0 = threadIdx.x / kP;
    const int64_t colj = pair_index(ra, rb, j);
    // A tiles move as four 32 x 32 sub-blocks (u, v), u in {pa, pb}, v in
    // {ra, rb}: a stored one (u <= v) directly, a stale one (u > v) as the
    // transpose of (v, u); thread t moves stored rows t / 32 + 8 k, column
    // t % 32 of each (coalesced 256-byte rows either way).  Each unordered
    // pair of column blocks belongs to one tile of the round, so every
    // stored sub-block is written once; in a diagonal tile (P == R) the
    // (pb, pa) sub-block duplicates (pa, pb) and is not written back.
    const int sr0 = threadIdx.x >> 5, sc = threadIdx.x & 31;
    auto sub_addr = [&](int ub, int vb, int k, int &ti, int &tj) -> int64_t {
        const int u = ub ? pb : pa, v = vb ? rb : ra, sr = sr0 + 8 * k;
        if (u <= v) {
            ti = ub * 32 + sr;
            tj = vb * 32 + sc;
            return (int64_t)(u * 32 + sr) * B.N + v * 32 + sc;
        }
        ti = ub * 32 + sc;
        tj = vb * 32 + sr;
        return (int64_t)(v * 32 + sr) * B.N + u * 32 + sc;
    };
    if (isA) {
        double v[16];
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                int ti, tj;
                v[s * 4 + k] = M[sub_addr(s >> 1, s & 1, k, ti, tj)];
            }
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                int ti, tj;
                (void)sub_addr(s >> 1, s & 1, k, ti, tj);
                T[ti * kLd + tj] = v[s * 4 + k];
            }
    } else {
        // V tiles: thread (i0, j) moves rows i0, i0 + 4, ... of column j, all
        // 16 loads issued before the first LDS write
        double v[kP * kP / kT];
#pragma unroll
        for (int u = 0; u < kP * kP / kT; u++)
            v[u] = M[(int64_t)(rt * kP + i0 + u * (kT / kP)) * B.N + colj];
#pragma unroll
        for (int u = 0; u < kP * kP / kT; u++) T[(i0 + u * (kT / kP)) * kLd + j] = v[u];
    }
    // The tile passes through LDS once.  Wave w forms the 64 x 16 column strip
    // w of U = T Q_R (A operand T from LDS, B operand Q_R from global memory)
    // in its MFMA accumulators: accumulator register r of row block bi holds
    // U[16 bi + 4 r + lane / 16][16 w + lane % 16], which is exactly the B
    // operand layout of the k-step 16 bi + 4 r of the next product, so the
    // strip of Q_P^T U is formed from registers.  Stored sub-blocks are then
    // written from registers (four 128-byte row pieces per store); stale
    // (transposed) ones go through LDS for coalesced writes.
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int jc = 16 * wave + li;
    v4d u[4];
    if (actR) {
        const double *QR = Qbuf + B.qoff + (int64_t)R * kP * kP;
#pragma unroll
        for (int bi = 0; bi < 4; bi++) u[bi] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
        for (int k0 = 0; k0 < kP; k0 += 4) {
            const double bv = QR[(k0 + lk) * kP + jc];
#pragma unroll
            for (int bi = 0; bi < 4; bi++) {
                const double av = T[(16 * bi + li) * kLd + k0 + lk];
                u[bi] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, u[bi], 0, 0, 0);
            }
        }
    } else {
#pragma unroll
        for (int bi = 0; bi < 4; bi++)
#pragma unroll
            for (int r = 0; r < 4; r++) u[bi][r] = T[(16 * bi + 4 * r + lk) * kLd + jc];
    }
    v4d o[4];
    if (actP) {
        const double *QP = Qbuf + B.qoff + (int64_t)P * kP * kP;
#pragma unroll
        for (int bi = 0; bi < 4; bi++) o[bi] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kb = 0; kb < 4; kb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int k = 16 * kb + 4 * r + lk;
                const double bv = u[kb][r];
#pragma unroll
                for (int bi = 0; bi < 4; bi++) {
                    const double av = QP[k * kP + 16 * bi + li];   // (Q_P^T)[row][k]
                    o[bi] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, o[bi], 0, 0, 0);
                }
            }
    } else {
#pragma unroll
        for (int bi = 0; bi < 4; bi++) o[bi] = u[bi];
    }
    if (!isA) {
        const int64_t col = pair_index(ra, rb, jc);
#pragma unroll
        for (int bi = 0; bi < 4; bi++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                M[(int64_t)(rt * kP + 16 * bi + 4 * r + lk) * B.N + col] = o[bi][r];
        return;
    }
    __syncthreads();   // every wave has read T
    const int vb = wave >> 1, vv = vb ? rb : ra;
#pragma unroll
    for (int ub = 0; ub < 2; ub++) {
        const int uu = ub ? pb : pa;
        if (P == R && ub == 1 && vb == 0) continue;   // (pb, pa) of a diagonal tile
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int bi = 2 * ub + h;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = 16 * bi + 4 * r + lk;
                if (uu <= vv)
                    M[(int64_t)(uu * 32 + i - 32 * ub) * B.N + vv * 32 + jc - 32 * vb] = o[bi][r];
                else
                    T[i * kLd + jc] = o[bi][r];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const int ub = s >> 1, vb2 = s & 1;
        const int uu = ub ? pb : pa, v2 = vb2 ? rb : ra;
        if (uu <= v2 || (P == R && s == 2)) continue;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int sr = sr0 + 8 * k;
            M[(int64_t)(v2 * 32 + sr) * B.N + uu * 32 + sc] = T[(ub * 32 + sc) * kLd + vb2 * 32 + sr];
        }
    }
}
// T[:, j] = V[:, j] sqrt(max(lambda_j, 0) / mu), lambda_j = A[j][j]
__global__ void __launch_bounds__(kT)
k_scale_vecs(const EigBlock *__restrict__ blocks, const double *__restrict__ Aw,
             const double *__restrict__ Vw, double *__restrict__ Tw, double mu0, EigBatch eb)
{
    const EigBlock B = blocks[blockIdx.y];
    if (eb.active && !eb.active[B.prob]) return;
    const double mu = eb.mus ? eb.mus[B.prob] : mu0;
    const int64_t NN = (int64_t)B.N * B.N;
    for (int64_t w = blockIdx.x * (int64_t)kT + threadIdx.x; w < NN; w += (int64_t)gridDim.x * kT) {
        const int j = (int)(w % B.N);
        const double l = Aw[B.woff + (int64_t)j * B.N + j];
        Tw[B.woff + w] = l > 0.0 ? Vw[B.woff + w] * sqrt(l / mu) : 0.0;
    }
}

__global__ void __launch_bounds__(kT)
k_z(const EigBlock *__restrict__ blocks, const double *__restrict__ x,
    const double *__restrict__ Mw, double *__restrict__ z, double mu0, EigBatch eb)
{
    const EigBlock B = blocks[blockIdx.y];
    if (eb.active && !eb.active[B.prob]) return;
    const double mu = eb.mus ? eb.mus[B.prob] : mu0;
    const int n = B.n;
    const int64_t nn = (int64_t)n * n;
    for (int64_t w = blockIdx.x * (int64_t)kT + threadIdx.x; w < nn; w += (int64_t)gridDim.x * kT) {
        const int p = (int)(w / n), q = (int)(w % n);
        z[B.xoff + w] = mu * x[B.xoff + w] - Mw[B.woff + (int64_t)p * B.N + q];
    }
}

}  // namespace

void eig_small(const EigBlock *blocks, int nblocks, const double *aty, const double *c, double *x,
               double *z, double mu, hipStream_t st, EigBatch eb)
{
    if (nblocks > 0)
        hipLaunchKernelGGL(k_small, dim3(nblocks), dim3(kT), 0, st, blocks, aty, c, x, z, mu, eb);
}

void eig_build_m(const EigBlock *blocks, int nblocks, int max_N, const double *aty, const double *c,
                 const double *x, double mu, double *Mw, double *Aw, double *Vw, int cold,
                 hipStream_t st, EigBatch eb)
{
    const int nt = max_N / kP;
    hipLaunchKernelGGL(k_build_m, dim3(nt * (nt + 1) / 2, nblocks), dim3(kT), 0, st, blocks, aty, c,
                       x, mu, Mw, Aw, Vw, cold, eb);
}

void eig_tol(const EigBlock *blocks, int nblocks, const double *Mw, double *tol, double *part,
             int tight, hipStream_t st, EigBatch eb)
{
    hipLaunchKernelGGL(k_fro_part, dim3(kFroParts, nblocks), dim3(kT), 0, st, blocks, Mw, part, eb);
    hipLaunchKernelGGL(k_fro_fin, dim3((nblocks + 63) / 64), dim3(64), 0, st, blocks, nblocks, part,
                       tol, tight);
}

void eig_pairs(const EigBlock *blocks, int nblocks, int maxpairs, int round, const double *Aw,
               double *Qbuf, int *active, const double *tol, int *counter, int inner_sweeps,
               int wide, int bip, hipStream_t st, EigBatch eb)
{
    if (wide)
        hipLaunchKernelGGL(k_pair_eig<1024>, dim3(maxpairs, nblocks), dim3(1024), 0, st, blocks, round,
                           Aw, Qbuf, active, tol, counter, inner_sweeps, bip, eb);
    else
        hipLaunchKernelGGL(k_pair_eig<256>, dim3(maxpairs, nblocks), dim3(256), 0, st, blocks, round,
                           Aw, Qbuf, active, tol, counter, inner_sweeps, bip, eb);
}

void eig_apply(const EigBlock *blocks, int nblocks, int64_t maxtiles, int round, double *Aw,
               double *Vw, const double *Qbuf, const int *active, hipStream_t st, EigBatch eb)
{
    hipLaunchKernelGGL(k_apply, dim3((unsigned)maxtiles, nblocks), dim3(kT), 0, st, blocks, round, Aw,
                       Vw, Qbuf, active, eb);
}

void eig_scale(const EigBlock *blocks, int nblocks, const double *Aw, const double *Vw, double *Tw,
               double mu, hipStream_t st, EigBatch eb)
{
    hipLaunchKernelGGL(k_scale_vecs, dim3(256, nblocks), dim3(kT), 0, st, blocks, Aw, Vw, Tw, mu, eb);
}

void eig_z(const EigBlock *blocks, int nblocks, const double *x, const double *Mw, double *z,
           double mu, hipStream_t st, EigBatch eb)
{
    hipLaunchKernelGGL(k_z, dim3(256, nblocks), dim3(kT), 0, st, blocks, x, Mw, z, mu, eb);
}

}  // namespace v2g
