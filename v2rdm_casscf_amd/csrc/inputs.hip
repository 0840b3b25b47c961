// inputs.hip -- the BPSDP path inputs built on the device (for batched runs,
// where host loops and uploads of b, c and the guess would dominate):
//
//   b     BuildConstraints (v2rdm_solver.cc:2518-3017): zero except the
//         trace rows, the D1 + Q1 = I diagonal, the S^2 target and the
//         -delta diagonal of the Q2 rows;
//   c     RepackIntegrals without core orbitals (tei.cc:144-228): (ik|jl) on
//         the D2ab blocks, (ik|jl) - (il|jk) on D2aa / D2bb, h on D1a / D1b;
//   guess TPDM_GUESS RANDOM (v2rdm_solver.cc:2479-2490): srand(0), then
//         rand() for x, z (interleaved) and y, each ((double) r / RAND_MAX
//         - 1) * 2.  glibc's rand() is the additive lagged-Fibonacci
//         generator r_i = r_(i-31) + r_(i-3) (mod 2^32), output r_i >> 1,
//         seeded by the Park-Miller sequence and with its first 310 outputs
//         discarded (glibc random_r.c, TYPE_3).  The recurrence is linear in
//         its last 31 words, so the state at the start of every chunk of
//         kGuessChunk outputs follows from a power of the 31 x 31 step matrix
//         (computed on the host, mod 2^32); each thread then runs one chunk.
//
// Every value equals the host version's bit for bit (same expressions).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "devlayout.hpp"
#include "inputs.hpp"

namespace v2g {

namespace {

constexpr int kT = 256;

__device__ __forceinline__ int find_start(const Segment *__restrict__ segs, int nseg, int64_t r)
{
    int lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (segs[mid].start <= r) lo = mid;
        else hi = mid;
    }
    return lo;
}

// b[r] from the row's family (segments sorted by start; rows of a segment
// are numbered in the family's own order, as the maps number them)
__global__ void __launch_bounds__(kT)
k_build_b(const DevLayout *__restrict__ Lp, const Segment *__restrict__ segs, int nseg, int64_t ncon,
          double *__restrict__ b)
{
    for (int64_t r = blockIdx.x * (int64_t)kT + threadIdx.x; r < ncon; r += (int64_t)gridDim.x * kT) {
        const Segment sg = segs[find_start(segs, nseg, r)];
        const DevLayout &L = Lp[sg.prob];
        const int64_t e = r - sg.start;
        double v = 0.0;
        switch (sg.kind) {
        case RF_TRACE:       // v2rdm_solver.cc:2530-2560
            if (sg.sub == 0) v = L.na * L.nb;
            else if (sg.sub == 1) v = L.na * (L.na - 1.0);
            else if (sg.sub == 2) v = L.nb * (L.nb - 1.0);
            else v = (L.na + L.nb) * (L.na + L.nb - 1.0);
            break;
        case RF_D1Q1: {      // :2562-2580, D1 + Q1 = I
            const int64_t n = sg.dim;
            v = (e / n == e % n) ? 1.0 : 0.0;
            break;
        }
        case RF_S2:          // :2630, <S^2> target
            v = 0.5 * (L.na + L.nb) + L.ms * L.ms - L.ms * (L.ms + 1.0);
            break;
        case RF_Q2AB:
        case RF_Q2AA: {      // :2669-2709, -delta_ik delta_jl on the diagonal
            const int64_t g = sg.dim;
            v = (e / g == e % g) ? -1.0 : 0.0;
            break;
        }
        default: break;
        }
        b[r] = v;
    }
}

__global__ void __launch_bounds__(kT) k_fill(int64_t n, double v, double *__restrict__ out)
{
    for (int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
        out[i] = v;
}

// c on the D2ab (blocks = 0), D2aa and D2bb (blocks = 1) entries of irrep
// blockIdx.y; eri is amo^4 in chemists' notation, row-major
__global__ void __launch_bounds__(kT)
k_build_c_d2(const DevLayout *__restrict__ Lp, const double *__restrict__ eri, int aa,
             double *__restrict__ c)
{
    const DevLayout &L = *Lp;
    const int h = blockIdx.y;
    const int64_t a = L.amo;
    const int64_t g = aa ? L.gems_aa[h] : L.gems_ab[h];
    const int *bp = aa ? L.baa + L.baa_off[h] : L.bab + L.bab_off[h];
    const int64_t gg = g * g;
    auto E = [&](int64_t p, int64_t q, int64_t r, int64_t s) { return eri[((p * a + q) * a + r) * a + s]; };
    for (int64_t w = blockIdx.x * (int64_t)kT + threadIdx.x; w < gg; w += (int64_t)gridDim.x * kT) {
        const int64_t ij = w / g, kl = w % g;
        const int i = bp[ij] & 0xffff, j = bp[ij] >> 16, k = bp[kl] & 0xffff, l = bp[kl] >> 16;
        if (!aa) {
            c[L.xoff[XB_D2AB][h] + w] = E(i, k, j, l);
        } else {
            const double v = E(i, k, j, l) - E(i, l, j, k);
            c[L.xoff[XB_D2AA][h] + w] = v;
            c[L.xoff[XB_D2BB][h] + w] = v;
        }
    }
}

__global__ void __launch_bounds__(kT)
k_build_c_d1(const DevLayout *__restrict__ Lp, const double *__restrict__ h1, double *__restrict__ c)
{
    const DevLayout &L = *Lp;
    const int h = blockIdx.y;
    const int64_t a = L.amo, m = L.amopi[h];
    for (int64_t w = blockIdx.x * (int64_t)kT + threadIdx.x; w < m * m; w += (int64_t)gridDim.x * kT) {
        const int64_t i = w / m, j = w % m;
        const double v = h1[(i + L.poff[h]) * a + (j + L.poff[h])];
        c[L.xoff[XB_D1A][h] + w] = v;
        c[L.xoff[XB_D1B][h] + w] = v;
    }
}

// one chunk of kGuessChunk glibc outputs per thread from its 31-word start
// state (oldest first); output k of the whole sequence goes to x[k/2] (k even)
// or z[k/2] (k odd) for k < 2 dimx, else to y[k - 2 dimx]
__global__ void __launch_bounds__(64)
k_guess(const uint32_t *__restrict__ starts, int64_t nchunks, int64_t total, int64_t dimx,
        double *__restrict__ x, double *__restrict__ z, double *__restrict__ y)
{
    const int64_t ch = blockIdx.x * (int64_t)64 + threadIdx.x;
    if (ch >= nchunks) return;
    uint32_t s[31];
#pragma unroll
    for (int i = 0; i < 31; i++) s[i] = starts[ch * 31 + i];
    const int64_t k0 = ch * kGuessChunk;
    const int64_t k1 = (k0 + kGuessChunk < total) ? k0 + kGuessChunk : total;
    auto put = [&](int64_t k, uint32_t r) {
        const double v = ((double)(r >> 1) / 2147483647.0 - 1.0) * 2.0;
        if (k < 2 * dimx) {
            if (k & 1) z[k >> 1] = v;
            else x[k >> 1] = v;
        } else {
            y[k - 2 * dimx] = v;
        }
    };
    // s[i] holds r_(t - 31 + i) before step t; unrolled by 31 so that every
    // register index is a constant: step u of a round overwrites slot u
    int64_t k = k0;
    while (k < k1) {
#pragma unroll
        for (int u = 0; u < 31; u++) {
            if (k < k1) {
                const uint32_t r = s[u] + s[(u + 28) % 31];
                s[u] = r;
                put(k, r);
                k++;
            }
        }
    }
}

// 31 x 31 matrices over Z / 2^32 for the step s -> s' of the recurrence,
// s = (r_(t-31), ..., r_(t-1)), r_t = r_(t-31) + r_(t-3)
using Mat = std::vector<uint32_t>;

Mat mat_mul(const Mat &A, const Mat &B)
{
    Mat C(31 * 31, 0u);
    for (int i = 0; i < 31; i++)
        for (int k = 0; k < 31; k++) {
            const uint32_t a = A[i * 31 + k];
            if (!a) continue;
            for (int j = 0; j < 31; j++) C[i * 31 + j] += a * B[k * 31 + j];
        }
    return C;
}

}  // namespace

void glibc_rand_state_after_srand(unsigned seed, uint32_t (&state)[31])
{
    // glibc __srandom_r (TYPE_3): seed 0 acts as 1; r_i = 16807 r_(i-1) mod
    // (2^31 - 1) by Schrage's method for i < 31; r_31..r_33 = r_0..r_2;
    // outputs start at r_344 (310 discarded after the 34 seeding words)
    if (seed == 0) seed = 1;
    std::vector<uint32_t> r(344 + 31);
    r[0] = seed;
    int64_t word = seed;
    for (int i = 1; i < 31; i++) {
        const int64_t hi = word / 127773, lo = word % 127773;
        word = 16807 * lo - 2836 * hi;
        if (word < 0) word += 2147483647;
        r[i] = (uint32_t)word;
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    for (int i = 34; i < 344; i++) r[i] = r[i - 31] + r[i - 3];
    for (int i = 0; i < 31; i++) state[i] = r[344 - 31 + i];
}

void build_b_device(const DevLayout *dlay, const Segment *rowsegs, int nseg, int64_t ncon, double *b,
                    hipStream_t st)
{
    if (ncon <= 0) return;
    hipLaunchKernelGGL(k_build_b, dim3(2048), dim3(kT), 0, st, dlay, rowsegs, nseg, ncon, b);
}

void build_c_device(const DevLayout *dlay, int nirrep, int max_gab, int max_gaa, int max_m,
                    int64_t dimx, const double *h1, const double *eri, double *c, hipStream_t st)
{
    if (dimx > 0) hipLaunchKernelGGL(k_fill, dim3(2048), dim3(kT), 0, st, dimx, 0.0, c);
    auto grid = [](int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, (n + kT - 1) / kT)); };
    if (max_gab > 0)
        hipLaunchKernelGGL(k_build_c_d2, dim3(grid((int64_t)max_gab * max_gab), nirrep), dim3(kT), 0, st,
                           dlay, eri, 0, c);
    if (max_gaa > 0)
        hipLaunchKernelGGL(k_build_c_d2, dim3(grid((int64_t)max_gaa * max_gaa), nirrep), dim3(kT), 0, st,
                           dlay, eri, 1, c);
    if (max_m > 0)
        hipLaunchKernelGGL(k_build_c_d1, dim3(grid((int64_t)max_m * max_m), nirrep), dim3(kT), 0, st, dlay,
                           h1, c);
}

std::vector<uint32_t> guess_chunk_starts(int64_t total)
{
    uint32_t s0[31];
    glibc_rand_state_after_srand(0, s0);
    // step matrix T: s' = T s
    Mat T(31 * 31, 0u);
    for (int i = 0; i < 30; i++) T[i * 31 + i + 1] = 1u;
    T[30 * 31 + 0] = 1u;
    T[30 * 31 + 28] = 1u;
    // T^kGuessChunk by squaring (kGuessChunk is a power of two)
    Mat P = T;
    for (int64_t e = 1; e < kGuessChunk; e *= 2) P = mat_mul(P, P);
    const int64_t nchunks = (total + kGuessChunk - 1) / kGuessChunk;
    std::vector<uint32_t> starts((size_t)nchunks * 31);
    std::vector<uint32_t> cur(s0, s0 + 31), nxt(31);
    for (int64_t c = 0; c < nchunks; c++) {
        std::copy(cur.begin(), cur.end(), starts.begin() + c * 31);
        for (int i = 0; i < 31; i++) {
            uint32_t acc = 0u;
            for (int j = 0; j < 31; j++) acc += P[i * 31 + j] * cur[j];
            nxt[i] = acc;
        }
        cur.swap(nxt);
    }
    return starts;
}

void guess_device(const uint32_t *starts, int64_t total, int64_t dimx, double *x, double *z, double *y,
                  hipStream_t st)
{
    const int64_t nchunks = (total + kGuessChunk - 1) / kGuessChunk;
    if (nchunks <= 0) return;
    hipLaunchKernelGGL(k_guess, dim3((unsigned)((nchunks + 63) / 64)), dim3(64), 0, st, starts, nchunks,
                       total, dimx, x, z, y);
}

}  // namespace v2g
