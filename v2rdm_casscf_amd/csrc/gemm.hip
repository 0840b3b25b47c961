// gemm.hip -- batched FP64 GEMM C = op(A) op(B) for the eigen split
// (warm-start transform V^T M V and the reconstruction x = W W^T).
// 64 x 64 output tile per workgroup, 32-deep K panels staged in LDS, the next
// panel prefetched into registers while the current one is multiplied; each
// wave computes a 32 x 32 sub-tile with v_mfma_f64_16x16x4f64 (A[i][k] from
// lane i + 16k, B[k][j] from lane j + 16k, D[i][j] in lane j + 16 (i % 4),
// element i / 4).  A descriptor with sym = 1 declares C symmetric: only the
// tiles on and above the diagonal are computed, and each is also written
// transposed.
#include <hip/hip_runtime.h>

#include "eig.hpp"

namespace v2g {

namespace {

constexpr int kT = 256, kTile = 64, kK = 32;
constexpr int kLoads = kK * kTile / kT;   // elements of each panel one thread moves
typedef double v4d __attribute__((ext_vector_type(4)));

// four waves per SIMD (the LDS panels allow four workgroups per CU; 126 VGPRs)
__global__ void __launch_bounds__(kT, 4)
k_gemm(const GemmDesc *__restrict__ descs)
{
    __shared__ double As[kK][kTile + 1];
    __shared__ double Bs[kK][kTile + 1];
    const GemmDesc d = descs[blockIdx.y];
    const int tiles_n = (d.N + kTile - 1) / kTile;
    const int tiles_m = (d.M + kTile - 1) / kTile;
    if ((int)blockIdx.x >= tiles_m * tiles_n) return;
    const int tm = blockIdx.x / tiles_n, tn = blockIdx.x % tiles_n;
    if (d.sym && tm > tn) return;
    const int m0 = tm * kTile, n0 = tn * kTile;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32, li = lane & 15, lk = lane >> 4;

    // element w of a panel: A panel (kK x kTile, As[kk][r]) and B panel
    // (kK x kTile, Bs[kk][c]); the walk makes consecutive threads read
    // consecutive memory for either orientation of the operand
    double ra[kLoads], rb[kLoads];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < kLoads; u++) {
            const int w = threadIdx.x + u * kT;
            int kk, r;
            if (d.ta) { r = w % kTile; kk = w / kTile; } else { kk = w % kK; r = w / kK; }
            const int gm = m0 + r, gk = k0 + kk;
            ra[u] = (gm < d.M && gk < d.K)
                        ? (d.ta ? d.A[(int64_t)gk * d.lda + gm] : d.A[(int64_t)gm * d.lda + gk])
                        : 0.0;
            int kk2, c2;
            if (d.tb) { kk2 = w % kK; c2 = w / kK; } else { c2 = w % kTile; kk2 = w / kTile; }
            const int gn = n0 + c2, gk2 = k0 + kk2;
            rb[u] = (gn < d.N && gk2 < d.K)
                        ? (d.tb ? d.B[(int64_t)gn * d.ldb + gk2] : d.B[(int64_t)gk2 * d.ldb + gn])
                        : 0.0;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int u = 0; u < kLoads; u++) {
            const int w = threadIdx.x + u * kT;
            int kk, r;
            if (d.ta) { r = w % kTile; kk = w / kTile; } else { kk = w % kK; r = w / kK; }
            As[kk][r] = ra[u];
            int kk2, c2;
            if (d.tb) { kk2 = w % kK; c2 = w / kK; } else { c2 = w % kTile; kk2 = w / kTile; }
            Bs[kk2][c2] = rb[u];
        }
    };

    v4d acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
    fetch(0);
    for (int k0 = 0; k0 < d.K; k0 += kK) {
        stash();
        __syncthreads();
        if (k0 + kK < d.K) fetch(k0 + kK);   // next panel in flight during the MFMAs
#pragma unroll
        for (int k4 = 0; k4 < kK; k4 += 4) {
            const int kk = k4 + lk;
            double av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; a++) av[a] = As[kk][wr + a * 16 + li];
#pragma unroll
            for (int b = 0; b < 2; b++) bv[b] = Bs[kk][wc + b * 16 + li];
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int r = 0; r < 4; r++) {
                const int gm = m0 + wr + a * 16 + 4 * r + lk, gn = n0 + wc + b * 16 + li;
                if (gm < d.M && gn < d.N) {
                    d.C[(int64_t)gm * d.ldc + gn] = acc[a][b][r];
                    if (d.sym && tm != tn) d.C[(int64_t)gn * d.ldc + gm] = acc[a][b][r];
                }
            }
}

}  // namespace

void gemm_batched(const GemmDesc *descs, int ndesc, int max_m, int max_n, hipStream_t st)
{
    if (ndesc <= 0) return;
    const int tiles = ((max_m + kTile - 1) / kTile) * ((max_n + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_gemm, dim3(tiles, ndesc), dim3(kT), 0, st, descs);
}

}  // namespace v2g
