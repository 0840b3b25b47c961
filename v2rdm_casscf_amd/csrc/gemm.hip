// gemm.hip -- batched FP64 GEMM C = op(A) op(B) for the eigen split
// (warm-start transform V^T M V and the reconstruction x = W W^T).
// 64 x 64 output tile per workgroup, 16-deep K panels staged in LDS,
// 4 x 4 outputs per thread.
#include <hip/hip_runtime.h>

#include "eig.hpp"

namespace v2g {

namespace {

constexpr int kT = 256, kTile = 64, kK = 16;

__global__ void __launch_bounds__(kT)
k_gemm(const GemmDesc *__restrict__ descs)
{
    __shared__ double As[kK][kTile + 1];
    __shared__ double Bs[kK][kTile + 1];
    const GemmDesc d = descs[blockIdx.y];
    const int tiles_n = (d.N + kTile - 1) / kTile;
    const int tiles_m = (d.M + kTile - 1) / kTile;
    if ((int)blockIdx.x >= tiles_m * tiles_n) return;
    const int m0 = (blockIdx.x / tiles_n) * kTile, n0 = (blockIdx.x % tiles_n) * kTile;
    const int ti = (threadIdx.x / 16) * 4, tj = (threadIdx.x % 16) * 4;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < d.K; k0 += kK) {
        for (int w = threadIdx.x; w < kK * kTile; w += kT) {
            int kk, r;
            // choose the walk so consecutive threads read consecutive memory
            if (d.ta) { r = w % kTile; kk = w / kTile; } else { kk = w % kK; r = w / kK; }
            const int gm = m0 + r, gk = k0 + kk;
            double v = 0.0;
            if (gm < d.M && gk < d.K)
                v = d.ta ? d.A[(int64_t)gk * d.lda + gm] : d.A[(int64_t)gm * d.lda + gk];
            As[kk][r] = v;
            int kk2, c2;
            if (d.tb) { kk2 = w % kK; c2 = w / kK; } else { c2 = w % kTile; kk2 = w / kTile; }
            const int gn = n0 + c2, gk2 = k0 + kk2;
            double u = 0.0;
            if (gn < d.N && gk2 < d.K)
                u = d.tb ? d.B[(int64_t)gn * d.ldb + gk2] : d.B[(int64_t)gk2 * d.ldb + gn];
            Bs[kk2][c2] = u;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kK; kk++) {
            double a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; u++) a[u] = As[kk][ti + u];
#pragma unroll
            for (int v = 0; v < 4; v++) b[v] = Bs[kk][tj + v];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[u][v] += a[u] * b[v];
        }
        __syncthreads();
    }
    for (int u = 0; u < 4; u++) {
        const int gm = m0 + ti + u;
        if (gm >= d.M) continue;
        for (int v = 0; v < 4; v++) {
            const int gn = n0 + tj + v;
            if (gn < d.N) d.C[(int64_t)gm * d.ldc + gn] = acc[u][v];
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
