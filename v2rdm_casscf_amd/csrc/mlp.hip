// mlp.hip -- the learned diagonal preconditioner of the CG on A A^T.
//
// A per-constraint-row MLP (14 inputs, two tanh layers of 64, one output g,
// FP32 weights) gives the diagonal M^-1_ii = exp(clamp(g_i)) (identity base:
// g = 0 is the reference's plain CG) or exp(clamp(g_i)) / diag(A A^T)_ii
// (Jacobi base), SPD for any weights.
// Inputs per row: one-hot constraint family (11), log diag(A A^T)_ii / 4,
// asinh(rhs_i / rms(rhs)), asinh(y_i / rms(y)) with y the previous solution.
//
// Kernels: family codes per row (once), inference over all rows, and the
// training step on a sampled batch: forward, backward (per-workgroup
// gradients accumulated as LDS outer products), a deterministic cross-block
// reduction and Adam.  No reference counterpart; the gradients are checked
// against the CPU oracle's orc_mlp_loss_grad.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "mlp.hpp"

namespace v2g {

namespace {

constexpr int kT = 256;

__device__ __forceinline__ int family_group(int kind)
{
    switch (kind) {
    case RF_TRACE: case RF_S2: return 0;
    case RF_D1Q1: return 1;
    case RF_CONTR: return 2;
    case RF_SPIN_D1: case RF_SPIN_D2: return 3;
    case RF_SPIN_ANTI: return 4;
    case RF_SPIN_D200: case RF_SPIN_D200_4: return 5;
    case RF_MAXSPIN: return 6;
    case RF_Q2AB: case RF_Q2AA: case RF_Q2SA: return 7;
    case RF_G2AB: case RF_G2SA: return 8;
    case RF_G2AA: return 9;
    default: return 10;
    }
}

__global__ void __launch_bounds__(kT)
k_family_codes(const Segment *__restrict__ segs, int nseg, unsigned char *__restrict__ code)
{
    const Segment s = segs[blockIdx.y];
    const unsigned char g = (unsigned char)family_group(s.kind);
    for (int64_t e = blockIdx.x * (int64_t)kT + threadIdx.x; e < s.count;
         e += (int64_t)gridDim.x * kT)
        code[s.start + e] = g;
}

// features of row i into f[kMlpIn]
__device__ __forceinline__ void features(int grp, double d, double rhs, double y, float rs,
                                         float ys, float *f)
{
#pragma unroll
    for (int k = 0; k < kMlpFam; k++) f[k] = (k == grp) ? 1.0f : 0.0f;
    f[kMlpFam] = logf((float)d > 1e-30f ? (float)d : 1e-30f) * 0.25f;
    f[kMlpFam + 1] = asinhf((float)rhs * rs);
    f[kMlpFam + 2] = asinhf((float)y * ys);
}

// out = w3 . tanh(W2^T tanh(W1^T f + b1) + b2) + b3 with W in LDS
__device__ __forceinline__ float forward(const float *W, const float *f)
{
    const float *W1 = W, *b1 = W + kMlpIn * kMlpH, *W2 = b1 + kMlpH, *b2 = W2 + kMlpH * kMlpH,
                *w3 = b2 + kMlpH;
    const float b3 = w3[kMlpH];
    float h1[kMlpH];
#pragma unroll
    for (int j = 0; j < kMlpH; j++) {
        float s = b1[j];
#pragma unroll
        for (int k = 0; k < kMlpIn; k++) s += W1[k * kMlpH + j] * f[k];
        h1[j] = tanhf(s);
    }
    float o = b3;
    for (int j = 0; j < kMlpH; j++) {
        float s = b2[j];
#pragma unroll
        for (int i = 0; i < kMlpH; i++) s += W2[i * kMlpH + j] * h1[i];
        o += w3[j] * tanhf(s);
    }
    return o;
}

__global__ void __launch_bounds__(kT)
k_mlp_minv(int64_t n, const unsigned char *__restrict__ code, const double *__restrict__ d,
           const double *__restrict__ rhs, const double *__restrict__ y,
           const float *__restrict__ Wg, const double *__restrict__ scales,
           double *__restrict__ minv, double clampv, int jacobi_base)
{
    __shared__ float W[kMlpParams];
    for (int i = threadIdx.x; i < kMlpParams; i += kT) W[i] = Wg[i];
    __syncthreads();
    const float rs = (float)scales[0], ys = (float)scales[1];
    for (int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kT) {
        float f[kMlpIn];
        features(code[i], d[i], rhs[i], y[i], rs, ys, f);
        float g = forward(W, f);
        g = fminf(fmaxf(g, (float)-clampv), (float)clampv);
        const double di = (jacobi_base && d[i] > 0.0) ? d[i] : 1.0;
        minv[i] = exp((double)g) / di;
    }
}

// batch for the spectral objective L = log rho(r_hi) - log rho(r_lo), with
// rho(r) = (Sr).K(Sr) / (r.Sr) the Rayleigh quotient of the preconditioned
// operator S K at z = S r: r_hi leans towards its largest eigenvalues (power
// iterations), r_lo towards its smallest (the final residual of a CG solve),
// so L estimates the log condition number.  The upstream gradient with
// respect to g_i (s_i = exp(g_i) / base_i) is
//   dL/dg_i = [2 s_i r_i (K S r)_i / b - s_i r_i^2 / a]_hi - [ ... ]_lo
// with a = r.Sr and b = (Sr).K(Sr) (scales[2..5] = a_hi, b_hi, a_lo, b_lo),
// plus pen g_i / ncon from the penalty (pen / 2) mean(g^2); zero where g is
// clamped.  `coef` = ncon / nb.
__global__ void __launch_bounds__(kT)
k_mlp_batch_spec(int nb, const int64_t *__restrict__ idx, const unsigned char *__restrict__ code,
                 const double *__restrict__ d, const double *__restrict__ rhs,
                 const double *__restrict__ yprev, const double *__restrict__ rhi,
                 const double *__restrict__ kdhi, const double *__restrict__ rlo,
                 const double *__restrict__ kdlo, const double *__restrict__ s,
                 const double *__restrict__ scales, float *__restrict__ feat,
                 float *__restrict__ dout, double clampv, double coef, int jacobi_base,
                 double pen)
{
    const int b = blockIdx.x * kT + threadIdx.x;
    if (b >= nb) return;
    const int64_t i = idx[b];
    float f[kMlpIn];
    features(code[i], d[i], rhs[i], yprev[i], (float)scales[0], (float)scales[1], f);
    for (int k = 0; k < kMlpIn; k++) feat[(int64_t)b * kMlpIn + k] = f[k];
    const double ahi = scales[2], bhi = scales[3], alo = scales[4], blo = scales[5];
    const double base = (jacobi_base && d[i] > 0.0) ? d[i] : 1.0;
    const double g = log(s[i] * base);
    double u = 0.0;
    if (fabs(g) < clampv - 1e-6 && ahi > 0.0 && bhi > 0.0 && alo > 0.0 && blo > 0.0) {
        const double uh = 2.0 * s[i] * rhi[i] * kdhi[i] / bhi - s[i] * rhi[i] * rhi[i] / ahi;
        const double ul = 2.0 * s[i] * rlo[i] * kdlo[i] / blo - s[i] * rlo[i] * rlo[i] / alo;
        // + d/dg_i of (pen / 2) mean(g^2): pulls the diagonal towards its base
        u = uh - ul + pen * g / (coef * nb);
    }
    dout[b] = (float)(coef * u);
}

constexpr int kS = 128;     // samples per workgroup in the training kernel
constexpr int kLdH = kMlpH + 1;

// loss = sum_b wt_b (out_b - tgt_b)^2 / nb; per-block gradient partials
__global__ void __launch_bounds__(kT)
k_mlp_grad(int nb, const float *__restrict__ feat, const float *__restrict__ tgt,
           const float *__restrict__ wt, const float *__restrict__ Wg, float *__restrict__ part,
           float *__restrict__ loss_part)
{
    __shared__ float W[kMlpParams];
    __shared__ float H1[kS * kLdH];
    __shared__ float D2[kS * kLdH];
    __shared__ float F[kS * kMlpIn];
    __shared__ float DO[kS];
    __shared__ float red[kT / 64];
    for (int i = threadIdx.x; i < kMlpParams; i += kT) W[i] = Wg[i];
    const int s0 = blockIdx.x * kS;
    for (int w = threadIdx.x; w < kS * kMlpIn; w += kT) {
        const int s = s0 + w / kMlpIn;
        F[w] = s < nb ? feat[(int64_t)s0 * kMlpIn + w] : 0.0f;
    }
    __syncthreads();
    const float *W1 = W, *b1 = W + kMlpIn * kMlpH, *W2 = b1 + kMlpH, *b2 = W2 + kMlpH * kMlpH,
                *w3 = b2 + kMlpH;
    const float b3 = w3[kMlpH];
    float lossv = 0.0f;
    const int t = threadIdx.x;
    if (t < kS) {
        const int s = s0 + t;
        const bool valid = s < nb;
        for (int j = 0; j < kMlpH; j++) {
            float a = b1[j];
            for (int k = 0; k < kMlpIn; k++) a += W1[k * kMlpH + j] * F[t * kMlpIn + k];
            H1[t * kLdH + j] = tanhf(a);
        }
        float o = b3;
        for (int j = 0; j < kMlpH; j++) {
            float a = b2[j];
            for (int i = 0; i < kMlpH; i++) a += W2[i * kMlpH + j] * H1[t * kLdH + i];
            const float h2 = tanhf(a);
            D2[t * kLdH + j] = h2;
            o += w3[j] * h2;
        }
        float dout = 0.0f;
        if (valid) {
            if (tgt) {
                const float r = o - tgt[s];
                lossv = wt[s] * r * r;
                dout = 2.0f * wt[s] * r / (float)nb;
            } else {
                dout = wt[s];  // upstream gradient d(loss)/d(out) given directly
            }
        }
        DO[t] = dout;
    }
    __syncthreads();
    float *P = part + (int64_t)blockIdx.x * kMlpParams;
    const int oW1 = 0, ob1 = kMlpIn * kMlpH, oW2 = ob1 + kMlpH, ob2 = oW2 + kMlpH * kMlpH,
              ow3 = ob2 + kMlpH, ob3 = ow3 + kMlpH;
    // gw3_j = sum_s dout_s h2_sj ; gb3 = sum_s dout_s
    if (t < kMlpH) {
        float g = 0.0f;
        for (int s = 0; s < kS; s++) g += DO[s] * D2[s * kLdH + t];
        P[ow3 + t] = g;
    }
    if (t == kMlpH) {
        float g = 0.0f;
        for (int s = 0; s < kS; s++) g += DO[s];
        P[ob3] = g;
    }
    __syncthreads();
    // delta2 = dout w3 (1 - h2^2), in place
    if (t < kS)
        for (int j = 0; j < kMlpH; j++) {
            const float h2 = D2[t * kLdH + j];
            D2[t * kLdH + j] = DO[t] * w3[j] * (1.0f - h2 * h2);
        }
    __syncthreads();
    // gW2[i][j] = sum_s h1_si delta2_sj : 16 outputs per thread
    {
        const int i = t >> 2, j0 = (t & 3) * 16;
        float g[16];
        for (int q = 0; q < 16; q++) g[q] = 0.0f;
        for (int s = 0; s < kS; s++) {
            const float h = H1[s * kLdH + i];
            for (int q = 0; q < 16; q++) g[q] += h * D2[s * kLdH + j0 + q];
        }
        for (int q = 0; q < 16; q++) P[oW2 + i * kMlpH + j0 + q] = g[q];
    }
    if (t < kMlpH) {
        float g = 0.0f;
        for (int s = 0; s < kS; s++) g += D2[s * kLdH + t];
        P[ob2 + t] = g;
    }
    __syncthreads();
    // delta1_i = (sum_j W2[i][j] delta2_j)(1 - h1_i^2), in place of h1
    if (t < kS) {
        float d1[kMlpH];
        for (int i = 0; i < kMlpH; i++) {
            float a = 0.0f;
            for (int j = 0; j < kMlpH; j++) a += W2[i * kMlpH + j] * D2[t * kLdH + j];
            const float h1 = H1[t * kLdH + i];
            d1[i] = a * (1.0f - h1 * h1);
        }
        for (int i = 0; i < kMlpH; i++) H1[t * kLdH + i] = d1[i];
    }
    __syncthreads();
    // gW1[k][i] = sum_s f_sk delta1_si ; gb1_i = sum_s delta1_si
    for (int w = t; w < kMlpIn * kMlpH; w += kT) {
        const int k = w / kMlpH, i = w % kMlpH;
        float g = 0.0f;
        for (int s = 0; s < kS; s++) g += F[s * kMlpIn + k] * H1[s * kLdH + i];
        P[oW1 + w] = g;
    }
    if (t < kMlpH) {
        float g = 0.0f;
        for (int s = 0; s < kS; s++) g += H1[s * kLdH + t];
        P[ob1 + t] = g;
    }
    // loss partial
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lossv += __shfl_xor(lossv, o, 64);
    if ((t & 63) == 0) red[t >> 6] = lossv;
    __syncthreads();
    if (t == 0) {
        float l = 0.0f;
        for (int i = 0; i < kT / 64; i++) l += red[i];
        loss_part[blockIdx.x] = l;
    }
}

// grad[p] = sum over blocks of the partials (FP64 accumulate, fixed order)
__global__ void __launch_bounds__(kT)
k_mlp_reduce(int nblk, const float *__restrict__ part, const float *__restrict__ loss_part, int nb,
             double *__restrict__ grad, double *__restrict__ loss)
{
    const int p = blockIdx.x * kT + threadIdx.x;
    if (p < kMlpParams) {
        double g = 0.0;
        for (int b = 0; b < nblk; b++) g += part[(int64_t)b * kMlpParams + p];
        grad[p] = g;
    }
    if (p == 0) {
        double l = 0.0;
        for (int b = 0; b < nblk; b++) l += loss_part[b];
        *loss = l / nb;
    }
}

__global__ void __launch_bounds__(kT)
k_adam(float *__restrict__ W, const double *__restrict__ grad, double *__restrict__ m,
       double *__restrict__ v, double lr, double b1c, double b2c)
{
    const int p = blockIdx.x * kT + threadIdx.x;
    if (p >= kMlpParams) return;
    const double g = grad[p];
    const double mm = 0.9 * m[p] + 0.1 * g;
    const double vv = 0.999 * v[p] + 0.001 * g * g;
    m[p] = mm;
    v[p] = vv;
    W[p] = (float)(W[p] - lr * (mm / b1c) / (sqrt(vv / b2c) + 1e-8));
}

}  // namespace

void mlp_family_codes(const Segment *segs_dev, const std::vector<Segment> &segs_host,
                      unsigned char *code, hipStream_t st)
{
    int64_t maxc = 1;
    for (const auto &s : segs_host) maxc = std::max(maxc, s.count);
    const int gx = (int)std::min<int64_t>(4096, (maxc + kT - 1) / kT);
    if (!segs_host.empty())
        hipLaunchKernelGGL(k_family_codes, dim3(gx, (unsigned)segs_host.size()), dim3(kT), 0, st,
                           segs_dev, (int)segs_host.size(), code);
}

void mlp_minv(int64_t n, const unsigned char *code, const double *d, const double *rhs,
              const double *y, const float *W, const double *scales, double *minv, double clampv,
              int jacobi_base, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_minv, dim3(2048), dim3(kT), 0, st, n, code, d, rhs, y, W, scales, minv,
                       clampv, jacobi_base);
}

void mlp_batch_spec(int nb, const int64_t *idx, const unsigned char *code, const double *d,
                    const double *rhs, const double *yprev, const double *rhi, const double *kdhi,
                    const double *rlo, const double *kdlo, const double *s, const double *scales,
                    float *feat, float *dout, double clampv, double coef, int jacobi_base,
                    double pen, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_batch_spec, dim3((nb + kT - 1) / kT), dim3(kT), 0, st, nb, idx, code, d,
                       rhs, yprev, rhi, kdhi, rlo, kdlo, s, scales, feat, dout, clampv, coef,
                       jacobi_base, pen);
}

int mlp_grad_blocks(int nb) { return (nb + kS - 1) / kS; }

void mlp_grad(int nb, const float *feat, const float *tgt, const float *wt, const float *W,
              float *part, float *loss_part, double *grad, double *loss, hipStream_t st)
{
    const int nblk = mlp_grad_blocks(nb);
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(kT), 0, st, nb, feat, tgt, wt, W, part,
                       loss_part);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + kT - 1) / kT), dim3(kT), 0, st, nblk, part,
                       loss_part, nb, grad, loss);
}

void mlp_adam(float *W, const double *grad, double *m, double *v, double lr, int step,
              hipStream_t st)
{
    const double b1c = 1.0 - std::pow(0.9, step), b2c = 1.0 - std::pow(0.999, step);
    hipLaunchKernelGGL(k_adam, dim3((kMlpParams + kT - 1) / kT), dim3(kT), 0, st, W, grad, m, v,
                       lr, b1c, b2c);
}

}  // namespace v2g
