// mlp.hpp -- learned diagonal preconditioner (mlp.hip): sizes and launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "layout.hpp"

namespace v2g {

constexpr int kMlpFam = 11;                 // constraint-family one-hot width
constexpr int kMlpIn = kMlpFam + 3;         // + log diag, rhs, previous y
constexpr int kMlpH = 64;                   // hidden width (two layers)
// W1[in][h], b1[h], W2[h][h], b2[h], w3[h], b3
constexpr int kMlpParams = kMlpIn * kMlpH + kMlpH + kMlpH * kMlpH + kMlpH + kMlpH + 1;

void mlp_family_codes(const Segment *segs_dev, const std::vector<Segment> &segs_host,
                      unsigned char *code, hipStream_t st);
// minv = exp(clamp(g)), divided by d when jacobi_base
void mlp_minv(int64_t n, const unsigned char *code, const double *d, const double *rhs,
              const double *y, const float *W, const double *scales, double *minv, double clampv,
              int jacobi_base, hipStream_t st);
// spectral objective: scales[2..5] = a_hi, b_hi, a_lo, b_lo
void mlp_batch_spec(int nb, const int64_t *idx, const unsigned char *code, const double *d,
                    const double *rhs, const double *yprev, const double *rhi, const double *kdhi,
                    const double *rlo, const double *kdlo, const double *s, const double *scales,
                    float *feat, float *dout, double clampv, double coef, int jacobi_base,
                    double pen, hipStream_t st);
int mlp_grad_blocks(int nb);
// tgt == nullptr: wt holds the upstream gradient d(loss)/d(out) per sample
void mlp_grad(int nb, const float *feat, const float *tgt, const float *wt, const float *W,
              float *part, float *loss_part, double *grad, double *loss, hipStream_t st);
void mlp_adam(float *W, const double *grad, double *m, double *v, double lr, int step,
              hipStream_t st);

}  // namespace v2g
