// nn_ops.h -- launchers of the warm-start network's ops (nn_ops.hip).
// Tensors NHWC, activations bf16 (uint16_t bits), parameters fp32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cvt {

// 3x3 conv, padding 1: out = act(bias + conv(in, w)); w [9][Cout][Cin] bf16;
// Cin % 32 == 0, Cout % 32 == 0
void nn_conv3x3(hipStream_t s, const uint16_t *in, const uint16_t *w, const float *bias,
                uint16_t *out, int N, int H, int W, int Cin, int Cout, int relu);
// weight gradient dw[9][Cout][Cin] (+= when accumulate) of out-gradient g;
// work holds nn_dw_splits(M, Cin, Cout) * 9 * Cout * Cin floats
void nn_conv3x3_dw(hipStream_t s, const uint16_t *in, const uint16_t *g, float *dw, float *work,
                   int N, int H, int W, int Cin, int Cout, int accumulate);
int nn_dw_splits(long M, int Cin, int Cout);
// db[c] = sum_m g[m][c]; work holds nn_colsum_splits(M) * C floats.  With
// relu_y, g is first masked in place by the ReLU derivative of the layer's
// output relu_y (g = 0 where relu_y <= 0).
// splits of the column sum over M rows (its workspace holds splits * C floats)
int nn_colsum_splits(long M);
void nn_colsum(hipStream_t s, uint16_t *g, long M, int C, float *db, float *work,
               const uint16_t *relu_y);
// fp32 weights -> bf16 forward copy and flipped/transposed data-gradient copy
void nn_weights_bf16(hipStream_t s, const float *w, int Cin, int Cout, uint16_t *wf, uint16_t *wb);
void nn_avgpool2(hipStream_t s, const uint16_t *in, uint16_t *out, int N, int H, int W, int C);
void nn_avgpool2_bwd(hipStream_t s, const uint16_t *gout, uint16_t *gin, int N, int H, int W, int C,
                     int accumulate);
// (N, H, W, C1) low upsampled x2 into channels [0, C1), skip (N, 2H, 2W, C2)
// into [C1, C1 + C2)
void nn_up_concat(hipStream_t s, const uint16_t *low, const uint16_t *skip, uint16_t *out, int N,
                  int H, int W, int C1, int C2);
void nn_up_concat_bwd(hipStream_t s, const uint16_t *gout, uint16_t *glow, uint16_t *gskip, int N,
                      int H, int W, int C1, int C2, int acc_skip);
// 1x1 head NN_HEAD_C -> 2 (fp32 out); dwh_db = [dwh (2C), dbh (2)]; work
// nn_head_splits(M) * (2C + 2) floats
constexpr int NN_HEAD_C = 32;
int nn_head_splits(long M);
void nn_head_fwd(hipStream_t s, const uint16_t *x, const float *wh, const float *bh, float *f,
                 long M);
void nn_head_bwd(hipStream_t s, const uint16_t *x, const float *gf, const float *wh, uint16_t *gx,
                 float *dwh_db, float *work, long M);
// bilinear sample of (nb, R, R, 2) at uv [nb][S][2]; loss partials per 256
// sites; gfield (zeroed by the caller) receives the gradient
void nn_sample_loss(hipStream_t s, const float *field, int R, const float *uv, const float *target,
                    int S, int nb, float scale, float *pred, float *gfield, float *loss_part);
void nn_adam(hipStream_t s, float *w, const float *g, float *m1, float *m2, long n, float lr,
             float b1, float b2, float eps, int t);
#define NN_SUM_BLOCKS 256
// 1 / mean of an n x n density into inv_mean[0] (device), 0 if the sum is 0;
// part is NN_SUM_BLOCKS doubles of scratch
void nn_density_inv_mean(hipStream_t s, const float *rho, int n, double *part, float *inv_mean);
// half-width h of the window that averages the site splat: round(R / sqrt(S))
// - 1, so a window holds about three sites on average
int nn_splat_halfwidth(int R, int S);
// network input (R, R, C) from an n x n density, its 1 / mean (device) and S
// sites (pixel units); count is an R*R float scratch
void nn_prepare(hipStream_t s, const float *rho, int n, const float *inv_mean, const float *xy,
                int S, int R, int C, uint16_t *out, float *count);
// sites += field(site / n) * n for free sites whose target pixel has density
void nn_apply_disp(hipStream_t s, float2 *sites, const uint8_t *fixed, int S, const float *field,
                   int R, const float *rho, int n);

}  // namespace cvt
