// nn_ops.hip -- neural-network ops of the learned warm start, for gfx950.
//
// Tensors are NHWC; activations bf16, accumulation fp32.  A 3x3 convolution
// (padding 1) is an implicit GEMM: rows m = (n, y, x) pixels, columns output
// channels, K = 9 taps x input channels, on v_mfma_f32_16x16x32_bf16 (A:
// lane i + 16 (k / 8) holds A[i][k], element k % 8; B: lane j + 16 (k / 8)
// holds B[k][j]; C: lane j + 16 (i / 4) holds C[i][j], element i % 4).
// Weights: [9][Cout][Cin] (tap = 3 ky + kx).  The data gradient of a conv is a
// forward conv of the output gradient with the flipped, transposed weights;
// the weight gradient is a split-K GEMM over pixels whose partials are summed
// in a fixed order (deterministic).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nn_ops.h"

#include <algorithm>

namespace cvt {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ __forceinline__ float bf2f(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }
__device__ __forceinline__ uint16_t f2bf(float f)
{
    // round to nearest even (no NaN inputs in this network)
    const uint32_t u = __float_as_uint(f);
    const uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}

#define LDA 40   // LDS row pitch in bf16 (32 + 8 padding)

// out[m][co] = act(bias[co] + sum_tap sum_ci in[m + off(tap)][ci] * w[tap][co][ci])
// grid (ceil(M / 64), Cout / (16 NB)), block 256; Cin % 32 == 0, Cout % (16 NB) == 0
template <int NB>
__global__ __launch_bounds__(256) void k_conv3x3(const uint16_t *__restrict__ in,
                                                 const uint16_t *__restrict__ w,
                                                 const float *__restrict__ bias,
                                                 uint16_t *__restrict__ out, int N, int H, int W,
                                                 int Cin, int Cout, int relu)
{
    __shared__ uint16_t sA[64 * LDA];
    __shared__ uint16_t sB[16 * NB * LDA];
    const long M = (long)N * H * W;
    const long m0 = (long)blockIdx.x * 64;
    const int co0 = blockIdx.y * 16 * NB;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    // the pixel this thread stages for A: m_local = t / 4, 8 channels at part * 8
    const int am = t >> 2, apart = t & 3;
    const long m = m0 + am;
    int an = 0, ay = 0, ax = 0;
    if (m < M) {
        an = (int)(m / ((long)H * W));
        const int r = (int)(m - (long)an * H * W);
        ay = r / W;
        ax = r - ay * W;
    }
    f32x4 acc[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 9; ++tap) {
        const int yy = ay + tap / 3 - 1, xx = ax + tap % 3 - 1;
        const bool inside = (m < M) && yy >= 0 && yy < H && xx >= 0 && xx < W;
        const long src = inside ? (((long)an * H + yy) * W + xx) * Cin : 0;
        for (int ci0 = 0; ci0 < Cin; ci0 += 32) {
            __syncthreads();
            uint4 va = make_uint4(0, 0, 0, 0);
            if (inside) va = *reinterpret_cast<const uint4 *>(in + src + ci0 + apart * 8);
            *reinterpret_cast<uint4 *>(&sA[am * LDA + apart * 8]) = va;
            if (t < 16 * NB * 4) {
                const int bco = t >> 2, bpart = t & 3;
                const uint4 vb = *reinterpret_cast<const uint4 *>(
                    w + ((long)tap * Cout + co0 + bco) * Cin + ci0 + bpart * 8);
                *reinterpret_cast<uint4 *>(&sB[bco * LDA + bpart * 8]) = vb;
            }
            __syncthreads();
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(
                &sA[(wave * 16 + (lane & 15)) * LDA + 8 * (lane >> 4)]);
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const bf16x8 bb = *reinterpret_cast<const bf16x8 *>(
                    &sB[(b * 16 + (lane & 15)) * LDA + 8 * (lane >> 4)]);
                acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[b], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int co = co0 + b * 16 + (lane & 15);
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const long mm = m0 + wave * 16 + 4 * (lane >> 4) + e;
            if (mm < M) {
                float v = acc[b][e] + bv;
                if (relu) v = fmaxf(v, 0.0f);
                out[mm * Cout + co] = f2bf(v);
            }
        }
    }
}

// partial weight gradient of one tap, 32 co x 32 ci, over pixels
// [split * chunk, (split + 1) * chunk): part[split][tap][co][ci]
// grid (9 * Cout/32 * Cin/32, splits), block 256 (4 waves = 2 x 2 blocks of 16x16)
__global__ __launch_bounds__(256) void k_conv3x3_dw(const uint16_t *__restrict__ in,
                                                    const uint16_t *__restrict__ g,
                                                    float *__restrict__ part, int N, int H, int W,
                                                    int Cin, int Cout, long chunk)
{
    __shared__ uint16_t sG[32 * LDA];   // [co][m]
    __shared__ uint16_t sX[32 * LDA];   // [ci][m]
    const int nco = Cout / 32, nci = Cin / 32;
    const int bx = blockIdx.x;
    const int tap = bx / (nco * nci);
    const int co0 = ((bx / nci) % nco) * 32, ci0 = (bx % nci) * 32;
    const long M = (long)N * H * W;
    const long mb = (long)blockIdx.y * chunk, me = min(M, mb + chunk);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int bi = wave >> 1, bj = wave & 1;      // this wave's 16x16 block
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const int ky = tap / 3 - 1, kx = tap % 3 - 1;
    // staging: thread t handles pixel row mr = t / 8 (32 rows), 4 channels at (t % 8) * 4
    const int mr = t >> 3, cpart = (t & 7) * 4;
    for (long mc = mb; mc < me; mc += 32) {
        __syncthreads();
        const long mm = mc + mr;
        uint2 vg = make_uint2(0, 0), vx = make_uint2(0, 0);
        if (mm < me) {
            vg = *reinterpret_cast<const uint2 *>(g + mm * Cout + co0 + cpart);
            const int n = (int)(mm / ((long)H * W));
            const int r = (int)(mm - (long)n * H * W);
            const int y = r / W, x = r - (r / W) * W;
            const int yy = y + ky, xx = x + kx;
            if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                vx = *reinterpret_cast<const uint2 *>(in + (((long)n * H + yy) * W + xx) * Cin + ci0 + cpart);
        }
        // transpose into [channel][m]
        const uint16_t *pg = reinterpret_cast<const uint16_t *>(&vg);
        const uint16_t *px = reinterpret_cast<const uint16_t *>(&vx);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            sG[(cpart + q) * LDA + mr] = pg[q];
            sX[(cpart + q) * LDA + mr] = px[q];
        }
        __syncthreads();
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&sG[(bi * 16 + (lane & 15)) * LDA + 8 * (lane >> 4)]);
        const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&sX[(bj * 16 + (lane & 15)) * LDA + 8 * (lane >> 4)]);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    const long tap_sz = (long)Cout * Cin;
    float *dst = part + (long)blockIdx.y * 9 * tap_sz + (long)tap * tap_sz;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int co = co0 + bi * 16 + 4 * (lane >> 4) + e;
        const int ci = ci0 + bj * 16 + (lane & 15);
        dst[(long)co * Cin + ci] = acc[e];
    }
}

// out[i] = sum_s part[s][i] in split order (deterministic)
__global__ void k_sum_splits(const float *__restrict__ part, int splits, long n,
                             float *__restrict__ out, int accumulate)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int k = 0; k < splits; ++k) s += part[(long)k * n + i];
    out[i] = accumulate ? out[i] + s : s;
}

// per-split column sums of g [M][C] -> part[split][C]
__global__ __launch_bounds__(256) void k_colsum(const uint16_t *__restrict__ g, long M, int C,
                                                long chunk, float *__restrict__ part)
{
    const long mb = (long)blockIdx.x * chunk, me = min(M, mb + chunk);
    for (int c = threadIdx.x; c < C; c += blockDim.x) {
        float s = 0.0f;
        for (long m = mb; m < me; ++m) s += bf2f(g[m * C + c]);
        part[(long)blockIdx.x * C + c] = s;
    }
}

// fp32 master [9][Cout][Cin] -> bf16 forward copy (same layout) and the
// flipped, transposed copy for the data gradient: [9][Cin][Cout], tap -> 8 - tap
__global__ void k_weights_bf16(const float *__restrict__ w, int Cin, int Cout,
                               uint16_t *__restrict__ wf, uint16_t *__restrict__ wb)
{
    const long n = 9L * Cin * Cout;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int tap = (int)(i / ((long)Cin * Cout));
    const int r = (int)(i - (long)tap * Cin * Cout);
    const int co = r / Cin, ci = r - (r / Cin) * Cin;
    const uint16_t v = f2bf(w[i]);
    wf[i] = v;
    wb[((long)(8 - tap) * Cin + ci) * Cout + co] = v;
}

// g *= (y > 0)
__global__ void k_relu_bwd(uint16_t *__restrict__ g, const uint16_t *__restrict__ y, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(bf2f(y[i]) > 0.0f)) g[i] = 0;
}

// 2x2 average pool, NHWC (H, W even)
__global__ void k_avgpool2(const uint16_t *__restrict__ in, uint16_t *__restrict__ out, int N,
                           int H, int W, int C)
{
    const int Ho = H / 2, Wo = W / 2;
    const long n = (long)N * Ho * Wo * C;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    const long p = i / C;
    const int x = (int)(p % Wo), y = (int)((p / Wo) % Ho), b = (int)(p / ((long)Wo * Ho));
    const long base = (((long)b * H + 2 * y) * W + 2 * x) * C + c;
    const float s = bf2f(in[base]) + bf2f(in[base + C]) + bf2f(in[base + (long)W * C]) +
                    bf2f(in[base + (long)W * C + C]);
    out[i] = f2bf(0.25f * s);
}

// gradient of the 2x2 average pool: each input pixel gets a quarter
__global__ void k_avgpool2_bwd(const uint16_t *__restrict__ gout, uint16_t *__restrict__ gin,
                               int N, int H, int W, int C, int accumulate)
{
    const long n = (long)N * H * W * C;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    const long p = i / C;
    const int x = (int)(p % W), y = (int)((p / W) % H), b = (int)(p / ((long)W * H));
    const float v = 0.25f * bf2f(gout[(((long)b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c]);
    gin[i] = f2bf(accumulate ? bf2f(gin[i]) + v : v);
}

// nearest 2x upsample of a (N, H, W, C1) tensor into channels [0, C1) of a
// (N, 2H, 2W, C1 + C2) tensor, and copy of skip (N, 2H, 2W, C2) into [C1, C1+C2)
__global__ void k_up_concat(const uint16_t *__restrict__ low, const uint16_t *__restrict__ skip,
                            uint16_t *__restrict__ out, int N, int H, int W, int C1, int C2)
{
    const int Ho = 2 * H, Wo = 2 * W, C = C1 + C2;
    const long n = (long)N * Ho * Wo * C;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = (int)(i % C);
    const long p = i / C;
    const int x = (int)(p % Wo), y = (int)((p / Wo) % Ho), b = (int)(p / ((long)Wo * Ho));
    out[i] = (c < C1) ? low[(((long)b * H + y / 2) * W + x / 2) * C1 + c]
                      : skip[p * C2 + (c - C1)];
}

// gradient of k_up_concat: 2x2 sums into the low tensor, and the skip part
// added into the skip gradient (accumulate) or written
__global__ void k_up_concat_bwd(const uint16_t *__restrict__ gout, uint16_t *__restrict__ glow,
                                uint16_t *__restrict__ gskip, int N, int H, int W, int C1, int C2,
                                int acc_skip)
{
    const int Ho = 2 * H, Wo = 2 * W, C = C1 + C2;
    const long nl = (long)N * H * W * C1;
    const long ns = (long)N * Ho * Wo * C2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nl) {
        const int c = (int)(i % C1);
        const long p = i / C1;
        const int x = (int)(p % W), y = (int)((p / W) % H), b = (int)(p / ((long)W * H));
        const long o = (((long)b * Ho + 2 * y) * Wo + 2 * x) * C + c;
        const float s = bf2f(gout[o]) + bf2f(gout[o + C]) + bf2f(gout[o + (long)Wo * C]) +
                        bf2f(gout[o + (long)Wo * C + C]);
        glow[i] = f2bf(s);
    }
    if (i < ns) {
        const int c = (int)(i % C2);
        const long p = i / C2;
        const float v = bf2f(gout[p * C + C1 + c]);
        gskip[i] = f2bf(acc_skip ? bf2f(gskip[i]) + v : v);
    }
}

// head: 1x1 conv C -> 2 in fp32: f[m][j] = b[j] + sum_c x[m][c] wh[j][c]
__global__ void k_head_fwd(const uint16_t *__restrict__ x, const float *__restrict__ wh,
                           const float *__restrict__ bh, float *__restrict__ f, long M, int C)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    float a = bh[0], b = bh[1];
    for (int c = 0; c < C; ++c) {
        const float v = bf2f(x[m * C + c]);
        a += v * wh[c];
        b += v * wh[C + c];
    }
    f[2 * m] = a;
    f[2 * m + 1] = b;
}

// head gradients: gx[m][c] = sum_j gf[m][j] wh[j][c]; per-split partials of
// dwh[j][c] = sum_m gf[m][j] x[m][c] and dbh[j] = sum_m gf[m][j]
__global__ __launch_bounds__(256) void k_head_bwd(const uint16_t *__restrict__ x,
                                                  const float *__restrict__ gf,
                                                  const float *__restrict__ wh,
                                                  uint16_t *__restrict__ gx, float *__restrict__ part,
                                                  long M, int C, long chunk)
{
    __shared__ float red[256];
    const long mb = (long)blockIdx.x * chunk, me = min(M, mb + chunk);
    float accw[2 * 32 + 2];
    for (int k = 0; k < 2 * C + 2; ++k) accw[k] = 0.0f;
    for (long m = mb + threadIdx.x; m < me; m += 256) {
        const float g0 = gf[2 * m], g1 = gf[2 * m + 1];
        for (int c = 0; c < C; ++c) {
            const float v = bf2f(x[m * C + c]);
            gx[m * C + c] = f2bf(g0 * wh[c] + g1 * wh[C + c]);
            accw[c] += g0 * v;
            accw[C + c] += g1 * v;
        }
        accw[2 * C] += g0;
        accw[2 * C + 1] += g1;
    }
    for (int k = 0; k < 2 * C + 2; ++k) {
        red[threadIdx.x] = accw[k];
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(long)blockIdx.x * (2 * C + 2) + k] = red[0];
        __syncthreads();
    }
}

// bilinear sample of the (N, R, R, 2) field at sites u in [0, 1)^2 (pixel
// centres at (i + 0.5) / R, clamped at the border)
__device__ __forceinline__ void bil_coords(float u, int R, int &i0, int &i1, float &t)
{
    float p = u * R - 0.5f;
    p = fminf(fmaxf(p, 0.0f), (float)(R - 1));
    i0 = (int)floorf(p);
    i1 = min(i0 + 1, R - 1);
    t = p - (float)i0;
}

// per site: pred = field(u), loss += |pred - target|^2 * scale, gfield +=
// d loss / d field (atomics; training only)
__global__ void k_sample_loss(const float *__restrict__ field, int R, const float *__restrict__ uv,
                              const float *__restrict__ target, int S, int nb, float scale,
                              float *__restrict__ pred, float *__restrict__ gfield,
                              float *__restrict__ loss_part)
{
    __shared__ float red[256];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < (long)nb * S) {
        const int b = (int)(i / S);
        int x0, x1, y0, y1;
        float tx, ty;
        bil_coords(uv[2 * i], R, x0, x1, tx);
        bil_coords(uv[2 * i + 1], R, y0, y1, ty);
        const float *F = field + (long)b * R * R * 2;
        const float w00 = (1 - tx) * (1 - ty), w10 = tx * (1 - ty), w01 = (1 - tx) * ty, w11 = tx * ty;
        const long p00 = (long)y0 * R + x0, p10 = (long)y0 * R + x1, p01 = (long)y1 * R + x0,
                   p11 = (long)y1 * R + x1;
        float pv[2];
        for (int j = 0; j < 2; ++j)
            pv[j] = w00 * F[2 * p00 + j] + w10 * F[2 * p10 + j] + w01 * F[2 * p01 + j] +
                    w11 * F[2 * p11 + j];
        if (pred) { pred[2 * i] = pv[0]; pred[2 * i + 1] = pv[1]; }
        if (target) {
            const float d0 = pv[0] - target[2 * i], d1 = pv[1] - target[2 * i + 1];
            l = (d0 * d0 + d1 * d1) * scale;
            if (gfield) {
                float *G = gfield + (long)b * R * R * 2;
                const float g0 = 2.0f * d0 * scale, g1 = 2.0f * d1 * scale;
                atomicAdd(&G[2 * p00], w00 * g0); atomicAdd(&G[2 * p00 + 1], w00 * g1);
                atomicAdd(&G[2 * p10], w10 * g0); atomicAdd(&G[2 * p10 + 1], w10 * g1);
                atomicAdd(&G[2 * p01], w01 * g0); atomicAdd(&G[2 * p01 + 1], w01 * g1);
                atomicAdd(&G[2 * p11], w11 * g0); atomicAdd(&G[2 * p11 + 1], w11 * g1);
            }
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss_part) loss_part[blockIdx.x] = red[0];
}

// Moves each free site of an n x n engine grid by the field sampled at its
// position (displacements in units of the grid side).  A site keeps its
// position when it is fixed, when the target lies off the grid, or when the
// target pixel has zero density (a site there would own no mass).
__global__ void k_apply_disp(float2 *__restrict__ sites, const uint8_t *__restrict__ fixed, int S,
                             const float *__restrict__ field, int R, const float *__restrict__ rho,
                             int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S || fixed[i]) return;
    const float2 s = sites[i];
    const float inv = 1.0f / (float)n;
    int x0, x1, y0, y1;
    float tx, ty;
    bil_coords(s.x * inv, R, x0, x1, tx);
    bil_coords(s.y * inv, R, y0, y1, ty);
    const float w00 = (1 - tx) * (1 - ty), w10 = tx * (1 - ty), w01 = (1 - tx) * ty, w11 = tx * ty;
    const long p00 = (long)y0 * R + x0, p10 = (long)y0 * R + x1, p01 = (long)y1 * R + x0,
               p11 = (long)y1 * R + x1;
    float d[2];
    for (int j = 0; j < 2; ++j)
        d[j] = w00 * field[2 * p00 + j] + w10 * field[2 * p10 + j] + w01 * field[2 * p01 + j] +
               w11 * field[2 * p11 + j];
    const float x = s.x + d[0] * (float)n, y = s.y + d[1] * (float)n;
    if (!(x >= 0.0f && x < (float)n && y >= 0.0f && y < (float)n)) return;   // NaN too
    if (rho[(long)(int)y * n + (int)x] > 0.0f) sites[i] = make_float2(x, y);
}

// Adam (bias-corrected) on fp32 parameters
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    w[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

__global__ void k_fill_f32(float *__restrict__ p, float v, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// network input at R x R: channel 0 = box-mean density / its mean, channel 1 =
// site count per cell * R^2 / S, channel 2 = 1 where the cell has density;
// channels 3 .. C-1 zero.  The density is n x n with n a multiple of R.
__global__ void k_prepare_density(const float *__restrict__ rho, int n, int R, int C,
                                  const float *__restrict__ inv_mean, uint16_t *__restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)R * R) return;
    const int x = (int)(i % R), y = (int)(i / R);
    const int f = n / R;
    float s = 0.0f;
    for (int yy = 0; yy < f; ++yy)
        for (int xx = 0; xx < f; ++xx) s += rho[(long)(y * f + yy) * n + (x * f + xx)];
    s /= (float)(f * f);
    uint16_t *o = out + i * C;
    o[0] = f2bf(s * inv_mean[0]);
    o[1] = 0;
    o[2] = f2bf(s > 0.0f ? 1.0f : 0.0f);
    for (int c = 3; c < C; ++c) o[c] = 0;
}

// float64 sum of the density: per-block partial sums over a fixed partition,
// then one block adds the partials in index order (deterministic)
__global__ __launch_bounds__(256) void k_rho_partial(const float *__restrict__ rho, long n,
                                                     double *__restrict__ part)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += 256L * NN_SUM_BLOCKS)
        acc += (double)rho[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_rho_inv_mean(const double *__restrict__ part, long n,
                                                      float *__restrict__ inv_mean)
{
    __shared__ double red[256];
    red[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) inv_mean[0] = red[0] > 0.0 ? (float)((double)n / red[0]) : 0.0f;
}

__global__ void k_prepare_sites(const float *__restrict__ xy, int S, int n, int R, int C,
                                float weight, float *__restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    const int x = min(max((int)(xy[2 * i] * R / n), 0), R - 1);
    const int y = min(max((int)(xy[2 * i + 1] * R / n), 0), R - 1);
    atomicAdd(&count[(long)y * R + x], weight);
}

// channel 1: the site counts averaged over a (2h+1)^2 window (zero outside
// the grid, divided by the full window area), so the channel holds a local
// site density of about the same noise level at any site count
__global__ void k_prepare_finish(const float *__restrict__ count, int R, int C, int h,
                                 uint16_t *__restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)R * R) return;
    const int x = (int)(i % R), y = (int)(i / R);
    float acc = 0.0f;
    for (int yy = max(y - h, 0); yy <= min(y + h, R - 1); ++yy)
        for (int xx = max(x - h, 0); xx <= min(x + h, R - 1); ++xx) acc += count[(long)yy * R + xx];
    out[i * C + 1] = f2bf(acc / (float)((2 * h + 1) * (2 * h + 1)));
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

void nn_conv3x3(hipStream_t s, const uint16_t *in, const uint16_t *w, const float *bias,
                uint16_t *out, int N, int H, int W, int Cin, int Cout, int relu)
{
    const long M = (long)N * H * W;
    if (Cout % 64 == 0)
        hipLaunchKernelGGL(k_conv3x3<4>, dim3(nblk(M, 64), Cout / 64), dim3(256), 0, s, in, w, bias,
                           out, N, H, W, Cin, Cout, relu);
    else
        hipLaunchKernelGGL(k_conv3x3<2>, dim3(nblk(M, 64), Cout / 32), dim3(256), 0, s, in, w, bias,
                           out, N, H, W, Cin, Cout, relu);
}

int nn_dw_splits(long M) { return (int)std::min<long>(64, std::max<long>(1, M / 4096)); }

void nn_conv3x3_dw(hipStream_t s, const uint16_t *in, const uint16_t *g, float *dw, float *work,
                   int N, int H, int W, int Cin, int Cout, int accumulate)
{
    const long M = (long)N * H * W;
    const int splits = nn_dw_splits(M);
    const long chunk = ((M + splits - 1) / splits + 31) / 32 * 32;
    hipLaunchKernelGGL(k_conv3x3_dw, dim3(9 * (Cout / 32) * (Cin / 32), splits), dim3(256), 0, s, in,
                       g, work, N, H, W, Cin, Cout, chunk);
    const long n = 9L * Cout * Cin;
    hipLaunchKernelGGL(k_sum_splits, dim3(nblk(n)), dim3(256), 0, s, work, splits, n, dw, accumulate);
}

void nn_colsum(hipStream_t s, const uint16_t *g, long M, int C, float *db, float *work)
{
    const int splits = nn_dw_splits(M);
    const long chunk = (M + splits - 1) / splits;
    hipLaunchKernelGGL(k_colsum, dim3(splits), dim3(256), 0, s, g, M, C, chunk, work);
    hipLaunchKernelGGL(k_sum_splits, dim3(nblk(C)), dim3(256), 0, s, work, splits, (long)C, db, 0);
}

void nn_weights_bf16(hipStream_t s, const float *w, int Cin, int Cout, uint16_t *wf, uint16_t *wb)
{
    hipLaunchKernelGGL(k_weights_bf16, dim3(nblk(9L * Cin * Cout)), dim3(256), 0, s, w, Cin, Cout, wf, wb);
}

void nn_relu_bwd(hipStream_t s, uint16_t *g, const uint16_t *y, long n)
{
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(n)), dim3(256), 0, s, g, y, n);
}

void nn_avgpool2(hipStream_t s, const uint16_t *in, uint16_t *out, int N, int H, int W, int C)
{
    hipLaunchKernelGGL(k_avgpool2, dim3(nblk((long)N * (H / 2) * (W / 2) * C)), dim3(256), 0, s, in,
                       out, N, H, W, C);
}

void nn_avgpool2_bwd(hipStream_t s, const uint16_t *gout, uint16_t *gin, int N, int H, int W, int C,
                     int accumulate)
{
    hipLaunchKernelGGL(k_avgpool2_bwd, dim3(nblk((long)N * H * W * C)), dim3(256), 0, s, gout, gin, N,
                       H, W, C, accumulate);
}

void nn_up_concat(hipStream_t s, const uint16_t *low, const uint16_t *skip, uint16_t *out, int N,
                  int H, int W, int C1, int C2)
{
    hipLaunchKernelGGL(k_up_concat, dim3(nblk((long)N * 4 * H * W * (C1 + C2))), dim3(256), 0, s,
                       low, skip, out, N, H, W, C1, C2);
}

void nn_up_concat_bwd(hipStream_t s, const uint16_t *gout, uint16_t *glow, uint16_t *gskip, int N,
                      int H, int W, int C1, int C2, int acc_skip)
{
    const long n = std::max((long)N * H * W * C1, (long)N * 4 * H * W * C2);
    hipLaunchKernelGGL(k_up_concat_bwd, dim3(nblk(n)), dim3(256), 0, s, gout, glow, gskip, N, H, W,
                       C1, C2, acc_skip);
}

void nn_head_fwd(hipStream_t s, const uint16_t *x, const float *wh, const float *bh, float *f,
                 long M, int C)
{
    hipLaunchKernelGGL(k_head_fwd, dim3(nblk(M)), dim3(256), 0, s, x, wh, bh, f, M, C);
}

void nn_head_bwd(hipStream_t s, const uint16_t *x, const float *gf, const float *wh, uint16_t *gx,
                 float *dwh_db, float *work, long M, int C)
{
    const int splits = (int)std::min<long>(256, std::max<long>(1, M / 1024));
    const long chunk = (M + splits - 1) / splits;
    hipLaunchKernelGGL(k_head_bwd, dim3(splits), dim3(256), 0, s, x, gf, wh, gx, work, M, C, chunk);
    hipLaunchKernelGGL(k_sum_splits, dim3(nblk(2 * C + 2)), dim3(256), 0, s, work, splits,
                       (long)(2 * C + 2), dwh_db, 0);
}

void nn_sample_loss(hipStream_t s, const float *field, int R, const float *uv, const float *target,
                    int S, int nb, float scale, float *pred, float *gfield, float *loss_part)
{
    hipLaunchKernelGGL(k_sample_loss, dim3(nblk((long)nb * S)), dim3(256), 0, s, field, R, uv, target,
                       S, nb, scale, pred, gfield, loss_part);
}

void nn_adam(hipStream_t s, float *w, const float *g, float *m1, float *m2, long n, float lr,
             float b1, float b2, float eps, int t)
{
    const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(256), 0, s, w, g, m1, m2, n, lr, b1, b2, eps, c1, c2);
}

int nn_splat_halfwidth(int R, int S)
{
    if (S <= 0) return 0;
    return std::max(0, (int)lround((double)R / sqrt((double)S)) - 1);
}

void nn_density_inv_mean(hipStream_t s, const float *rho, int n, double *part, float *inv_mean)
{
    hipLaunchKernelGGL(k_rho_partial, dim3(NN_SUM_BLOCKS), dim3(256), 0, s, rho, (long)n * n, part);
    hipLaunchKernelGGL(k_rho_inv_mean, dim3(1), dim3(256), 0, s, part, (long)n * n, inv_mean);
}

void nn_prepare(hipStream_t s, const float *rho, int n, const float *inv_mean, const float *xy,
                int S, int R, int C, uint16_t *out, float *count)
{
    hipLaunchKernelGGL(k_prepare_density, dim3(nblk((long)R * R)), dim3(256), 0, s, rho, n, R, C,
                       inv_mean, out);
    hipLaunchKernelGGL(k_fill_f32, dim3(nblk((long)R * R)), dim3(256), 0, s, count, 0.0f,
                       (long)R * R);
    if (S > 0)
        hipLaunchKernelGGL(k_prepare_sites, dim3(nblk(S)), dim3(256), 0, s, xy, S, n, R, C,
                           (float)R * R / (float)S, count);
    hipLaunchKernelGGL(k_prepare_finish, dim3(nblk((long)R * R)), dim3(256), 0, s, count, R, C,
                       nn_splat_halfwidth(R, S), out);
}

void nn_apply_disp(hipStream_t s, float2 *sites, const uint8_t *fixed, int S, const float *field,
                   int R, const float *rho, int n)
{
    if (S > 0)
        hipLaunchKernelGGL(k_apply_disp, dim3(nblk(S)), dim3(256), 0, s, sites, fixed, S, field, R,
                           rho, n);
}

}  // namespace cvt
