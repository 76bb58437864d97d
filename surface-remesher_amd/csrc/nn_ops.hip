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

// Implicit-GEMM 3x3 convolution, NHWC bf16, fp32 accumulation on MFMA:
// out[m][co] = act(bias[co] + sum_tap sum_ci in[m + off(tap)][ci] * w[tap][co][ci])
// A block computes 128 pixels x BN output channels (4 waves x 32 pixels;
// each wave 2 x BN/16 MFMA tiles of 16 x 16); the reduction runs over
// (tap, 32-channel slice) stages, double-buffered in LDS with the next
// stage's global loads issued before the current stage's MFMAs.
// grid (ceil(M / 128), Cout / BN), block 256; Cin % 32 == 0, Cout % BN == 0.
template <int BN>
__global__ __launch_bounds__(256) void k_conv3x3_t(const uint16_t *__restrict__ in,
                                                   const uint16_t *__restrict__ w,
                                                   const float *__restrict__ bias,
                                                   uint16_t *__restrict__ out, int N, int H, int W,
                                                   int Cin, int Cout, int relu)
{
    constexpr int BM = 128, LK = 40;               // pixels per block, LDS row pitch (bf16)
    __shared__ uint16_t sA[2][BM * LK];
    __shared__ uint16_t sB[2][BN * LK];
    const long M = (long)N * H * W;
    const long m0 = (long)blockIdx.x * BM;
    const int co0 = blockIdx.y * BN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    // A staging: pixel am = t / 2, channels [16 * (t % 2), +16): two 16-byte words
    const int am = t >> 1, ahalf = t & 1;
    const long m = m0 + am;
    int an = 0, ay = 0, ax = 0;
    if (m < M) {
        an = (int)(m / ((long)H * W));
        const int r = (int)(m - (long)an * H * W);
        ay = r / W;
        ax = r - ay * W;
    }
    // B staging: output channel bco = t / 4, 8 channels at 8 * (t % 4)
    const bool bstage = t < BN * 4;
    const int bco = t >> 2, bpart = t & 3;
    const int nci = Cin / 32, nstage = 9 * nci;
    uint4 va0, va1, vb;
    auto load = [&](int st) {
        const int tap = st / nci, ci0 = (st - tap * nci) * 32;
        const int yy = ay + tap / 3 - 1, xx = ax + tap % 3 - 1;
        va0 = va1 = make_uint4(0, 0, 0, 0);
        if (m < M && yy >= 0 && yy < H && xx >= 0 && xx < W) {
            const uint4 *src = reinterpret_cast<const uint4 *>(
                in + (((long)an * H + yy) * W + xx) * Cin + ci0 + 16 * ahalf);
            va0 = src[0];
            va1 = src[1];
        }
        if (bstage)
            vb = *reinterpret_cast<const uint4 *>(w + ((long)tap * Cout + co0 + bco) * Cin + ci0 +
                                                  8 * bpart);
    };
    auto store = [&](int buf) {
        uint4 *da = reinterpret_cast<uint4 *>(&sA[buf][am * LK + 16 * ahalf]);
        da[0] = va0;
        da[1] = va1;
        if (bstage) *reinterpret_cast<uint4 *>(&sB[buf][bco * LK + 8 * bpart]) = vb;
    };
    f32x4 acc[2][BN / 16];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int b = 0; b < BN / 16; ++b) acc[i][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    load(0);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int st = 0; st < nstage; ++st) {
        const bool more = st + 1 < nstage;
        if (more) load(st + 1);
        bf16x8 a[2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
            a[i] = *reinterpret_cast<const bf16x8 *>(
                &sA[buf][(wave * 32 + i * 16 + (lane & 15)) * LK + 8 * (lane >> 4)]);
#pragma unroll
        for (int b = 0; b < BN / 16; ++b) {
            const bf16x8 bb = *reinterpret_cast<const bf16x8 *>(
                &sB[buf][(b * 16 + (lane & 15)) * LK + 8 * (lane >> 4)]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
                acc[i][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], bb, acc[i][b], 0, 0, 0);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int b = 0; b < BN / 16; ++b) {
        const int co = co0 + b * 16 + (lane & 15);
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const long mm = m0 + wave * 32 + i * 16 + 4 * (lane >> 4) + e;
                if (mm < M) {
                    float v = acc[i][b][e] + bv;
                    if (relu) v = fmaxf(v, 0.0f);
                    out[mm * Cout + co] = f2bf(v);
                }
            }
    }
}

// Weight-gradient partials of TG taps on a BCO x BCI tile of (co, ci):
// part[split][tap][co][ci] = sum over the split's pixels m of
// g[m][co] * in[m + off(tap)][ci].  grid (9 / TG * Cout/BCO * Cin/BCI,
// splits), block 256; TG = 3 takes the three taps of one kernel row, which
// share the g tile and read the same input rows.  The reduction runs over
// pixels in stages of KM; each stage's g tile and TG tap-shifted input tiles
// are staged in LDS transposed to [channel][pixel] (the MFMA operands take 8
// consecutive pixels per lane): a staging item is 4 pixels x 4 channels,
// written as 4 eight-byte rows, dealt to the 256 threads in turn.  Each wave
// then runs KM / 32 MFMA k-steps on each of its tiles, in pixel order, so
// the sums depend on neither KM nor TG.  The next stage's loads are issued
// before this stage's MFMAs (two LDS buffers, one barrier per stage).
// Requires W % 4 == 0 and chunk % 4 == 0.
template <int BCO, int BCI, int KM, int TG>
__global__ __launch_bounds__(256) void k_conv3x3_dw2(const uint16_t *__restrict__ in,
                                                     const uint16_t *__restrict__ g,
                                                     float *__restrict__ part, int N, int H, int W,
                                                     int Cin, int Cout, long chunk)
{
    constexpr int LP = KM + 8;                     // LDS row pitch (bf16)
    constexpr int NTT = (BCO / 16) * (BCI / 16);   // 16x16 output tiles per tap
    constexpr int TPW = TG * NTT / 4;              // tiles per wave
    constexpr int NGI = (KM / 4) * (BCO / 4);      // g items per stage
    constexpr int NXT = (KM / 4) * (BCI / 4);      // input items per tap and stage
    constexpr int IG = (NGI + 255) / 256, IX = (TG * NXT + 255) / 256;
    static_assert(TPW * 4 == TG * NTT, "tiles must fill the waves");
    __shared__ uint16_t sG[2][BCO * LP];
    __shared__ uint16_t sX[2][TG][BCI * LP];
    const int nco = Cout / BCO, nci = Cin / BCI;
    const int bx = blockIdx.x;
    const int tap0 = (bx / (nco * nci)) * TG;
    const int co0 = ((bx / nci) % nco) * BCO, ci0 = (bx % nci) * BCI;
    const long M = (long)N * H * W;
    const long mb = (long)blockIdx.y * chunk, me = min(M, mb + chunk);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    // g item it: pixel group it / (BCO / 4), channel group it % (BCO / 4);
    // input item it: tap tap0 + it / NXT, then the same split of it % NXT
    uint2 vg[IG][4], vx[IX][4];
    auto load = [&](long mc) {
#pragma unroll
        for (int r = 0; r < IG; ++r) {
            const int it = t + 256 * r, ch = 4 * (it % (BCO / 4));
            const long m = mc + 4 * (it / (BCO / 4));
#pragma unroll
            for (int p = 0; p < 4; ++p)
                vg[r][p] = (it < NGI && m + p < me)
                               ? *reinterpret_cast<const uint2 *>(g + (m + p) * Cout + co0 + ch)
                               : make_uint2(0, 0);
        }
#pragma unroll
        for (int r = 0; r < IX; ++r) {
            const int it = t + 256 * r, j = it / NXT, rest = it % NXT;
            const int ch = 4 * (rest % (BCI / 4));
            const long m = mc + 4 * (rest / (BCI / 4));   // 4 pixels of one row
#pragma unroll
            for (int p = 0; p < 4; ++p) vx[r][p] = make_uint2(0, 0);
            if (j >= TG || m >= me) continue;
            const int tap = tap0 + j, ky = tap / 3 - 1, kx = tap % 3 - 1;
            const int n = (int)(m / ((long)H * W));
            const int rr = (int)(m - (long)n * H * W);
            const int y = rr / W, x = rr - (rr / W) * W;
            const int yy = y + ky;
            if (yy < 0 || yy >= H) continue;
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const int xx = x + p + kx;
                if (m + p < me && xx >= 0 && xx < W)
                    vx[r][p] = *reinterpret_cast<const uint2 *>(
                        in + (((long)n * H + yy) * W + xx) * Cin + ci0 + ch);
            }
        }
    };
    // channel q of an item: its values at the 4 pixels, one 8-byte row piece
    auto put = [&](uint16_t *dst, const uint2 (&v)[4], int ch, int pg) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t e0 = ((q < 2 ? v[0].x : v[0].y) >> (16 * (q & 1))) & 0xffffu;
            const uint32_t e1 = ((q < 2 ? v[1].x : v[1].y) >> (16 * (q & 1))) & 0xffffu;
            const uint32_t e2 = ((q < 2 ? v[2].x : v[2].y) >> (16 * (q & 1))) & 0xffffu;
            const uint32_t e3 = ((q < 2 ? v[3].x : v[3].y) >> (16 * (q & 1))) & 0xffffu;
            *reinterpret_cast<uint2 *>(&dst[(ch + q) * LP + 4 * pg]) =
                make_uint2(e0 | (e1 << 16), e2 | (e3 << 16));
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int r = 0; r < IG; ++r) {
            const int it = t + 256 * r;
            if (it < NGI) put(sG[buf], vg[r], 4 * (it % (BCO / 4)), it / (BCO / 4));
        }
#pragma unroll
        for (int r = 0; r < IX; ++r) {
            const int it = t + 256 * r, j = it / NXT, rest = it % NXT;
            if (j < TG) put(sX[buf][j], vx[r], 4 * (rest % (BCI / 4)), rest / (BCI / 4));
        }
    };
    f32x4 acc[TPW];
#pragma unroll
    for (int k = 0; k < TPW; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};
    int buf = 0;
    load(mb);
    store(0);
    __syncthreads();
    for (long mc = mb; mc < me; mc += KM) {
        const bool more = mc + KM < me;
        if (more) load(mc + KM);
#pragma unroll
        for (int kk = 0; kk < KM / 32; ++kk)
#pragma unroll
            for (int k = 0; k < TPW; ++k) {
                const int tile = wave + 4 * k, j = tile / NTT, rem = tile % NTT;
                const int tco = rem / (BCI / 16), tci = rem % (BCI / 16);
                const bf16x8 a = *reinterpret_cast<const bf16x8 *>(
                    &sG[buf][(tco * 16 + (lane & 15)) * LP + 32 * kk + 8 * (lane >> 4)]);
                const bf16x8 b = *reinterpret_cast<const bf16x8 *>(
                    &sX[buf][j][(tci * 16 + (lane & 15)) * LP + 32 * kk + 8 * (lane >> 4)]);
                acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[k], 0, 0, 0);
            }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    const long tap_sz = (long)Cout * Cin;
#pragma unroll
    for (int k = 0; k < TPW; ++k) {
        const int tile = wave + 4 * k, j = tile / NTT, rem = tile % NTT;
        const int tco = rem / (BCI / 16), tci = rem % (BCI / 16);
        float *dst = part + (long)blockIdx.y * 9 * tap_sz + (long)(tap0 + j) * tap_sz;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int co = co0 + tco * 16 + 4 * (lane >> 4) + e;
            const int ci = ci0 + tci * 16 + (lane & 15);
            dst[(long)co * Cin + ci] = acc[k][e];
        }
    }
}

// out[c] = sum_s part[s][c] for few columns and many splits: one block per
// column, threads take every 256th split, then a fixed-order tree
__global__ __launch_bounds__(256) void k_sum_cols(const float *__restrict__ part, int splits, int C,
                                                  float *__restrict__ out)
{
    __shared__ float red[256];
    const int c = blockIdx.x;
    float acc = 0.0f;
    for (int k = threadIdx.x; k < splits; k += 256) acc += part[(long)k * C + c];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

// out[i] = sum_s part[s][i] in a fixed order (deterministic): block 256
// takes 64 consecutive i; its four waves sum the splits s = q, q + 4, ...
// (q = wave) in order, then the four sums are added in wave order.
// grid ceil(n / 64).
__global__ __launch_bounds__(256) void k_sum_splits(const float *__restrict__ part, int splits,
                                                    long n, float *__restrict__ out,
                                                    int accumulate)
{
    __shared__ float red[4][64];
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const long i = (long)blockIdx.x * 64 + lane;
    float s = 0.0f;
    if (i < n) {
#pragma unroll 8
        for (int k = q; k < splits; k += 4) s += part[(long)k * n + i];
    }
    red[q][lane] = s;
    __syncthreads();
    if (q == 0 && i < n) {
        const float t = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
        out[i] = accumulate ? out[i] + t : t;
    }
}

// per-split column sums of g [M][C] (bf16, C % 8 == 0, C <= 256) ->
// part[split][C]: a thread owns 8 consecutive channels (one 16-byte load per
// row) and every (256 / (C / 8))-th row of the split; the threads of one
// channel group are then added in thread order (deterministic).  With RELU,
// g is first masked in place by the ReLU derivative of the layer's output y
// (g = 0 where y <= 0), the backward of the activation, and the masked
// values are summed.
template <bool RELU>
__global__ __launch_bounds__(256) void k_colsum(uint16_t *__restrict__ g,
                                                const uint16_t *__restrict__ y, long M, int C,
                                                long chunk, float *__restrict__ part)
{
    __shared__ float red[256][9];
    const long mb = (long)blockIdx.x * chunk, me = min(M, mb + chunk);
    const int cg = C >> 3;                    // channel groups of 8
    const int rows = 256 / cg;                // rows per block step
    const int t = threadIdx.x;
    const int grp = t % cg, r0 = t / cg;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (r0 < rows) {
        for (long m = mb + r0; m < me; m += rows) {
            uint4 *gp = reinterpret_cast<uint4 *>(g + m * C + grp * 8);
            const uint4 v = *gp;
            uint32_t w[4] = {v.x, v.y, v.z, v.w};
            if constexpr (RELU) {
                const uint4 yv = *reinterpret_cast<const uint4 *>(y + m * C + grp * 8);
                const uint32_t yw[4] = {yv.x, yv.y, yv.z, yv.w};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (!(__uint_as_float(yw[k] << 16) > 0.0f)) w[k] &= 0xffff0000u;
                    if (!(__uint_as_float(yw[k] & 0xffff0000u) > 0.0f)) w[k] &= 0x0000ffffu;
                }
                *gp = make_uint4(w[0], w[1], w[2], w[3]);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                acc[2 * k] += __uint_as_float(w[k] << 16);
                acc[2 * k + 1] += __uint_as_float(w[k] & 0xffff0000u);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) red[t][k] = acc[k];
    __syncthreads();
    if (t < C) {
        const int cgp = t >> 3, k = t & 7;
        float s = 0.0f;
        for (int r = 0; r < rows; ++r) s += red[r * cg + cgp][k];
        part[(long)blockIdx.x * C + t] = s;
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
// Elementwise layers work on 8 channels (one 16-byte word) per thread with
// 32-bit indices: channel counts are multiples of 8 and every tensor of the
// network has fewer than 2^31 elements.
__device__ __forceinline__ void unpack8(const uint4 &v, float f[8])
{
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        f[2 * k] = __uint_as_float(w[k] << 16);
        f[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
    }
}
__device__ __forceinline__ uint4 pack8(const float f[8])
{
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = (uint32_t)f2bf(f[2 * k]) | ((uint32_t)f2bf(f[2 * k + 1]) << 16);
    return make_uint4(w[0], w[1], w[2], w[3]);
}

// 2x2 average pool, NHWC (H, W even)
__global__ void k_avgpool2(const uint16_t *__restrict__ in, uint16_t *__restrict__ out, int N,
                           int H, int W, int C)
{
    const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    const int n8 = N * Ho * Wo * C8;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    const int c8 = i % C8;
    const int p = i / C8;
    const int x = p % Wo, y = (p / Wo) % Ho, b = p / (Wo * Ho);
    const uint4 *src = reinterpret_cast<const uint4 *>(in);
    const int base = ((b * H + 2 * y) * W + 2 * x) * C8 + c8;
    float f0[8], f1[8], f2[8], f3[8], o[8];
    unpack8(src[base], f0);
    unpack8(src[base + C8], f1);
    unpack8(src[base + W * C8], f2);
    unpack8(src[base + W * C8 + C8], f3);
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = 0.25f * (f0[k] + f1[k] + f2[k] + f3[k]);
    reinterpret_cast<uint4 *>(out)[i] = pack8(o);
}

__global__ void k_avgpool2_bwd(const uint16_t *__restrict__ gout, uint16_t *__restrict__ gin,
                               int N, int H, int W, int C, int accumulate)
{
    const int C8 = C / 8;
    const int n8 = N * H * W * C8;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    const int c8 = i % C8;
    const int p = i / C8;
    const int x = p % W, y = (p / W) % H, b = p / (W * H);
    float v[8], o[8];
    unpack8(reinterpret_cast<const uint4 *>(gout)[((b * (H / 2) + y / 2) * (W / 2) + x / 2) * C8 + c8], v);
    if (accumulate) {
        float g0[8];
        unpack8(reinterpret_cast<const uint4 *>(gin)[i], g0);
#pragma unroll
        for (int k = 0; k < 8; ++k) o[k] = g0[k] + 0.25f * v[k];
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) o[k] = 0.25f * v[k];
    }
    reinterpret_cast<uint4 *>(gin)[i] = pack8(o);
}

// nearest 2x upsample of a (N, H, W, C1) tensor into channels [0, C1) of a
// (N, 2H, 2W, C1 + C2) tensor, and copy of skip (N, 2H, 2W, C2) into [C1, C1+C2)
__global__ void k_up_concat(const uint16_t *__restrict__ low, const uint16_t *__restrict__ skip,
                            uint16_t *__restrict__ out, int N, int H, int W, int C1, int C2)
{
    const int Ho = 2 * H, Wo = 2 * W, C8 = (C1 + C2) / 8, L8 = C1 / 8, S8 = C2 / 8;
    const int n8 = N * Ho * Wo * C8;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    const int c8 = i % C8;
    const int p = i / C8;
    const int x = p % Wo, y = (p / Wo) % Ho, b = p / (Wo * Ho);
    reinterpret_cast<uint4 *>(out)[i] =
        (c8 < L8) ? reinterpret_cast<const uint4 *>(low)[((b * H + y / 2) * W + x / 2) * L8 + c8]
                  : reinterpret_cast<const uint4 *>(skip)[p * S8 + (c8 - L8)];
}

// gradient of k_up_concat: 2x2 sums into the low tensor, and the skip part
// added into the skip gradient (accumulate) or written
__global__ void k_up_concat_bwd(const uint16_t *__restrict__ gout, uint16_t *__restrict__ glow,
                                uint16_t *__restrict__ gskip, int N, int H, int W, int C1, int C2,
                                int acc_skip)
{
    const int Ho = 2 * H, Wo = 2 * W, C8 = (C1 + C2) / 8, L8 = C1 / 8, S8 = C2 / 8;
    const int nl = N * H * W * L8;
    const int ns = N * Ho * Wo * S8;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint4 *go = reinterpret_cast<const uint4 *>(gout);
    if (i < nl) {
        const int c8 = i % L8;
        const int p = i / L8;
        const int x = p % W, y = (p / W) % H, b = p / (W * H);
        const int o = ((b * Ho + 2 * y) * Wo + 2 * x) * C8 + c8;
        float f0[8], f1[8], f2[8], f3[8], r[8];
        unpack8(go[o], f0);
        unpack8(go[o + C8], f1);
        unpack8(go[o + Wo * C8], f2);
        unpack8(go[o + Wo * C8 + C8], f3);
#pragma unroll
        for (int k = 0; k < 8; ++k) r[k] = f0[k] + f1[k] + f2[k] + f3[k];
        reinterpret_cast<uint4 *>(glow)[i] = pack8(r);
    }
    if (i < ns) {
        const int c8 = i % S8;
        const int p = i / S8;
        float v[8];
        unpack8(go[p * C8 + L8 + c8], v);
        if (acc_skip) {
            float g0[8];
            unpack8(reinterpret_cast<const uint4 *>(gskip)[i], g0);
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = g0[k] + v[k];
        }
        reinterpret_cast<uint4 *>(gskip)[i] = pack8(v);
    }
}

// 1x1 head, one pixel per thread: f[m] = bh + sum_c x[m][c] * wh[.][c], the
// pixel's NN_HEAD_C bf16 channels read as 16-byte words, summed in channel
// order
__global__ void k_head_fwd(const uint16_t *__restrict__ x, const float *__restrict__ wh,
                           const float *__restrict__ bh, float *__restrict__ f, long M)
{
    constexpr int C = NN_HEAD_C;
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    float a = bh[0], b = bh[1];
    const uint4 *src = reinterpret_cast<const uint4 *>(x + m * C);
#pragma unroll
    for (int q = 0; q < C / 8; ++q) {
        const uint4 u = src[q];
        const uint32_t wd[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int c = 8 * q + e;
            const float v = bf2f((uint16_t)(wd[e >> 1] >> (16 * (e & 1))));
            a += v * wh[c];
            b += v * wh[C + c];
        }
    }
    f[2 * m] = a;
    f[2 * m + 1] = b;
}

// head gradients: gx[m][c] = sum_j gf[m][j] wh[j][c]; per-split partials of
// dwh[j][c] = sum_m gf[m][j] x[m][c] and dbh[j] = sum_m gf[m][j].  A thread
// takes every 256th pixel of the block's split (16-byte loads and stores of
// the pixel's channels), then each partial is summed by a fixed butterfly
// over the wave and the four waves in order (deterministic).
__global__ __launch_bounds__(256) void k_head_bwd(const uint16_t *__restrict__ x,
                                                  const float *__restrict__ gf,
                                                  const float *__restrict__ wh,
                                                  uint16_t *__restrict__ gx, float *__restrict__ part,
                                                  long M, long chunk)
{
    constexpr int C = NN_HEAD_C, NP = 2 * C + 2;
    __shared__ float red[4][NP];
    const long mb = (long)blockIdx.x * chunk, me = min(M, mb + chunk);
    float accw[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) accw[k] = 0.0f;
    for (long m = mb + threadIdx.x; m < me; m += 256) {
        const float g0 = gf[2 * m], g1 = gf[2 * m + 1];
        const uint4 *src = reinterpret_cast<const uint4 *>(x + m * C);
        uint4 *dst = reinterpret_cast<uint4 *>(gx + m * C);
#pragma unroll
        for (int q = 0; q < C / 8; ++q) {
            const uint4 u = src[q];
            const uint32_t wd[4] = {u.x, u.y, u.z, u.w};
            uint32_t o[4];
#pragma unroll
            for (int e = 0; e < 8; e += 2) {
                const int c = 8 * q + e;
                const float v0 = bf2f((uint16_t)(wd[e >> 1] & 0xffffu));
                const float v1 = bf2f((uint16_t)(wd[e >> 1] >> 16));
                o[e >> 1] = (uint32_t)f2bf(g0 * wh[c] + g1 * wh[C + c]) |
                            ((uint32_t)f2bf(g0 * wh[c + 1] + g1 * wh[C + c + 1]) << 16);
                accw[c] += g0 * v0;
                accw[C + c] += g1 * v0;
                accw[c + 1] += g0 * v1;
                accw[C + c + 1] += g1 * v1;
            }
            dst[q] = make_uint4(o[0], o[1], o[2], o[3]);
        }
        accw[2 * C] += g0;
        accw[2 * C + 1] += g1;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        float v = accw[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < NP) {
        const int k = threadIdx.x;
        part[(long)blockIdx.x * NP + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
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
        hipLaunchKernelGGL(k_conv3x3_t<64>, dim3(nblk(M, 128), Cout / 64), dim3(256), 0, s, in, w,
                           bias, out, N, H, W, Cin, Cout, relu);
    else
        hipLaunchKernelGGL(k_conv3x3_t<32>, dim3(nblk(M, 128), Cout / 32), dim3(256), 0, s, in, w,
                           bias, out, N, H, W, Cin, Cout, relu);
}

// pixel splits of a weight gradient: about 8192 blocks per launch, at most
// 256 splits and at least 1024 pixels per split
int nn_dw_splits(long M, int Cin, int Cout)
{
    const int bco = (Cout % 64 == 0) ? 64 : 32, bci = (Cin % 64 == 0) ? 64 : 32;
    const long per_split = 9L * (Cout / bco) * (Cin / bci);
    const long want = std::max<long>(1, 8192 / per_split);
    return (int)std::max<long>(1, std::min<long>({256, want, M / 1024}));
}

void nn_conv3x3_dw(hipStream_t s, const uint16_t *in, const uint16_t *g, float *dw, float *work,
                   int N, int H, int W, int Cin, int Cout, int accumulate)
{
    const long M = (long)N * H * W;
    const int splits = nn_dw_splits(M, Cin, Cout);
    const int bco = (Cout % 64 == 0) ? 64 : 32, bci = (Cin % 64 == 0) ? 64 : 32;
    // 64 pixels per stage; a kernel row of three taps per block, but one tap
    // for the 64 x 64 tile (LDS for four blocks per CU)
    const long chunk = ((M + splits - 1) / splits + 63) / 64 * 64;
    const int tg = (bco == 64 && bci == 64) ? 1 : 3;
    const dim3 grid(9 / tg * (Cout / bco) * (Cin / bci), splits);
    if (bco == 64 && bci == 64)
        hipLaunchKernelGGL((k_conv3x3_dw2<64, 64, 64, 1>), grid, dim3(256), 0, s, in, g, work, N, H,
                           W, Cin, Cout, chunk);
    else if (bco == 64)
        hipLaunchKernelGGL((k_conv3x3_dw2<64, 32, 64, 3>), grid, dim3(256), 0, s, in, g, work, N, H,
                           W, Cin, Cout, chunk);
    else if (bci == 64)
        hipLaunchKernelGGL((k_conv3x3_dw2<32, 64, 64, 3>), grid, dim3(256), 0, s, in, g, work, N, H,
                           W, Cin, Cout, chunk);
    else
        hipLaunchKernelGGL((k_conv3x3_dw2<32, 32, 64, 3>), grid, dim3(256), 0, s, in, g, work, N, H,
                           W, Cin, Cout, chunk);
    const long n = 9L * Cout * Cin;
    hipLaunchKernelGGL(k_sum_splits, dim3(nblk(n, 64)), dim3(256), 0, s, work, splits, n, dw,
                       accumulate);
}

int nn_colsum_splits(long M) { return (int)std::min<long>(1024, std::max<long>(1, M / 1024)); }

void nn_colsum(hipStream_t s, uint16_t *g, long M, int C, float *db, float *work,
               const uint16_t *relu_y)
{
    const int splits = nn_colsum_splits(M);
    const long chunk = (M + splits - 1) / splits;
    if (relu_y)
        hipLaunchKernelGGL(k_colsum<true>, dim3(splits), dim3(256), 0, s, g, relu_y, M, C, chunk,
                           work);
    else
        hipLaunchKernelGGL(k_colsum<false>, dim3(splits), dim3(256), 0, s, g, nullptr, M, C, chunk,
                           work);
    hipLaunchKernelGGL(k_sum_cols, dim3(C), dim3(256), 0, s, work, splits, C, db);
}

void nn_weights_bf16(hipStream_t s, const float *w, int Cin, int Cout, uint16_t *wf, uint16_t *wb)
{
    hipLaunchKernelGGL(k_weights_bf16, dim3(nblk(9L * Cin * Cout)), dim3(256), 0, s, w, Cin, Cout, wf, wb);
}

void nn_avgpool2(hipStream_t s, const uint16_t *in, uint16_t *out, int N, int H, int W, int C)
{
    hipLaunchKernelGGL(k_avgpool2, dim3(nblk((long)N * (H / 2) * (W / 2) * C / 8)), dim3(256), 0, s, in,
                       out, N, H, W, C);
}

void nn_avgpool2_bwd(hipStream_t s, const uint16_t *gout, uint16_t *gin, int N, int H, int W, int C,
                     int accumulate)
{
    hipLaunchKernelGGL(k_avgpool2_bwd, dim3(nblk((long)N * H * W * C / 8)), dim3(256), 0, s, gout, gin, N,
                       H, W, C, accumulate);
}

void nn_up_concat(hipStream_t s, const uint16_t *low, const uint16_t *skip, uint16_t *out, int N,
                  int H, int W, int C1, int C2)
{
    hipLaunchKernelGGL(k_up_concat, dim3(nblk((long)N * 4 * H * W * (C1 + C2) / 8)), dim3(256), 0, s,
                       low, skip, out, N, H, W, C1, C2);
}

void nn_up_concat_bwd(hipStream_t s, const uint16_t *gout, uint16_t *glow, uint16_t *gskip, int N,
                      int H, int W, int C1, int C2, int acc_skip)
{
    const long n = std::max((long)N * H * W * C1, (long)N * 4 * H * W * C2) / 8;
    hipLaunchKernelGGL(k_up_concat_bwd, dim3(nblk(n)), dim3(256), 0, s, gout, glow, gskip, N, H, W,
                       C1, C2, acc_skip);
}

void nn_head_fwd(hipStream_t s, const uint16_t *x, const float *wh, const float *bh, float *f,
                 long M)
{
    hipLaunchKernelGGL(k_head_fwd, dim3(nblk(M)), dim3(256), 0, s, x, wh, bh, f, M);
}

int nn_head_splits(long M) { return (int)std::min<long>(1024, std::max<long>(1, M / 1024)); }

void nn_head_bwd(hipStream_t s, const uint16_t *x, const float *gf, const float *wh, uint16_t *gx,
                 float *dwh_db, float *work, long M)
{
    const int splits = nn_head_splits(M);
    const long chunk = (M + splits - 1) / splits;
    constexpr long NP = 2 * NN_HEAD_C + 2;
    hipLaunchKernelGGL(k_head_bwd, dim3(splits), dim3(256), 0, s, x, gf, wh, gx, work, M, chunk);
    hipLaunchKernelGGL(k_sum_splits, dim3(nblk(NP, 64)), dim3(256), 0, s, work, splits, NP, dwh_db,
                       0);
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
