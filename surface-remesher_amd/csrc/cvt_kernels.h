// cvt_kernels.h -- CDNA4 (gfx950) kernels of the CVT Lloyd iteration.
//
// Data layout in HBM (one grid of W x H pixels, S sites):
//   rho      fp32 [H][W]      density, unnormalised (scaled by rho_scale = 2^-k
//                             in-kernel so that every scaled value is <= 1)
//   seed     uint32 [H][W]    seed map: label of the site on that pixel, `none`
//                             elsewhere; read by JFA pass 0, cleared by pass 1,
//                             re-seeded by the site update
//   lab[2]   uint32 [H][W]    JFA ping-pong label maps
//   sites    float2 [capS+1]  site table, pixel units; sites[capS] = (inf, inf)
//                             is the sentinel
// A label is the byte offset of its site in the table (8 * index) and `none`
// = 8 * capS, so the JFA gather needs no address arithmetic; labels leave the
// engine as int32 indices (CVT_NONE for none).
//   slots    Slot [S][9]      per-site partial moments, one per tile of the 3x3
//                             tile window around the site's home tile (unique
//                             writer per slot, tagged with the epoch)
//   ovf      OvfRec [tiles][cap]  partial moments of tiles outside that window
//
// The moment reduction is deterministic and uses no global atomics: per
// 64-pixel row segment a wavefront-level segmented scan sums each run of equal
// labels; run sums enter an LDS hash table as 64-bit fixed point (integer
// addition is associative, so the LDS atomics' order cannot change the
// result); the table is flushed once per tile to unique-writer slots.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define CVT_NONE 0x7fffffff

// Fixed-point moment sums of one (tile, site) pair: W = sum rho', DX = sum
// rho' (px - sx), DY = sum rho' (py - sy), E = sum rho' |p - s|^2, scaled by
// 2^MOM_FW, 2^MOM_FX, 2^MOM_FX, 2^MOM_FE.  rho' = rho * 2^-k <= 1, a tile has
// at most 256^2 = 2^16 pixels and |p - s| < 2^15 (sites stay inside
// [-W, 2W] x [-H, 2H], W, H <= 8192), so each tile sum fits in int64; the
// per-site totals are summed in 128 bits.  Integer sums are exact, hence
// independent of order and grouping.
#define MOM_FW 46
#define MOM_FX 31
#define MOM_FE 15

struct __align__(8) Slot {
    long long w, dx, dy, e;
    int epoch;
    int pad;
};

struct __align__(8) OvfRec {
    long long w, dx, dy, e;
    int label;
    int pad;
};

struct __align__(16) Acc128 {   // one 128-bit accumulator per moment
    __int128 v[4];
};

// device flags
enum { FLAG_OVF = 0, FLAG_FULL = 1, FLAG_BAD = 2, NUM_FLAGS = 4 };
// device state words
enum { ST_EPOCH = 0, ST_HIST = 1, NUM_STATE = 4 };

// ---------------------------------------------------------------------------
// distance order (must round exactly like the oracle's orc_dist / orc_better)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float cvt_dist(float2 s, float px, float py)
{
    const float dx = s.x - px;
    const float dy = s.y - py;
    return __builtin_fmaf(dx, dx, dy * dy);
}

// ---------------------------------------------------------------------------
// fills, seeding
// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t v, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__device__ __forceinline__ int cvt_pixel_of(float v, int n)
{
    float f = floorf(v);
    f = fminf(fmaxf(f, 0.0f), (float)(n - 1));   // NaN -> 0
    return (int)f;
}

// Sites must be finite and inside [-W, 2W] x [-H, 2H] (keeps |p - s| < 2^15
// for the fixed-point moments); a violation raises FLAG_BAD.
__global__ void k_site_check(const float2 *__restrict__ sites, int S, int W, int H,
                             int *__restrict__ dflags)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    const float2 s = sites[i];
    const bool ok = s.x >= -(float)W && s.x <= 2.0f * W && s.y >= -(float)H && s.y <= 2.0f * H;
    if (!ok) dflags[FLAG_BAD] = 1;
}

// Every site claims its pixel; the lowest index wins a shared pixel.  The
// integer atomicMin makes the result independent of thread order.
__global__ void k_seed(const float2 *__restrict__ sites, int S, uint32_t *__restrict__ seed,
                       int W, int H)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    float2 s = sites[i];
    int x = cvt_pixel_of(s.x, W), y = cvt_pixel_of(s.y, H);
    atomicMin(&seed[(long)y * W + x], 8u * (uint32_t)i);
}

// engine labels (byte offsets) -> int32 indices, CVT_NONE for none
__global__ void k_export_labels(const uint32_t *__restrict__ lab, int32_t *__restrict__ out, long n,
                                uint32_t none)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const uint32_t l = lab[i];
        out[i] = (l >= none) ? CVT_NONE : (int32_t)(l >> 3);
    }
}

// ---------------------------------------------------------------------------
// moments (reference replacement for pbaCVDComputeCentroid + the energy
// texture, source/gcvt.cu:562-732, :788-802)
// ---------------------------------------------------------------------------
template <int T> struct TileCfg {
    static constexpr int LOG = (T == 32) ? 5 : (T == 64) ? 6 : (T == 128) ? 7 : 8;
    static constexpr int SEG = T < 64 ? T : 64;       // pixels per row segment
    static constexpr int RPW = 64 / SEG;              // rows per wave step
    static constexpr int SPR = T / SEG;               // segments per row
    static constexpr int GROUPS = T * T / 64;         // wave steps per tile
};

#define CVT_HASH_CAP 1024

__device__ __forceinline__ long long cvt_fix(float v, int f)
{
    return __double2ll_rn(ldexp((double)v, f));
}

template <int T>
__global__ __launch_bounds__(256) void k_moments(
    const uint32_t *__restrict__ lab, const float *__restrict__ rho,
    const float2 *__restrict__ sites, int S, int W, int H, float rho_scale, int tilesX,
    int tilesY, Slot *__restrict__ slots, OvfRec *__restrict__ ovf, int ovf_cap,
    int *__restrict__ ovf_count, int *__restrict__ dflags, const int *__restrict__ dstate)
{
    using C = TileCfg<T>;
    __shared__ int keys[CVT_HASH_CAP];
    __shared__ unsigned long long acc[CVT_HASH_CAP][4];
    __shared__ int n_over, full;

    const int tile = blockIdx.x;
    const int tX = tile % tilesX, tY = tile / tilesX;
    const int tx0 = tX * T, ty0 = tY * T;
    for (int j = threadIdx.x; j < CVT_HASH_CAP; j += 256) {
        keys[j] = -1;
        acc[j][0] = acc[j][1] = acc[j][2] = acc[j][3] = 0ull;
    }
    if (threadIdx.x == 0) { n_over = 0; full = 0; }
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int seg_lane = lane % C::SEG;

    for (int g = wave; g < C::GROUPS; g += 4) {
        const int r = (g / C::SPR) * C::RPW + lane / C::SEG;
        const int c = (g % C::SPR) * C::SEG + seg_lane;
        const int x = tx0 + c, y = ty0 + r;
        int key = -1;
        float a = 0.f, b = 0.f, cc = 0.f, e = 0.f;
        if (x < W && y < H) {
            const long p = (long)y * W + x;
            const int l = (int)(lab[p] >> 3);
            const float rv = rho[p];
            if ((unsigned)l < (unsigned)S && rv != 0.0f) {
                key = l;
                const float2 st = sites[l];
                const float rs = rv * rho_scale;
                const float dx = ((float)x + 0.5f) - st.x;
                const float dy = ((float)y + 0.5f) - st.y;
                a = rs;
                b = rs * dx;
                cc = rs * dy;
                e = rs * __builtin_fmaf(dx, dx, dy * dy);
            }
        }
        // segmented inclusive scan over runs of equal key within the segment
        const int kprev = __shfl_up(key, 1);
        const bool head = (seg_lane == 0) || (kprev != key);
        int start = head ? lane : 0;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int o = __shfl_up(start, off);
            if (lane >= off) start = max(start, o);
        }
#pragma unroll
        for (int off = 1; off < C::SEG; off <<= 1) {
            const float ua = __shfl_up(a, off), ub = __shfl_up(b, off);
            const float uc = __shfl_up(cc, off), ue = __shfl_up(e, off);
            if (lane - off >= start) { a += ua; b += ub; cc += uc; e += ue; }
        }
        const int knext = __shfl_down(key, 1);
        const bool tail = (seg_lane == C::SEG - 1) || (knext != key);
        if (tail && key >= 0) {
            const unsigned hsh = ((unsigned)key * 2654435761u) >> (32 - 10);
            int slot = -1;
            for (int probe = 0; probe < CVT_HASH_CAP; ++probe) {
                const int h = (hsh + probe) & (CVT_HASH_CAP - 1);
                const int old = atomicCAS(&keys[h], -1, key);
                if (old == -1 || old == key) { slot = h; break; }
            }
            if (slot >= 0) {
                atomicAdd(&acc[slot][0], (unsigned long long)cvt_fix(a, MOM_FW));
                atomicAdd(&acc[slot][1], (unsigned long long)cvt_fix(b, MOM_FX));
                atomicAdd(&acc[slot][2], (unsigned long long)cvt_fix(cc, MOM_FX));
                atomicAdd(&acc[slot][3], (unsigned long long)cvt_fix(e, MOM_FE));
            } else {
                full = 1;
            }
        }
    }
    __syncthreads();

    const int epoch = dstate[ST_EPOCH];
    for (int j = threadIdx.x; j < CVT_HASH_CAP; j += 256) {
        const int l = keys[j];
        if (l < 0) continue;
        const float2 s = sites[l];
        const long long wt = (long long)acc[j][0], dx = (long long)acc[j][1];
        const long long dy = (long long)acc[j][2], ee = (long long)acc[j][3];
        const int hx = min(max((int)floorf(s.x / (float)T), 0), tilesX - 1);
        const int hy = min(max((int)floorf(s.y / (float)T), 0), tilesY - 1);
        const int rx = tX - hx, ry = tY - hy;
        if (rx >= -1 && rx <= 1 && ry >= -1 && ry <= 1) {
            Slot sl;
            sl.w = wt; sl.dx = dx; sl.dy = dy; sl.e = ee; sl.epoch = epoch; sl.pad = 0;
            slots[(long)l * 9 + (ry + 1) * 3 + (rx + 1)] = sl;
        } else {
            const int k = atomicAdd(&n_over, 1);
            if (k < ovf_cap) {
                OvfRec o;
                o.w = wt; o.dx = dx; o.dy = dy; o.e = ee; o.label = l; o.pad = 0;
                ovf[(long)tile * ovf_cap + k] = o;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ovf_count[tile] = n_over;
        if (n_over) dflags[FLAG_OVF] = 1;
        if (full || n_over > ovf_cap) dflags[FLAG_FULL] = 1;
    }
}

// Overflow records (cells reaching beyond the 3x3 window of their home tile)
// are added to ovf_acc in tile order by a single workgroup: deterministic, and
// free when no tile overflowed.
__global__ __launch_bounds__(256) void k_overflow(const OvfRec *__restrict__ ovf, int ovf_cap,
                                                  const int *__restrict__ ovf_count, int ntiles,
                                                  Acc128 *__restrict__ ovf_acc,
                                                  const int *__restrict__ dflags)
{
    if (dflags[FLAG_OVF] == 0) return;
    for (int t = 0; t < ntiles; ++t) {
        const int n = min(ovf_count[t], ovf_cap);
        // records of one tile carry distinct labels: no two threads collide
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            const OvfRec o = ovf[(long)t * ovf_cap + k];
            Acc128 &a = ovf_acc[o.label];
            a.v[0] += o.w; a.v[1] += o.dx; a.v[2] += o.dy; a.v[3] += o.e;
        }
        __syncthreads();
    }
}

// exact 128-bit fixed point -> double (value * 2^-f)
__device__ __forceinline__ double cvt_i128_to_double(__int128 v, int f)
{
    const long long hi = (long long)(v >> 64);
    const unsigned long long lo = (unsigned long long)v;
    return ldexp((double)hi, 64 - f) + ldexp((double)lo, -f);
}

// ---------------------------------------------------------------------------
// site update (reference replacement for pbaCVDUpdateSites,
// source/gcvt.cu:753-781, continuous form) + energy / gradient partials
// ---------------------------------------------------------------------------
enum { UPD_MOVE = 1, UPD_SEED = 2 };

__global__ __launch_bounds__(256) void k_update(
    float2 *__restrict__ sites, const uint8_t *__restrict__ fixed, int S,
    const Slot *__restrict__ slots, Acc128 *__restrict__ ovf_acc,
    const int *__restrict__ dflags, const int *__restrict__ dstate, float omega, int mode,
    double rho_unscale, double *__restrict__ mom_out, double *__restrict__ blockpart,
    uint32_t *__restrict__ seed, int W, int H)
{
    __shared__ double red_e[256], red_g[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0, g = 0.0;
    if (i < S) {
        const int epoch = dstate[ST_EPOCH];
        __int128 sw = 0, sx = 0, sy = 0, se = 0;
        const Slot *sl = slots + (long)i * 9;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if (sl[k].epoch == epoch) {
                sw += sl[k].w; sx += sl[k].dx; sy += sl[k].dy; se += sl[k].e;
            }
        }
        if (dflags[FLAG_OVF]) {
            Acc128 &a = ovf_acc[i];
            sw += a.v[0]; sx += a.v[1]; sy += a.v[2]; se += a.v[3];
            a.v[0] = a.v[1] = a.v[2] = a.v[3] = 0;
        }
        const double w = cvt_i128_to_double(sw, MOM_FW);
        const double dx = cvt_i128_to_double(sx, MOM_FX);
        const double dy = cvt_i128_to_double(sy, MOM_FX);
        const double ee = cvt_i128_to_double(se, MOM_FE);
        e = ee;
        g = dx * dx + dy * dy;
        if (mom_out) {
            mom_out[4L * i + 0] = w * rho_unscale;
            mom_out[4L * i + 1] = dx * rho_unscale;
            mom_out[4L * i + 2] = dy * rho_unscale;
            mom_out[4L * i + 3] = ee * rho_unscale;
        }
        float2 s = sites[i];
        if ((mode & UPD_MOVE) && w > 0.0 && !(fixed && fixed[i])) {
            s.x = (float)((double)s.x + (double)omega * (dx / w));
            s.y = (float)((double)s.y + (double)omega * (dy / w));
            sites[i] = s;
        }
        if (mode & UPD_SEED) {
            const int x = cvt_pixel_of(s.x, W), y = cvt_pixel_of(s.y, H);
            atomicMin(&seed[(long)y * W + x], 8u * (uint32_t)i);
        }
    }
    red_e[threadIdx.x] = e;
    red_g[threadIdx.x] = g;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            red_e[threadIdx.x] += red_e[threadIdx.x + o];
            red_g[threadIdx.x] += red_g[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        blockpart[2 * blockIdx.x] = red_e[0];
        blockpart[2 * blockIdx.x + 1] = red_g[0];
    }
}

// Sums the block partials in a fixed order, appends (E, |grad E|) to the
// history when record != 0, advances the epoch and clears the flags.
__global__ __launch_bounds__(256) void k_finalize(const double *__restrict__ blockpart,
                                                  int nblocks, int *__restrict__ dstate,
                                                  int *__restrict__ dflags,
                                                  double *__restrict__ hist, int hist_cap,
                                                  int record, double e_scale, double g_scale,
                                                  double *__restrict__ out_eg)
{
    __shared__ double re[256], rg[256];
    double e = 0.0, g = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256) {
        e += blockpart[2 * b];
        g += blockpart[2 * b + 1];
    }
    re[threadIdx.x] = e;
    rg[threadIdx.x] = g;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            re[threadIdx.x] += re[threadIdx.x + o];
            rg[threadIdx.x] += rg[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double E = re[0] * e_scale;
        const double G = sqrt(rg[0]) * g_scale;
        if (record) {
            const int h = dstate[ST_HIST];
            if (h < hist_cap) { hist[2 * h] = E; hist[2 * h + 1] = G; }
            dstate[ST_HIST] = h + 1;
        }
        if (out_eg) { out_eg[0] = E; out_eg[1] = G; }
        dstate[ST_EPOCH] = dstate[ST_EPOCH] + 1;
        if (dflags[FLAG_OVF]) dflags[FLAG_OVF] = 0;
    }
}

// ---------------------------------------------------------------------------
// float64 energy over pixels (the fp64 convergence check)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_energy64(const uint32_t *__restrict__ lab,
                                                  const float *__restrict__ rho,
                                                  const float2 *__restrict__ sites, int S,
                                                  int W, int H, double *__restrict__ blockpart)
{
    __shared__ double red[256];
    const long n = (long)W * H;
    double e = 0.0;
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long)gridDim.x * 256) {
        const int l = (int)(lab[p] >> 3);
        const float r = rho[p];
        if ((unsigned)l < (unsigned)S && r != 0.0f) {
            const int y = (int)(p / W), x = (int)(p - (long)y * W);
            const float2 s = sites[l];
            const double dx = ((double)x + 0.5) - (double)s.x;
            const double dy = ((double)y + 0.5) - (double)s.y;
            e += (double)r * (dx * dx + dy * dy);
        }
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) { blockpart[2 * blockIdx.x] = red[0]; blockpart[2 * blockIdx.x + 1] = 0.0; }
}

// density check: max over rho and a flag for NaN / inf / negative entries
__global__ __launch_bounds__(256) void k_density_stats(const float *__restrict__ rho, long n,
                                                       float *__restrict__ block_max,
                                                       int *__restrict__ dflags)
{
    __shared__ float red[256];
    float m = 0.0f;
    bool bad = false;
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long)gridDim.x * 256) {
        const float r = rho[p];
        if (!(r >= 0.0f) || r == __builtin_inff()) bad = true;
        else m = fmaxf(m, r);
    }
    if (bad) dflags[FLAG_BAD] = 1;
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) block_max[blockIdx.x] = red[0];
}
