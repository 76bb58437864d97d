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
//   ovf      OvfRec [cap]     partial moments of tiles outside that window,
//                             chained per site from ovf_head[S]
//
// The moment reduction is deterministic and uses no global atomics: per
// 64-pixel row segment a wavefront-level segmented scan sums each run of equal
// labels; run sums enter an LDS hash table as 64-bit fixed point (integer
// addition is associative, so the LDS atomics' order cannot change the
// result); the table is flushed once per tile to unique-writer slots.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cvt_types.h"

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
static __global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t v, long n)
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
static __global__ void k_site_check(const float2 *__restrict__ sites, int S, int W, int H,
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
static __global__ void k_seed(const float2 *__restrict__ sites, int S, uint32_t *__restrict__ seed,
                       int W, int H)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    float2 s = sites[i];
    int x = cvt_pixel_of(s.x, W), y = cvt_pixel_of(s.y, H);
    atomicMin(&seed[(long)y * W + x], 8u * (uint32_t)i);
}

// engine labels (byte offsets) -> int32 indices, CVT_NONE for none
static __global__ void k_export_labels(const uint32_t *__restrict__ lab, int32_t *__restrict__ out, long n,
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
    static constexpr int SEG = T < 64 ? T : 64;       // pixels per row segment
    static constexpr int RPW = 64 / SEG;              // rows per wave step
    static constexpr int SPR = T / SEG;               // segments per row
    static constexpr int GROUPS = T * T / 64;         // wave steps per tile
};

#define CVT_HASH_CAP 1024

__device__ __forceinline__ long long cvt_fix(double v, int f)
{
    return __double2ll_rn(ldexp(v, f));
}

// --- cross-lane moves (DPP, no LDS traffic) --------------------------------
// value of lane - 1 over the whole wave (lane 0 gets fill)
__device__ __forceinline__ int dpp_wave_shr1(int v, int fill)
{
    return __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xf, 0xf, false);
}
// value of lane + 1 over the whole wave (lane 63 gets fill)
__device__ __forceinline__ int dpp_wave_shl1(int v, int fill)
{
    return __builtin_amdgcn_update_dpp(fill, v, 0x130, 0xf, 0xf, false);
}
// value of lane - K inside the 16-lane row (0 where that leaves the row)
template <int K> __device__ __forceinline__ int dpp_row_shr(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, 0x110 + K, 0xf, 0xf, false);
}
// lane 15 of rows 0 / 2 to every lane of rows 1 / 3 (0 elsewhere)
__device__ __forceinline__ int dpp_bcast15(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
}
// lane 31 to every lane of rows 2 and 3 (0 elsewhere)
__device__ __forceinline__ int dpp_bcast31(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
}
__device__ __forceinline__ float fbits(int v) { return __int_as_float(v); }
__device__ __forceinline__ int ibits(float v) { return __float_as_int(v); }

// One step of a segmented inclusive scan: add the value from lane `src` when
// that lane lies in this lane's run (src >= start).  t is the moved value.
__device__ __forceinline__ float seg_add(float v, int t, int src, int start)
{
    return (src >= start) ? v + fbits(t) : v;
}

template <int T>
static __global__ __launch_bounds__(256) void k_moments(
    const uint32_t *__restrict__ lab, const float *__restrict__ rho,
    const float2 *__restrict__ sites, int S, int W, int H, float rho_scale, int tilesX,
    int tilesY, Slot *__restrict__ slots, OvfRec *__restrict__ ovf, int ovf_cap,
    int *__restrict__ ovf_head, int *__restrict__ dflags, int *__restrict__ dstate)
{
    using C = TileCfg<T>;
    __shared__ int keys[CVT_HASH_CAP];
    __shared__ unsigned long long acc[CVT_HASH_CAP][4];
    __shared__ int full;

    const int tile = blockIdx.x;
    const int tX = tile % tilesX, tY = tile / tilesX;
    const int tx0 = tX * T, ty0 = tY * T;
    for (int j = threadIdx.x; j < CVT_HASH_CAP; j += 256) {
        keys[j] = -1;
        acc[j][0] = acc[j][1] = acc[j][2] = acc[j][3] = 0ull;
    }
    if (threadIdx.x == 0) full = 0;
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int seg_lane = lane % C::SEG;
    const int row16 = lane & 15;

    for (int g = wave; g < C::GROUPS; g += 4) {
        const int r = (g / C::SPR) * C::RPW + lane / C::SEG;
        const int c0 = (g % C::SPR) * C::SEG;
        const int x = tx0 + c0 + seg_lane, y = ty0 + r;
        int key = -1;
        float rs = 0.0f;
        if (x < W && y < H) {
            const long p = (long)y * W + x;
            const int l = (int)(lab[p] >> 3);
            const float rv = rho[p];
            if ((unsigned)l < (unsigned)S && rv != 0.0f) {
                key = l;
                rs = rv * rho_scale;
            }
        }
        // runs of equal key inside the segment; start = first lane of the run
        const int kprev = dpp_wave_shr1(key, -2);
        const bool head = (seg_lane == 0) || (kprev != key);
        int start = head ? lane : 0;
        start = max(start, dpp_row_shr<1>(start));
        start = max(start, dpp_row_shr<2>(start));
        start = max(start, dpp_row_shr<4>(start));
        start = max(start, dpp_row_shr<8>(start));
        start = max(start, dpp_bcast15(start));
        start = max(start, dpp_bcast31(start));
        // per pixel, relative to the run's first pixel: u = lane - start
        const float u = (float)(lane - start);
        float a = rs, b = rs * u, cc = rs * (u * u);
        // segmented inclusive scan of (a, b, cc)
#define CVT_SCAN_STEP(MOVE, SRC)                                                   \
        {                                                                          \
            const int ta = MOVE(ibits(a)), tb = MOVE(ibits(b)), tc = MOVE(ibits(cc)); \
            a = seg_add(a, ta, SRC, start);                                        \
            b = seg_add(b, tb, SRC, start);                                        \
            cc = seg_add(cc, tc, SRC, start);                                      \
        }
        CVT_SCAN_STEP(dpp_row_shr<1>, row16 >= 1 ? lane - 1 : -1)
        CVT_SCAN_STEP(dpp_row_shr<2>, row16 >= 2 ? lane - 2 : -1)
        CVT_SCAN_STEP(dpp_row_shr<4>, row16 >= 4 ? lane - 4 : -1)
        CVT_SCAN_STEP(dpp_row_shr<8>, row16 >= 8 ? lane - 8 : -1)
        CVT_SCAN_STEP(dpp_bcast15, ((lane >> 4) & 1) ? (lane & ~15) - 1 : -1)
        CVT_SCAN_STEP(dpp_bcast31, (lane >= 32) ? 31 : -1)
#undef CVT_SCAN_STEP
        const int knext = dpp_wave_shl1(key, -2);
        const bool tail = (seg_lane == C::SEG - 1) || (knext != key);
        if (tail && key >= 0) {
            // run -> site-relative sums in float64 (one site read per run)
            const float2 st = sites[key];
            const double A = a, B = b, Cc = cc;
            const double ox = ((double)(x - (lane - start)) + 0.5) - (double)st.x;
            const double oy = ((double)y + 0.5) - (double)st.y;
            const double W_ = A;
            const double DX = B + A * ox;
            const double DY = A * oy;
            const double E = Cc + 2.0 * ox * B + A * (ox * ox + oy * oy);
            const unsigned hsh = ((unsigned)key * 2654435761u) >> (32 - 10);
            int slot = -1;
            for (int probe = 0; probe < CVT_HASH_CAP; ++probe) {
                const int h = (hsh + probe) & (CVT_HASH_CAP - 1);
                const int old = atomicCAS(&keys[h], -1, key);
                if (old == -1 || old == key) { slot = h; break; }
            }
            if (slot >= 0) {
                atomicAdd(&acc[slot][0], (unsigned long long)cvt_fix(W_, MOM_FW));
                atomicAdd(&acc[slot][1], (unsigned long long)cvt_fix(DX, MOM_FX));
                atomicAdd(&acc[slot][2], (unsigned long long)cvt_fix(DY, MOM_FX));
                atomicAdd(&acc[slot][3], (unsigned long long)cvt_fix(E, MOM_FE));
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
            // a cell reaching beyond its home tile's 3x3 window: chain a record
            // to the site (integer sums: the chain order cannot change them)
            const int k = atomicAdd(&dstate[ST_OVFN], 1);
            if (k < ovf_cap) {
                OvfRec o;
                o.w = wt; o.dx = dx; o.dy = dy; o.e = ee;
                o.next = atomicExch(&ovf_head[l], k);
                o.pad = 0;
                ovf[k] = o;
            } else {
                full = 1;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && full) dflags[FLAG_FULL] = 1;
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
enum {
    UPD_MOVE = 1,     // continuous update s += omega (c - s)
    UPD_SEED = 2,     // scatter the (new) site into the seed map
    UPD_COMPAT = 4,   // integer-site rule of the reference: targets into tgt / owner
    UPD_DEV_S = 8,    // the live site count is dstate[ST_NSITES] (<= S)
};

// Moment totals of site i: the 9 window slots of this epoch plus its overflow
// chain, summed exactly in 128 bits.  Consumes the chain.
__device__ __forceinline__ void cvt_site_totals(int i, int epoch, const Slot *__restrict__ slots,
                                                const OvfRec *__restrict__ ovf,
                                                int *__restrict__ ovf_head, double &w, double &dx,
                                                double &dy, double &ee)
{
    __int128 sw = 0, sx = 0, sy = 0, se = 0;
    const Slot *sl = slots + (long)i * 9;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        if (sl[k].epoch == epoch) {
            sw += sl[k].w; sx += sl[k].dx; sy += sl[k].dy; se += sl[k].e;
        }
    }
    for (int k = ovf_head[i]; k >= 0; k = ovf[k].next) {
        sw += ovf[k].w; sx += ovf[k].dx; sy += ovf[k].dy; se += ovf[k].e;
    }
    ovf_head[i] = -1;
    w = cvt_i128_to_double(sw, MOM_FW);
    dx = cvt_i128_to_double(sx, MOM_FX);
    dy = cvt_i128_to_double(sy, MOM_FX);
    ee = cvt_i128_to_double(se, MOM_FE);
}



static __global__ __launch_bounds__(256) void k_update(
    float2 *__restrict__ sites, const uint8_t *__restrict__ fixed, int S,
    const Slot *__restrict__ slots, const OvfRec *__restrict__ ovf, int *__restrict__ ovf_head,
    const int *__restrict__ dflags, int *__restrict__ dstate, float omega_arg,
    const float *__restrict__ domega, int mode, double rho_unscale, double *__restrict__ mom_out,
    double *__restrict__ blockpart, uint32_t *__restrict__ seed, int W, int H, CompatArgs ca)
{
    __shared__ double red_e[256], red_g[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int live = (mode & UPD_DEV_S) ? dstate[ST_NSITES] : S;
    const float omega = domega ? *domega : omega_arg;
    double e = 0.0, g = 0.0;
    if (i < live) {
        double w, dx, dy, ee;
        cvt_site_totals(i, dstate[ST_EPOCH], slots, ovf, ovf_head, w, dx, dy, ee);
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
        if (mode & UPD_COMPAT) {
            // source/gcvt.cu:753-781: target trunc(s + omega (c - s) + 0.5)
            // (toward zero) per axis, clamped; a site on a mask pixel, or a
            // target on rho == 0, stays.  Sites sit at pixel centres.
            const int sx = (int)floorf(s.x), sy = (int)floorf(s.y);
            int nx = sx, ny = sy;
            const bool fx = ca.mask && ca.mask[(long)sy * W + sx];
            if (!fx && w > 0.0) {
                const double vx = (double)sx + (double)omega * (dx / w) + 0.5;
                const double vy = (double)sy + (double)omega * (dy / w) + 0.5;
                int cx = (int)vx, cy = (int)vy;
                cx = min(max(cx, 0), W - 1);
                cy = min(max(cy, 0), H - 1);
                if (ca.rho[(long)cy * W + cx] != 0.0f) { nx = cx; ny = cy; }
            }
            const int t = ny * W + nx;
            ca.tgt[i] = t;
            atomicMin(&ca.owner[t], (uint32_t)i);
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
    // every chain was consumed above: the record pool restarts empty
    if (blockIdx.x == 0 && threadIdx.x == 0) dstate[ST_OVFN] = 0;
}

// Integer-site merge (one workgroup): sites whose target pixel they own
// survive, in their old relative order; survivors become the new site table
// (pixel centres) and their seeds; the owner map is cleared again.
static __global__ __launch_bounds__(1024) void k_compat_compact(float2 *__restrict__ sites,
                                                         const int *__restrict__ tgt,
                                                         uint32_t *__restrict__ owner,
                                                         uint32_t *__restrict__ seed, int W,
                                                         int *__restrict__ dstate)
{
    __shared__ int part[1024];
    const int S = dstate[ST_NSITES];
    const int per = (S + 1023) / 1024;
    const int i0 = min(S, (int)threadIdx.x * per), i1 = min(S, i0 + per);
    int cnt = 0;
    for (int i = i0; i < i1; ++i) cnt += (owner[tgt[i]] == (uint32_t)i);
    part[threadIdx.x] = cnt;
    __syncthreads();
    // inclusive scan of the per-thread counts (Hillis-Steele)
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = (threadIdx.x >= (unsigned)off) ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int pos = part[threadIdx.x] - cnt;
    const int total = part[1023];
    for (int i = i0; i < i1; ++i) {
        const int t = tgt[i];
        if (owner[t] == (uint32_t)i) {
            const int x = t % W, y = t / W;
            sites[pos] = make_float2((float)x + 0.5f, (float)y + 0.5f);
            seed[t] = 8u * (uint32_t)pos;
            ++pos;
        }
    }
    __syncthreads();
    for (int i = i0; i < i1; ++i) owner[tgt[i]] = CVT_NONE;
    if (threadIdx.x == 0) dstate[ST_NSITES] = total;
}

// labels -> the reference's coordinate map: interleaved (x, y) int16 of the
// pixel's site, MARKER (-32768) where no site reaches
static __global__ void k_labels_to_coords(const uint32_t *__restrict__ lab,
                                   const float2 *__restrict__ sites, uint32_t none, long n,
                                   short2 *__restrict__ out)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const uint32_t l = lab[i];
        if (l >= none) { out[i] = make_short2(-32768, -32768); continue; }
        const float2 s = sites[l >> 3];
        out[i] = make_short2((short)floorf(s.x), (short)floorf(s.y));
    }
}

// Sums the block partials in a fixed order, appends (E, |grad E|) to the
// history when record != 0, advances the epoch and clears the flags.
static __global__ __launch_bounds__(256) void k_finalize(const double *__restrict__ blockpart,
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
    }
}

// ---------------------------------------------------------------------------
// float64 energy over pixels (the fp64 convergence check)
// ---------------------------------------------------------------------------
static __global__ __launch_bounds__(256) void k_energy64(const uint32_t *__restrict__ lab,
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
static __global__ __launch_bounds__(256) void k_density_stats(const float *__restrict__ rho, long n,
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
