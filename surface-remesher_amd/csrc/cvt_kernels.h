// cvt_kernels.h -- CDNA4 (gfx950) kernels of the CVT Lloyd iteration.
//
// Data layout in HBM (one grid of W x H pixels, S sites):
//   rho      fp32 [H][W]      density, unnormalised (scaled by rho_scale = 2^-k
//                             in-kernel so that every scaled value is <= 1)
//   seed     uint32 [H][W]    seed map: label of the site on that pixel, `none`
//                             elsewhere; read by JFA pass 0, cleared by pass 1,
//                             re-seeded by the site update
//   first    uint32 [H][W]    output of JFA pass 0 (the seeds scattered to
//                             their receivers); `none` elsewhere between JFAs
//   lab[2]   uint32 [H][W]    JFA ping-pong label maps
//   sites    float2 [capS+1]  site table, pixel units; sites[capS] = (inf, inf)
//                             is the sentinel
// A label is the byte offset of its site in the table (8 * index) and `none`
// = 8 * capS, so the JFA gather needs no address arithmetic; labels leave the
// engine as int32 indices (CVT_NONE for none).
//   slots    Slot [9][capS]   per-site partial moments, one per tile of the 3x3
//                             tile window around the site's home tile (unique
//                             writer per slot, tagged with the epoch); slot-major,
//                             so slot k of consecutive sites is contiguous
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

#include "cvt_finalize.h"
#include "cvt_types.h"

// Each translation unit that includes this header launches only some of its
// kernels.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wunused-function"

// ---------------------------------------------------------------------------
// distance order (must round exactly like the oracle's orc_dist / orc_better)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float cvt_dist(float2 s, float px, float py)
{
    const float dx = s.x - px;
    const float dy = s.y - py;
    return __builtin_fmaf(dy, dy, dx * dx);
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

// ---------------------------------------------------------------------------
// Tile moments.  A lane owns 4 consecutive pixels of a row segment of SEG
// pixels (SEG = 64 for every tile of side >= 64, so all those tiles cut runs
// at the same places and give bit-identical sums; SEG = 32 for 32-pixel
// tiles).  A segment is SEG/4 lanes inside one 16-lane DPP row.  Runs of one
// key (label with density) are summed relative to the run's first pixel
// (u = pixel - run start) in fp32: in-lane partial sums, then a segmented
// scan of the lanes' carries (4 DPP steps).  Each run end becomes a record in
// the wave's LDS queue (ballot + mbcnt compaction); the wave then converts
// the records to site-relative fp64 sums and adds them, as int64 fixed point,
// to the tile's LDS hash with integer atomics (exact, order-free), whenever
// its queue could overflow and at the end of each step.
// ---------------------------------------------------------------------------
struct TailRec {
    int key;
    float a, b, c;
    int xy;          // run start x | y << 16
};
// queued records per wave: the queue is drained into the hash before a
// sub-step (64 lanes, at most one record each) could overflow it
#define CVT_TAILQ 128

// LDS of one tile: a hash of HC labels (a power of two <= CVT_HASH_CAP) with
// their fixed-point sums, and the waves' run-end queues
template <int HC> struct TileLds {
    int keys[HC];
    unsigned long long acc[HC][4];
    TailRec q[8][CVT_TAILQ];   // one queue per wave, up to 8 waves
    int full;
};

// hash capacity of a T x T tile kernel: enough for one label per 16 pixels
// at T = 64 and 32 pixels at T = 128 (denser sites take a smaller tile)
template <int T> struct TileHash {
    static constexpr int HC = T == 64 ? 256 : T == 128 ? 512 : CVT_HASH_CAP;
};

template <int HC>
__device__ __forceinline__ void tile_lds_init(TileLds<HC> &L)
{
    for (int j = threadIdx.x; j < HC; j += blockDim.x) {
        L.keys[j] = -1;
        L.acc[j][0] = L.acc[j][1] = L.acc[j][2] = L.acc[j][3] = 0ull;
    }
    if (threadIdx.x == 0) L.full = 0;
    __syncthreads();
}

__device__ __forceinline__ unsigned lane_rank(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// the labels and densities of a lane's four pixels in one wave step
struct MomLoad {
    uint32_t lv[4], rv[4];
};

// Adds the runs of the tile [tx0, tx0 + T) x [ty0, ty0 + T) to the hash
// (which it clears first).  lab: labels (byte offsets); a label's site index
// is (lab >> 3) - key_base, valid when below S; sites: the S sites of this
// map.  A wave's first label and density loads are issued before the hash is
// cleared.  stop (may be null): the stop word of cvt_lloyd_until, read with
// those loads; when it is set the routine returns true before adding
// anything.
template <int SEG, int HC>
__device__ __forceinline__ bool tile_moments(TileLds<HC> &L, const uint32_t *__restrict__ lab,
                                             const float *__restrict__ rho,
                                             const float2 *__restrict__ sites, int key_base,
                                             int S, int W, int H, float rho_scale, int tx0,
                                             int ty0, int T, const int *__restrict__ stop)
{
    constexpr int SL = SEG / 4;                    // lanes per segment
    constexpr int SPW = 64 / SL;                   // segments per wave step
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int seg_lane = lane % SL;
    const int spr = T / SEG;                       // segments per tile row
    const int steps = T * spr / SPW;               // wave steps per tile
    const uint32_t map_bytes = 4u * (uint32_t)W * (uint32_t)H;
    const __amdgpu_buffer_rsrc_t rl = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint32_t *>(lab), (short)0, (int)map_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float *>(rho), (short)0, (int)map_bytes, 0x00020000);
    TailRec *q = L.q[wave];
    // the pixels of step g: row y, columns x0 .. x0 + 3 of this lane
    auto step_pos = [&](int g, int &x0, int &y) {
        const int seg = g * SPW + lane / SL;
        const int r = seg / spr;
        x0 = tx0 + (seg - r * spr) * SEG + 4 * seg_lane;
        y = ty0 + r;
    };
    auto step_load = [&](int g, MomLoad &ld) {
        int x0, y;
        step_pos(g, x0, y);
        const uint32_t off = 4u * ((uint32_t)min(y, H - 1) * (uint32_t)W + (uint32_t)x0);
        if ((W & 3) == 0) {
            typedef unsigned u4 __attribute__((ext_vector_type(4)));
            const u4 a4 = __builtin_amdgcn_raw_buffer_load_b128(rl, off, 0, 0);
            const u4 b4 = __builtin_amdgcn_raw_buffer_load_b128(rr, off, 0, 0);
            ld.lv[0] = a4.x; ld.lv[1] = a4.y; ld.lv[2] = a4.z; ld.lv[3] = a4.w;
            ld.rv[0] = b4.x; ld.rv[1] = b4.y; ld.rv[2] = b4.z; ld.rv[3] = b4.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                ld.lv[j] = __builtin_amdgcn_raw_buffer_load_b32(rl, off + 4u * j, 0, 0);
                ld.rv[j] = __builtin_amdgcn_raw_buffer_load_b32(rr, off + 4u * j, 0, 0);
            }
        }
    };

    // a vector load, so it returns in order with the step loads (a null stop
    // is an empty buffer range, which reads 0)
    const int halt = (int)__builtin_amdgcn_raw_buffer_load_b32(
        __builtin_amdgcn_make_buffer_rsrc(const_cast<int *>(stop), (short)0, stop ? 4 : 0,
                                          0x00020000),
        0, 0, 0);
    MomLoad ld;
    if (wave < steps) step_load(wave, ld);
    tile_lds_init(L);
    if (halt) return true;

    const int nw = blockDim.x >> 6;
    for (int g = wave; g < steps; g += nw) {
        if (g > wave) step_load(g, ld);
        int x0, y;
        step_pos(g, x0, y);
        int key[4];
        float rs[4];
        {
            const bool rowok = y < H;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int l = (int)(ld.lv[j] >> 3) - key_base;
                const float dv = __uint_as_float(ld.rv[j]);
                const bool ok = rowok && (x0 + j < W) && (unsigned)l < (unsigned)S && dv != 0.0f;
                key[j] = ok ? l : -1;
                rs[j] = ok ? dv * rho_scale : 0.0f;
            }
        }
        // run heads and run starts (positions 4 * lane + j within the wave)
        const int kprev = dpp_wave_shr1(key[3], -2);
        bool hd[4];
        hd[0] = (seg_lane == 0) || (kprev != key[0]);
        hd[1] = key[1] != key[0];
        hd[2] = key[2] != key[1];
        hd[3] = key[3] != key[2];
        const int p0 = 4 * lane;
        int lasth = hd[3] ? p0 + 3 : hd[2] ? p0 + 2 : hd[1] ? p0 + 1 : hd[0] ? p0 : -1;
        // inclusive max over the lanes of the segment (a segment's first lane
        // always has a head, and positions grow with the lane)
        lasth = max(lasth, dpp_row_shr<1>(lasth));
        lasth = max(lasth, dpp_row_shr<2>(lasth));
        if (SL > 4) lasth = max(lasth, dpp_row_shr<4>(lasth));
        if (SL > 8) lasth = max(lasth, dpp_row_shr<8>(lasth));
        const int prevh = dpp_row_shr<1>(lasth);          // start carried into pixel 0
        int st[4];
        st[0] = hd[0] ? p0 : prevh;
        st[1] = hd[1] ? p0 + 1 : st[0];
        st[2] = hd[2] ? p0 + 2 : st[1];
        st[3] = hd[3] ? p0 + 3 : st[2];
        // in-lane running sums of the current piece, reset at heads
        float pa[4], pb[4], pc[4];
        {
            float ra = 0.0f, rb = 0.0f, rc = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float u = (float)(p0 + j - st[j]);
                if (hd[j]) { ra = 0.0f; rb = 0.0f; rc = 0.0f; }
                ra += rs[j];
                rb += rs[j] * u;
                rc += rs[j] * (u * u);
                pa[j] = ra; pb[j] = rb; pc[j] = rc;
            }
        }
        // segmented inclusive scan of the lanes' carries (their last piece);
        // a lane's chain starts at the lane of its last run head
        const int hlane = lasth >> 2;
        float ca = pa[3], cb = pb[3], cc = pc[3];
#define CVT_CARRY_STEP(K)                                                                \
        {                                                                                \
            const int ta = dpp_row_shr<K>(ibits(ca)), tb = dpp_row_shr<K>(ibits(cb)),     \
                      tc = dpp_row_shr<K>(ibits(cc));                                    \
            const int src = ((lane & 15) >= K) ? lane - K : -1;                           \
            ca = seg_add(ca, ta, src, hlane);                                            \
            cb = seg_add(cb, tb, src, hlane);                                            \
            cc = seg_add(cc, tc, src, hlane);                                            \
        }
        CVT_CARRY_STEP(1)
        CVT_CARRY_STEP(2)
        if (SL > 4) CVT_CARRY_STEP(4)
        if (SL > 8) CVT_CARRY_STEP(8)
#undef CVT_CARRY_STEP
        // carry into this lane's first piece (only when pixel 0 continues a run)
        const float ia = fbits(dpp_row_shr<1>(ibits(ca)));
        const float ib = fbits(dpp_row_shr<1>(ibits(cb)));
        const float ic = fbits(dpp_row_shr<1>(ibits(cc)));
        // run ends
        const int knext = dpp_wave_shl1(key[0], -2);
        bool tl[4];
        tl[0] = key[1] != key[0];
        tl[1] = key[2] != key[1];
        tl[2] = key[3] != key[2];
        tl[3] = (seg_lane == SL - 1) || (knext != key[3]);
        auto drain = [&](int nrec) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // one lane per record: site-relative fp64 sums into the hash
            for (int e = lane; e < nrec; e += 64) {
                const TailRec t = q[e];
                const float2 sv = sites[t.key];
                const double A = t.a, B = t.b, Cc = t.c;
                const double ox = ((double)(t.xy & 0xffff) + 0.5) - (double)sv.x;
                const double oy = ((double)(t.xy >> 16) + 0.5) - (double)sv.y;
                const double DX = B + A * ox;
                const double DY = A * oy;
                const double E = Cc + 2.0 * ox * B + A * (ox * ox + oy * oy);
                const unsigned hsh = ((unsigned)t.key * 2654435761u) >> (32 - __builtin_ctz(HC));
                int slot = -1;
                for (int probe = 0; probe < HC; ++probe) {
                    const int h = (hsh + probe) & (HC - 1);
                    const int old = atomicCAS(&L.keys[h], -1, t.key);
                    if (old == -1 || old == t.key) { slot = h; break; }
                }
                if (slot >= 0) {
                    atomicAdd(&L.acc[slot][0], (unsigned long long)cvt_fix(A, MOM_FW));
                    atomicAdd(&L.acc[slot][1], (unsigned long long)cvt_fix(DX, MOM_FX));
                    atomicAdd(&L.acc[slot][2], (unsigned long long)cvt_fix(DY, MOM_FX));
                    atomicAdd(&L.acc[slot][3], (unsigned long long)cvt_fix(E, MOM_FE));
                } else {
                    L.full = 1;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        };
        // queue the run ends: records ordered by (j, lane)
        int base = 0;
        bool first_piece = true;                 // no head yet in this lane at <= j
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (hd[j]) first_piece = false;
            if (base > CVT_TAILQ - 64) {   // uniform: base is the wave's count
                drain(base);
                base = 0;
            }
            const bool m = tl[j] && key[j] >= 0;
            const unsigned long long bal = __ballot(m);
            if (m) {
                const int idx = base + (int)lane_rank(bal);
                TailRec t;
                t.key = key[j];
                const bool carry = first_piece;  // the run began in an earlier lane
                t.a = carry ? pa[j] + ia : pa[j];
                t.b = carry ? pb[j] + ib : pb[j];
                t.c = carry ? pc[j] + ic : pc[j];
                const int xs = x0 + (st[j] - p0);
                t.xy = xs | (y << 16);
                q[idx] = t;
            }
            base += __popcll(bal);
        }
        drain(base);
    }
    return false;
}

// Writes the tile's hash entries: a site whose home tile (the tile holding
// its position) is within one tile of this one gets a unique-writer slot of
// its 3 x 3 window, tagged with the epoch; otherwise a record is chained to
// the site (integer atomics only to allocate records).  site_off: index of
// this map's first site in slots / ovf_head / sites.
template <int HC>
__device__ __forceinline__ void tile_flush(TileLds<HC> &L, const float2 *__restrict__ sites,
                                           int site_off, int T, int tX, int tY, int tilesX,
                                           int tilesY, int epoch, Slot *__restrict__ slots, long nrows,
                                           OvfRec *__restrict__ ovf, int ovf_cap,
                                           int *__restrict__ ovf_head, int *__restrict__ dstate,
                                           int *__restrict__ dflags)
{
    __syncthreads();
    for (int j = threadIdx.x; j < HC; j += blockDim.x) {
        const int l = L.keys[j];
        if (l < 0) continue;
        const int gl = site_off + l;
        const float2 s = sites[gl];
        const long long wt = (long long)L.acc[j][0], dx = (long long)L.acc[j][1];
        const long long dy = (long long)L.acc[j][2], ee = (long long)L.acc[j][3];
        const int hx = min(max((int)floorf(s.x / (float)T), 0), tilesX - 1);
        const int hy = min(max((int)floorf(s.y / (float)T), 0), tilesY - 1);
        const int rx = tX - hx, ry = tY - hy;
        if (rx >= -1 && rx <= 1 && ry >= -1 && ry <= 1) {
            Slot sl;
            sl.w = wt; sl.dx = dx; sl.dy = dy; sl.e = ee; sl.epoch = epoch; sl.pad = 0;
            slots[(long)((ry + 1) * 3 + (rx + 1)) * nrows + gl] = sl;
        } else {
            const int k = atomicAdd(&dstate[ST_OVFN], 1);
            if (k < ovf_cap) {
                OvfRec o;
                o.w = wt; o.dx = dx; o.dy = dy; o.e = ee;
                o.next = atomicExch(&ovf_head[gl], k);
                o.pad = 0;
                ovf[k] = o;
            } else {
                L.full = 1;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && L.full) dflags[FLAG_FULL] = 1;
}

// moments (reference replacement for pbaCVDComputeCentroid + the energy
// texture, source/gcvt.cu:562-732, :788-802): one workgroup per T x T tile,
// eight waves for tiles of 64 pixels or more (four for 32: a 32-pixel tile
// has four wave steps)
template <int T>
static __global__ __launch_bounds__(512) void k_moments(
    const uint32_t *__restrict__ lab, const float *__restrict__ rho,
    const float2 *__restrict__ sites, int S, int W, int H, float rho_scale, int tilesX,
    int tilesY, Slot *__restrict__ slots, long nrows, OvfRec *__restrict__ ovf, int ovf_cap,
    int *__restrict__ ovf_head, int *__restrict__ dflags, int *__restrict__ dstate)
{
    __shared__ TileLds<TileHash<T>::HC> L;
    const int tile = blockIdx.x;
    const int tX = tile % tilesX, tY = tile / tilesX;
    if (tile_moments<(T < 64 ? T : 64)>(L, lab, rho, sites, 0, S, W, H, rho_scale, tX * T, tY * T,
                                        T, dstate + ST_STOP))
        return;   // a stop rule held (cvt_lloyd_until)
    tile_flush(L, sites, 0, T, tX, tY, tilesX, tilesY, dstate[ST_EPOCH], slots, nrows, ovf, ovf_cap,
               ovf_head, dstate, dflags);
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
// chain, summed exactly in 128 bits.  Returns the chain's head (-1: none); the
// caller consumes the chain (ovf_head[i] = -1) once it may write.
__device__ __forceinline__ int cvt_site_totals(int i, int epoch, const Slot *__restrict__ slots,
                                               long nrows, const OvfRec *__restrict__ ovf,
                                               const int *__restrict__ ovf_head, double &w,
                                               double &dx, double &dy, double &ee)
{
    __int128 sw = 0, sx = 0, sy = 0, se = 0;
    // the chain head and all nine slots are loaded before any is used, so
    // the loads overlap (slot k of consecutive sites is contiguous: a wave's
    // loads coalesce)
    const int head = ovf_head[i];
    Slot v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = slots[(long)k * nrows + i];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const bool on = v[k].epoch == epoch;
        sw += on ? v[k].w : 0;
        sx += on ? v[k].dx : 0;
        sy += on ? v[k].dy : 0;
        se += on ? v[k].e : 0;
    }
    for (int k = head; k >= 0; k = ovf[k].next) {
        sw += ovf[k].w; sx += ovf[k].dx; sy += ovf[k].dy; se += ovf[k].e;
    }
    w = cvt_i128_to_double(sw, MOM_FW);
    dx = cvt_i128_to_double(sx, MOM_FX);
    dy = cvt_i128_to_double(sy, MOM_FX);
    ee = cvt_i128_to_double(se, MOM_FE);
    return head;
}



static __global__ __launch_bounds__(256) void k_update(
    float2 *__restrict__ sites, const uint8_t *__restrict__ fixed, int S,
    const Slot *__restrict__ slots, long nrows, const OvfRec *__restrict__ ovf,
    int *__restrict__ ovf_head, const int *__restrict__ dflags, int *__restrict__ dstate, float omega_arg,
    const float *__restrict__ domega, int mode, double rho_unscale, double *__restrict__ mom_out,
    double *__restrict__ blockpart, uint32_t *__restrict__ seed, int W, int H, CompatArgs ca,
    FinalArgs fa)
{
    __shared__ double red_e[256], red_g[256];
    __shared__ int last;
    // the stop word of cvt_lloyd_until is loaded with the other state words
    // and tested once this thread's loads are issued: a stopped update
    // returns before its first write
    const int halt = dstate[ST_STOP];
    const int epoch = dstate[ST_EPOCH];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int live = (mode & UPD_DEV_S) ? dstate[ST_NSITES] : S;
    const float omega = domega ? *domega : omega_arg;
    double e = 0.0, g = 0.0;
    // the site and its fixed flag (the context's flags, never null) are
    // loaded with the moment slots; nothing waits for them until used
    float2 s = make_float2(0.0f, 0.0f);
    int pinned = 0, head = -1;
    double w = 0.0, dx = 0.0, dy = 0.0, ee = 0.0;
    if (i < live) {
        s = sites[i];
        pinned = fixed[i];
        head = cvt_site_totals(i, epoch, slots, nrows, ovf, ovf_head, w, dx, dy, ee);
    }
    if (halt) return;   // uniform over the grid
    if (i < live) {
        if (head >= 0) ovf_head[i] = -1;
        e = ee;
        g = dx * dx + dy * dy;
        if (mom_out) {
            mom_out[4L * i + 0] = w * rho_unscale;
            mom_out[4L * i + 1] = dx * rho_unscale;
            mom_out[4L * i + 2] = dy * rho_unscale;
            mom_out[4L * i + 3] = ee * rho_unscale;
        }
        if ((mode & UPD_MOVE) && w > 0.0 && pinned == 0) {
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
    // block sums: tree steps 128 and 64 in LDS, 32 .. 1 in wave 0
    red_e[threadIdx.x] = e;
    red_g[threadIdx.x] = g;
    __syncthreads();
    for (int o = 128; o >= 64; o >>= 1) {
        if (threadIdx.x < o) {
            red_e[threadIdx.x] += red_e[threadIdx.x + o];
            red_g[threadIdx.x] += red_g[threadIdx.x + o];
        }
        __syncthreads();
    }
    double be = 0.0, bg = 0.0;
    if (threadIdx.x < 64) {
        be = wave_tree_sum(red_e[threadIdx.x]);
        bg = wave_tree_sum(red_g[threadIdx.x]);
    }
    // every chain was consumed above: the record pool restarts empty
    if (blockIdx.x == 0 && threadIdx.x == 0) dstate[ST_OVFN] = 0;
    if (fa.defer) {
        // the partials wait for the next seed pass (or a finalize launch)
        if (threadIdx.x == 0) {
            blockpart[2 * blockIdx.x] = be;
            blockpart[2 * blockIdx.x + 1] = bg;
            if (blockIdx.x == 0) dstate[ST_PEND] = (int)gridDim.x;
        }
        return;
    }
    // publish this block's partials; the last block to finish adds all of
    // them in block order (the release / acquire pair at agent scope makes
    // every block's partials visible to it, whichever XCD wrote them)
    if (threadIdx.x == 0) {
        blockpart[2 * blockIdx.x] = be;
        blockpart[2 * blockIdx.x + 1] = bg;
        const int done = __hip_atomic_fetch_add(&dstate[ST_DONE], 1, __ATOMIC_ACQ_REL,
                                                __HIP_MEMORY_SCOPE_AGENT);
        last = (done == (int)gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    cvt_finalize_block(blockpart, (int)gridDim.x, dstate, fa, red_e, red_g);
}

// A deferred finalize (one 256-thread workgroup), when one is pending.
static __global__ __launch_bounds__(256) void k_finalize_pending(const double *__restrict__ blockpart,
                                                                 int *__restrict__ dstate,
                                                                 FinalArgs fa)
{
    __shared__ double red_e[256], red_g[256];
    const int nb = dstate[ST_PEND];
    if (nb > 0) cvt_finalize_block(blockpart, nb, dstate, fa, red_e, red_g);
}

// cvt_lloyd_until: arm the stop rule (relative decrease over `window`
// iterations below rel_tol, or energy at or below target) from the current
// history length, or disarm it and report the iterations the call ran
static __global__ void k_until(int *__restrict__ dstate, double *__restrict__ until, int arm,
                               double rel_tol, double target, int window, int *__restrict__ ran)
{
    if (threadIdx.x != 0) return;
    if (arm) {
        until[0] = rel_tol;
        until[1] = target;
        dstate[ST_UBASE] = dstate[ST_HIST];
        dstate[ST_UWIN] = window;
        dstate[ST_STOP] = 0;
        dstate[ST_UACTIVE] = 1;
    } else {
        *ran = dstate[ST_HIST] - dstate[ST_UBASE];
        dstate[ST_UACTIVE] = 0;
        dstate[ST_STOP] = 0;
    }
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

#pragma clang diagnostic pop
