// jfa.hip -- the JFA(+1) pass, the hot kernel of the Lloyd iteration
// (replaces the reference's PBA+ pba2DCompute, source/gcvt.cu:967-978).
//
// Built with -mno-amdgpu-ieee -fno-honor-nans (see the Makefile): no value in
// this file is ever NaN, and with IEEE mode off v_min_f64 needs no
// canonicalising v_max_f64 in front of it.
//
// Label encoding inside the engine: a label is the byte offset of its site in
// the site table (8 * index), "none" is the byte offset of the sentinel site
// (inf, inf) at index capS.  The gather then needs no address arithmetic, and
// the order of byte offsets is the order of indices.
//
// Candidate order: (d, index) lexicographic with d = fmaf(dy, dy, dx*dx) >= 0
// (dx*dx rounded, then dy*dy + dx*dx rounded once).  A label read from a tap
// row serves the pixels of up to three output rows in the same column, so
// dx*dx is computed once per label and each use costs one subtraction (dy),
// one fma and one min.
// The pair is packed into one 64-bit word, hi = bits(d), lo = label, and read
// as a double: for d >= 0 (including +inf) the hi word is below 0x7ff00000, so
// the double is finite, not NaN, and double order equals (d, label) order
// (f64 denormals are preserved, so d = 0 still orders by label).  One
// v_min_f64 per candidate does the compare and the select.
//
// Out-of-grid taps are redirected to the centre row / column: such a tap
// becomes a duplicate of a candidate the pixel already has, which cannot
// change a minimum, so the pass needs no bounds selects.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "jfa.h"
#include "cvt_finalize.h"

#include <algorithm>

namespace cvt {

// Cache policy of the map stores: device scope (sc1).  The pass's output is
// read by every XCD in the next pass; device-scope stores write the lines
// through to the device's point of coherence while the pass runs, instead of
// each XCD's L2 writing its dirty lines back when the kernel ends.  Measured
// at 2048^2 / 16K sites: 14.0 us per pass against 14.9 with default stores
// (non-temporal stores were slower, 15.5).
#define CVT_MAP_STORE_POLICY 16

__device__ __forceinline__ double jfa_key(float d, uint32_t label)
{
    return __hiloint2double(__float_as_int(d), (int)label);
}

// Raw buffer resources: the label map's base and size live in scalar
// registers, a tap load is buffer_load_dword with the row's byte offset in
// soffset (uniform) and the column's byte offset in voffset (per lane), so a
// tap costs no vector ALU.  dword3 0x00020000 is the raw 32-bit format of the
// gfx9 family.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                             0x00020000);
}

// grid: 1-D, one block per 64 x (4 * ROWS) pixel tile; wave w of the block
// takes rows w, w + 4, ..., so the per-wave setup (tile position, tap columns)
// is paid once per ROWS rows.  Tiles are numbered so that each XCD (blocks
// are dealt to the 8 XCDs round robin) gets a contiguous band of rows,
// keeping small-step tap rows inside one XCD's L2.
template <int ROWS>
__global__ __launch_bounds__(256) void k_jfa_pass(const uint32_t *__restrict__ in,
                                                  uint32_t *__restrict__ out,
                                                  const float2 *__restrict__ sites, int W,
                                                  int H, int tilesX, int ntiles, int step,
                                                  uint32_t *__restrict__ clear, uint32_t none,
                                                  const int *__restrict__ stop)
{
    if (stop && *stop) return;
    // XCD-aware tile order (wave-uniform values live in scalar registers)
    const int b = blockIdx.x;
    const int q = ntiles >> 3, r = ntiles & 7;
    const int xcd = b & 7, pos = b >> 3;
    const int tile = __builtin_amdgcn_readfirstlane(xcd * q + min(xcd, r) + pos);
    const int tY = __builtin_amdgcn_readfirstlane(tile / tilesX);
    const int tX = __builtin_amdgcn_readfirstlane(tile - tY * tilesX);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int x = tX * 64 + lane;
    const int xc = min(x, W - 1);

    // tap columns as byte offsets within a row, clamped to the centre column
    // when outside the grid
    const uint32_t o1 = 4u * (uint32_t)xc;
    const uint32_t o0 = (xc - step >= 0) ? o1 - 4u * (uint32_t)step : o1;
    const uint32_t o2 = (xc + step < W) ? o1 + 4u * (uint32_t)step : o1;
    const float px = (float)xc + 0.5f;
    const char *sb = reinterpret_cast<const char *>(sites);
    const uint32_t map_bytes = 4u * (uint32_t)W * (uint32_t)H;
    const __amdgpu_buffer_rsrc_t rin = buf_rsrc(in, map_bytes);
    const __amdgpu_buffer_rsrc_t rout = buf_rsrc(out, map_bytes);
    const uint32_t row_bytes = 4u * (uint32_t)W;

#pragma unroll
    for (int k = 0; k < ROWS; ++k) {
        const int y = __builtin_amdgcn_readfirstlane(tY * 4 * ROWS + wave + 4 * k);
        if (y >= H) break;
        // tap rows (uniform), clamped to the centre row when outside
        const int r0 = (y - step >= 0) ? y - step : y;
        const int r2 = (y + step < H) ? y + step : y;
        const uint32_t s0 = __builtin_amdgcn_readfirstlane((uint32_t)r0 * row_bytes);
        const uint32_t s1 = __builtin_amdgcn_readfirstlane((uint32_t)y * row_bytes);
        const uint32_t s2 = __builtin_amdgcn_readfirstlane((uint32_t)r2 * row_bytes);

        uint32_t l[9];
        l[0] = __builtin_amdgcn_raw_buffer_load_b32(rin, o0, s0, 0);
        l[1] = __builtin_amdgcn_raw_buffer_load_b32(rin, o1, s0, 0);
        l[2] = __builtin_amdgcn_raw_buffer_load_b32(rin, o2, s0, 0);
        l[3] = __builtin_amdgcn_raw_buffer_load_b32(rin, o0, s1, 0);
        l[4] = __builtin_amdgcn_raw_buffer_load_b32(rin, o1, s1, 0);
        l[5] = __builtin_amdgcn_raw_buffer_load_b32(rin, o2, s1, 0);
        l[6] = __builtin_amdgcn_raw_buffer_load_b32(rin, o0, s2, 0);
        l[7] = __builtin_amdgcn_raw_buffer_load_b32(rin, o1, s2, 0);
        l[8] = __builtin_amdgcn_raw_buffer_load_b32(rin, o2, s2, 0);

        const float py = (float)y + 0.5f;
        double best;
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const float2 s = *reinterpret_cast<const float2 *>(sb + l[t]);
            const float dx = s.x - px;
            const float dy = s.y - py;
            const float d = __builtin_fmaf(dy, dy, dx * dx);
            const double key = jfa_key(d, l[t]);
            best = (t == 0) ? key : __builtin_fmin(best, key);
        }
        if (x < W) {
            __builtin_amdgcn_raw_buffer_store_b32((uint32_t)__double2loint(best), rout, o1, s1,
                                                  CVT_MAP_STORE_POLICY);
            if (clear)
                __builtin_amdgcn_raw_buffer_store_b32(none, buf_rsrc(clear, map_bytes), o1, s1,
                                                      CVT_MAP_STORE_POLICY);
        }
    }
}

// ---------------------------------------------------------------------------
// Quad pass.  Each lane owns 4 consecutive pixels of a row, a wave takes one
// 256-pixel row segment per item, and persistent workgroups own a share of
// the items: a band of rows (quad_rows), or runs of four rows one step apart
// whose items share tap rows (quad_runs, for the small steps).  The code is
// specialised on the step:
//   STEP_ANY (>= 4): nine tap quads; every (pixel, tap) pair is distinct.
//   STEP_1 / STEP_2: the three tap columns of a row overlap, so a tap row
//                    holds 6 (8) distinct positions; their site lookups and
//                    dy^2 are shared by the 4 pixels.
// The next item's label loads are issued before the current item is
// evaluated, so each wave keeps one item's loads in flight while it computes.
// ---------------------------------------------------------------------------
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(3))) char lds_cchar;
typedef const __attribute__((address_space(3))) f32x2 lds_cf32x2;
typedef __attribute__((address_space(3))) f32x4 lds_f32x4;

enum { STEP_ANY = 0, STEP_1 = 1, STEP_2 = 2 };
enum { TAB_LDS0 = 0, TAB_LDSB = 1, TAB_GLOBAL = 2, TAB_SPLIT = 3 };

// Site lookup by label (a byte offset into the site table).
//   TAB_LDS0:   the table is staged at LDS byte 0, so the label is the LDS
//               address and a lookup is one ds_read_b64;
//   TAB_LDSB:   one patch's table is staged in LDS; `base` is the label of
//               its first entry;
//   TAB_GLOBAL: a buffer load from the global table;
//   TAB_SPLIT:  the table's first entries (labels below `split`) are staged
//               at LDS byte 0, followed by a zero entry at byte `split`; the
//               other labels are read from a buffer that starts at the
//               first entry not staged (`g`, offset l - split).
template <int TAB> struct SiteTab {
    lds_cchar *lds;
    __amdgpu_buffer_rsrc_t g;
    uint32_t base;
    uint32_t split;       // TAB_SPLIT: the first label not staged in LDS (bytes)
    const f32x4 *src;     // staging source (LDS kinds)
    int n4;               // float4 count to stage

    __device__ __forceinline__ float2 operator()(uint32_t l) const
    {
        if constexpr (TAB == TAB_LDS0) {
            const f32x2 v = *(lds_cf32x2 *)(lds + l);
            return make_float2(v.x, v.y);
        } else if constexpr (TAB == TAB_LDSB) {
            const f32x2 v = *(lds_cf32x2 *)(lds + (l - base));
            return make_float2(v.x, v.y);
        } else if constexpr (TAB == TAB_SPLIT) {
            // branch-free: every lane reads LDS, clamped to the zero entry,
            // and issues the gather; for a label below the split, l - split
            // wraps past the buffer's range, so the gather reads 0 and
            // fetches nothing.  One of the two is 0: OR combines them.
            const f32x2 a = *(lds_cf32x2 *)(lds + min(l, split));
            const u32x2 b = __builtin_amdgcn_raw_buffer_load_b64(g, l - split, 0, 0);
            return make_float2(__uint_as_float(__float_as_uint(a.x) | b.x),
                               __uint_as_float(__float_as_uint(a.y) | b.y));
        } else {
            const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(g, l, 0, 0);
            return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
        }
    }
    // copy the table into LDS (all threads of the block take part); each
    // thread issues all of its loads (8 per round, range-checked buffer loads)
    // before its first LDS store, so the round trips overlap
    __device__ __forceinline__ void stage() const
    {
        if constexpr (TAB != TAB_GLOBAL) {
            typedef unsigned u4 __attribute__((ext_vector_type(4)));
            lds_f32x4 *dst = (lds_f32x4 *)lds;
            const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<f32x4 *>(src), (short)0, 16 * n4, 0x00020000);
            const int stride = blockDim.x;
            for (int i0 = threadIdx.x; i0 < n4; i0 += 8 * stride) {
                u4 v[8];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    v[k] = __builtin_amdgcn_raw_buffer_load_b128(r, 16u * (uint32_t)(i0 + k * stride),
                                                                 0, 0);
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (i0 + k * stride < n4) {
                        const f32x4 f = {__uint_as_float(v[k].x), __uint_as_float(v[k].y),
                                         __uint_as_float(v[k].z), __uint_as_float(v[k].w)};
                        dst[i0 + k * stride] = f;
                    }
            }
            // the split table's zero entry (and the unused one beside it)
            if (TAB == TAB_SPLIT && threadIdx.x == 0) dst[n4] = f32x4{0.f, 0.f, 0.f, 0.f};
            __syncthreads();
        }
    }
};

// label loads of one item
template <int MODE> struct Item;
template <> struct Item<STEP_ANY> { u32x4 q[9]; };
template <> struct Item<STEP_1> { u32x4 c[3]; uint32_t l[3], r[3]; };
template <> struct Item<STEP_2> { u32x4 c[3]; u32x2 l[3], r[3]; };

// per-pass constants of a map
struct MapGeo {
    __amdgpu_buffer_rsrc_t rin, rout, rclr;
    uint32_t row_bytes;
    int W, H, step;
    bool clear;
    uint32_t none;
};

// (y is wave-uniform: the product is taken in scalar registers)
__device__ __forceinline__ uint32_t row_off(const MapGeo &m, int y)
{
    return (uint32_t)__builtin_amdgcn_readfirstlane(y) * m.row_bytes;
}

// the labels one tap row contributes to a segment: the left, centre and
// right quads (step >= 4), or the centre quad and its side words
template <int MODE> struct RowData;
template <> struct RowData<STEP_ANY> { u32x4 l, c, r; };
template <> struct RowData<STEP_1> { u32x4 c; uint32_t l, r; };
template <> struct RowData<STEP_2> { u32x4 c; u32x2 l, r; };

// loads tap row y (already redirected onto the grid) for segment seg
template <int MODE>
__device__ __forceinline__ void load_row(RowData<MODE> &d, const MapGeo &m, int y, int seg,
                                         int lane4)
{
    const int xc = min(seg * 256 + lane4, m.W - 4);
    const uint32_t o1 = 4u * (uint32_t)xc;
    const uint32_t s = row_off(m, y);
    if constexpr (MODE == STEP_ANY) {
        // a tap column outside the grid is redirected to the centre column
        // (for step >= 4 the whole quad is in or out together)
        const int st = m.step;
        const uint32_t o0 = (xc - st >= 0) ? o1 - 4u * (uint32_t)st : o1;
        const uint32_t o2 = (xc + 3 + st < m.W) ? o1 + 4u * (uint32_t)st : o1;
        d.l = __builtin_amdgcn_raw_buffer_load_b128(m.rin, o0, s, 0);
        d.c = __builtin_amdgcn_raw_buffer_load_b128(m.rin, o1, s, 0);
        d.r = __builtin_amdgcn_raw_buffer_load_b128(m.rin, o2, s, 0);
    } else {
        // side words: the one (two) labels left of the quad and right of it;
        // unused (and replaced by the centre labels) at the grid's edges
        const uint32_t oL = (xc > 0) ? o1 - 4u * MODE : o1;
        const uint32_t oR = (xc + 4 < m.W) ? o1 + 16u : o1;
        d.c = __builtin_amdgcn_raw_buffer_load_b128(m.rin, o1, s, 0);
        if constexpr (MODE == STEP_1) {
            d.l = __builtin_amdgcn_raw_buffer_load_b32(m.rin, oL, s, 0);
            d.r = __builtin_amdgcn_raw_buffer_load_b32(m.rin, oR, s, 0);
        } else {
            d.l = __builtin_amdgcn_raw_buffer_load_b64(m.rin, oL, s, 0);
            d.r = __builtin_amdgcn_raw_buffer_load_b64(m.rin, oR, s, 0);
        }
    }
}

// the item of centre row y from its up, centre and down tap rows
template <int MODE>
__device__ __forceinline__ Item<MODE> item_of(const RowData<MODE> &u, const RowData<MODE> &c,
                                              const RowData<MODE> &d)
{
    Item<MODE> it;
    if constexpr (MODE == STEP_ANY) {
        it.q[0] = u.l; it.q[1] = u.c; it.q[2] = u.r;
        it.q[3] = c.l; it.q[4] = c.c; it.q[5] = c.r;
        it.q[6] = d.l; it.q[7] = d.c; it.q[8] = d.r;
    } else {
        it.c[0] = u.c; it.l[0] = u.l; it.r[0] = u.r;
        it.c[1] = c.c; it.l[1] = c.l; it.r[1] = c.r;
        it.c[2] = d.c; it.l[2] = d.l; it.r[2] = d.r;
    }
    return it;
}

// tap rows of centre row y: off-grid rows are redirected to the centre row
__device__ __forceinline__ int up_row(const MapGeo &m, int y) { return (y - m.step >= 0) ? y - m.step : y; }
__device__ __forceinline__ int down_row(const MapGeo &m, int y) { return (y + m.step < m.H) ? y + m.step : y; }

template <int MODE>
__device__ __forceinline__ void load_item(Item<MODE> &it, const MapGeo &m, int y, int seg, int lane4)
{
    RowData<MODE> u, c, d;
    load_row<MODE>(u, m, up_row(m, y), seg, lane4);
    load_row<MODE>(c, m, y, seg, lane4);
    load_row<MODE>(d, m, down_row(m, y), seg, lane4);
    it = item_of<MODE>(u, c, d);
}

__device__ __forceinline__ uint32_t q_el(const u32x4 &q, int j)
{
    return j == 0 ? q.x : j == 1 ? q.y : j == 2 ? q.z : q.w;
}

__device__ __forceinline__ double cand(float2 s, uint32_t l, float px, float py)
{
    const float dx = s.x - px;
    const float dy = s.y - py;
    return jfa_key(__builtin_fmaf(dy, dy, dx * dx), l);
}

template <int MODE, int TAB>
__device__ __forceinline__ void eval_item(const Item<MODE> &it, const MapGeo &m,
                                          const SiteTab<TAB> &tab, int y, int seg, int lane4)
{
    const int xb = seg * 256 + lane4;
    const int xc = min(xb, m.W - 4);
    const float px0 = (float)xc + 0.5f;
    const float py = (float)y + 0.5f;
    uint32_t res[4];
    if constexpr (MODE == STEP_ANY) {
        // a tap row or column that lies off the grid for every lane of the
        // wave was redirected to the centre: its candidates duplicate the
        // centre's and cannot change a minimum, so they are not evaluated
        // (the test is per item; items with every side on the grid, almost
        // all of them below step 256, take the branch-free path)
        const bool up = y - m.step >= 0, down = y + m.step < m.H;
        const bool left = __ballot(xc - m.step >= 0) != 0ull;
        const bool right = __ballot(xc + 3 + m.step < m.W) != 0ull;
        if (up && down && left && right) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float px = px0 + (float)j;
                double best = cand(tab(q_el(it.q[4], j)), q_el(it.q[4], j), px, py);
#pragma unroll
                for (int t = 0; t < 9; ++t) {
                    if (t == 4) continue;
                    const uint32_t l = q_el(it.q[t], j);
                    best = __builtin_fmin(best, cand(tab(l), l, px, py));
                }
                res[j] = (uint32_t)__double2loint(best);
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float px = px0 + (float)j;
                double best = cand(tab(q_el(it.q[4], j)), q_el(it.q[4], j), px, py);
#pragma unroll
                for (int t = 0; t < 9; ++t) {
                    if (t == 4) continue;
                    const int tr = t / 3, tc = t % 3;
                    if ((tr == 0 && !up) || (tr == 2 && !down) || (tc == 0 && !left) ||
                        (tc == 2 && !right))
                        continue;
                    const uint32_t l = q_el(it.q[t], j);
                    best = __builtin_fmin(best, cand(tab(l), l, px, py));
                }
                res[j] = (uint32_t)__double2loint(best);
            }
        }
    } else {
        constexpr int NP = (MODE == STEP_1) ? 6 : 8;
        const bool lok = xc > 0, rok = xc + 4 < m.W;
        double best[4];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const u32x4 &c = it.c[r];
            uint32_t P[NP];
            if constexpr (MODE == STEP_1) {
                P[0] = lok ? it.l[r] : c.x;
                P[1] = c.x; P[2] = c.y; P[3] = c.z; P[4] = c.w;
                P[5] = rok ? it.r[r] : c.w;
            } else {
                P[0] = lok ? it.l[r].x : c.x;
                P[1] = lok ? it.l[r].y : c.y;
                P[2] = c.x; P[3] = c.y; P[4] = c.z; P[5] = c.w;
                P[6] = rok ? it.r[r].x : c.z;
                P[7] = rok ? it.r[r].y : c.w;
            }
            float sx[NP], dyr[NP];
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                const float2 s = tab(P[k]);
                sx[k] = s.x;
                dyr[k] = s.y - py;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float px = px0 + (float)j;
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const int k = j + t * MODE;      // left, centre, right
                    const float dx = sx[k] - px;
                    const double key =
                        jfa_key(__builtin_fmaf(dyr[k], dyr[k], dx * dx), P[k]);
                    best[j] = (r == 0 && t == 0) ? key : __builtin_fmin(best[j], key);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) res[j] = (uint32_t)__double2loint(best[j]);
    }
    if (xb < m.W) {
        const uint32_t o1 = 4u * (uint32_t)xc;
        const uint32_t s1 = row_off(m, y);
        const u32x4 r4 = {res[0], res[1], res[2], res[3]};
        __builtin_amdgcn_raw_buffer_store_b128(r4, m.rout, o1, s1, CVT_MAP_STORE_POLICY);
        if (m.clear) {
            const u32x4 n4 = {m.none, m.none, m.none, m.none};
            __builtin_amdgcn_raw_buffer_store_b128(n4, m.rclr, o1, s1, CVT_MAP_STORE_POLICY);
        }
    }
}

// Work a pass does once its first label loads and the table staging are
// issued: PRO() returns true when the pass must stop there.  NoPrologue is
// the batched pass's; SparseReset is the single-map pass's (below).
struct NoPrologue {
    __device__ bool operator()() const { return false; }
};

// Rows [y0, y1) of one map by all waves of the block.  Items (row, segment)
// are dealt to the waves round robin; a wave's first load is issued before
// the site table is staged, so the two latencies overlap.
template <int MODE, int TAB, class PRO>
__device__ __forceinline__ void quad_rows(const MapGeo &m, const SiteTab<TAB> &tab, int y0, int y1,
                                          const PRO &pro)
{
    const int nseg = (m.W + 255) >> 8;
    const int nwaves = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane4 = 4 * (threadIdx.x & 63);
    const int last = max(y1 - 1, 0);
    int row = y0 + wave / nseg, seg = wave % nseg;
    Item<MODE> A, B;
    load_item<MODE>(A, m, min(row, last), seg, lane4);
    tab.stage();
    if (pro()) return;
    while (row < y1) {
        int r2 = row, g2 = seg + nwaves;
        while (g2 >= nseg) { g2 -= nseg; ++r2; }
        // loads past the band are clamped to its last row (never used), so
        // every iteration has the same number of loads in flight
        load_item<MODE>(B, m, min(r2, last), g2, lane4);
        eval_item<MODE, TAB>(A, m, tab, row, seg, lane4);
        if (r2 >= y1) break;
        int r3 = r2, g3 = g2 + nwaves;
        while (g3 >= nseg) { g3 -= nseg; ++r3; }
        load_item<MODE>(A, m, min(r3, last), g3, lane4);
        eval_item<MODE, TAB>(B, m, tab, r2, g2, lane4);
        row = r3;
        seg = g3;
    }
}

__device__ __forceinline__ MapGeo map_geo(const uint32_t *in, uint32_t *out, uint32_t *clear,
                                          int W, int H, int step, uint32_t none)
{
    MapGeo m;
    const uint32_t bytes = 4u * (uint32_t)W * (uint32_t)H;
    m.rin = buf_rsrc(in, bytes);
    m.rout = buf_rsrc(out, bytes);
    m.rclr = buf_rsrc(clear ? clear : out, bytes);
    m.row_bytes = 4u * (uint32_t)W;
    m.W = W;
    m.H = H;
    m.step = step;
    m.clear = clear != nullptr;
    m.none = none;
    return m;
}

// share [a, b) of N work units for workgroup g of G: shares are dealt so
// that each XCD (blocks go to the 8 XCDs round robin) owns a contiguous
// stretch of units
__device__ __forceinline__ void band_of(int g, int G, int N, int &a, int &b)
{
    const int q8 = G >> 3, r8 = G & 7, xcd = g & 7;
    const int band = xcd * q8 + min(xcd, r8) + (g >> 3);
    // N * (band + 1) < 2^32: N is at most a map's run or row count (<= 2^16)
    a = (int)((uint32_t)N * (uint32_t)band / (uint32_t)G);
    b = (int)((uint32_t)N * (uint32_t)(band + 1) / (uint32_t)G);
}

// Runs.  The rows of a map fall into chains of rows one step apart
// (y = c, c + k, c + 2k, ... for c < min(k, H)); consecutive rows of a chain
// share two of their three tap rows.  A run is four consecutive rows of a
// chain in one 256-pixel segment, so its items need 6 tap rows instead of 12.
// Runs are numbered by position in the chain, then chain, then segment: the
// runs that start together cover neighbouring rows (a contiguous share of
// runs is a compact band of first rows, as in quad_rows).
struct RunGeo {
    int nseg, nq, k, chains;  // segments per row, runs per chain, step, chains
    int count;                // runs in the map
};

__device__ __forceinline__ RunGeo run_geo(const MapGeo &m)
{
    RunGeo g;
    g.k = m.step;
    g.nseg = (m.W + 255) >> 8;
    g.chains = min(m.step, m.H);
    // longest chain (steps are powers of two)
    const int n = (m.H + m.step - 1) >> __builtin_ctz((unsigned)m.step);
    g.nq = (n + 3) >> 2;
    g.count = g.chains * g.nq * g.nseg;
    return g;
}

// u / d and u % d for wave-uniform u and d: shifts when d is a power of two
// (the segment count of the usual grids, the chain count of every step up to
// the grid's height), a division otherwise
__device__ __forceinline__ void divmod_u(int u, int d, int &quo, int &rem)
{
    if ((d & (d - 1)) == 0) {
        quo = u >> __builtin_ctz((unsigned)d);
        rem = u & (d - 1);
    } else {
        quo = u / d;
        rem = u - quo * d;
    }
}

// first centre row and segment of run u (wave-uniform)
__device__ __forceinline__ void run_of(const RunGeo &g, int u, int &y, int &seg)
{
    u = __builtin_amdgcn_readfirstlane(u);
    int qc, c, q;
    divmod_u(u, g.nseg, qc, seg);
    divmod_u(qc, g.chains, q, c);
    y = __builtin_amdgcn_readfirstlane(c + 4 * q * g.k);
    seg = __builtin_amdgcn_readfirstlane(seg);
}

// One tap row of a run, evaluated as it streams in.  A run's output rows are
// y + j*k (j = 0..3) and its tap rows y + (i - 1)*k (i = 0..5); tap row I is
// the down tap of output row I - 2, the centre of I - 1 and the up tap of I.
// Each label of the row is looked up once, its dx*dx computed once per pixel
// column, and then folded into the (up to three) output rows that use it:
// dy, fma and min per use.  Off-grid tap rows are never loaded (their
// redirected candidates duplicate the centre row's); a column quad that is
// off the grid for every lane of the wave is skipped for the same reason,
// and one that is off for some lanes holds the centre quad's labels there.
template <int MODE, int TAB, int I>
__device__ __forceinline__ void run_row(const RowData<MODE> &R, const SiteTab<TAB> &tab,
                                        const float (&px)[4], const float (&py)[4],
                                        double (&best)[4][4], bool lok, bool rok, bool left,
                                        bool right)
{
    constexpr int j0 = I - 2 < 0 ? 0 : I - 2;
    constexpr int j1 = I > 3 ? 3 : I;
    if constexpr (MODE == STEP_ANY) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (q == 0 && !left) continue;
            if (q == 2 && !right) continue;
            const u32x4 &Q = q == 0 ? R.l : q == 1 ? R.c : R.r;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const uint32_t l = q_el(Q, jj);
                const float2 st = tab(l);
                const float dx = st.x - px[jj];
                const float dx2 = dx * dx;
                // one key register per label: lo = label, hi = each use's d
                u32x2 kv;
                kv.x = l;
#pragma unroll
                for (int j = j0; j <= j1; ++j) {
                    const float dy = st.y - py[j];
                    kv.y = __float_as_uint(__builtin_fmaf(dy, dy, dx2));
                    best[j][jj] = __builtin_fmin(best[j][jj], __builtin_bit_cast(double, kv));
                }
            }
        }
    } else {
        // tap positions x - MODE .. x + 3 + MODE of the row; pixel jj takes
        // positions jj, jj + MODE, jj + 2 * MODE (left, centre, right)
        constexpr int NP = (MODE == STEP_1) ? 6 : 8;
        const u32x4 &c = R.c;
        uint32_t P[NP];
        if constexpr (MODE == STEP_1) {
            P[0] = lok ? R.l : c.x;
            P[1] = c.x; P[2] = c.y; P[3] = c.z; P[4] = c.w;
            P[5] = rok ? R.r : c.w;
        } else {
            P[0] = lok ? R.l.x : c.x;
            P[1] = lok ? R.l.y : c.y;
            P[2] = c.x; P[3] = c.y; P[4] = c.z; P[5] = c.w;
            P[6] = rok ? R.r.x : c.z;
            P[7] = rok ? R.r.y : c.w;
        }
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const float2 st = tab(P[k]);
            float dy[4];
#pragma unroll
            for (int j = j0; j <= j1; ++j) dy[j] = st.y - py[j];
            u32x2 kv;
            kv.x = P[k];
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const int jj = k - t * MODE;
                if (jj < 0 || jj > 3) continue;
                const float dx = st.x - px[jj];
                const float dx2 = dx * dx;
#pragma unroll
                for (int j = j0; j <= j1; ++j) {
                    kv.y = __float_as_uint(__builtin_fmaf(dy[j], dy[j], dx2));
                    best[j][jj] = __builtin_fmin(best[j][jj], __builtin_bit_cast(double, kv));
                }
            }
        }
    }
}

// the largest key: +inf distance with an all-ones label, above every
// candidate (the sentinel's key is (+inf, none) with none < 0xffffffff)
__device__ __forceinline__ double key_max() { return __hiloint2double(0x7f800000, -1); }

// Runs u, u + stride, ... below u_end of one map, by one wave.  The wave
// streams a run's six tap rows through three row buffers, two rows ahead of
// the one being evaluated; during the last two rows it loads the next run's
// first two.  Rows past the grid end their chain: their outputs are not
// stored and their loads not issued.  With `stage` the site table is staged
// after the wave's first loads are issued, so the two latencies overlap.
template <int MODE, int TAB, class PRO = NoPrologue>
__device__ __forceinline__ void wave_runs(const MapGeo &m, const SiteTab<TAB> &tab,
                                          const RunGeo &g, int u, int u_end, int stride,
                                          bool stage, const PRO &pro = PRO())
{
    const int lane4 = 4 * (threadIdx.x & 63);
    const int k = m.step, H = m.H, W = m.W;
    RowData<MODE> R0, R1, R2;
    int y = 0, seg = 0;
    if (u < u_end) {
        run_of(g, u, y, seg);
        if (y - k >= 0) load_row<MODE>(R0, m, y - k, seg, lane4);
        if (y < H) load_row<MODE>(R1, m, y, seg, lane4);
    }
    if (stage) tab.stage();
    if (pro()) return;
    while (u < u_end) {
        const int un = u + stride;
        int yn = 0, sn = 0;
        if (un < u_end) run_of(g, un, yn, sn);
        const int xb = seg * 256 + lane4;
        const int xc = min(xb, W - 4);
        float px[4], py[4];
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) px[jj] = (float)(xc + jj) + 0.5f;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            py[j] = __int_as_float(
                __builtin_amdgcn_readfirstlane(__float_as_int((float)(y + j * k) + 0.5f)));
        const bool lok = xc > 0, rok = xc + 4 < W;
        bool left = true, right = true;
        if constexpr (MODE == STEP_ANY) {
            left = __ballot(xc - k >= 0) != 0ull;
            right = __ballot(xc + 3 + k < W) != 0ull;
        }
        double best[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) best[j][jj] = key_max();

        if (y + k < H) load_row<MODE>(R2, m, y + k, seg, lane4);
        if (y - k >= 0) run_row<MODE, TAB, 0>(R0, tab, px, py, best, lok, rok, left, right);
        if (y + 2 * k < H) load_row<MODE>(R0, m, y + 2 * k, seg, lane4);
        if (y < H) run_row<MODE, TAB, 1>(R1, tab, px, py, best, lok, rok, left, right);
        if (y + 3 * k < H) load_row<MODE>(R1, m, y + 3 * k, seg, lane4);
        if (y + k < H) run_row<MODE, TAB, 2>(R2, tab, px, py, best, lok, rok, left, right);
        if (y + 4 * k < H) load_row<MODE>(R2, m, y + 4 * k, seg, lane4);
        if (y + 2 * k < H) run_row<MODE, TAB, 3>(R0, tab, px, py, best, lok, rok, left, right);
        if (un < u_end && yn - k >= 0) load_row<MODE>(R0, m, yn - k, sn, lane4);
        if (y + 3 * k < H) run_row<MODE, TAB, 4>(R1, tab, px, py, best, lok, rok, left, right);
        if (un < u_end && yn < H) load_row<MODE>(R1, m, yn, sn, lane4);
        if (y + 4 * k < H) run_row<MODE, TAB, 5>(R2, tab, px, py, best, lok, rok, left, right);

        if (xb < W) {
            const uint32_t o1 = 4u * (uint32_t)xc;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int yj = y + j * k;
                if (yj >= H) break;
                const uint32_t s1 = row_off(m, yj);
                const u32x4 r4 = {(uint32_t)__double2loint(best[j][0]),
                                  (uint32_t)__double2loint(best[j][1]),
                                  (uint32_t)__double2loint(best[j][2]),
                                  (uint32_t)__double2loint(best[j][3])};
                __builtin_amdgcn_raw_buffer_store_b128(r4, m.rout, o1, s1, CVT_MAP_STORE_POLICY);
                if (m.clear) {
                    const u32x4 n4 = {m.none, m.none, m.none, m.none};
                    __builtin_amdgcn_raw_buffer_store_b128(n4, m.rclr, o1, s1, CVT_MAP_STORE_POLICY);
                }
            }
        }
        u = un;
        y = yn;
        seg = sn;
    }
}

// Runs [u0, u1) of one map by all waves of the block, dealt round robin.
template <int MODE, int TAB, class PRO>
__device__ __forceinline__ void quad_runs(const MapGeo &m, const SiteTab<TAB> &tab, int u0, int u1,
                                          const PRO &pro)
{
    const int nwaves = blockDim.x >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    wave_runs<MODE, TAB>(m, tab, run_geo(m), u0 + wave, u1, nwaves, true, pro);
}

// Single map, site table staged at LDS byte 0 in a static array of ENTRIES
// float2 (the kernel's only LDS), gathered from global memory when
// ENTRIES == 0, or split when ENTRIES < 0: the first -ENTRIES sites staged
// in LDS, the rest gathered.  RUNS picks the work layout: runs down chains
// of rows, or bands of rows (quad_rows).
#define JFA_THREADS(E) ((E) > 5120 || (E) < 0 ? 1024 : 256)
#define JFA_MINB(E) ((E) > 5120 || (E) < 0 ? 1 : 4)

template <int MODE, int TAB, class PRO>
__device__ __forceinline__ void quad_layout(const MapGeo &m, const SiteTab<TAB> &tab, bool runs,
                                            const PRO &pro)
{
    int a, b;
    if (runs) {
        band_of(blockIdx.x, gridDim.x, run_geo(m).count, a, b);
        quad_runs<MODE, TAB>(m, tab, a, b, pro);
    } else {
        band_of(blockIdx.x, gridDim.x, m.H, a, b);
        quad_rows<MODE, TAB>(m, tab, a, b, pro);
    }
}

// the pixel a site seeds (clamped to the grid): the seed scatter's rule
__device__ __forceinline__ int seed_pixel(float v, int n)
{
    float f = floorf(v);
    f = fminf(fmaxf(f, 0.0f), (float)(n - 1));
    return (int)f;
}

// The stop word (0 without one), as a vector load: it returns in order with
// the pass's first label loads, so testing it after them costs no wait of its
// own (a scalar load would be waited for with the kernel arguments).  A null
// stop gives an empty buffer range, whose load returns 0.
__device__ __forceinline__ int stop_word(const int *stop)
{
    return (int)__builtin_amdgcn_raw_buffer_load_b32(buf_rsrc(stop, stop ? 4u : 0u), 0, 0, 0);
}

// The single-map pass's prologue.  halt: the stop word of cvt_lloyd_until,
// loaded at kernel entry and tested here, so its latency overlaps the first
// loads (a stopped pass writes nothing).  nclear > 0: clear is a map whose
// only entries sit at the seed pixels q of sites [0, nclear) (rstep == 0),
// or at q + rstep * {-1, 0, 1}^2 (rstep > 0, the receivers of a seed pass);
// they are reset one store per pixel of that set instead of W * H stores
// (the pass does not read that map).
struct SparseReset {
    int halt;
    uint32_t *clear;
    const float2 *sites;
    int W, H, nclear, rstep;
    uint32_t none;
    __device__ bool operator()() const
    {
        if (halt) return true;
        if (clear && nclear > 0) {
            const __amdgpu_buffer_rsrc_t rc = buf_rsrc(clear, 4u * (uint32_t)W * (uint32_t)H);
            const int R = rstep > 0 ? 9 : 1;
            for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < R * nclear;
                 t += gridDim.x * blockDim.x) {
                const int i = t / R, r = t - R * (t / R);
                const float2 v = sites[i];
                int x = seed_pixel(v.x, W), y = seed_pixel(v.y, H);
                if (R == 9) {
                    x += (r % 3 - 1) * rstep;
                    y += (r / 3 - 1) * rstep;
                }
                if (x >= 0 && x < W && y >= 0 && y < H)
                    __builtin_amdgcn_raw_buffer_store_b32(none, rc, 4u * (uint32_t)(y * W + x), 0,
                                                          0);
            }
        }
        return false;
    }
};

template <int MODE, int ENTRIES, bool RUNS>
__global__ __launch_bounds__(JFA_THREADS(ENTRIES), JFA_MINB(ENTRIES)) void k_jfa_quad(
    const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
    const float2 *__restrict__ sites, int table_f4, int W, int H, int step,
    uint32_t *__restrict__ clear, uint32_t none, int nclear, int rstep, const int *__restrict__ stop)
{
    SparseReset pro;
    pro.halt = stop_word(stop);
    pro.clear = clear;
    pro.sites = sites;
    pro.W = W;
    pro.H = H;
    pro.nclear = nclear;
    pro.rstep = rstep;
    pro.none = none;
    const MapGeo m = map_geo(in, out, nclear > 0 ? nullptr : clear, W, H, step, none);
    if constexpr (ENTRIES > 0) {
        __shared__ float4 lds_tab[ENTRIES / 2];
        SiteTab<TAB_LDS0> tab;
        tab.lds = (lds_cchar *)lds_tab;
        tab.src = reinterpret_cast<const f32x4 *>(sites);
        tab.n4 = table_f4;
        tab.base = 0;
        quad_layout<MODE, TAB_LDS0>(m, tab, RUNS, pro);
    } else if constexpr (ENTRIES < 0) {
        __shared__ float4 lds_tab[-ENTRIES / 2];
        // entries 0 .. -ENTRIES - 3 staged, then a zero pair; the rest of the
        // table (from entry -ENTRIES - 2) is gathered
        constexpr int kept = -ENTRIES - 2;
        SiteTab<TAB_SPLIT> tab;
        tab.lds = (lds_cchar *)lds_tab;
        tab.src = reinterpret_cast<const f32x4 *>(sites);
        tab.n4 = kept / 2;
        tab.g = buf_rsrc(sites + kept, 8u * (uint32_t)(2 * table_f4 - kept));
        tab.split = 8u * (uint32_t)kept;
        tab.base = 0;
        quad_layout<MODE, TAB_SPLIT>(m, tab, RUNS, pro);
    } else {
        SiteTab<TAB_GLOBAL> tab;
        tab.g = buf_rsrc(sites, 16u * (uint32_t)table_f4);
        tab.base = 0;
        tab.n4 = 0;
        quad_layout<MODE, TAB_GLOBAL>(m, tab, RUNS, pro);
    }
}

// ---------------------------------------------------------------------------
// Fused tail: the last five passes (steps 8, 4, 2, 1 and the extra 1) of one
// map in one launch, for site tables that fit in LDS beside the tile.  Each
// workgroup owns a kTailT x kTailT tile of the output.  It loads the tile and
// a kTailHalo-pixel halo of the step-16 map into LDS and runs the five
// passes there, each over the tile grown by the steps still to come (8, 4,
// 2, 1, 0 pixels), so every tap a pass reads was written by the pass before
// it.  Taps off the grid are redirected to the pixel's own row or column,
// as in the dense pass, so the grid edges need no halo.  Each pixel takes
// the (d, label) minimum over its nine taps with the dense pass's distance
// formula: the map equals five dense passes bit for bit.
// ---------------------------------------------------------------------------
static constexpr int kTailT = 32, kTailHalo = 16, kTailP = kTailT + 2 * kTailHalo;

template <int ENTRIES>
__global__ __launch_bounds__(1024) void k_jfa_tail(const uint32_t *__restrict__ in,
                                                   uint32_t *__restrict__ out,
                                                   const float2 *__restrict__ sites, int table_f4,
                                                   int W, int H, int tilesX,
                                                   const int *__restrict__ stop)
{
    // the site table at LDS byte 0, then the two tile buffers
    __shared__ float4 lds_all[ENTRIES / 2 + kTailP * kTailP / 2];
    uint32_t *const buf0 = reinterpret_cast<uint32_t *>(lds_all + ENTRIES / 2);
    uint32_t *const buf1 = buf0 + kTailP * kTailP;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int tY = blockIdx.x / tilesX, tX = blockIdx.x - tY * tilesX;
    const int X0 = tX * kTailT, Y0 = tY * kTailT;
    const int bx = X0 - kTailHalo, by = Y0 - kTailHalo;   // the buffers' origin on the grid
    const int halt = stop_word(stop);
    {
        const int x0 = max(bx, 0), x1 = min(X0 + kTailT + kTailHalo, W);
        const int y0 = max(by, 0), y1 = min(Y0 + kTailT + kTailHalo, H);
        const int w = x1 - x0, n = w * (y1 - y0);
        for (int i = tid; i < n; i += nt) {
            const int r = i / w, x = x0 + (i - r * w), y = y0 + r;
            buf0[(y - by) * kTailP + (x - bx)] = in[(long)y * W + x];
        }
    }
    SiteTab<TAB_LDS0> tab;
    tab.lds = (lds_cchar *)lds_all;
    tab.src = reinterpret_cast<const f32x4 *>(sites);
    tab.n4 = table_f4;
    tab.base = 0;
    tab.stage();                     // ends with a barrier: buf0 is complete too
    if (halt) return;
    uint32_t *cur = buf0, *nxt = buf1;
#pragma unroll 1
    for (int k = 0; k < 5; ++k) {
        const int s = k == 0 ? 8 : k == 1 ? 4 : k == 2 ? 2 : 1;
        const int g = k == 0 ? 8 : k == 1 ? 4 : k == 2 ? 2 : k == 3 ? 1 : 0;   // halo after
        const int x0 = max(X0 - g, 0), x1 = min(X0 + kTailT + g, W);
        const int y0 = max(Y0 - g, 0), y1 = min(Y0 + kTailT + g, H);
        const int w = x1 - x0, n = w * (y1 - y0);
        for (int i = tid; i < n; i += nt) {
            const int r = i / w, x = x0 + (i - r * w), y = y0 + r;
            // tap columns and rows, off-grid ones redirected to the pixel's own
            const int cx[3] = {x - s >= 0 ? x - s : x, x, x + s < W ? x + s : x};
            const int cy[3] = {y - s >= 0 ? y - s : y, y, y + s < H ? y + s : y};
            const float px = (float)x + 0.5f, py = (float)y + 0.5f;
            double best = 0.0;
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const uint32_t l = cur[(cy[t / 3] - by) * kTailP + (cx[t % 3] - bx)];
                const double key = cand(tab(l), l, px, py);
                best = (t == 0) ? key : __builtin_fmin(best, key);
            }
            nxt[(y - by) * kTailP + (x - bx)] = (uint32_t)__double2loint(best);
        }
        __syncthreads();
        uint32_t *const t = cur;
        cur = nxt;
        nxt = t;
    }
    const int x1 = min(X0 + kTailT, W), y1 = min(Y0 + kTailT, H);
    const int w = x1 - X0, n = w * (y1 - Y0);
    for (int i = tid; i < n; i += nt) {
        const int r = i / w, x = X0 + (i - r * w), y = Y0 + r;
        out[(long)y * W + x] = cur[(y - by) * kTailP + (x - bx)];
    }
}

bool jfa_tail_ok(const JfaPlan &p, int passes)
{
    // the dense passes 1 and 2 (which reset the seed and first-pass maps)
    // stay before the tail; the table fits beside the tile buffers
    return passes >= 8 && (p.lds_entries == 2048 || p.lds_entries == 5120);
}

void launch_jfa_tail(hipStream_t st, const JfaPlan &p, const uint32_t *in, uint32_t *out,
                     const float2 *sites, int W, int H, const int *stop)
{
    const int tx = (W + kTailT - 1) / kTailT, ty = (H + kTailT - 1) / kTailT;
    if (p.lds_entries == 2048)
        hipLaunchKernelGGL(k_jfa_tail<2048>, dim3(tx * ty), dim3(1024), 0, st, in, out, sites,
                           p.table_f4, W, H, tx, stop);
    else
        hipLaunchKernelGGL(k_jfa_tail<5120>, dim3(tx * ty), dim3(1024), 0, st, in, out, sites,
                           p.table_f4, W, H, tx, stop);
}

// ---------------------------------------------------------------------------
// Batched pass: one launch over many patches.  Each workgroup takes one work
// unit (patch, band of rows); patches run their passes right-aligned, so pass
// k of the batch is pass k - (Pmax - P) of a patch with P passes (skipped
// when negative) and every patch ends its JFA in the same launch.
// ---------------------------------------------------------------------------
template <bool SITES_IN_LDS, int MODE>
__global__ __launch_bounds__(1024) void k_jfa_batch(const BatchPatch *__restrict__ patches,
                                                    const BatchUnit *__restrict__ units,
                                                    uint32_t *__restrict__ seed,
                                                    uint32_t *__restrict__ first,
                                                    uint32_t *__restrict__ lab0,
                                                    uint32_t *__restrict__ lab1,
                                                    const float2 *__restrict__ sites, int pass,
                                                    int pmax)
{
    extern __shared__ float4 lds_dyn[];
    const BatchUnit u = units[blockIdx.x];
    const BatchPatch P = patches[u.patch];
    const int local = pass - (pmax - P.passes);
    if (local < 0) return;                                   // uniform per block
    const bool scatter = P.passes >= 3;
    if (scatter && local == 0) return;                       // the seed scatter's pass
    const int step = (local < P.passes - 1) ? (P.first_step >> local) : 1;
    if ((step == 1 ? STEP_1 : step == 2 ? STEP_2 : STEP_ANY) != MODE) return;
    uint32_t *bufs[2] = {lab0 + P.pix_off, lab1 + P.pix_off};
    const uint32_t *in = (local == 0)                 ? seed + P.pix_off
                         : (local == 1 && scatter) ? first + P.pix_off
                                                   : bufs[(local - 1) & 1];
    uint32_t *out = bufs[local & 1];
    uint32_t *clear = (local == 1 && !scatter) ? seed + P.pix_off : nullptr;
    const uint32_t none = 8u * (uint32_t)(P.site_off + P.cap);
    if (scatter && (local == 1 || local == 2)) {
        // this unit's share of the sparse reset: the seed pixels (pass 1) or
        // the seeds' receivers at the first step (pass 2); neither map is
        // read or written by this pass
        uint32_t *map = (local == 1 ? seed : first) + P.pix_off;
        const int R = local == 2 ? 9 : 1;
        const int s0 = (int)((long)P.S * u.b / u.nb), s1 = (int)((long)P.S * (u.b + 1) / u.nb);
        for (int t = threadIdx.x; t < R * (s1 - s0); t += blockDim.x) {
            const int i = s0 + t / R, r = t - R * (t / R);
            const float2 v = sites[P.site_off + i];
            int x = seed_pixel(v.x, P.W), y = seed_pixel(v.y, P.H);
            if (R == 9) {
                x += (r % 3 - 1) * P.first_step;
                y += (r / 3 - 1) * P.first_step;
            }
            if (x >= 0 && x < P.W && y >= 0 && y < P.H) map[(long)y * P.W + x] = none;
        }
    }
    const MapGeo m = map_geo(in, out, clear, P.W, P.H, step, none);
    // part b of the pass's runs (the planner gives a patch at most one part
    // per 16 runs, one per wave)
    const RunGeo g = run_geo(m);
    const int nwaves = blockDim.x >> 6;
    const int u0 = (int)((long)g.count * u.b / u.nb), u1 = (int)((long)g.count * (u.b + 1) / u.nb);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (SITES_IN_LDS) {
        SiteTab<TAB_LDSB> tab;
        tab.lds = (lds_cchar *)lds_dyn;
        tab.src = reinterpret_cast<const f32x4 *>(sites + P.site_off);
        tab.n4 = (P.cap + 2) / 2;
        tab.base = 8u * (uint32_t)P.site_off;
        wave_runs<MODE, TAB_LDSB>(m, tab, g, u0 + wave, u1, nwaves, true);
    } else {
        SiteTab<TAB_GLOBAL> tab;
        tab.g = buf_rsrc(sites, 0xffffffffu);
        tab.base = 0;
        tab.n4 = 0;
        wave_runs<MODE, TAB_GLOBAL>(m, tab, g, u0 + wave, u1, nwaves, false);
    }
}

template <bool LDS, int MODE>
static void launch_batch_mode(hipStream_t st, const BatchPatch *patches, const BatchUnit *units,
                              int nunits, uint32_t *seed, uint32_t *first, uint32_t *lab0,
                              uint32_t *lab1, const float2 *sites, int pass, int pmax,
                              size_t lds_bytes)
{
    if (LDS) {
        static bool attr = false;
        if (!attr) {
            (void)hipFuncSetAttribute((const void *)k_jfa_batch<LDS, MODE>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            attr = true;
        }
    }
    hipLaunchKernelGGL((k_jfa_batch<LDS, MODE>), dim3(nunits), dim3(1024), LDS ? lds_bytes : 0, st,
                       patches, units, seed, first, lab0, lab1, sites, pass, pmax);
}

void launch_jfa_batch(hipStream_t st, const BatchPatch *patches, const BatchUnit *units, int nunits,
                      uint32_t *seed, uint32_t *first, uint32_t *lab0, uint32_t *lab1,
                      const float2 *sites, int pass, int pmax, size_t lds_bytes, unsigned modes)
{
#define CVT_BATCH_MODE(M)                                                                      \
    if (modes & (1u << (M))) {                                                                 \
        if (lds_bytes)                                                                         \
            launch_batch_mode<true, M>(st, patches, units, nunits, seed, first, lab0, lab1,    \
                                       sites, pass, pmax, lds_bytes);                          \
        else                                                                                   \
            launch_batch_mode<false, M>(st, patches, units, nunits, seed, first, lab0, lab1,   \
                                        sites, pass, pmax, 0);                                 \
    }
    CVT_BATCH_MODE(STEP_ANY)
    CVT_BATCH_MODE(STEP_1)
    CVT_BATCH_MODE(STEP_2)
#undef CVT_BATCH_MODE
}

#ifndef CVT_JFA_ROWS
#define CVT_JFA_ROWS 4
#endif

static int g_num_cu = 0;

// static LDS table sizes of the single-map kernel, in float2 entries
static const int kLdsEntries[3] = {2048, 5120, 20480};

JfaPlan plan_jfa(int W, int H, int capS, int runs_mask, int blocks_per_cu)
{
    if (!g_num_cu) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        hipDeviceProp_t prop;
        g_num_cu = (hipGetDeviceProperties(&prop, dev) == hipSuccess) ? prop.multiProcessorCount
                                                                      : 256;
    }
    JfaPlan p;
    p.quad = (W % 4 == 0) && W >= 8;
    // sites and the sentinel (index capS), as whole float4 pairs; the site
    // buffer holds capS + 2 entries, so the last pair is always inside it
    p.table_f4 = (capS + 2) / 2;
    const int entries = 2 * p.table_f4;
    p.lds_entries = 0;
    if (p.quad) {
        for (int e : kLdsEntries)
            if (entries <= e) { p.lds_entries = e; break; }
        // too large for LDS: the first 20480 entries there, the rest gathered
        if (p.lds_entries == 0) p.lds_entries = -kLdsEntries[2];
    }
    p.lds = p.lds_entries > 0;
    p.lds_bytes = p.lds_entries != 0 ? (size_t)std::abs(p.lds_entries) * 8 : 0;
    const long items = (long)H * ((W + 255) / 256);
    if (p.lds_entries > 5120 || p.lds_entries < 0) {
        p.threads = 1024;                               // one workgroup per CU
        p.grid = g_num_cu;
    } else {
        p.threads = 256;
        p.grid = g_num_cu * 4;
    }
    p.runs_mask = runs_mask < 0 ? JFA_AUTO : runs_mask;
    p.blocks_per_cu = blocks_per_cu;
    if (blocks_per_cu > 0) p.grid = g_num_cu * blocks_per_cu;
    // at least one item per wave
    const long waves = (long)p.grid * (p.threads / 64);
    if (waves > items) p.grid = (int)std::max<long>(1, items / (p.threads / 64));
    if (p.grid > H) p.grid = H;
    return p;
}

template <int MODE, bool RUNS>
static void launch_quad(hipStream_t st, const JfaPlan &p, const uint32_t *in, uint32_t *out,
                        const float2 *sites, int W, int H, int step, uint32_t *clear, uint32_t none,
                        int nclear, int rstep, const int *stop)
{
    const dim3 g(p.grid), b(p.threads);
    switch (p.lds_entries) {
    case 2048:
        hipLaunchKernelGGL((k_jfa_quad<MODE, 2048, RUNS>), g, b, 0, st, in, out, sites, p.table_f4,
                           W, H, step, clear, none, nclear, rstep, stop);
        break;
    case 5120:
        hipLaunchKernelGGL((k_jfa_quad<MODE, 5120, RUNS>), g, b, 0, st, in, out, sites, p.table_f4,
                           W, H, step, clear, none, nclear, rstep, stop);
        break;
    case 20480:
        hipLaunchKernelGGL((k_jfa_quad<MODE, 20480, RUNS>), g, b, 0, st, in, out, sites,
                           p.table_f4, W, H, step, clear, none, nclear, rstep, stop);
        break;
    case -20480:
        hipLaunchKernelGGL((k_jfa_quad<MODE, -20480, RUNS>), g, b, 0, st, in, out, sites,
                           p.table_f4, W, H, step, clear, none, nclear, rstep, stop);
        break;
    default:
        hipLaunchKernelGGL((k_jfa_quad<MODE, 0, RUNS>), g, b, 0, st, in, out, sites, p.table_f4, W,
                           H, step, clear, none, nclear, rstep, stop);
        break;
    }
}

// Layout by step.  Runs are evaluated as their tap rows stream in (each
// label's dx^2 serves up to three output rows), so they take fewer VALU
// operations per pixel than bands at every step: at 2048^2 / 16K sites the
// banded steps 128 and 64 took 13.3 us against 10.8 for the runs of the other
// steps >= 4 (rocprof means), and runs at every step measured 6547 against
// 6510 iter/s.  Runs need work for every wave, though: a run is four rows of
// a 256-pixel segment, so on small grids (512^2: 256 runs at step 1 for 1024
// waves) bands keep more waves busy (15 870 against 13 170 iter/s with runs
// whenever they fill a quarter of the waves).
static bool use_runs(const JfaPlan &p, int W, int H, int step)
{
    if (p.runs_mask >= 0) return (p.runs_mask >> __builtin_ctz((unsigned)step)) & 1;
    const long n = (H + step - 1) / step;
    const long runs = (long)std::min(step, H) * ((n + 3) / 4) * ((W + 255) / 256);
    return runs >= (long)p.grid * (p.threads / 64);
}

int jfa_runs_mask(const JfaPlan &p, int W, int H)
{
    // steps 1, 2, 4 .. the first step (the largest power of two below max(W, H))
    int mask = 0;
    if (p.quad)
        for (int k = 0;; ++k) {
            if (use_runs(p, W, H, 1 << k)) mask |= 1 << k;
            if ((2 << k) >= std::max(W, H)) break;
        }
    return mask;
}

void launch_jfa_pass(hipStream_t st, const JfaPlan &p, const uint32_t *in, uint32_t *out,
                     const float2 *sites, int W, int H, int step, uint32_t *clear, uint32_t none,
                     int nclear, int rstep, const int *stop)
{
    if (p.quad) {
        const bool runs = use_runs(p, W, H, step);
        if (step == 1 && runs)
            launch_quad<STEP_1, true>(st, p, in, out, sites, W, H, step, clear, none, nclear, rstep, stop);
        else if (step == 1)
            launch_quad<STEP_1, false>(st, p, in, out, sites, W, H, step, clear, none, nclear, rstep, stop);
        // (step 2 with the all-global table keeps bands: six buffers of it
        // spill)
        else if (step == 2 && runs && p.lds_entries != 0)
            launch_quad<STEP_2, true>(st, p, in, out, sites, W, H, step, clear, none, nclear, rstep, stop);
        else if (step == 2) launch_quad<STEP_2, false>(st, p, in, out, sites, W, H, step, clear, none, nclear, rstep, stop);
        else if (runs)
            launch_quad<STEP_ANY, true>(st, p, in, out, sites, W, H, step, clear, none, nclear, rstep, stop);
        else launch_quad<STEP_ANY, false>(st, p, in, out, sites, W, H, step, clear, none, nclear, rstep, stop);
        return;
    }
    constexpr int R = CVT_JFA_ROWS;
    const int tilesX = (W + 63) / 64, tilesY = (H + 4 * R - 1) / (4 * R);
    const int ntiles = tilesX * tilesY;
    hipLaunchKernelGGL(k_jfa_pass<R>, dim3(ntiles), dim3(256), 0, st, in, out, sites, W, H, tilesX,
                       ntiles, step, clear, none, stop);
}


// ---------------------------------------------------------------------------
// First pass from a seed map, as a scatter of the seeds.  The seed map holds
// one label at the seed pixel of each winning site and none elsewhere, so the
// pass's output is none except at the pixels whose taps hit a seed: the
// receivers p = q - t of each seed pixel q (t in step * {-1, 0, 1}^2; the
// off-grid tap redirection only ever adds the centre, which is already in the
// set).  The output map must hold none everywhere; one thread per (site,
// offset) evaluates its receiver exactly as the dense pass would (all nine
// taps, same candidate order) and writes it.  A receiver of several seeds
// (or of a pixel several sites share) is written with the same value by
// each.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_jfa_seeds(const uint32_t *__restrict__ seed,
                                                   uint32_t *__restrict__ out,
                                                   const float2 *__restrict__ sites, int S, int W,
                                                   int H, int step, const int *__restrict__ stop,
                                                   int fin, const double *__restrict__ blockpart,
                                                   int *__restrict__ dstate, FinalArgs fa)
{
    if (fin && blockIdx.x == 0) {
        // the previous iteration's deferred finalize, beside the scatter
        __shared__ double red_e[256], red_g[256];
        const int nb = dstate[ST_PEND];
        if (nb > 0) cvt_finalize_block(blockpart, nb, dstate, fa, red_e, red_g);
        return;
    }
    const int halt = stop_word(stop);   // tested once the site's load is in flight
    const int t = (blockIdx.x - fin) * blockDim.x + threadIdx.x;
    const int i = t / 9, r = t - 9 * (t / 9);
    if (i >= S) return;
    const float2 s = sites[i];
    if (halt) return;
    const int qx = seed_pixel(s.x, W), qy = seed_pixel(s.y, H);
    const int px = qx - (r % 3 - 1) * step, py = qy - (r / 3 - 1) * step;
    if (px < 0 || px >= W || py < 0 || py >= H) return;
    const char *sb = reinterpret_cast<const char *>(sites);
    const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
    double best = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int ox = (k % 3 - 1) * step, oy = (k / 3 - 1) * step;
        const int tx = (px + ox >= 0 && px + ox < W) ? px + ox : px;
        const int ty = (py + oy >= 0 && py + oy < H) ? py + oy : py;
        const uint32_t l = seed[(long)ty * W + tx];
        const double key = cand(*reinterpret_cast<const float2 *>(sb + l), l, fx, fy);
        best = (k == 0) ? key : __builtin_fmin(best, key);
    }
    out[(long)py * W + px] = (uint32_t)__double2loint(best);
}

// pass 0 of the patches whose own pass 0 falls at batched pass `pass`: the
// seed scatter of k_jfa_seeds, one thread per (site slot, receiver)
__global__ __launch_bounds__(256) void k_jfa_batch_seeds(const BatchPatch *__restrict__ patches,
                                                         const int *__restrict__ site_patch,
                                                         int total, const uint32_t *__restrict__ seed,
                                                         uint32_t *__restrict__ first,
                                                         const float2 *__restrict__ sites,
                                                         int pass, int pmax)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int gi = t / 9, r = t - 9 * (t / 9);
    if (gi >= total) return;
    const int pid = site_patch[gi];
    if (pid < 0) return;
    const BatchPatch P = patches[pid];
    if (pass - (pmax - P.passes) != 0 || P.passes < 3 || gi - P.site_off >= P.S) return;
    const int W = P.W, H = P.H, step = P.first_step;
    const float2 s = sites[gi];
    const int qx = seed_pixel(s.x, W), qy = seed_pixel(s.y, H);
    const int px = qx - (r % 3 - 1) * step, py = qy - (r / 3 - 1) * step;
    if (px < 0 || px >= W || py < 0 || py >= H) return;
    const uint32_t *sm = seed + P.pix_off;
    const char *sb = reinterpret_cast<const char *>(sites);
    const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
    double best = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int ox = (k % 3 - 1) * step, oy = (k / 3 - 1) * step;
        const int tx = (px + ox >= 0 && px + ox < W) ? px + ox : px;
        const int ty = (py + oy >= 0 && py + oy < H) ? py + oy : py;
        const uint32_t l = sm[(long)ty * W + tx];
        const double key = cand(*reinterpret_cast<const float2 *>(sb + l), l, fx, fy);
        best = (k == 0) ? key : __builtin_fmin(best, key);
    }
    first[P.pix_off + (long)py * W + px] = (uint32_t)__double2loint(best);
}

void launch_jfa_batch_seeds(hipStream_t st, const BatchPatch *patches, const int *site_patch,
                            int total_sites, const uint32_t *seed, uint32_t *first,
                            const float2 *sites, int pass, int pmax)
{
    if (total_sites > 0)
        hipLaunchKernelGGL(k_jfa_batch_seeds, dim3((9 * total_sites + 255) / 256), dim3(256), 0,
                           st, patches, site_patch, total_sites, seed, first, sites, pass, pmax);
}

void launch_jfa_seed_pass(hipStream_t st, const uint32_t *seed, uint32_t *out,
                          const float2 *sites, int S, int W, int H, int step, const int *stop,
                          const double *blockpart, int *dstate, const FinalArgs *fa)
{
    const int fin = fa ? 1 : 0;
    const int nb = (9 * S + 255) / 256 + fin;
    if (nb > 0)
        hipLaunchKernelGGL(k_jfa_seeds, dim3(nb), dim3(256), 0, st, seed, out, sites, S, W, H, step,
                           stop, fin, blockpart, dstate, fa ? *fa : FinalArgs());
}

}  // namespace cvt
