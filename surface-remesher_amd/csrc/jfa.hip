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
// Candidate order: (d, index) lexicographic with d = fmaf(dx, dx, dy*dy) >= 0.
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

#include <algorithm>

namespace cvt {

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
                                                  uint32_t *__restrict__ clear, uint32_t none)
{
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
            const float d = __builtin_fmaf(dx, dx, dy * dy);
            const double key = jfa_key(d, l[t]);
            best = (t == 0) ? key : __builtin_fmin(best, key);
        }
        if (x < W) {
            __builtin_amdgcn_raw_buffer_store_b32((uint32_t)__double2loint(best), rout, o1, s1, 0);
            if (clear)
                __builtin_amdgcn_raw_buffer_store_b32(none, buf_rsrc(clear, map_bytes), o1, s1, 0);
        }
    }
}

// ---------------------------------------------------------------------------
// Quad pass: each lane owns 4 consecutive pixels (one dwordx4 per tap row and
// column), workgroups are persistent and own a band of rows.  With
// SITES_IN_LDS the whole site table is first copied to LDS (<= 160 KB) and
// every candidate lookup is an LDS read at the label's byte offset; otherwise
// the lookups gather from the global table.
// ---------------------------------------------------------------------------
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t q_el(const u32x4 &q, int j)
{
    return j == 0 ? q.x : j == 1 ? q.y : j == 2 ? q.z : q.w;
}

// labels of the left (or right) tap of pixels 0..3 of a quad at step < 4:
// elements of the neighbouring quad nb and the centre quad c, i.e. the labels
// at x + j - s (left) or x + j + s (right); an out-of-grid neighbour quad
// yields the pixel's own centre label (a duplicate candidate)
__device__ __forceinline__ uint32_t small_left(const u32x4 &nb, const u32x4 &c, int j, int s,
                                               bool nb_ok)
{
    const int k = j - s;   // -3 .. 2
    if (k >= 0) return q_el(c, k);
    return nb_ok ? q_el(nb, 4 + k) : q_el(c, j);
}
__device__ __forceinline__ uint32_t small_right(const u32x4 &nb, const u32x4 &c, int j, int s,
                                                bool nb_ok)
{
    const int k = j + s;   // 1 .. 6
    if (k < 4) return q_el(c, k);
    return nb_ok ? q_el(nb, k - 4) : q_el(c, j);
}

// Rows [y0, y1) of one W x H map, all waves of the block.  tab is the site
// table (LDS or global) and tab_base the byte offset its first entry has in
// label space (labels are byte offsets into the full table).
__device__ __forceinline__ void jfa_quad_rows(const uint32_t *__restrict__ in,
                                              uint32_t *__restrict__ out,
                                              uint32_t *__restrict__ clear, const char *tab,
                                              uint32_t tab_base, int W, int H, int y0, int y1,
                                              int step, uint32_t none)
{
    tab -= tab_base;
    const int nseg = (W + 255) >> 8;              // 256-pixel segments per row
    const int nitems = (y1 - y0) * nseg;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nwaves = blockDim.x >> 6;
    const int lane = threadIdx.x & 63;
    const uint32_t map_bytes = 4u * (uint32_t)W * (uint32_t)H;
    const __amdgpu_buffer_rsrc_t rin = buf_rsrc(in, map_bytes);
    const __amdgpu_buffer_rsrc_t rout = buf_rsrc(out, map_bytes);
    const uint32_t row_bytes = 4u * (uint32_t)W;
    const bool small = step < 4;
    const int cstep = small ? 4 : step;           // column offset of the side quads

    for (int it = wave; it < nitems; it += nwaves) {
        const int row = it / nseg;
        const int y = __builtin_amdgcn_readfirstlane(y0 + row);
        const int seg = __builtin_amdgcn_readfirstlane(it - row * nseg);
        const int xb = seg * 256 + lane * 4;
        const bool live = xb < W;
        const int xc = min(xb, W - 4);
        const uint32_t o1 = 4u * (uint32_t)xc;
        const bool lok = xc - cstep >= 0, rok = xc + 3 + cstep < W;
        const uint32_t o0 = lok ? o1 - 4u * (uint32_t)cstep : o1;
        const uint32_t o2 = rok ? o1 + 4u * (uint32_t)cstep : o1;
        const int r0 = (y - step >= 0) ? y - step : y;
        const int r2 = (y + step < H) ? y + step : y;
        const uint32_t s0 = __builtin_amdgcn_readfirstlane((uint32_t)r0 * row_bytes);
        const uint32_t s1 = __builtin_amdgcn_readfirstlane((uint32_t)y * row_bytes);
        const uint32_t s2 = __builtin_amdgcn_readfirstlane((uint32_t)r2 * row_bytes);
        u32x4 q[9];
        q[0] = __builtin_amdgcn_raw_buffer_load_b128(rin, o0, s0, 0);
        q[1] = __builtin_amdgcn_raw_buffer_load_b128(rin, o1, s0, 0);
        q[2] = __builtin_amdgcn_raw_buffer_load_b128(rin, o2, s0, 0);
        q[3] = __builtin_amdgcn_raw_buffer_load_b128(rin, o0, s1, 0);
        q[4] = __builtin_amdgcn_raw_buffer_load_b128(rin, o1, s1, 0);
        q[5] = __builtin_amdgcn_raw_buffer_load_b128(rin, o2, s1, 0);
        q[6] = __builtin_amdgcn_raw_buffer_load_b128(rin, o0, s2, 0);
        q[7] = __builtin_amdgcn_raw_buffer_load_b128(rin, o1, s2, 0);
        q[8] = __builtin_amdgcn_raw_buffer_load_b128(rin, o2, s2, 0);
        const float py = (float)y + 0.5f;
        uint32_t res[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float px = (float)(xc + j) + 0.5f;
            uint32_t l[9];
#pragma unroll
            for (int rr = 0; rr < 3; ++rr) {
                const u32x4 &L = q[3 * rr], &Cq = q[3 * rr + 1], &R = q[3 * rr + 2];
                l[3 * rr + 1] = q_el(Cq, j);
                if (small) {
                    l[3 * rr] = small_left(L, Cq, j, step, lok);
                    l[3 * rr + 2] = small_right(R, Cq, j, step, rok);
                } else {
                    l[3 * rr] = q_el(L, j);
                    l[3 * rr + 2] = q_el(R, j);
                }
            }
            double best;
#pragma unroll
            for (int t = 0; t < 9; ++t) {
                const float2 sv = *reinterpret_cast<const float2 *>(tab + l[t]);
                const float dx = sv.x - px;
                const float dy = sv.y - py;
                const float d = __builtin_fmaf(dx, dx, dy * dy);
                const double key = jfa_key(d, l[t]);
                best = (t == 0) ? key : __builtin_fmin(best, key);
            }
            res[j] = (uint32_t)__double2loint(best);
        }
        if (live) {
            const u32x4 r4 = {res[0], res[1], res[2], res[3]};
            __builtin_amdgcn_raw_buffer_store_b128(r4, rout, o1, s1, 0);
            if (clear) {
                const u32x4 n4 = {none, none, none, none};
                __builtin_amdgcn_raw_buffer_store_b128(n4, buf_rsrc(clear, map_bytes), o1, s1, 0);
            }
        }
    }
}

template <bool SITES_IN_LDS>
__global__ __launch_bounds__(1024) void k_jfa_quad(const uint32_t *__restrict__ in,
                                                   uint32_t *__restrict__ out,
                                                   const float2 *__restrict__ sites,
                                                   int table_f4, int W, int H, int step,
                                                   uint32_t *__restrict__ clear, uint32_t none)
{
    extern __shared__ float4 lds_sites[];
    if (SITES_IN_LDS) {
        const float4 *g4 = reinterpret_cast<const float4 *>(sites);
        for (int i = threadIdx.x; i < table_f4; i += blockDim.x) lds_sites[i] = g4[i];
        __syncthreads();
    }
    const char *tab = SITES_IN_LDS ? reinterpret_cast<const char *>(lds_sites)
                                   : reinterpret_cast<const char *>(sites);
    // band of rows of this workgroup; bands are dealt so that each XCD (blocks
    // go to the 8 XCDs round robin) owns a contiguous stretch of rows
    const int G = gridDim.x, g = blockIdx.x;
    const int q8 = G >> 3, r8 = G & 7, xcd = g & 7;
    const int band = xcd * q8 + min(xcd, r8) + (g >> 3);
    const int y0 = (int)((long)H * band / G), y1 = (int)((long)H * (band + 1) / G);
    jfa_quad_rows(in, out, clear, tab, 0u, W, H, y0, y1, step, none);
}

// ---------------------------------------------------------------------------
// batched pass: one launch over many patches.  Each workgroup takes one work
// unit (patch, band of rows); patches run their passes right-aligned, so pass
// k of the batch is pass k - (Pmax - P) of a patch with P passes (skipped
// when negative) and every patch ends its JFA in the same launch.
// ---------------------------------------------------------------------------
template <bool SITES_IN_LDS>
__global__ __launch_bounds__(1024) void k_jfa_batch(const BatchPatch *__restrict__ patches,
                                                    const BatchUnit *__restrict__ units,
                                                    uint32_t *__restrict__ seed,
                                                    uint32_t *__restrict__ lab0,
                                                    uint32_t *__restrict__ lab1,
                                                    const float2 *__restrict__ sites, int pass,
                                                    int pmax)
{
    extern __shared__ float4 lds_sites[];
    const BatchUnit u = units[blockIdx.x];
    const BatchPatch P = patches[u.patch];
    const int local = pass - (pmax - P.passes);
    if (local < 0) return;                                   // uniform per block
    const int step = (local < P.passes - 1) ? (P.first_step >> local) : 1;
    const char *tab;
    uint32_t base;
    if (SITES_IN_LDS) {
        const float4 *g4 = reinterpret_cast<const float4 *>(sites + P.site_off);
        const int nf4 = (P.cap + 2) / 2;
        for (int i = threadIdx.x; i < nf4; i += blockDim.x) lds_sites[i] = g4[i];
        __syncthreads();
        tab = reinterpret_cast<const char *>(lds_sites);
        base = 8u * (uint32_t)P.site_off;
    } else {
        tab = reinterpret_cast<const char *>(sites);
        base = 0u;
    }
    uint32_t *bufs[2] = {lab0 + P.pix_off, lab1 + P.pix_off};
    const uint32_t *in = (local == 0) ? seed + P.pix_off : bufs[(local - 1) & 1];
    uint32_t *out = bufs[local & 1];
    uint32_t *clear = (local == 1) ? seed + P.pix_off : nullptr;
    const uint32_t none = 8u * (uint32_t)(P.site_off + P.cap);
    jfa_quad_rows(in, out, clear, tab, base, P.W, P.H, u.y0, u.y1, step, none);
}

void launch_jfa_batch(hipStream_t st, const BatchPatch *patches, const BatchUnit *units, int nunits,
                      uint32_t *seed, uint32_t *lab0, uint32_t *lab1, const float2 *sites, int pass,
                      int pmax, size_t lds_bytes)
{
    if (lds_bytes) {
        static bool attr = false;
        if (!attr) {
            (void)hipFuncSetAttribute((const void *)k_jfa_batch<true>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            attr = true;
        }
        hipLaunchKernelGGL(k_jfa_batch<true>, dim3(nunits), dim3(1024), lds_bytes, st, patches,
                           units, seed, lab0, lab1, sites, pass, pmax);
    } else {
        hipLaunchKernelGGL(k_jfa_batch<false>, dim3(nunits), dim3(1024), 0, st, patches, units,
                           seed, lab0, lab1, sites, pass, pmax);
    }
}

#ifndef CVT_JFA_ROWS
#define CVT_JFA_ROWS 4
#endif

static int g_num_cu = 0;
static size_t g_lds_max = 0;

JfaPlan plan_jfa(int W, int H, int capS)
{
    if (!g_num_cu) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
            g_num_cu = prop.multiProcessorCount;
            g_lds_max = prop.sharedMemPerBlock;
            if (prop.maxSharedMemoryPerMultiProcessor > g_lds_max)
                g_lds_max = prop.maxSharedMemoryPerMultiProcessor;
        } else {
            g_num_cu = 256;
            g_lds_max = 65536;
        }
        (void)hipFuncSetAttribute((const void *)k_jfa_quad<true>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)g_lds_max);
    }
    JfaPlan p;
    p.quad = (W % 4 == 0) && W >= 8;
    const size_t table = (size_t)(capS + 2) / 2 * 16;   // sites incl. sentinel, float4 units
    p.table_f4 = (int)(table / 16);
    p.lds = p.quad && table <= g_lds_max;
    p.lds_bytes = p.lds ? table : 0;
    const long items = (long)H * ((W + 255) / 256);
    if (p.lds && table > 40 * 1024) {
        p.threads = 1024;                               // one workgroup per CU
        p.grid = g_num_cu;
    } else {
        p.threads = 256;
        int per_cu = p.lds ? (int)std::min<size_t>(8, g_lds_max / std::max<size_t>(table, 1)) : 8;
        p.grid = g_num_cu * std::max(per_cu, 1);
    }
    // at least one item per wave
    const long waves = (long)p.grid * (p.threads / 64);
    if (waves > items) p.grid = (int)std::max<long>(1, items / (p.threads / 64));
    if (p.grid > H) p.grid = H;
    return p;
}

void launch_jfa_pass(hipStream_t st, const JfaPlan &p, const uint32_t *in, uint32_t *out,
                     const float2 *sites, int W, int H, int step, uint32_t *clear, uint32_t none)
{
    if (p.quad) {
        if (p.lds)
            hipLaunchKernelGGL(k_jfa_quad<true>, dim3(p.grid), dim3(p.threads), p.lds_bytes, st, in,
                               out, sites, p.table_f4, W, H, step, clear, none);
        else
            hipLaunchKernelGGL(k_jfa_quad<false>, dim3(p.grid), dim3(p.threads), 0, st, in, out,
                               sites, p.table_f4, W, H, step, clear, none);
        return;
    }
    constexpr int R = CVT_JFA_ROWS;
    const int tilesX = (W + 63) / 64, tilesY = (H + 4 * R - 1) / (4 * R);
    const int ntiles = tilesX * tilesY;
    hipLaunchKernelGGL(k_jfa_pass<R>, dim3(ntiles), dim3(256), 0, st, in, out, sites, W, H, tilesX,
                       ntiles, step, clear, none);
}

}  // namespace cvt
