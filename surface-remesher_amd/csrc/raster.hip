// raster.hip -- density rasterizer (replaces the reference's discretization_d,
// source/discretization.cu:57-120).
//
// Rule (the reference's, source/discretization.cu:57-85): pixel (tx, ty)
// samples the point (tx * scale, ty * scale) of the shifted UV frame; the first
// triangle in index order whose three float64 barycentric weights are all >= 0
// interpolates the vertex weights, rounded once to float; no hit gives 0.
// Degenerate triangles (zero denominator) never hit.
//
// Instead of testing every triangle at every pixel (O(n^2 F)), triangles are
// binned into 16 x 16 pixel tiles by a conservative bounding box (a counting
// sort: count, prefix sum, fill; linear in the number of (triangle, tile)
// pairs).  A pixel tests its tile's triangles and keeps the lowest-index hit,
// which is the first hit in index order: every triangle holding the sample
// point lists it in that tile.  A triangle so thin that rounding could
// accept points well outside its box (sin^2 of its angle at vertex 1 below
// 1e-12) goes to an "always" list that every pixel tests.  The barycentric
// arithmetic is the oracle's, operation for operation (contraction is off
// for this file).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "../../include/cvt.h"
#include "cvt_engine.h"

namespace cvt {

#define RT 16   // tile side in pixels

struct TriBox {
    int x0, x1, y0, y1;   // inclusive pixel-index range the triangle may touch; x0 > x1 = none
};

__device__ __forceinline__ void bary(double x1, double y1, double x2, double y2, double x3,
                                     double y3, double x0, double y0, double &b1, double &b2,
                                     double &b3)
{
    const double v0x = x2 - x1, v0y = y2 - y1;
    const double v1x = x3 - x1, v1y = y3 - y1;
    const double v2x = x0 - x1, v2y = y0 - y1;
    const double d00 = v0x * v0x + v0y * v0y;
    const double d01 = v0x * v1x + v0y * v1y;
    const double d11 = v1x * v1x + v1y * v1y;
    const double d20 = v2x * v0x + v2y * v0y;
    const double d21 = v2x * v1x + v2y * v1y;
    const double den = d00 * d11 - d01 * d01;
    if (den == 0) { b1 = b2 = b3 = -1; return; }
    const double u = (d11 * d20 - d01 * d21) / den;
    const double v = (d00 * d21 - d01 * d20) / den;
    b2 = u;
    b3 = v;
    b1 = 1.0 - u - v;
}

// per triangle: index validation, the conservative pixel box and the
// ill-conditioned flag
__global__ void k_tri_boxes(const double *__restrict__ pts, int nv, const int *__restrict__ tri,
                            int nt, double scale, int n, TriBox *__restrict__ box,
                            uint8_t *__restrict__ always, int *__restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const int a = tri[3 * i], b = tri[3 * i + 1], c = tri[3 * i + 2];
    TriBox bx = {1, 0, 1, 0};
    always[i] = 0;
    if ((unsigned)a >= (unsigned)nv || (unsigned)b >= (unsigned)nv || (unsigned)c >= (unsigned)nv) {
        *bad = 1;
        box[i] = bx;
        return;
    }
    const double x1 = pts[2 * a], y1 = pts[2 * a + 1];
    const double x2 = pts[2 * b], y2 = pts[2 * b + 1];
    const double x3 = pts[2 * c], y3 = pts[2 * c + 1];
    const double v0x = x2 - x1, v0y = y2 - y1, v1x = x3 - x1, v1y = y3 - y1;
    const double d00 = v0x * v0x + v0y * v0y, d11 = v1x * v1x + v1y * v1y;
    const double d01 = v0x * v1x + v0y * v1y;
    const double den = d00 * d11 - d01 * d01;
    if (den == 0) { box[i] = bx; return; }          // never hits
    if (!(den > 1e-12 * d00 * d11)) {                // thin or non-finite: test everywhere
        always[i] = 1;
        box[i] = bx;
        return;
    }
    const double xmin = fmin(x1, fmin(x2, x3)), xmax = fmax(x1, fmax(x2, x3));
    const double ymin = fmin(y1, fmin(y2, y3)), ymax = fmax(y1, fmax(y2, y3));
    const double eps = 1e-9 * fmax(xmax - xmin, ymax - ymin) + 1e-300;
    const double lo_x = (xmin - eps) / scale, hi_x = (xmax + eps) / scale;
    const double lo_y = (ymin - eps) / scale, hi_y = (ymax + eps) / scale;
    bx.x0 = (int)fmax(0.0, floor(lo_x));
    bx.x1 = (int)fmin((double)(n - 1), ceil(hi_x));
    bx.y0 = (int)fmax(0.0, floor(lo_y));
    bx.y1 = (int)fmin((double)(n - 1), ceil(hi_y));
    if (!(hi_x >= 0.0) || !(hi_y >= 0.0) || !(lo_x <= n - 1.0) || !(lo_y <= n - 1.0))
        bx = TriBox{1, 0, 1, 0};
    box[i] = bx;
}

// (triangle, tile) pairs of a pixel box, as tile ranges
__device__ __forceinline__ void tile_range(const TriBox &b, int &tx0, int &tx1, int &ty0, int &ty1)
{
    tx0 = b.x0 / RT; tx1 = b.x1 / RT;
    ty0 = b.y0 / RT; ty1 = b.y1 / RT;
}

// (triangle, tile) pairs: count per tile
__global__ void k_tile_count(const TriBox *__restrict__ box, int nt, int tiles_x,
                             int *__restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const TriBox b = box[i];
    if (b.x0 > b.x1 || b.y0 > b.y1) return;
    int tx0, tx1, ty0, ty1;
    tile_range(b, tx0, tx1, ty0, ty1);
    for (int y = ty0; y <= ty1; ++y)
        for (int x = tx0; x <= tx1; ++x) atomicAdd(&count[y * tiles_x + x], 1);
}

// (triangle, tile) pairs: each triangle writes its index into its tiles' lists.
// With a list of `cap` entries (cap >= 0) and more pairs than that, nothing
// is written and *overflow is raised.
__global__ void k_tile_fill(const TriBox *__restrict__ box, int nt, int tiles_x, int ntiles,
                            const int *__restrict__ start, int *__restrict__ cursor,
                            int *__restrict__ list, int cap, int *__restrict__ overflow)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    if (cap >= 0 && start[ntiles] > cap) {
        if (i == 0) *overflow = 1;
        return;
    }
    const TriBox b = box[i];
    if (b.x0 > b.x1 || b.y0 > b.y1) return;
    int tx0, tx1, ty0, ty1;
    tile_range(b, tx0, tx1, ty0, ty1);
    for (int y = ty0; y <= ty1; ++y)
        for (int x = tx0; x <= tx1; ++x) {
            const int tile = y * tiles_x + x;
            list[start[tile] + atomicAdd(&cursor[tile], 1)] = i;
        }
}

// one workgroup per tile, one thread per pixel: the lowest-index triangle of
// the tile list and of the "always" list holding the sample point
__global__ __launch_bounds__(256) void k_raster(const double *__restrict__ pts,
                                                const double *__restrict__ wt,
                                                const int *__restrict__ tri,
                                                const int *__restrict__ list,
                                                const int *__restrict__ start,
                                                const int *__restrict__ alw, int nalw, int tiles_x,
                                                double scale, int n, float *__restrict__ density,
                                                const int *__restrict__ dev_nalw,
                                                const int *__restrict__ overflow)
{
    if (overflow && *overflow) return;
    if (dev_nalw) nalw = *dev_nalw;
    const int tile = blockIdx.x;
    const int tx = (tile % tiles_x) * RT + (threadIdx.x & (RT - 1));
    const int ty = (tile / tiles_x) * RT + (threadIdx.x / RT);
    if (tx >= n || ty >= n) return;
    const int a = start[tile], len = start[tile + 1] - a;
    const double x = tx * scale, y = ty * scale;
    float res = 0.0f;
    int hit = -1;
    for (int j = 0; j < len + nalw; ++j) {
        const int i = (j < len) ? list[a + j] : alw[j - len];
        if (hit >= 0 && i > hit) continue;          // a lower-index hit is known
        const int p1 = tri[3 * i], p2 = tri[3 * i + 1], p3 = tri[3 * i + 2];
        double b1, b2, b3;
        bary(pts[2 * p1], pts[2 * p1 + 1], pts[2 * p2], pts[2 * p2 + 1], pts[2 * p3],
             pts[2 * p3 + 1], x, y, b1, b2, b3);
        if (b1 < 0 || b2 < 0 || b3 < 0) continue;
        res = (float)(b1 * wt[p1] + b2 * wt[p2] + b3 * wt[p3]);
        hit = i;
    }
    density[(long)ty * n + tx] = res;
}

}  // namespace cvt

using namespace cvt;

#define RCHK(call, what)                                                               \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) { rc = fail(CVT_ERR_HIP, "%s: %s", what, hipGetErrorString(e_)); goto done; } \
    } while (0)

extern "C" int cvt_rasterize(const double *points, const double *weight, int num_point,
                             const int *triangle, int num_tri, float *density, double scale, int n)
{
    clear_error();
    if (n < 1 || n > 16384) return fail(CVT_ERR_ARG, "n = %d outside [1, 16384]", n);
    if (num_point < 0 || num_tri < 0) return fail(CVT_ERR_ARG, "negative sizes");
    if (!density || (num_point > 0 && (!points || !weight)) || (num_tri > 0 && !triangle))
        return fail(CVT_ERR_ARG, "null array");
    if (!(scale > 0.0) || !isfinite(scale)) return fail(CVT_ERR_ARG, "scale %g", scale);
    const int tiles_x = (n + RT - 1) / RT, ntiles = tiles_x * tiles_x;
    int rc = CVT_OK;
    hipStream_t st = nullptr;   // the legacy default stream, as the reference's calls use
    double *d_pts = nullptr, *d_wt = nullptr;
    int *d_tri = nullptr, *d_list = nullptr, *d_count = nullptr, *d_start = nullptr,
        *d_cursor = nullptr, *d_alw = nullptr, *d_bad = nullptr;
    TriBox *d_box = nullptr;
    uint8_t *d_always = nullptr;
    float *d_den = nullptr;
    int bad = 0, nalw = 0, total = 0;
    std::vector<uint8_t> always;
    std::vector<int> alw;
    const long npx = (long)n * n;
    RCHK(hipMalloc(&d_pts, sizeof(double) * 2 * std::max(num_point, 1)), "alloc");
    RCHK(hipMalloc(&d_wt, sizeof(double) * std::max(num_point, 1)), "alloc");
    RCHK(hipMalloc(&d_tri, sizeof(int) * 3 * std::max(num_tri, 1)), "alloc");
    RCHK(hipMalloc(&d_box, sizeof(TriBox) * std::max(num_tri, 1)), "alloc");
    RCHK(hipMalloc(&d_always, std::max(num_tri, 1)), "alloc");
    RCHK(hipMalloc(&d_alw, sizeof(int) * std::max(num_tri, 1)), "alloc");
    RCHK(hipMalloc(&d_count, sizeof(int) * ntiles), "alloc");
    RCHK(hipMalloc(&d_cursor, sizeof(int) * ntiles), "alloc");
    RCHK(hipMalloc(&d_start, sizeof(int) * ((size_t)ntiles + 1)), "alloc");
    RCHK(hipMalloc(&d_bad, sizeof(int)), "alloc");
    RCHK(hipMalloc(&d_den, sizeof(float) * npx), "alloc");
    if (num_point > 0) {
        RCHK(hipMemcpyAsync(d_pts, points, sizeof(double) * 2 * num_point, hipMemcpyDefault, st), "upload");
        RCHK(hipMemcpyAsync(d_wt, weight, sizeof(double) * num_point, hipMemcpyDefault, st), "upload");
    }
    if (num_tri > 0)
        RCHK(hipMemcpyAsync(d_tri, triangle, sizeof(int) * 3 * num_tri, hipMemcpyDefault, st), "upload");
    RCHK(hipMemsetAsync(d_bad, 0, sizeof(int), st), "memset");
    RCHK(hipMemsetAsync(d_count, 0, sizeof(int) * ntiles, st), "memset");
    RCHK(hipMemsetAsync(d_cursor, 0, sizeof(int) * ntiles, st), "memset");
    if (num_tri > 0) {
        hipLaunchKernelGGL(k_tri_boxes, dim3((num_tri + 255) / 256), dim3(256), 0, st, d_pts,
                           num_point, d_tri, num_tri, scale, n, d_box, d_always, d_bad);
        hipLaunchKernelGGL(k_tile_count, dim3((num_tri + 255) / 256), dim3(256), 0, st, d_box,
                           num_tri, tiles_x, d_count);
    }
    launch_exclusive_scan(st, d_count, ntiles, d_start);
    RCHK(hipMemcpyAsync(&total, d_start + ntiles, sizeof(int), hipMemcpyDeviceToHost, st), "total");
    RCHK(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, st), "flag");
    always.resize(std::max(num_tri, 1));
    if (num_tri > 0)
        RCHK(hipMemcpyAsync(always.data(), d_always, num_tri, hipMemcpyDeviceToHost, st), "flags");
    RCHK(hipStreamSynchronize(st), "boxes");
    if (bad) { rc = fail(CVT_ERR_ARG, "a triangle indexes a vertex outside [0, %d)", num_point); goto done; }
    for (int i = 0; i < num_tri; ++i)
        if (always[i]) alw.push_back(i);
    nalw = (int)alw.size();
    if (nalw) RCHK(hipMemcpyAsync(d_alw, alw.data(), sizeof(int) * nalw, hipMemcpyHostToDevice, st), "upload");
    RCHK(hipMalloc(&d_list, sizeof(int) * (size_t)std::max(total, 1)), "alloc");
    if (num_tri > 0)
        hipLaunchKernelGGL(k_tile_fill, dim3((num_tri + 255) / 256), dim3(256), 0, st, d_box,
                           num_tri, tiles_x, ntiles, d_start, d_cursor, d_list, -1, nullptr);
    hipLaunchKernelGGL(k_raster, dim3(ntiles), dim3(256), 0, st, d_pts, d_wt, d_tri, d_list,
                       d_start, d_alw, nalw, tiles_x, scale, n, d_den, nullptr, nullptr);
    RCHK(hipGetLastError(), "raster launch");
    RCHK(hipMemcpyAsync(density, d_den, sizeof(float) * npx, hipMemcpyDefault, st), "download");
    RCHK(hipStreamSynchronize(st), "raster");
done:
    void *bufs[] = {d_pts, d_wt, d_tri, d_box, d_always, d_alw, d_list, d_count, d_start,
                    d_cursor, d_bad, d_den};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    return rc;
}

// ---------------------------------------------------------------------------
// cvt_rasterize_noalloc: cvt_rasterize for callers that cannot allocate.  All
// of its working memory (triangle boxes, the thin-triangle list, per-tile
// counts and list starts, the (triangle, tile) lists and three status words)
// lives in one static device array, so the call allocates nothing.  Each
// device holds its own copy of the array; its address is looked up once per
// device with hipGetSymbolAddress.  Every call on a device shares that
// device's array: calls that overlap on one device race.
// ---------------------------------------------------------------------------
namespace cvt {

#define NOALLOC_MAX_N 4096
#define NOALLOC_MAX_TRI (1 << 18)
#define NOALLOC_MAX_PAIRS (1 << 21)
#define NOALLOC_TILES ((NOALLOC_MAX_N / RT) * (NOALLOC_MAX_N / RT))

// byte offsets of the parts (each a multiple of 256)
enum : size_t {
    NA_BOX = 0,
    NA_ALWAYS = NA_BOX + sizeof(TriBox) * (size_t)NOALLOC_MAX_TRI,
    NA_ALW = NA_ALWAYS + (size_t)NOALLOC_MAX_TRI,
    NA_COUNT = NA_ALW + sizeof(int) * (size_t)NOALLOC_MAX_TRI,
    NA_CURSOR = NA_COUNT + sizeof(int) * (size_t)NOALLOC_TILES,
    NA_START = NA_CURSOR + sizeof(int) * (size_t)NOALLOC_TILES,
    NA_LIST = NA_START + sizeof(int) * ((size_t)NOALLOC_TILES + 64),
    NA_STATUS = NA_LIST + sizeof(int) * (size_t)NOALLOC_MAX_PAIRS,   // bad, overflow, nalw
    NA_BYTES = NA_STATUS + 256,
};
static_assert(NA_ALWAYS % 256 == 0 && NA_ALW % 256 == 0 && NA_COUNT % 256 == 0 &&
                  NA_CURSOR % 256 == 0 && NA_START % 256 == 0 && NA_LIST % 256 == 0 &&
                  NA_STATUS % 256 == 0,
              "scratch parts must stay 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

#define NOALLOC_MAX_DEVICES 64
static unsigned char *g_scratch_ptr[NOALLOC_MAX_DEVICES];

// the device address of this device's scratch (looked up on its first use)
static int scratch_of(int dev, unsigned char **out)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(CVT_ERR_ARG, "device %d beyond the %d the no-alloc path supports", dev,
                    NOALLOC_MAX_DEVICES);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(CVT_ERR_HIP, "scratch address: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[dev];
    return CVT_OK;
}

// the thin triangles, appended in any order (a pixel keeps the lowest-index
// hit whatever the order of its candidates)
static __global__ void k_always_append(const uint8_t *__restrict__ always, int nt,
                                       int *__restrict__ alw, int *__restrict__ nalw)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nt && always[i]) alw[atomicAdd(nalw, 1)] = i;
}

static __global__ void k_zero_i32(int *__restrict__ p, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        p[i] = 0;
}

// true when p is memory the device can read and write (device, pinned or
// managed memory); pageable host memory is not
static bool device_accessible(const void *p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.devicePointer != nullptr;
}

}  // namespace cvt

extern "C" int cvt_rasterize_noalloc(const double *points, const double *weight, int num_point,
                                     const int *triangle, int num_tri, float *density,
                                     double scale, int n)
{
    clear_error();
    if (n < 1 || n > NOALLOC_MAX_N)
        return fail(CVT_ERR_ARG, "n = %d outside [1, %d] (the no-alloc limit)", n, NOALLOC_MAX_N);
    if (num_point < 0 || num_tri < 0) return fail(CVT_ERR_ARG, "negative sizes");
    if (num_tri > NOALLOC_MAX_TRI)
        return fail(CVT_ERR_ARG, "%d triangles above the no-alloc limit %d", num_tri,
                    NOALLOC_MAX_TRI);
    if (!density || (num_point > 0 && (!points || !weight)) || (num_tri > 0 && !triangle))
        return fail(CVT_ERR_ARG, "null array");
    if (!(scale > 0.0) || !isfinite(scale)) return fail(CVT_ERR_ARG, "scale %g", scale);
    if (!device_accessible(density) ||
        (num_point > 0 && (!device_accessible(points) || !device_accessible(weight))) ||
        (num_tri > 0 && !device_accessible(triangle)))
        return fail(CVT_ERR_ARG, "the no-alloc rasterizer needs device-accessible buffers");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(CVT_ERR_HIP, "no current device");
    unsigned char *base = nullptr;
    int rc = scratch_of(dev, &base);
    if (rc) return rc;
    TriBox *box = reinterpret_cast<TriBox *>(base + NA_BOX);
    uint8_t *always = base + NA_ALWAYS;
    int *alw = reinterpret_cast<int *>(base + NA_ALW);
    int *count = reinterpret_cast<int *>(base + NA_COUNT);
    int *cursor = reinterpret_cast<int *>(base + NA_CURSOR);
    int *start = reinterpret_cast<int *>(base + NA_START);
    int *list = reinterpret_cast<int *>(base + NA_LIST);
    int *status = reinterpret_cast<int *>(base + NA_STATUS);   // bad, overflow, nalw
    const int tiles_x = (n + RT - 1) / RT, ntiles = tiles_x * tiles_x;
    hipStream_t st = nullptr;   // the legacy default stream, as cvt_rasterize
    // the counts, cursors and status words left by the last call start at zero
    hipLaunchKernelGGL(k_zero_i32, dim3(256), dim3(256), 0, st, count, (long)ntiles);
    hipLaunchKernelGGL(k_zero_i32, dim3(256), dim3(256), 0, st, cursor, (long)ntiles);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, status, 4L);
    if (num_tri > 0) {
        const dim3 g((num_tri + 255) / 256), b(256);
        hipLaunchKernelGGL(k_tri_boxes, g, b, 0, st, points, num_point, triangle, num_tri, scale,
                           n, box, always, status);
        hipLaunchKernelGGL(k_always_append, g, b, 0, st, always, num_tri, alw, status + 2);
        hipLaunchKernelGGL(k_tile_count, g, b, 0, st, box, num_tri, tiles_x, count);
    }
    launch_exclusive_scan(st, count, ntiles, start);
    if (num_tri > 0)
        hipLaunchKernelGGL(k_tile_fill, dim3((num_tri + 255) / 256), dim3(256), 0, st, box,
                           num_tri, tiles_x, ntiles, start, cursor, list, NOALLOC_MAX_PAIRS,
                           status + 1);
    hipLaunchKernelGGL(k_raster, dim3(ntiles), dim3(256), 0, st, points, weight, triangle, list,
                       start, alw, 0, tiles_x, scale, n, density, status + 2, status + 1);
    int h[3] = {0, 0, 0};
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h, status, sizeof h, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(CVT_ERR_HIP, "no-alloc raster: %s", hipGetErrorString(e));
    if (h[0]) return fail(CVT_ERR_ARG, "a triangle indexes a vertex outside [0, %d)", num_point);
    if (h[1])
        return fail(CVT_ERR_CAPACITY, "more than %d (triangle, tile) pairs (the no-alloc limit)",
                    NOALLOC_MAX_PAIRS);
    return CVT_OK;
}

