// locate.hip -- point location in a 2D triangle mesh (replaces the brute-force
// face scan of the reference's recover(), source/recover.h:63-81, run for
// every site at :109-125 and for every output triangle's centroid at
// :134-141).
//
// Rule (the reference's locate(), source/recover.h:63-81): a point lies in the
// first face, in face order, whose three float64 barycentric weights (the
// formula of source/recover.h:30-53) are all >= 0; a face with a zero
// denominator never holds a point.  The weights are returned in the face's
// vertex order: vertex 0 gets 1 - w1 - w2, vertices 1 and 2 get w1 and w2
// (source/recover.h:79, the argument order of barycentric()).  The 3D point
// of a located site is sum_j w_j * P3[v_j], summed in that order from 0.0
// (source/recover.h:117-123).  A non-finite query point is in no face (the
// reference would accept it in the first non-degenerate face: NaN weights
// fail no "< 0" test).
//
// Instead of testing every face for every point (O(Q F)), faces are binned
// into a G x G grid over the vertices' bounding box by a conservative box (a
// counting sort: count, prefix sum, fill; linear in the number of (face, bin)
// pairs).  A point tests the faces of its bin and keeps the lowest-index hit,
// which is the first hit in face order: every face holding the point lists
// it in that bin.  Faces so thin that rounding could accept points well
// outside their box go to an "always" list that every point tests.
// Contraction is off for this file: every expression rounds as written, as
// in the oracle.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "../../include/cvt.h"
#include "cvt_engine.h"

namespace cvt {

struct FaceBox {
    int x0, x1, y0, y1;   // inclusive bin range; x0 > x1 = in no bin
};

// the bin grid: G x G bins over [ox, ox + G / inv) x [oy, oy + G / inv)
struct BinGrid {
    double ox, oy, inv;
    int G;
};

// bin of a coordinate: monotone in v (so a point inside a box maps inside the
// box's bin range) and clamped to the grid, NaN to bin 0
__device__ __forceinline__ int bin_of(double v, double o, double inv, int G)
{
    double t = (v - o) * inv;
    t = fmax(0.0, fmin(t, (double)(G - 1)));
    return (int)floor(t);
}

// source/recover.h:30-53: w1, w2 from the dot products of v0 = p2 - p1,
// v1 = p3 - p1, v2 = q - p1; a zero denominator gives -1 weights
__device__ __forceinline__ void rec_bary(double x1, double y1, double x2, double y2, double x3,
                                         double y3, double x0, double y0, double &w0,
                                         double &w1, double &w2)
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
    if (den == 0) { w0 = w1 = w2 = -1; return; }
    w1 = (d11 * d20 - d01 * d21) / den;
    w2 = (d00 * d21 - d01 * d20) / den;
    w0 = 1.0 - w1 - w2;
}

// bounding box of the vertices (one workgroup): bb = {xmin, ymin, xmax, ymax}
__global__ __launch_bounds__(256) void k_vertex_bounds(const double *__restrict__ uv, int nv,
                                                       double *__restrict__ bb)
{
    __shared__ double r[4][256];
    double a = INFINITY, b = INFINITY, c = -INFINITY, d = -INFINITY;
    for (int i = threadIdx.x; i < nv; i += 256) {
        const double x = uv[2 * i], y = uv[2 * i + 1];
        a = fmin(a, x); b = fmin(b, y); c = fmax(c, x); d = fmax(d, y);
    }
    r[0][threadIdx.x] = a; r[1][threadIdx.x] = b; r[2][threadIdx.x] = c; r[3][threadIdx.x] = d;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            r[0][threadIdx.x] = fmin(r[0][threadIdx.x], r[0][threadIdx.x + o]);
            r[1][threadIdx.x] = fmin(r[1][threadIdx.x], r[1][threadIdx.x + o]);
            r[2][threadIdx.x] = fmax(r[2][threadIdx.x], r[2][threadIdx.x + o]);
            r[3][threadIdx.x] = fmax(r[3][threadIdx.x], r[3][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) bb[threadIdx.x] = r[threadIdx.x][0];
}

// per face: index validation, the conservative bin box and the
// ill-conditioned flag
__global__ void k_face_boxes(const double *__restrict__ uv, int nv, const int *__restrict__ tri,
                             int nf, BinGrid g, FaceBox *__restrict__ box,
                             uint8_t *__restrict__ always, int *__restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const int a = tri[3 * i], b = tri[3 * i + 1], c = tri[3 * i + 2];
    FaceBox bx = {1, 0, 1, 0};
    always[i] = 0;
    if ((unsigned)a >= (unsigned)nv || (unsigned)b >= (unsigned)nv || (unsigned)c >= (unsigned)nv) {
        *bad = 1;
        box[i] = bx;
        return;
    }
    const double x1 = uv[2 * a], y1 = uv[2 * a + 1];
    const double x2 = uv[2 * b], y2 = uv[2 * b + 1];
    const double x3 = uv[2 * c], y3 = uv[2 * c + 1];
    const double v0x = x2 - x1, v0y = y2 - y1, v1x = x3 - x1, v1y = y3 - y1;
    const double d00 = v0x * v0x + v0y * v0y, d11 = v1x * v1x + v1y * v1y;
    const double d01 = v0x * v1x + v0y * v1y;
    const double den = d00 * d11 - d01 * d01;
    if (den == 0) { box[i] = bx; return; }          // holds no point
    if (!(den > 1e-12 * d00 * d11)) {                // thin or non-finite: test everywhere
        always[i] = 1;
        box[i] = bx;
        return;
    }
    const double xmin = fmin(x1, fmin(x2, x3)), xmax = fmax(x1, fmax(x2, x3));
    const double ymin = fmin(y1, fmin(y2, y3)), ymax = fmax(y1, fmax(y2, y3));
    const double eps = 1e-9 * fmax(xmax - xmin, ymax - ymin) + 1e-300;
    bx.x0 = bin_of(xmin - eps, g.ox, g.inv, g.G);
    bx.x1 = bin_of(xmax + eps, g.ox, g.inv, g.G);
    bx.y0 = bin_of(ymin - eps, g.oy, g.inv, g.G);
    bx.y1 = bin_of(ymax + eps, g.oy, g.inv, g.G);
    box[i] = bx;
}

// (face, bin) pairs: count per bin
__global__ void k_bin_count(const FaceBox *__restrict__ box, int nf, int G,
                            int *__restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const FaceBox b = box[i];
    for (int y = b.y0; y <= b.y1; ++y)
        for (int x = b.x0; x <= b.x1; ++x) atomicAdd(&count[y * G + x], 1);
}

// Exclusive prefix sum of n counts into start[0 .. n], in three launches
// over tiles of 4096 counts (1024 threads x 4 consecutive counts):
//   k_scan_tiles        each tile's total, parked in start[tile * 4096];
//   k_scan_tile_totals  one workgroup turns the parked totals into the tiles'
//                       offsets in place and writes start[n];
//   k_scan_apply        each tile reads its offset, then writes its starts
//                       (its first entry, the parked offset, last).
constexpr int kScanTile = 4096;

// exclusive prefix sum over a 1024-thread workgroup of one value per thread;
// total receives the workgroup's sum (lds: 32 ints)
__device__ __forceinline__ int wg_exclusive_scan(int v, int *lds, int &total)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    __syncthreads();                       // lds is free (a previous call has read it)
    if (lane == 63) lds[wv] = x;
    __syncthreads();
    if (wv == 0) {
        int t = (lane < 16) ? lds[lane] : 0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const int y = __shfl_up(t, o, 64);
            if (lane >= o) t += y;
        }
        if (lane < 16) lds[16 + lane] = t;
    }
    __syncthreads();
    total = lds[31];
    return (wv > 0 ? lds[15 + wv] : 0) + x - v;
}

// the thread's 4 counts (0 past n)
__device__ __forceinline__ int4 tile_counts(const int *__restrict__ count, long base, int n)
{
    if (base + 3 < n) return *reinterpret_cast<const int4 *>(count + base);
    int4 c = {0, 0, 0, 0};
    if (base < n) c.x = count[base];
    if (base + 1 < n) c.y = count[base + 1];
    if (base + 2 < n) c.z = count[base + 2];
    return c;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(const int *__restrict__ count, int n,
                                                     int *__restrict__ start)
{
    __shared__ int lds[32];
    const long tile0 = (long)blockIdx.x * kScanTile;
    const int4 c = tile_counts(count, tile0 + 4 * threadIdx.x, n);
    int total;
    wg_exclusive_scan(c.x + c.y + c.z + c.w, lds, total);
    if (threadIdx.x == 0) start[tile0] = total;
}

__global__ __launch_bounds__(1024) void k_scan_tile_totals(int *__restrict__ start, int n,
                                                           int ntiles)
{
    __shared__ int lds[32];
    int carry = 0;
    for (int c = 0; c < ntiles; c += 1024) {
        const int b = c + (int)threadIdx.x;
        const long at = (long)b * kScanTile;
        const int v = (b < ntiles) ? start[at] : 0;
        int total;
        const int ex = wg_exclusive_scan(v, lds, total);
        if (b < ntiles) start[at] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) start[n] = carry;
}

__global__ __launch_bounds__(1024) void k_scan_apply(const int *__restrict__ count, int n,
                                                     int *__restrict__ start)
{
    __shared__ int lds[32];
    const long tile0 = (long)blockIdx.x * kScanTile;
    const int off = start[tile0];          // read by every thread before the scan's barriers
    const long base = tile0 + 4 * threadIdx.x;
    const int4 c = tile_counts(count, base, n);
    int total;
    int run = off + wg_exclusive_scan(c.x + c.y + c.z + c.w, lds, total);
    const int v[4] = {c.x, c.y, c.z, c.w};
    for (int k = 0; k < 4; ++k)
        if (base + k < n) {
            start[base + k] = run;
            run += v[k];
        }
}

// (face, bin) pairs: each face writes its index into the lists of its bins
__global__ void k_bin_fill(const FaceBox *__restrict__ box, int nf, int G,
                           const int *__restrict__ start, int *__restrict__ cursor,
                           int *__restrict__ list)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const FaceBox b = box[i];
    for (int y = b.y0; y <= b.y1; ++y)
        for (int x = b.x0; x <= b.x1; ++x) {
            const int bin = y * G + x;
            list[start[bin] + atomicAdd(&cursor[bin], 1)] = i;
        }
}

// one thread per query point: the lowest-index face of its bin list and of
// the "always" list that holds it; face -1 when none does
__global__ __launch_bounds__(256) void k_locate(const double *__restrict__ uv,
                                                const double *__restrict__ p3,
                                                const int *__restrict__ tri,
                                                const int *__restrict__ list,
                                                const int *__restrict__ start,
                                                const int *__restrict__ alw, int nalw, BinGrid g,
                                                const double *__restrict__ q, int nq,
                                                int *__restrict__ face, double *__restrict__ w,
                                                double *__restrict__ out3)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nq) return;
    const double x = q[2 * k], y = q[2 * k + 1];
    int hit = -1;
    double w0 = -1, w1 = -1, w2 = -1;
    // a non-finite point is in no face
    if (isfinite(x) && isfinite(y)) {
        const int bin = bin_of(y, g.oy, g.inv, g.G) * g.G + bin_of(x, g.ox, g.inv, g.G);
        const int a = start[bin], n = start[bin + 1] - a;
        for (int j = 0; j < n + nalw; ++j) {
            const int i = (j < n) ? list[a + j] : alw[j - n];
            if (hit >= 0 && i > hit) continue;      // a lower-index hit is known
            const int p0 = tri[3 * i], p1 = tri[3 * i + 1], p2 = tri[3 * i + 2];
            double b0, b1, b2;
            rec_bary(uv[2 * p0], uv[2 * p0 + 1], uv[2 * p1], uv[2 * p1 + 1], uv[2 * p2],
                     uv[2 * p2 + 1], x, y, b0, b1, b2);
            if (b0 < 0 || b1 < 0 || b2 < 0) continue;
            hit = i;
            w0 = b0; w1 = b1; w2 = b2;
        }
    }
    face[k] = hit;
    if (w) {
        w[3 * k] = w0;
        w[3 * k + 1] = w1;
        w[3 * k + 2] = w2;
    }
    if (out3) {
        double r[3] = {0.0, 0.0, 0.0};
        if (hit >= 0) {
            const int v[3] = {tri[3 * hit], tri[3 * hit + 1], tri[3 * hit + 2]};
            const double wt[3] = {w0, w1, w2};
            for (int c = 0; c < 3; ++c)
                for (int j = 0; j < 3; ++j) r[c] += p3[3 * v[j] + c] * wt[j];
        }
        out3[3 * k] = r[0];
        out3[3 * k + 1] = r[1];
        out3[3 * k + 2] = r[2];
    }
}

void launch_exclusive_scan(hipStream_t st, const int *count, int n, int *start)
{
    const int ntiles = (int)(((long)n + kScanTile - 1) / kScanTile);
    if (ntiles > 0)
        hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(1024), 0, st, count, n, start);
    hipLaunchKernelGGL(k_scan_tile_totals, dim3(1), dim3(1024), 0, st, start, n, ntiles);
    if (ntiles > 0)
        hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(1024), 0, st, count, n, start);
}

}  // namespace cvt

using namespace cvt;

#define LCHK(call, what)                                                               \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) { rc = fail(CVT_ERR_HIP, "%s: %s", what, hipGetErrorString(e_)); goto done; } \
    } while (0)

extern "C" int cvt_locate_points(const double *uv, const double *p3, int nv, const int *tri,
                                 int nf, const double *q, int nq, int *face, double *w,
                                 double *out3)
{
    clear_error();
    if (nv < 0 || nf < 0 || nq < 0) return fail(CVT_ERR_ARG, "negative sizes");
    if ((nv > 0 && !uv) || (nf > 0 && !tri) || (nq > 0 && (!q || !face)))
        return fail(CVT_ERR_ARG, "null array");
    if (out3 && nv > 0 && !p3) return fail(CVT_ERR_ARG, "out3 needs the 3D vertices p3");
    if (nq == 0) return CVT_OK;
    // about two faces per bin, at most 1024 x 1024 bins
    const int G = std::max(1, std::min(1024, (int)ceil(sqrt((double)std::max(nf, 1) / 2.0))));
    int rc = CVT_OK;
    hipStream_t st = nullptr;   // the legacy default stream, as the reference's calls use
    double *d_uv = nullptr, *d_p3 = nullptr, *d_q = nullptr, *d_w = nullptr, *d_out3 = nullptr,
           *d_bb = nullptr;
    int *d_tri = nullptr, *d_list = nullptr, *d_count = nullptr, *d_start = nullptr,
        *d_cursor = nullptr, *d_alw = nullptr, *d_bad = nullptr, *d_face = nullptr;
    FaceBox *d_box = nullptr;
    uint8_t *d_always = nullptr;
    char *arena = nullptr;
    int bad = 0, nalw = 0, total = 0, nbins = 0;
    double bb[4] = {0, 0, 0, 0};
    BinGrid g;
    std::vector<uint8_t> always;
    std::vector<int> alw;
    const size_t nva = (size_t)std::max(nv, 1), nfa = (size_t)std::max(nf, 1);
    const size_t nb = (size_t)G * G;
    // the call's fixed-size buffers in one allocation, each at a 256-byte
    // boundary (the bin lists, sized by the count, get their own)
    size_t at = 0;
    auto carve = [&](auto *&p, size_t n) {
        at = (at + 255) & ~(size_t)255;
        p = reinterpret_cast<std::remove_reference_t<decltype(p)>>(arena + at);
        at += n * sizeof(*p);
    };
    auto layout = [&]() {
        at = 0;
        carve(d_uv, 2 * nva); carve(d_tri, 3 * nfa); carve(d_q, 2 * (size_t)nq);
        carve(d_face, (size_t)nq); carve(d_box, nfa); carve(d_always, nfa); carve(d_alw, nfa);
        carve(d_count, nb); carve(d_cursor, nb); carve(d_start, nb + 1); carve(d_bad, 1);
        carve(d_bb, 4);
        if (w) carve(d_w, 3 * (size_t)nq);
        if (out3) { carve(d_out3, 3 * (size_t)nq); carve(d_p3, 3 * nva); }
    };
    layout();
    LCHK(hipMalloc(&arena, at), "alloc");
    layout();
    if (out3 && nv > 0)
        LCHK(hipMemcpyAsync(d_p3, p3, sizeof(double) * 3 * nv, hipMemcpyDefault, st), "upload");
    if (nv > 0) LCHK(hipMemcpyAsync(d_uv, uv, sizeof(double) * 2 * nv, hipMemcpyDefault, st), "upload");
    if (nf > 0) LCHK(hipMemcpyAsync(d_tri, tri, sizeof(int) * 3 * nf, hipMemcpyDefault, st), "upload");
    LCHK(hipMemcpyAsync(d_q, q, sizeof(double) * 2 * nq, hipMemcpyDefault, st), "upload");
    LCHK(hipMemsetAsync(d_bad, 0, sizeof(int), st), "memset");
    if (nv > 0) {
        hipLaunchKernelGGL(k_vertex_bounds, dim3(1), dim3(256), 0, st, d_uv, nv, d_bb);
        LCHK(hipMemcpyAsync(bb, d_bb, sizeof(bb), hipMemcpyDeviceToHost, st), "bounds");
    }
    LCHK(hipStreamSynchronize(st), "bounds");
    {
        // square bins over the vertex box (a zero-size or non-finite box: one bin)
        const double ext = std::max(bb[2] - bb[0], bb[3] - bb[1]);
        g.G = G;
        g.ox = std::isfinite(bb[0]) ? bb[0] : 0.0;
        g.oy = std::isfinite(bb[1]) ? bb[1] : 0.0;
        g.inv = (ext > 0.0 && std::isfinite(ext)) ? (double)G / ext : 0.0;
        if (g.inv == 0.0) g.G = 1;
    }
    nbins = g.G * g.G;
    LCHK(hipMemsetAsync(d_count, 0, sizeof(int) * nbins, st), "memset");
    LCHK(hipMemsetAsync(d_cursor, 0, sizeof(int) * nbins, st), "memset");
    if (nf > 0) {
        hipLaunchKernelGGL(k_face_boxes, dim3((nf + 255) / 256), dim3(256), 0, st, d_uv, nv, d_tri,
                           nf, g, d_box, d_always, d_bad);
        hipLaunchKernelGGL(k_bin_count, dim3((nf + 255) / 256), dim3(256), 0, st, d_box, nf, g.G,
                           d_count);
    }
    launch_exclusive_scan(st, d_count, nbins, d_start);
    LCHK(hipMemcpyAsync(&total, d_start + nbins, sizeof(int), hipMemcpyDeviceToHost, st), "total");
    LCHK(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, st), "flag");
    always.resize(std::max(nf, 1));
    if (nf > 0) LCHK(hipMemcpyAsync(always.data(), d_always, nf, hipMemcpyDeviceToHost, st), "flags");
    LCHK(hipStreamSynchronize(st), "boxes");
    if (bad) { rc = fail(CVT_ERR_ARG, "a face indexes a vertex outside [0, %d)", nv); goto done; }
    for (int i = 0; i < nf; ++i)
        if (always[i]) alw.push_back(i);
    nalw = (int)alw.size();
    if (nalw) LCHK(hipMemcpyAsync(d_alw, alw.data(), sizeof(int) * nalw, hipMemcpyHostToDevice, st), "upload");
    LCHK(hipMalloc(&d_list, sizeof(int) * (size_t)std::max(total, 1)), "alloc");
    if (nf > 0)
        hipLaunchKernelGGL(k_bin_fill, dim3((nf + 255) / 256), dim3(256), 0, st, d_box, nf, g.G,
                           d_start, d_cursor, d_list);
    hipLaunchKernelGGL(k_locate, dim3((nq + 255) / 256), dim3(256), 0, st, d_uv, d_p3, d_tri,
                       d_list, d_start, d_alw, nalw, g, d_q, nq, d_face, d_w, d_out3);
    LCHK(hipGetLastError(), "locate launch");
    LCHK(hipMemcpyAsync(face, d_face, sizeof(int) * nq, hipMemcpyDefault, st), "download");
    if (w) LCHK(hipMemcpyAsync(w, d_w, sizeof(double) * 3 * nq, hipMemcpyDefault, st), "download");
    if (out3)
        LCHK(hipMemcpyAsync(out3, d_out3, sizeof(double) * 3 * nq, hipMemcpyDefault, st), "download");
    LCHK(hipStreamSynchronize(st), "locate");
done:
    if (d_list) (void)hipFree(d_list);
    if (arena) (void)hipFree(arena);
    return rc;
}
