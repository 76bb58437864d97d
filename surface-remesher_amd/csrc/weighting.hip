// weighting.hip -- per-vertex density weights of the parameterized mesh: the
// ratio of the 3D area to the 2D area of each vertex's faces (replaces the
// host loops of the reference's weighting_vertices_with_area,
// source/weighting.h:37-97; its output is the rasterizer's vertex weight,
// source/main.cpp:199-204).
//
// Rule (source/weighting.h):
//   * every face adds its 3D area and its 2D area to the sums of each of its
//     three corners, faces taken in order (:53-74);
//   * the 2D area is get_area(A, B, C) = |cross(B - A, C - A)| / 2 with
//     cross(P, Q) = P.x Q.y - Q.x P.y over the face's vertices in order
//     (:29-35);
//   * the 3D area is CGAL 5.0.2's PMP::face_area: the square root of the
//     squared length of (q - p) x (r - p), over 4, on the points source(h),
//     target(h), target(next(h)) of the face's halfedge h.  A face is listed
//     as vertices_around_face(h) = (target(h), target(next(h)), source(h)) =
//     (t0, t1, t2), so (p, q, r) = (t2, t0, t1);
//   * weight = area3 / area2 (:79); a weight above 1e10 is replaced by the
//     largest other weight, which starts at 0 (:76-92).  A vertex in no face
//     gets 0 / 0 = NaN, which takes part in neither test.
//
// The reference's per-vertex sums are sequential in face order, so the
// rounding of a sum depends on that order.  Here the (vertex, corner) pairs
// are grouped by vertex with a counting sort (count, prefix sum, fill); the
// fill's atomics leave each vertex's list in arbitrary order, so each vertex
// sorts its own list of corner keys 3 f + c and then adds the face areas in
// that order, which is the reference's.  The largest weight is an atomic max
// on the bit pattern: weights are +0 or positive, where bit order is value
// order.  Contraction is off for this file: every expression rounds as
// written, as in the oracle.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "../../include/cvt.h"
#include "cvt_engine.h"

namespace cvt {

constexpr double kWeightCap = 1e10;   // source/weighting.h:76

// per face: index validation, the 3D and 2D areas, corner counts per vertex
__global__ __launch_bounds__(256) void k_face_areas(const double *__restrict__ uv,
                                                    const double *__restrict__ p3, int nv,
                                                    const int *__restrict__ tri, int nf,
                                                    double *__restrict__ a3,
                                                    double *__restrict__ a2,
                                                    int *__restrict__ count,
                                                    int *__restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const int t0 = tri[3 * i], t1 = tri[3 * i + 1], t2 = tri[3 * i + 2];
    if ((unsigned)t0 >= (unsigned)nv || (unsigned)t1 >= (unsigned)nv ||
        (unsigned)t2 >= (unsigned)nv) {
        *bad = 1;
        return;
    }
    {
        const double *p = p3 + 3 * t2, *q = p3 + 3 * t0, *r = p3 + 3 * t1;
        const double ux = q[0] - p[0], uy = q[1] - p[1], uz = q[2] - p[2];
        const double vx = r[0] - p[0], vy = r[1] - p[1], vz = r[2] - p[2];
        const double cx = uy * vz - uz * vy;
        const double cy = uz * vx - ux * vz;
        const double cz = ux * vy - uy * vx;
        a3[i] = sqrt((cx * cx + cy * cy + cz * cz) / 4);
    }
    {
        const double *a = uv + 2 * t0, *b = uv + 2 * t1, *c = uv + 2 * t2;
        const double px = b[0] - a[0], py = b[1] - a[1];
        const double qx = c[0] - a[0], qy = c[1] - a[1];
        a2[i] = fabs(px * qy - qx * py) / 2.;
    }
    atomicAdd(&count[t0], 1);
    atomicAdd(&count[t1], 1);
    atomicAdd(&count[t2], 1);
}

// each corner writes its key 3 f + c into its vertex's list (a face with an
// index outside [0, nv) was not counted and is skipped; the call fails)
__global__ __launch_bounds__(256) void k_corner_fill(const int *__restrict__ tri, int nv, int nf,
                                                     const int *__restrict__ start,
                                                     int *__restrict__ cursor,
                                                     int *__restrict__ list)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const int t[3] = {tri[3 * i], tri[3 * i + 1], tri[3 * i + 2]};
    if ((unsigned)t[0] >= (unsigned)nv || (unsigned)t[1] >= (unsigned)nv ||
        (unsigned)t[2] >= (unsigned)nv)
        return;
    for (int c = 0; c < 3; ++c) list[start[t[c]] + atomicAdd(&cursor[t[c]], 1)] = 3 * i + c;
}

// ascending heapsort of a[0 .. n), in place (a thread's own list)
__device__ void sift_down(int *a, int i, int n)
{
    const int v = a[i];
    for (;;) {
        int c = 2 * i + 1;
        if (c >= n) break;
        if (c + 1 < n && a[c + 1] > a[c]) ++c;
        if (a[c] <= v) break;
        a[i] = a[c];
        i = c;
    }
    a[i] = v;
}

__device__ void heap_sort(int *a, int n)
{
    for (int i = n / 2 - 1; i >= 0; --i) sift_down(a, i, n);
    for (int e = n - 1; e > 0; --e) {
        const int t = a[0];
        a[0] = a[e];
        a[e] = t;
        sift_down(a, 0, e);
    }
}

// one thread per vertex: its corners in face order, both area sums, the
// weight, and the running maximum of the weights at or below the cap
__global__ __launch_bounds__(256) void k_vertex_weight(const int *__restrict__ start,
                                                       int *__restrict__ list,
                                                       const double *__restrict__ a3,
                                                       const double *__restrict__ a2, int nv,
                                                       double *__restrict__ weight,
                                                       unsigned long long *__restrict__ mx)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int b = start[v], n = start[v + 1] - b;
    heap_sort(list + b, n);
    double s3 = 0.0, s2 = 0.0;
    for (int k = 0; k < n; ++k) {
        const int f = list[b + k] / 3;
        s3 += a3[f];
        s2 += a2[f];
    }
    const double w = s3 / s2;
    weight[v] = w;
    if (w <= kWeightCap) atomicMax(mx, (unsigned long long)__double_as_longlong(w));
}

__global__ __launch_bounds__(256) void k_weight_cap(double *__restrict__ weight, int nv,
                                                    const unsigned long long *__restrict__ mx)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    if (weight[v] > kWeightCap) weight[v] = __longlong_as_double((long long)*mx);
}

}  // namespace cvt

using namespace cvt;

#define WCHK(call, what)                                                               \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) { rc = fail(CVT_ERR_HIP, "%s: %s", what, hipGetErrorString(e_)); goto done; } \
    } while (0)

extern "C" int cvt_vertex_weights(const double *uv, const double *p3, int nv, const int *tri,
                                  int nf, double *weight)
{
    clear_error();
    if (nv < 0 || nf < 0) return fail(CVT_ERR_ARG, "negative sizes");
    if ((nv > 0 && (!uv || !p3 || !weight)) || (nf > 0 && !tri))
        return fail(CVT_ERR_ARG, "null array");
    if (nv == 0 && nf > 0) return fail(CVT_ERR_ARG, "faces without vertices");
    if (nv == 0) return CVT_OK;
    if ((long)nf * 3 > INT32_MAX) return fail(CVT_ERR_ARG, "too many faces");
    int rc = CVT_OK, bad = 0;
    hipStream_t st = nullptr;   // the legacy default stream, as the reference's calls use
    double *d_uv = nullptr, *d_p3 = nullptr, *d_a3 = nullptr, *d_a2 = nullptr, *d_w = nullptr;
    int *d_tri = nullptr, *d_count = nullptr, *d_start = nullptr, *d_cursor = nullptr,
        *d_list = nullptr, *d_bad = nullptr;
    unsigned long long *d_mx = nullptr;
    char *arena = nullptr;
    const size_t nfa = (size_t)std::max(nf, 1), nvs = (size_t)nv;
    // every buffer of the call in one allocation, each at a 256-byte boundary
    size_t at = 0;
    auto carve = [&](auto *&p, size_t n) {
        at = (at + 255) & ~(size_t)255;
        p = reinterpret_cast<std::remove_reference_t<decltype(p)>>(arena + at);
        at += n * sizeof(*p);
    };
    auto layout = [&]() {
        at = 0;
        carve(d_uv, 2 * nvs); carve(d_p3, 3 * nvs); carve(d_w, nvs);
        carve(d_tri, 3 * nfa); carve(d_a3, nfa); carve(d_a2, nfa); carve(d_list, 3 * nfa);
        carve(d_count, nvs); carve(d_cursor, nvs); carve(d_start, nvs + 1);
        carve(d_bad, 1); carve(d_mx, 1);
    };
    layout();
    WCHK(hipMalloc(&arena, at), "alloc");
    layout();
    WCHK(hipMemcpyAsync(d_uv, uv, sizeof(double) * 2 * nv, hipMemcpyDefault, st), "upload");
    WCHK(hipMemcpyAsync(d_p3, p3, sizeof(double) * 3 * nv, hipMemcpyDefault, st), "upload");
    if (nf > 0) WCHK(hipMemcpyAsync(d_tri, tri, sizeof(int) * 3 * nf, hipMemcpyDefault, st), "upload");
    WCHK(hipMemsetAsync(d_count, 0, sizeof(int) * nv, st), "memset");
    WCHK(hipMemsetAsync(d_cursor, 0, sizeof(int) * nv, st), "memset");
    WCHK(hipMemsetAsync(d_bad, 0, sizeof(int), st), "memset");
    WCHK(hipMemsetAsync(d_mx, 0, sizeof(unsigned long long), st), "memset");
    if (nf > 0)
        hipLaunchKernelGGL(k_face_areas, dim3((nf + 255) / 256), dim3(256), 0, st, d_uv, d_p3, nv,
                           d_tri, nf, d_a3, d_a2, d_count, d_bad);
    launch_exclusive_scan(st, d_count, nv, d_start);
    if (nf > 0)
        hipLaunchKernelGGL(k_corner_fill, dim3((nf + 255) / 256), dim3(256), 0, st, d_tri, nv,
                           nf, d_start, d_cursor, d_list);
    hipLaunchKernelGGL(k_vertex_weight, dim3((nv + 255) / 256), dim3(256), 0, st, d_start, d_list,
                       d_a3, d_a2, nv, d_w, d_mx);
    hipLaunchKernelGGL(k_weight_cap, dim3((nv + 255) / 256), dim3(256), 0, st, d_w, nv, d_mx);
    WCHK(hipGetLastError(), "weights launch");
    WCHK(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, st), "flag");
    WCHK(hipStreamSynchronize(st), "weights");
    if (bad) { rc = fail(CVT_ERR_ARG, "a face indexes a vertex outside [0, %d)", nv); goto done; }
    WCHK(hipMemcpyAsync(weight, d_w, sizeof(double) * nv, hipMemcpyDefault, st), "download");
    WCHK(hipStreamSynchronize(st), "weights");
done:
    if (arena) (void)hipFree(arena);
    return rc;
}
