// compat.hip -- the reference-compatible integer-site mode behind gCVT, the
// CDT point extraction of delaunayInput, and the C++ shims that keep the
// reference's two entry-point signatures.
//
// Integer-site mode (source/gcvt.cu:753-781, :1087-1156):
//   * sites are the pixels whose (x, y) label equals their own coordinates,
//     in row-major pixel order; a site on a mask pixel is fixed;
//   * the Voronoi map is this engine's JFA(+1) over the sites' pixel centres
//     (the reference uses exact PBA+; ties are implementation-defined there);
//   * moments are taken over the cells, the update moves each free site to
//     trunc(s + omega (c - s) + 0.5) (toward zero), clamped, unless the
//     target has rho == 0; sites landing on one pixel merge (lowest index
//     survives, survivors keep their order);
//   * omega = 2; the energy of iterations 0, 10, 20, ... (before their
//     update) drives omega = min(2, 1 + dE) and the stop rule dE / 10 < 1e-5
//     after every 10th update; at most maxIter iterations (at least one);
//   * the energy is the exact float64 value, not the reference's biased
//     float32 reduction (source/gcvt.cu:1070-1078);
//   * one final Voronoi pass; the output is the full (x, y) label map.
// depth > 1 runs the coarse-to-fine loop of source/gcvt.cu:1087-1147 on the
// 2x2-mean density pyramid: the initial sites are read from the first nc^2
// pairs of the caller's map as an nc x nc map (the coarsest side; the
// reference copies exactly those), each level reads the caller's mask with
// its own row stride, the iteration count, omega, the sampled energy and
// lastE run on across levels (a level whose first decade check falls mid-way
// uses the coarser level's last sample), the coarse stop threshold is 0.3,
// every level runs at least one iteration, and site coordinates double
// between levels.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <vector>

#include "../../include/cvt.h"
#include "../../include/gcvt_compat.h"
#include "cvt_engine.h"
#include "cvt_kernels.h"

namespace cvt {

// 2x2 mean of a (2n x 2n) density into n x n (source/gcvt.cu:497-511)
static __global__ void k_density_half(const float *__restrict__ in, float *__restrict__ out, int n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * n) return;
    const int x = (int)(i % n), y = (int)(i / n);
    const long w = 2L * n;
    float d = 0.0f;
    d += in[(2L * y) * w + 2 * x];
    d += in[(2L * y + 1) * w + 2 * x];
    d += in[(2L * y) * w + 2 * x + 1];
    d += in[(2L * y + 1) * w + 2 * x + 1];
    out[i] = d / 4.0f;
}

// State of gCVT's schedule shared by all levels (source/gcvt.cu:1105-1108):
// the iteration count, omega, the last sampled energy and the energy of the
// last decade check that went on.
struct Schedule {
    int it = 0;
    float omega = 2.0f;
    double E = 0.0, lastE = 1e18;
};

// One level of the integer-site Lloyd schedule on ctx (density and sites
// already set), continuing sc.  mask: the caller's mask, read with this
// level's row stride (the reference indexes it with the level's TOID,
// source/gcvt.cu:1033).  The level runs at least one iteration (do ...
// while, source/gcvt.cu:1111-1142).
static int run_level(Ctx *c, const float *rho, const uint8_t *mask, int max_iter, bool finest,
                     Schedule &sc, int S0, float *domega, int *tgt, uint32_t *owner, double escale)
{
    const double threshold = finest ? 1e-5 : 3e-1;   // source/gcvt.cu:1131-1137
    int nst[1] = {S0};
    if (hipMemcpyAsync(c->dstate + ST_NSITES, nst, sizeof(int), hipMemcpyHostToDevice, c->stream) !=
        hipSuccess)
        return fail(CVT_ERR_HIP, "state upload");
    if (hipMemcpyAsync(domega, &sc.omega, sizeof sc.omega, hipMemcpyHostToDevice, c->stream) !=
        hipSuccess)
        return fail(CVT_ERR_HIP, "omega upload");
    // one iteration, captured once
    CompatArgs ca{mask, rho, tgt, owner};
    hipGraph_t g = nullptr;
    hipGraphExec_t gx = nullptr;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return fail(CVT_ERR_HIP, "capture");
    int rc = c->enqueue_jfa();
    if (!rc) rc = c->enqueue_moments();
    if (!rc) rc = c->enqueue_update(UPD_COMPAT | UPD_DEV_S, sc.omega, nullptr, 1, nullptr, domega, ca);
    if (!rc)
        hipLaunchKernelGGL(k_compat_compact, dim3(1), dim3(1024), 0, c->stream, c->sites, tgt, owner,
                           c->seed, c->W, c->dstate);
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess || hipGraphInstantiate(&gx, g, nullptr, nullptr, 0) != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(CVT_ERR_HIP, "graph");
    }
    // the device history of this context holds one energy per iteration of
    // this level, before its update: entry it - start
    const int start = sc.it;
    double h[2];
    auto sample = [&](int k) -> int {
        if (hipMemcpyAsync(h, c->hist + 2 * k, 2 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return fail(CVT_ERR_HIP, "energy readback");
        sc.E = h[0] * escale;   // coarse levels: x 4^level (source/gcvt.cu:1082)
        return CVT_OK;
    };
    do {
        if (hipGraphLaunch(gx, c->stream) != hipSuccess) { rc = fail(CVT_ERR_HIP, "launch"); break; }
        ++sc.it;
        if (sc.it % 10 == 0) {
            // E is the energy of iteration it - 10; when that iteration ran on
            // a coarser level, sc.E already holds it
            const int k = (sc.it - 10) - start;
            if (k >= 0 && (rc = sample(k))) break;
            const double dE = sc.lastE - sc.E;
            sc.omega = (float)std::min(2.0, 1.0 + dE);
            if (dE / 10.0 < threshold) break;
            sc.lastE = sc.E;
            if (hipMemcpyAsync(domega, &sc.omega, sizeof sc.omega, hipMemcpyHostToDevice,
                               c->stream) != hipSuccess) { rc = fail(CVT_ERR_HIP, "omega upload"); break; }
        }
    } while (sc.it < max_iter);
    // the level's last sample (the last multiple of 10 it ran) is the energy
    // the next level's first decade check reads, as the reference's Energy
    // variable carries it
    const int last = ((sc.it - 1) / 10) * 10;
    if (!rc && last >= start) rc = sample(last - start);
    (void)hipGraphExecDestroy(gx);
    (void)hipGraphDestroy(g);
    if (!rc) rc = c->check_flags();
    return rc;
}

// Pyramid levels for a side and a requested depth (source/gcvt.cu:1091):
// levels stop before a side below 256; for sides the reference does not take
// (not a power of two >= 256) a level is added only while the side halves
// exactly, and a side below 256 runs one level.
static int compat_levels(int n, int depth)
{
    int levels = 1;
    while (levels < depth && (n % (1 << levels)) == 0 && (n >> levels) >= 256) ++levels;
    return levels;
}

}  // namespace cvt

using namespace cvt;

#define CCHK(call, what)                                                               \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) { rc = fail(CVT_ERR_HIP, "%s: %s", what, hipGetErrorString(e_)); goto done; } \
    } while (0)

extern "C" int cvt_compat_gcvt(int16_t *voronoi, const float *density, const uint8_t *mask,
                               int size, int depth, int max_iter, int *iterations, int *num_sites)
{
    clear_error();
    if (!voronoi || !density) return fail(CVT_ERR_ARG, "null array");
    if (size < 2 || size > 8192) return fail(CVT_ERR_ARG, "size %d outside [2, 8192]", size);
    if (depth < 1) return fail(CVT_ERR_ARG, "depth %d < 1", depth);
    const int levels = compat_levels(size, depth);
    const int n = size;
    const long np = (long)n * n;
    // the coarsest level's side: its initial map is the first nc^2 (x, y)
    // pairs of the caller's map, read with row stride nc (the reference
    // copies exactly those into its nc x nc texture, source/gcvt.cu:1099-1103;
    // for one level that is the whole map)
    const int nc = n >> (levels - 1);
    const long npc = (long)nc * nc;
    int rc = CVT_OK;
    int S = 0;
    Schedule sc;
    cvt_ctx *h = nullptr;
    Ctx *c = nullptr;
    std::vector<float *> lrho(levels, nullptr);
    uint8_t *dmask = nullptr;
    float *domega = nullptr;
    int *tgt = nullptr;
    uint32_t *owner = nullptr;
    short2 *dcoords = nullptr;
    std::vector<float> cur;
    // sites: the self-labelled pixels of that map, in row-major order; the
    // map may live on either side
    std::vector<int16_t> vin(2 * (size_t)npc);
    if (hipMemcpy(vin.data(), voronoi, sizeof(int16_t) * 2 * npc, hipMemcpyDefault) != hipSuccess)
        return fail(CVT_ERR_HIP, "label map download");
    for (int y = 0; y < nc; ++y)
        for (int x = 0; x < nc; ++x) {
            const long p = (long)y * nc + x;
            if (vin[2 * p] == x && vin[2 * p + 1] == y) {
                cur.push_back((float)x + 0.5f);
                cur.push_back((float)y + 0.5f);
            }
        }
    S = (int)(cur.size() / 2);
    // the finest level's context (sites only merge, so S bounds every level)
    rc = cvt_create(&h, n, n, std::max(S, 1));
    if (rc) return rc;
    c = reinterpret_cast<Ctx *>(h);
    c->use_graph = false;
    rc = cvt_set_density(h, density);
    if (rc) goto done;
    // the density pyramid (source/gcvt.cu:985-993) and the caller's mask,
    // which every level reads with its own row stride
    lrho[0] = c->rho;
    for (int l = 1; l < levels; ++l) {
        const long m = (long)(n >> l) * (n >> l);
        CCHK(hipMalloc(&lrho[l], sizeof(float) * m), "alloc");
        hipLaunchKernelGGL(k_density_half, dim3((m + 255) / 256), dim3(256), 0, c->stream,
                           lrho[l - 1], lrho[l], n >> l);
    }
    CCHK(hipMalloc(&dmask, np), "alloc");
    if (mask) CCHK(hipMemcpyAsync(dmask, mask, np, hipMemcpyDefault, c->stream), "mask upload");
    else CCHK(hipMemsetAsync(dmask, 0, np, c->stream), "mask clear");
    CCHK(hipMalloc(&domega, sizeof(float)), "alloc");
    CCHK(hipMalloc(&tgt, sizeof(int) * std::max(S, 1)), "alloc");
    CCHK(hipMalloc(&owner, sizeof(uint32_t) * np), "alloc");
    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, c->stream, owner, (uint32_t)CVT_NONE, np);
    CCHK(hipStreamSynchronize(c->stream), "setup");

    // coarse to fine (source/gcvt.cu:1110-1147); between levels the site
    // coordinates double (kernelZoomIn, source/gcvt.cu:485-495)
    for (int l = levels - 1; l >= 0; --l) {
        const int m = n >> l;
        cvt_ctx *hl = h;
        if (l > 0) {
            rc = cvt_create(&hl, m, m, std::max(S, 1));
            if (rc) goto done;
            reinterpret_cast<Ctx *>(hl)->use_graph = false;
            rc = cvt_set_density(hl, lrho[l]);
        }
        Ctx *cl = reinterpret_cast<Ctx *>(hl);
        const int Sl = (int)(cur.size() / 2);
        if (!rc) rc = cvt_set_sites(hl, cur.data(), nullptr, Sl);
        if (!rc)
            rc = run_level(cl, cl->rho, dmask, max_iter, l == 0, sc, Sl, domega, tgt, owner,
                           ldexp(1.0, 2 * l));
        if (!rc && l > 0) {
            int nst = 0;
            if (hipMemcpy(&nst, cl->dstate + ST_NSITES, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
                rc = fail(CVT_ERR_HIP, "readback");
            cur.resize(2 * (size_t)nst);
            if (!rc && nst > 0 && hipMemcpy(cur.data(), cl->sites, sizeof(float2) * nst,
                                            hipMemcpyDeviceToHost) != hipSuccess)
                rc = fail(CVT_ERR_HIP, "readback");
            for (size_t k = 0; k < cur.size(); ++k) cur[k] = (float)(2 * (int)cur[k]) + 0.5f;
        }
        if (l > 0) (void)cvt_destroy(hl);
        if (rc) goto done;
    }
    // final Voronoi of the surviving sites, as (x, y) coordinates
    {
        int nst = 0;
        CCHK(hipMemcpyAsync(&nst, c->dstate + ST_NSITES, sizeof(int), hipMemcpyDeviceToHost,
                            c->stream), "readback");
        CCHK(hipStreamSynchronize(c->stream), "readback");
        c->seeds_valid = true;   // the last merge wrote the seeds of the survivors
        rc = c->enqueue_jfa();
        if (rc) goto done;
        CCHK(hipMalloc(&dcoords, sizeof(short2) * np), "alloc");
        hipLaunchKernelGGL(k_labels_to_coords, dim3(1024), dim3(256), 0, c->stream,
                           c->labels_map(), c->sites, c->none(), np, dcoords);
        CCHK(hipMemcpyAsync(voronoi, dcoords, sizeof(short2) * np, hipMemcpyDefault, c->stream),
             "download");
        CCHK(hipStreamSynchronize(c->stream), "final pass");
        if (num_sites) *num_sites = nst;
        if (iterations) *iterations = sc.it;
    }
done:
    for (int l = 1; l < levels; ++l)
        if (lrho[l]) (void)hipFree(lrho[l]);
    if (dmask) (void)hipFree(dmask);
    if (domega) (void)hipFree(domega);
    if (tgt) (void)hipFree(tgt);
    if (owner) (void)hipFree(owner);
    if (dcoords) (void)hipFree(dcoords);
    if (h) (void)cvt_destroy(h);
    return rc;
}

// ---------------------------------------------------------------------------
// CDT input points (source/delaunay.h:48-58)
// ---------------------------------------------------------------------------
namespace cvt {

// per column x: number of self-labelled, non-mask pixels (pass 0) or their
// points written in y order from the column's offset (pass 1)
static __global__ void k_extract(const short2 *__restrict__ vor, const uint8_t *__restrict__ mask,
                                 int n, int *__restrict__ colcount, const long *__restrict__ coloff,
                                 double scale, double l, double b, double *__restrict__ out, long cap)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    long k = coloff ? coloff[x] : 0;
    int cnt = 0;
    for (int y = 0; y < n; ++y) {
        const long p = (long)y * n + x;
        const short2 v = vor[p];
        if (v.x == x && v.y == y && !(mask && mask[p])) {
            if (coloff && k < cap) {
                out[2 * k] = x * scale + l;
                out[2 * k + 1] = y * scale + b;
            }
            ++k;
            ++cnt;
        }
    }
    if (!coloff) colcount[x] = cnt;
}

}  // namespace cvt

extern "C" int cvt_extract_points(const int16_t *voronoi, const uint8_t *mask, int n, double scale,
                                  double l, double b, double *out, long capacity, long *count)
{
    clear_error();
    if (!voronoi || !count || n < 1 || n > 8192 || capacity < 0 || (capacity > 0 && !out))
        return fail(CVT_ERR_ARG, "bad arguments");
    const long np = (long)n * n;
    int rc = CVT_OK;
    short2 *dv = nullptr;
    uint8_t *dm = nullptr;
    int *dcnt = nullptr;
    long *doff = nullptr;
    double *dout = nullptr;
    hipStream_t st = nullptr;   // the legacy default stream, as the reference's calls use
    std::vector<int> cnt(n);
    std::vector<long> off(n);
    long total = 0;
    CCHK(hipMalloc(&dv, sizeof(short2) * np), "alloc");
    CCHK(hipMalloc(&dcnt, sizeof(int) * n), "alloc");
    CCHK(hipMalloc(&doff, sizeof(long) * n), "alloc");
    CCHK(hipMemcpyAsync(dv, voronoi, sizeof(short2) * np, hipMemcpyDefault, st), "upload");
    if (mask) {
        CCHK(hipMalloc(&dm, np), "alloc");
        CCHK(hipMemcpyAsync(dm, mask, np, hipMemcpyDefault, st), "upload");
    }
    hipLaunchKernelGGL(k_extract, dim3((n + 63) / 64), dim3(64), 0, st, dv, dm, n, dcnt, nullptr,
                       scale, l, b, nullptr, 0L);
    CCHK(hipMemcpyAsync(cnt.data(), dcnt, sizeof(int) * n, hipMemcpyDeviceToHost, st), "counts");
    CCHK(hipStreamSynchronize(st), "count pass");
    for (int x = 0; x < n; ++x) { off[x] = total; total += cnt[x]; }
    *count = total;
    if (total > 0 && capacity > 0) {
        CCHK(hipMalloc(&dout, sizeof(double) * 2 * std::min(total, capacity)), "alloc");
        CCHK(hipMemcpyAsync(doff, off.data(), sizeof(long) * n, hipMemcpyHostToDevice, st), "upload");
        hipLaunchKernelGGL(k_extract, dim3((n + 63) / 64), dim3(64), 0, st, dv, dm, n, dcnt, doff,
                           scale, l, b, dout, capacity);
        CCHK(hipMemcpyAsync(out, dout, sizeof(double) * 2 * std::min(total, capacity),
                            hipMemcpyDefault, st), "download");
        CCHK(hipStreamSynchronize(st), "write pass");
    }
    // out == NULL is a count query; otherwise every point must fit
    if (out && total > capacity) rc = fail(CVT_ERR_CAPACITY, "%ld points, capacity %ld", total, capacity);
done:
    void *bufs[] = {dv, dm, dcnt, doff, dout};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    return rc;
}

// ---------------------------------------------------------------------------
// C++ shims with the reference's signatures (source/gcvt.h:29,
// source/discretization.h:66-67).  Like the reference they return nothing
// and end the process on an error (source/gcvt.cu:40-47).
// ---------------------------------------------------------------------------
void gCVT(short *Voronoi, float *density_d, bool *mask, int size, int depth, int maxIter)
{
    static_assert(sizeof(bool) == 1, "the mask is one byte per pixel");
    const int rc = cvt_compat_gcvt(reinterpret_cast<int16_t *>(Voronoi), density_d,
                                   reinterpret_cast<const uint8_t *>(mask), size, depth, maxIter,
                                   nullptr, nullptr);
    if (rc) {
        fprintf(stderr, "gCVT: %s\n", cvt_last_error());
        exit(1);
    }
}

void discretization_d(double *points, double *weight, int num_point, int *triangle, int num_tri,
                      float *density, double scale, int n)
{
    const int rc = cvt_rasterize(points, weight, num_point, triangle, num_tri, density, scale, n);
    if (rc) {
        fprintf(stderr, "discretization_d: %s\n", cvt_last_error());
        exit(1);
    }
}
