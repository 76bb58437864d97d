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
// depth > 1 runs the coarse-to-fine schedule of source/gcvt.cu:1110-1147 on
// a 2x2-mean density pyramid (see run_level below).
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

// mask pooled to a coarser level: a coarse pixel is fixed if any of its
// fine pixels is
static __global__ void k_mask_half(const uint8_t *__restrict__ in, uint8_t *__restrict__ out, int n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * n) return;
    const int x = (int)(i % n), y = (int)(i / n);
    const long w = 2L * n;
    out[i] = in[(2L * y) * w + 2 * x] | in[(2L * y + 1) * w + 2 * x] |
             in[(2L * y) * w + 2 * x + 1] | in[(2L * y + 1) * w + 2 * x + 1];
}

struct Level {
    int n = 0;
    float *rho = nullptr;     // device, n x n
    uint8_t *mask = nullptr;  // device, n x n
};

// One level of the integer-site Lloyd schedule on ctx (density and sites
// already set).  Returns CVT_OK and advances *it.
static int run_level(Ctx *c, const Level &lv, int max_iter, bool finest, int *it, int S0,
                     float *domega, int *tgt, uint32_t *owner, double escale)
{
    const double threshold = finest ? 1e-5 : 3e-1;   // source/gcvt.cu:1131-1137
    int nst[1] = {S0};
    if (hipMemcpyAsync(c->dstate + ST_NSITES, nst, sizeof(int), hipMemcpyHostToDevice, c->stream) !=
        hipSuccess)
        return fail(CVT_ERR_HIP, "state upload");
    float omega = 2.0f;
    if (hipMemcpyAsync(domega, &omega, sizeof omega, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return fail(CVT_ERR_HIP, "omega upload");
    // one iteration, captured once
    CompatArgs ca{lv.mask, lv.rho, tgt, owner};
    hipGraph_t g = nullptr;
    hipGraphExec_t gx = nullptr;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return fail(CVT_ERR_HIP, "capture");
    int rc = c->enqueue_jfa();
    if (!rc) rc = c->enqueue_moments();
    if (!rc) rc = c->enqueue_update(UPD_COMPAT | UPD_DEV_S, omega, nullptr, 1, nullptr, domega, ca);
    if (!rc)
        hipLaunchKernelGGL(k_compat_compact, dim3(1), dim3(1024), 0, c->stream, c->sites, tgt, owner,
                           c->seed, lv.n, c->dstate);
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess || hipGraphInstantiate(&gx, g, nullptr, nullptr, 0) != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(CVT_ERR_HIP, "graph");
    }
    double lastE = 1e18;
    const int start = *it;
    std::vector<double> h(2);
    do {
        if (hipGraphLaunch(gx, c->stream) != hipSuccess) { rc = fail(CVT_ERR_HIP, "launch"); break; }
        ++*it;
        if ((*it) % 10 == 0) {
            // energy sampled before the update of iteration it - 10
            const int k = (*it - 10) - start;
            if (k < 0) continue;      // the level started mid-decade: no sample yet
            if (hipMemcpyAsync(h.data(), c->hist + 2 * k, 2 * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess) {
                rc = fail(CVT_ERR_HIP, "energy readback");
                break;
            }
            const double E = h[0] * escale;   // coarse levels: x 4^level (source/gcvt.cu:1082)
            const double dE = lastE - E;
            omega = (float)std::min(2.0, 1.0 + dE);
            if (dE / 10.0 < threshold) break;
            lastE = E;
            if (hipMemcpyAsync(domega, &omega, sizeof omega, hipMemcpyHostToDevice, c->stream) !=
                hipSuccess) { rc = fail(CVT_ERR_HIP, "omega upload"); break; }
        }
    } while (*it < max_iter);
    (void)hipGraphExecDestroy(gx);
    (void)hipGraphDestroy(g);
    if (!rc) rc = c->check_flags();
    return rc;
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
    // levels: like the reference, coarse levels stop at 256 (source/gcvt.cu:1091),
    // and each level halves the side exactly
    int levels = 1;
    while (levels < depth && (size % (1 << levels)) == 0 && (size >> levels) >= 256) ++levels;

    const int n = size;
    const long np = (long)n * n;
    int rc = CVT_OK;
    int it = 0, S = 0;
    cvt_ctx *h = nullptr;
    Ctx *c = nullptr;
    std::vector<Level> lv(levels);
    float *domega = nullptr;
    int *tgt = nullptr;
    uint32_t *owner = nullptr;
    short2 *dcoords = nullptr;
    std::vector<float> xy;
    // sites from the input map (row-major); the map may live on either side
    std::vector<int16_t> vin(2 * (size_t)np);
    if (hipMemcpy(vin.data(), voronoi, sizeof(int16_t) * 2 * np, hipMemcpyDefault) != hipSuccess)
        return fail(CVT_ERR_HIP, "label map download");
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const long p = (long)y * n + x;
            if (vin[2 * p] == x && vin[2 * p + 1] == y) {
                xy.push_back((float)x + 0.5f);
                xy.push_back((float)y + 0.5f);
            }
        }
    S = (int)(xy.size() / 2);
    rc = cvt_create(&h, n, n, std::max(S, 1));
    if (rc) return rc;
    c = reinterpret_cast<Ctx *>(h);
    c->use_graph = false;
    // device copies of the level densities and masks
    lv[0].n = n;
    lv[0].rho = c->rho;
    CCHK(hipMalloc(&lv[0].mask, np), "alloc");
    if (mask) CCHK(hipMemcpyAsync(lv[0].mask, mask, np, hipMemcpyDefault, c->stream), "mask upload");
    else CCHK(hipMemsetAsync(lv[0].mask, 0, np, c->stream), "mask clear");
    rc = cvt_set_density(h, density);
    if (rc) goto done;
    for (int l = 1; l < levels; ++l) {
        lv[l].n = n >> l;
        const long m = (long)lv[l].n * lv[l].n;
        CCHK(hipMalloc(&lv[l].rho, sizeof(float) * m), "alloc");
        CCHK(hipMalloc(&lv[l].mask, m), "alloc");
        hipLaunchKernelGGL(k_density_half, dim3((m + 255) / 256), dim3(256), 0, c->stream,
                           lv[l - 1].rho, lv[l].rho, lv[l].n);
        hipLaunchKernelGGL(k_mask_half, dim3((m + 255) / 256), dim3(256), 0, c->stream,
                           lv[l - 1].mask, lv[l].mask, lv[l].n);
    }
    CCHK(hipMalloc(&domega, sizeof(float)), "alloc");
    CCHK(hipMalloc(&tgt, sizeof(int) * std::max(S, 1)), "alloc");
    CCHK(hipMalloc(&owner, sizeof(uint32_t) * np), "alloc");
    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, c->stream, owner, (uint32_t)CVT_NONE, np);
    CCHK(hipStreamSynchronize(c->stream), "setup");

    if (levels == 1) {
        rc = cvt_set_sites(h, xy.data(), nullptr, S);
        if (rc) goto done;
        rc = run_level(c, lv[0], std::max(max_iter, 1), true, &it, S, domega, tgt, owner, 1.0);
        if (rc) goto done;
    } else {
        // coarse to fine: sites at (x >> l, y >> l) merged per coarse pixel;
        // after each coarser level the site coordinates double (source/gcvt.cu:485-495)
        std::vector<float> cur;
        for (int l = levels - 1; l >= 0; --l) {
            const int ln = lv[l].n;
            std::vector<float> lxy;
            if (l == levels - 1) {
                std::vector<uint8_t> seen((size_t)ln * ln, 0);
                for (int k = 0; k < S; ++k) {
                    const int x = (int)xy[2 * k] >> l, y = (int)xy[2 * k + 1] >> l;
                    if (seen[(size_t)y * ln + x]) continue;
                    seen[(size_t)y * ln + x] = 1;
                    lxy.push_back((float)x + 0.5f);
                    lxy.push_back((float)y + 0.5f);
                }
            } else {
                for (size_t k = 0; k + 1 < cur.size(); k += 2) {
                    lxy.push_back((float)(2 * (int)cur[k]) + 0.5f);
                    lxy.push_back((float)(2 * (int)cur[k + 1]) + 0.5f);
                }
            }
            // re-point the context at this level's grid
            cvt_ctx *hl = nullptr;
            const int Sl = (int)(lxy.size() / 2);
            rc = cvt_create(&hl, ln, ln, std::max(Sl, 1));
            if (rc) goto done;
            Ctx *cl = reinterpret_cast<Ctx *>(hl);
            cl->use_graph = false;
            if (l == 0) rc = cvt_set_density(hl, density);
            else rc = cvt_set_density(hl, lv[l].rho);
            Level use = lv[l];
            if (l == 0) use.rho = cl->rho;
            if (!rc) rc = cvt_set_sites(hl, lxy.data(), nullptr, Sl);
            if (!rc)
                rc = run_level(cl, use, std::max(max_iter, 1), l == 0, &it, Sl, domega, tgt, owner,
                               ldexp(1.0, 2 * l));
            if (!rc) {
                int nst = 0;
                if (hipMemcpy(&nst, cl->dstate + ST_NSITES, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
                    rc = fail(CVT_ERR_HIP, "readback");
                cur.resize(2 * (size_t)nst);
                if (!rc && nst > 0 && hipMemcpy(cur.data(), cl->sites, sizeof(float2) * nst,
                                                hipMemcpyDeviceToHost) != hipSuccess)
                    rc = fail(CVT_ERR_HIP, "readback");
            }
            if (l == 0 && !rc) {
                // keep the finest context for the final pass
                (void)cvt_destroy(h);
                h = hl;
                c = cl;
            } else {
                (void)cvt_destroy(hl);
            }
            if (rc) goto done;
            if (it >= std::max(max_iter, 1) && l > 0) {
                // iteration budget spent on a coarse level: zoom the rest
                for (int z = l - 1; z >= 0; --z) {
                    std::vector<float> nxt;
                    for (size_t k = 0; k + 1 < cur.size(); k += 2) {
                        nxt.push_back((float)(2 * (int)cur[k]) + 0.5f);
                        nxt.push_back((float)(2 * (int)cur[k + 1]) + 0.5f);
                    }
                    cur.swap(nxt);
                }
                cvt_ctx *hf = nullptr;
                rc = cvt_create(&hf, n, n, std::max((int)(cur.size() / 2), 1));
                if (rc) goto done;
                (void)cvt_destroy(h);
                h = hf;
                c = reinterpret_cast<Ctx *>(hf);
                rc = cvt_set_density(h, density);
                if (!rc) rc = cvt_set_sites(h, cur.data(), nullptr, (int)(cur.size() / 2));
                if (rc) goto done;
                int nst = (int)(cur.size() / 2);
                CCHK(hipMemcpy(c->dstate + ST_NSITES, &nst, sizeof(int), hipMemcpyHostToDevice), "state");
                break;
            }
        }
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
                           c->lab[c->final_buf], c->sites, c->none(), np, dcoords);
        CCHK(hipMemcpyAsync(voronoi, dcoords, sizeof(short2) * np, hipMemcpyDefault, c->stream),
             "download");
        CCHK(hipStreamSynchronize(c->stream), "final pass");
        if (num_sites) *num_sites = nst;
        if (iterations) *iterations = it;
    }
done:
    for (int l = 0; l < levels; ++l) {
        if (lv[l].mask) (void)hipFree(lv[l].mask);
        if (l > 0 && lv[l].rho) (void)hipFree(lv[l].rho);
    }
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
