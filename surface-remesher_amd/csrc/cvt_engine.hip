// cvt_engine.hip -- host side of the C ABI declared in include/cvt.h.
//
// One cvt_ctx owns every device buffer of one grid for its whole life (the
// reference allocates and frees about eleven N^2 buffers per gCVT call,
// source/gcvt.cu:840-903).  A Lloyd iteration is P JFA launches (the first a
// scatter of the seeds), one moment launch and one update launch (whose last
// block also adds the energy partials and appends to the history); it is
// captured once into a hipGraph
// and replayed, so the host does no work per iteration and never
// synchronises inside the loop (the reference copies the energy back every
// 10th iteration, source/gcvt.cu:1080).
#include "cvt_engine.h"
#include "cvt_kernels.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace cvt {

static int launch_moments_on(Ctx *c, const uint32_t *lab);
static thread_local char g_err[512];

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

void clear_error() { g_err[0] = 0; }
const char *last_error() { return g_err; }

int jfa_first_step(int W, int H)
{
    int n = std::max(W, H), s = 1;
    while (s * 2 < n) s *= 2;
    return s;
}

int jfa_num_passes(int W, int H)
{
    int s = jfa_first_step(W, H), p = 1;
    while (s > 1) { s >>= 1; ++p; }
    return p + 1;
}

static int choose_tile(int W, int H, int S)
{
    // the 3x3 tile window around a site's home tile must hold its cell: take a
    // tile of at least four mean cell diameters, at most 256
    double cell = sqrt((double)W * H / std::max(S, 1));
    int t = 32;
    while (t < 256 && t < 4.0 * cell) t *= 2;
    return t;
}

int Ctx::enqueue_moments() { return launch_moments_on(this, labels_map()); }

static int launch_moments_on(Ctx *c, const uint32_t *lab)
{
    const float rs = (float)ldexp(1.0, -c->rho_exp);
    dim3 grid(c->tilesX * c->tilesY), block(c->tile >= 64 ? 512 : 256);
    switch (c->tile) {
    case 32:
        hipLaunchKernelGGL(k_moments<32>, grid, block, 0, c->stream, lab, c->rho, c->sites, c->S,
                           c->W, c->H, rs, c->tilesX, c->tilesY, c->slots, (long)c->capS, c->ovf, c->ovf_cap,
                           c->ovf_head, c->dflags, c->dstate);
        break;
    case 64:
        hipLaunchKernelGGL(k_moments<64>, grid, block, 0, c->stream, lab, c->rho, c->sites, c->S,
                           c->W, c->H, rs, c->tilesX, c->tilesY, c->slots, (long)c->capS, c->ovf, c->ovf_cap,
                           c->ovf_head, c->dflags, c->dstate);
        break;
    case 128:
        hipLaunchKernelGGL(k_moments<128>, grid, block, 0, c->stream, lab, c->rho, c->sites, c->S,
                           c->W, c->H, rs, c->tilesX, c->tilesY, c->slots, (long)c->capS, c->ovf, c->ovf_cap,
                           c->ovf_head, c->dflags, c->dstate);
        break;
    default:
        hipLaunchKernelGGL(k_moments<256>, grid, block, 0, c->stream, lab, c->rho, c->sites, c->S,
                           c->W, c->H, rs, c->tilesX, c->tilesY, c->slots, (long)c->capS, c->ovf, c->ovf_cap,
                           c->ovf_head, c->dflags, c->dstate);
        break;
    }
    return hipGetLastError() == hipSuccess ? CVT_OK : fail(CVT_ERR_HIP, "moment launch failed");
}

// moment tile side (0: from the site density) and the overflow pool it needs
static int set_tile_internal(Ctx *c, int tile)
{
    if (tile == 0) tile = choose_tile(c->W, c->H, c->S);
    if (tile != 32 && tile != 64 && tile != 128 && tile != 256)
        return fail(CVT_ERR_ARG, "tile must be 32, 64, 128 or 256 (got %d)", tile);
    if (tile == c->tile && c->slots) return CVT_OK;
    c->tile = tile;
    c->tilesX = (c->W + tile - 1) / tile;
    c->tilesY = (c->H + tile - 1) / tile;
    const long ntiles = (long)c->tilesX * c->tilesY;
    // a tile holds at most CVT_HASH_CAP labels, so at most that many records
    // per tile can overflow (more raises CVT_ERR_CAPACITY)
    c->ovf_cap = (int)std::min<long>(ntiles * (long)CVT_HASH_CAP, 1L << 26);
    if (c->ovf) hipFree(c->ovf);
    c->ovf = nullptr;
    if (hipMalloc(&c->ovf, sizeof(OvfRec) * (size_t)c->ovf_cap) != hipSuccess)
        return fail(CVT_ERR_NOMEM, "overflow buffers (%ld tiles)", ntiles);
    c->graph_valid = false;
    return CVT_OK;
}

int Ctx::check_flags()
{
    int f[NUM_FLAGS];
    if (hipMemcpyAsync(f, dflags, sizeof f, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return fail(CVT_ERR_HIP, "flag readback: %s", hipGetErrorString(hipGetLastError()));
    if (f[FLAG_FULL]) {
        int z = 0;
        hipMemcpyAsync(dflags + FLAG_FULL, &z, sizeof z, hipMemcpyHostToDevice, stream);
        hipStreamSynchronize(stream);
        const int cap = tile == 64 ? TileHash<64>::HC : tile == 128 ? TileHash<128>::HC : CVT_HASH_CAP;
        return fail(CVT_ERR_CAPACITY,
                    "more than %d distinct sites met one %dx%d moment tile; use a smaller tile",
                    cap, tile, tile);
    }
    return CVT_OK;
}

// ---------------------------------------------------------------------------
// one Lloyd iteration (or its parts) on the stream
// ---------------------------------------------------------------------------
int Ctx::enqueue_seed()
{
    if (S > 0)
        hipLaunchKernelGGL(k_seed, dim3((S + 255) / 256), dim3(256), 0, stream, sites, S, seed, W,
                           H);
    seeds_valid = true;
    return CVT_OK;
}

void Ctx::clear_seed_map()
{
    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, stream, seed, none(), (long)W * H);
}

// JFA(+1) of the current sites; the labels end in lab[final_buf]
int Ctx::enqueue_jfa(hipEvent_t after_first)
{
    const int np = jfa_num_passes(W, H);
    int step = jfa_first_step(W, H);
    const uint32_t *in = seed;
    int cur = 0;
    // Pass 0 reads only the seed map: with three passes or more it scatters
    // the seeds into `first`, which holds none except at their receivers
    // (measured: 7.6 against 13.5 us at 2048^2, 6.9 against 7.8 at 512^2).
    // Pass 1 resets the seed map it no longer needs, pass 2 resets `first`:
    // both hold entries only around the seed pixels of sites [0, S) (the
    // sites do not move during a JFA; compaction only shrinks the count).
    const bool scatter = np >= 3;
    const int step0 = step;
    // A deferred finalize of the previous iteration (the Lloyd iteration's
    // site update leaves it pending) runs as one more workgroup of the seed
    // scatter, or as its own launch when there is no scatter; either is a
    // no-op when nothing is pending.
    const FinalArgs fa = final_args(1, nullptr);
    if (!scatter) enqueue_finalize();
    // with a small site table the last five passes run fused (launch_jfa_tail)
    const bool tail = use_tail && jfa_tail_ok(jplan, np);
    for (int k = 0; k < np; ++k) {
        if (tail && k == np - 5) {
            launch_jfa_tail(stream, jplan, in, lab[cur], sites, W, H, dstate + ST_STOP);
            cur ^= 1;
            break;
        }
        if (k == 0 && scatter) {
            launch_jfa_seed_pass(stream, seed, first, sites, S, W, H, step, dstate + ST_STOP,
                                 blockpart, dstate, &fa);
            if (after_first) hipEventRecord(after_first, stream);
            in = first;
            if (step > 1) step >>= 1;
            continue;
        }
        uint32_t *out = lab[cur];
        uint32_t *clear = (k == 1) ? seed : (k == 2 && scatter) ? first : nullptr;
        launch_jfa_pass(stream, jplan, in, out, sites, W, H, step, clear, none(),
                        clear ? S : 0, k == 2 ? step0 : 0, dstate + ST_STOP);
        if (k == 0 && after_first) hipEventRecord(after_first, stream);
        in = out;
        cur ^= 1;
        if (step > 1) step >>= 1;
    }
    final_buf = cur ^ 1;
    seeds_valid = false;
    return hipGetLastError() == hipSuccess ? CVT_OK : fail(CVT_ERR_HIP, "jfa launch failed");
}

FinalArgs Ctx::final_args(int record, double *out_eg) const
{
    const double a = (double)W * H;
    FinalArgs fa;
    fa.hist = hist;
    fa.hist_cap = hist_cap;
    fa.record = record;
    fa.e_scale = ldexp(1.0, rho_exp) / a;
    fa.g_scale = 2.0 * ldexp(1.0, rho_exp) / a;
    fa.out_eg = out_eg;
    fa.until = until;
    fa.defer = 0;
    return fa;
}

void Ctx::enqueue_finalize()
{
    hipLaunchKernelGGL(k_finalize_pending, dim3(1), dim3(256), 0, stream, blockpart, dstate,
                       final_args(1, nullptr));
}

int Ctx::enqueue_update(int mode, float omega, double *mom_out, int record, double *out_eg,
                        const float *domega, CompatArgs ca, int defer)
{
    const int nb = std::max(1, (S + 255) / 256);
    FinalArgs fa = final_args(record, out_eg);
    fa.defer = defer;
    hipLaunchKernelGGL(k_update, dim3(nb), dim3(256), 0, stream, sites, fixed, S, slots, (long)capS, ovf,
                       ovf_head, dflags, dstate, omega, domega, mode, ldexp(1.0, rho_exp), mom_out,
                       blockpart, seed, W, H, ca, fa);
    if (mode & UPD_SEED) seeds_valid = true;
    return hipGetLastError() == hipSuccess ? CVT_OK : fail(CVT_ERR_HIP, "update launch failed");
}

int Ctx::enqueue_iteration(float omega)
{
    int rc;
    if ((rc = enqueue_jfa())) return rc;
    if ((rc = enqueue_moments())) return rc;
    return enqueue_update(UPD_MOVE | UPD_SEED, omega, nullptr, 1, nullptr, nullptr, CompatArgs{},
                          1);
}

// Iterations per launch of the long graph.  Between two graph launches the
// device idles for the launch's own overhead (~3.8 us at 2048^2); a graph of
// eight iterations pays it once per eight (C2 6807 -> 7026, C1 16 960 ->
// 18 800 iter/s; sixteen gained 0.2% more).  It equals cvt_lloyd_until's
// chunk, which then takes one launch.
static constexpr int kGraphIters = 8;

// captures `iters` Lloyd iterations into g / x
static int capture_iterations(Ctx *c, float omega, int iters, hipGraph_t &g, hipGraphExec_t &x)
{
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return fail(CVT_ERR_HIP, "stream capture could not start");
    const bool sv = c->seeds_valid;
    int rc = CVT_OK;
    for (int k = 0; k < iters && !rc; ++k) rc = c->enqueue_iteration(omega);
    c->seeds_valid = sv;
    hipGraph_t cg = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &cg);
    if (rc) { if (cg) hipGraphDestroy(cg); return rc; }
    if (e != hipSuccess || !cg) return fail(CVT_ERR_HIP, "stream capture failed");
    if (hipGraphInstantiate(&x, cg, nullptr, nullptr, 0) != hipSuccess) {
        hipGraphDestroy(cg);
        return fail(CVT_ERR_HIP, "graph instantiation failed");
    }
    g = cg;
    return CVT_OK;
}

int Ctx::build_graph(float omega)
{
    if (graph_valid && graph_omega == omega) return CVT_OK;
    if (graph_exec) { hipGraphExecDestroy(graph_exec); graph_exec = nullptr; }
    if (graph) { hipGraphDestroy(graph); graph = nullptr; }
    if (graph_exec_n) { hipGraphExecDestroy(graph_exec_n); graph_exec_n = nullptr; }
    if (graph_n) { hipGraphDestroy(graph_n); graph_n = nullptr; }
    graph_valid = false;
    int rc = capture_iterations(this, omega, 1, graph, graph_exec);
    if (!rc) rc = capture_iterations(this, omega, kGraphIters, graph_n, graph_exec_n);
    if (rc) return rc;
    graph_valid = true;
    graph_omega = omega;
    return CVT_OK;
}

int Ctx::lloyd_until(int max_iter, float omega, double rel_tol, int window, double target,
                     int *ran_out)
{
    if (!have_density) return fail(CVT_ERR_STATE, "cvt_lloyd_until before cvt_set_density");
    if (!until_host) {
        if (hipHostMalloc((void **)&until_host, 2 * sizeof(int)) != hipSuccess ||
            hipEventCreateWithFlags(&until_ev[0], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&until_ev[1], hipEventDisableTiming) != hipSuccess)
            return fail(CVT_ERR_HIP, "stop-flag buffers");
    }
    hipLaunchKernelGGL(k_until, dim3(1), dim3(64), 0, stream, dstate, until, 1, rel_tol, target,
                       window, ran);
    // Chunks of iterations are enqueued with a copy of the stop flag behind
    // each; before enqueuing chunk j + 1 the host waits for chunk j - 1 only
    // (chunk j is still queued, so the device never idles) and stops
    // enqueuing once a copy reads set.  The first chunk alone is waited for
    // before the second is enqueued: a run that stops within it (a warm
    // start close to convergence) then leaves no chunk of no-op iterations
    // behind.  Iterations enqueued after the stop write nothing.
    constexpr int kChunk = kGraphIters;
    int rc = CVT_OK, done = 0;
    for (int j = 0; done < max_iter; ++j) {
        const int n = std::min(kChunk, max_iter - done);
        if ((rc = lloyd(n, omega))) break;
        done += n;
        if (hipMemcpyAsync(until_host + (j & 1), dstate + ST_STOP, sizeof(int),
                           hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipEventRecord(until_ev[j & 1], stream) != hipSuccess) {
            rc = fail(CVT_ERR_HIP, "stop-flag copy");
            break;
        }
        if (j != 1) {   // chunk 0 alone, then chunk j - 1 while chunk j is queued
            const int w = j == 0 ? 0 : j - 1;
            if (hipEventSynchronize(until_ev[w & 1]) != hipSuccess) {
                rc = fail(CVT_ERR_HIP, "stop-flag wait");
                break;
            }
            if (until_host[w & 1]) break;
        }
    }
    hipLaunchKernelGGL(k_until, dim3(1), dim3(64), 0, stream, dstate, until, 0, rel_tol, target,
                       window, ran);
    // A stop raised by the finalize inside a seed scatter can leave part of
    // that scatter's receivers in `first`, which holds none between JFAs:
    // clear it.
    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, stream, first, none(), (long)W * H);
    int n = 0;
    if (hipMemcpyAsync(&n, ran, sizeof n, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return rc ? rc : fail(CVT_ERR_HIP, "iteration count readback");
    if (ran_out) *ran_out = n;
    return rc ? rc : check_flags();
}

int Ctx::lloyd(int iterations, float omega)
{
    if (!have_density) return fail(CVT_ERR_STATE, "cvt_lloyd before cvt_set_density");
    if (iterations <= 0) return CVT_OK;
    if (!seeds_valid) enqueue_seed();
    int rc;
    if (use_graph) {
        if ((rc = build_graph(omega))) return rc;
        int k = 0;
        for (; k + kGraphIters <= iterations; k += kGraphIters)
            if (hipGraphLaunch(graph_exec_n, stream) != hipSuccess)
                return fail(CVT_ERR_HIP, "graph launch failed");
        for (; k < iterations; ++k)
            if (hipGraphLaunch(graph_exec, stream) != hipSuccess)
                return fail(CVT_ERR_HIP, "graph launch failed");
        seeds_valid = true;
    } else {
        for (int k = 0; k < iterations; ++k)
            if ((rc = enqueue_iteration(omega))) return rc;
    }
    // the last iteration's finalize
    enqueue_finalize();
    labels_fresh = true;
    return CVT_OK;
}

Ctx::~Ctx()
{
    if (until_host) {
        hipHostFree(until_host);
        hipEventDestroy(until_ev[0]);
        hipEventDestroy(until_ev[1]);
    }
    if (graph_exec) hipGraphExecDestroy(graph_exec);
    if (graph) hipGraphDestroy(graph);
    if (graph_exec_n) hipGraphExecDestroy(graph_exec_n);
    if (graph_n) hipGraphDestroy(graph_n);
    void *bufs[] = {rho, seed, lab[0], lab[1], first, sites, fixed, slots, ovf, ovf_head,
                    blockpart, hist, dstate, dflags, scratch, eg, until, ran};
    for (void *b : bufs)
        if (b) hipFree(b);
    if (own_stream && own) hipStreamDestroy(own_stream);
}

}  // namespace cvt

using namespace cvt;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
#define CHECK_CTX(c) \
    do { if (!(c)) return fail(CVT_ERR_ARG, "null context"); clear_error(); } while (0)
#define HIPCHK(call, what) \
    do { hipError_t e_ = (call); if (e_ != hipSuccess) \
        return fail(CVT_ERR_HIP, "%s: %s", what, hipGetErrorString(e_)); } while (0)

extern "C" {

const char *cvt_last_error(void) { return last_error(); }
int cvt_version(void) { return 100; }

int cvt_create(cvt_ctx **out, int W, int H, int max_sites)
{
    clear_error();
    if (!out) return fail(CVT_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (W < 1 || H < 1 || W > 8192 || H > 8192)
        return fail(CVT_ERR_ARG, "grid %dx%d outside [1, 8192]^2", W, H);
    if (max_sites < 0 || max_sites > (1 << 28)) return fail(CVT_ERR_ARG, "max_sites %d", max_sites);
    Ctx *c = new Ctx();
    c->W = W; c->H = H; c->maxS = max_sites; c->S = 0;
    c->capS = std::max(max_sites, 1);
    c->jplan = plan_jfa(W, H, c->capS);
    hipGetDevice(&c->device);
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(CVT_ERR_HIP, "stream creation failed");
    }
    c->own = true;
    c->stream = c->own_stream;
    const long n = (long)W * H;
    const long ms = std::max(max_sites, 1);
    c->hist_cap = 1 << 16;
    const int nb = (int)std::max(1L, (ms + 255) / 256);
    bool ok = hipMalloc(&c->rho, sizeof(float) * n) == hipSuccess &&
              hipMalloc(&c->seed, sizeof(uint32_t) * n) == hipSuccess &&
              hipMalloc(&c->lab[0], sizeof(uint32_t) * n) == hipSuccess &&
              hipMalloc(&c->lab[1], sizeof(uint32_t) * n) == hipSuccess &&
              hipMalloc(&c->first, sizeof(uint32_t) * n) == hipSuccess &&
              hipMalloc(&c->sites, sizeof(float2) * (c->capS + 2)) == hipSuccess &&
              hipMalloc(&c->fixed, ms) == hipSuccess &&
              hipMalloc(&c->slots, sizeof(Slot) * 9 * ms) == hipSuccess &&
              hipMalloc(&c->ovf_head, sizeof(int) * ms) == hipSuccess &&
              hipMalloc(&c->blockpart, sizeof(double) * 2 * std::max(nb, 4096)) == hipSuccess &&
              hipMalloc(&c->hist, sizeof(double) * 2 * c->hist_cap) == hipSuccess &&
              hipMalloc(&c->dstate, sizeof(int) * NUM_STATE) == hipSuccess &&
              hipMalloc(&c->dflags, sizeof(int) * NUM_FLAGS) == hipSuccess &&
              hipMalloc(&c->scratch, sizeof(float) * 4096) == hipSuccess &&
              hipMalloc(&c->eg, sizeof(double) * 4) == hipSuccess &&
              hipMalloc(&c->until, sizeof(double) * 2) == hipSuccess &&
              hipMalloc(&c->ran, sizeof(int)) == hipSuccess;
    if (!ok) {
        delete c;
        return fail(CVT_ERR_NOMEM, "device allocation for a %dx%d grid, %d sites", W, H, max_sites);
    }
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, s, c->seed, c->none(), n);
    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, s, c->lab[0], c->none(), n);
    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, s, c->lab[1], c->none(), n);
    hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, s, c->first, c->none(), n);
    hipMemsetAsync(c->rho, 0, sizeof(float) * n, s);
    hipMemsetAsync(c->fixed, 0, ms, s);
    hipMemsetAsync(c->slots, 0, sizeof(Slot) * 9 * ms, s);
    hipMemsetAsync(c->ovf_head, 0xff, sizeof(int) * ms, s);   // -1: empty chains
    hipMemsetAsync(c->dflags, 0, sizeof(int) * NUM_FLAGS, s);
    int st[NUM_STATE] = {1, 0, 0, 0};
    hipMemcpyAsync(c->dstate, st, sizeof st, hipMemcpyHostToDevice, s);
    float2 sentinel = make_float2(__builtin_inff(), __builtin_inff());
    hipMemcpyAsync(c->sites + c->capS, &sentinel, sizeof sentinel, hipMemcpyHostToDevice, s);
    int rc = set_tile_internal(c, 64);
    if (rc == CVT_OK && hipStreamSynchronize(s) != hipSuccess)
        rc = fail(CVT_ERR_HIP, "initialisation failed");
    if (rc) { delete c; return rc; }
    c->tile_auto = true;
    *out = reinterpret_cast<cvt_ctx *>(c);
    return CVT_OK;
}

int cvt_destroy(cvt_ctx *ctx)
{
    clear_error();
    if (!ctx) return CVT_OK;
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    hipStreamSynchronize(c->stream);
    delete c;
    return CVT_OK;
}

int cvt_set_stream(cvt_ctx *ctx, void *stream)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    hipStreamSynchronize(c->stream);
    c->stream = stream ? (hipStream_t)stream : c->own_stream;
    c->graph_valid = false;
    return CVT_OK;
}

int cvt_set_density(cvt_ctx *ctx, const float *density)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!density) return fail(CVT_ERR_ARG, "null density");
    const long n = (long)c->W * c->H;
    HIPCHK(hipMemcpyAsync(c->rho, density, sizeof(float) * n, hipMemcpyDefault, c->stream),
           "density upload");
    int zero = 0;
    HIPCHK(hipMemcpyAsync(c->dflags + FLAG_BAD, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream),
           "flag reset");
    const int nb = 1024;
    hipLaunchKernelGGL(k_density_stats, dim3(nb), dim3(256), 0, c->stream, c->rho, n, c->scratch,
                       c->dflags);
    std::vector<float> bm(nb);
    int bad = 0;
    HIPCHK(hipMemcpyAsync(bm.data(), c->scratch, sizeof(float) * nb, hipMemcpyDeviceToHost, c->stream),
           "density stats");
    HIPCHK(hipMemcpyAsync(&bad, c->dflags + FLAG_BAD, sizeof bad, hipMemcpyDeviceToHost, c->stream),
           "density flag");
    HIPCHK(hipStreamSynchronize(c->stream), "density stats");
    if (bad) {
        c->have_density = false;
        return fail(CVT_ERR_ARG, "density holds a negative, NaN or infinite value");
    }
    float mx = 0.0f;
    for (float v : bm) mx = std::max(mx, v);
    // rho' = rho * 2^-rho_exp <= 1 keeps the fixed-point moment sums in range
    int ex = 0;
    if (mx > 0.0f) {
        frexp((double)mx, &ex);   // mx in [2^(ex-1), 2^ex)
    }
    c->rho_exp = ex;
    c->have_density = true;
    c->graph_valid = false;
    return CVT_OK;
}

int cvt_set_sites(cvt_ctx *ctx, const float *xy, const uint8_t *fixed, int num_sites)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (num_sites < 0 || num_sites > c->maxS)
        return fail(CVT_ERR_ARG, "num_sites %d outside [0, %d]", num_sites, c->maxS);
    if (num_sites > 0 && !xy) return fail(CVT_ERR_ARG, "null site array");
    // the seed map is all CVT_NONE unless seeds are pending: clear them first
    if (c->seeds_valid && c->S > 0)
        hipLaunchKernelGGL(k_fill_u32, dim3(1024), dim3(256), 0, c->stream, c->seed, c->none(),
                           (long)c->W * c->H);
    if (num_sites > 0) {
        HIPCHK(hipMemcpyAsync(c->sites, xy, sizeof(float2) * num_sites, hipMemcpyDefault, c->stream),
               "site upload");
        if (fixed)
            HIPCHK(hipMemcpyAsync(c->fixed, fixed, num_sites, hipMemcpyDefault, c->stream),
                   "fixed upload");
        else
            HIPCHK(hipMemsetAsync(c->fixed, 0, num_sites, c->stream), "fixed clear");
    }
    int bad = 0;
    HIPCHK(hipMemcpyAsync(c->dflags + FLAG_BAD, &bad, sizeof bad, hipMemcpyHostToDevice, c->stream),
           "flag reset");
    if (num_sites > 0)
        hipLaunchKernelGGL(k_site_check, dim3((num_sites + 255) / 256), dim3(256), 0, c->stream,
                           c->sites, num_sites, c->W, c->H, c->dflags);
    HIPCHK(hipMemcpyAsync(&bad, c->dflags + FLAG_BAD, sizeof bad, hipMemcpyDeviceToHost, c->stream),
           "site check");
    HIPCHK(hipStreamSynchronize(c->stream), "site upload");
    if (bad) {
        c->S = 0;
        c->seeds_valid = false;
        c->labels_fresh = false;
        c->graph_valid = false;
        return fail(CVT_ERR_ARG, "a site is not finite or lies outside [-W, 2W] x [-H, 2H]");
    }
    const int oldS = c->S;
    c->S = num_sites;
    c->seeds_valid = false;
    c->enqueue_seed();
    if (c->tile_auto) {
        int rc = set_tile_internal(c, 0);
        if (rc) return rc;
    }
    if (oldS != num_sites) c->graph_valid = false;
    c->labels_fresh = false;
    HIPCHK(hipGetLastError(), "seed launch");
    return CVT_OK;
}

int cvt_get_sites(cvt_ctx *ctx, float *xy)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (c->S == 0) return CVT_OK;
    if (!xy) return fail(CVT_ERR_ARG, "null output");
    HIPCHK(hipMemcpyAsync(xy, c->sites, sizeof(float2) * c->S, hipMemcpyDefault, c->stream),
           "site download");
    HIPCHK(hipStreamSynchronize(c->stream), "site download");
    return CVT_OK;
}

int cvt_num_sites(cvt_ctx *ctx)
{
    CHECK_CTX(ctx);
    return reinterpret_cast<Ctx *>(ctx)->S;
}

int cvt_voronoi(cvt_ctx *ctx)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c->seeds_valid) c->enqueue_seed();
    int rc = c->enqueue_jfa();
    if (rc) return rc;
    c->labels_fresh = true;
    HIPCHK(hipGetLastError(), "voronoi");
    return CVT_OK;
}

int cvt_get_labels(cvt_ctx *ctx, int32_t *labels)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!labels) return fail(CVT_ERR_ARG, "null output");
    if (!c->labels_fresh) return fail(CVT_ERR_STATE, "no label map yet (call cvt_voronoi)");
    const long n = (long)c->W * c->H;
    int32_t *tmp = reinterpret_cast<int32_t *>(c->spare_map());
    hipLaunchKernelGGL(k_export_labels, dim3(1024), dim3(256), 0, c->stream, c->labels_map(),
                       tmp, n, c->none());
    HIPCHK(hipMemcpyAsync(labels, tmp, sizeof(int32_t) * n, hipMemcpyDefault, c->stream),
           "label download");
    HIPCHK(hipStreamSynchronize(c->stream), "label download");
    return CVT_OK;
}

int cvt_moments(cvt_ctx *ctx, double *m)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c->have_density) return fail(CVT_ERR_STATE, "cvt_moments before cvt_set_density");
    if (!c->labels_fresh) return fail(CVT_ERR_STATE, "no label map yet (call cvt_voronoi)");
    if (!m) return fail(CVT_ERR_ARG, "null output");
    double *dm = nullptr;
    HIPCHK(hipMalloc(&dm, sizeof(double) * 4 * std::max(c->S, 1)), "moment buffer");
    int rc = launch_moments_on(c, c->labels_map());
    if (!rc) rc = c->enqueue_update(0, 1.0f, dm, 0, nullptr);
    if (!rc && c->S > 0 &&
        hipMemcpyAsync(m, dm, sizeof(double) * 4 * c->S, hipMemcpyDefault, c->stream) != hipSuccess)
        rc = fail(CVT_ERR_HIP, "moment download");
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(CVT_ERR_HIP, "moments");
    hipFree(dm);
    if (!rc) rc = c->check_flags();
    return rc;
}

int cvt_lloyd(cvt_ctx *ctx, int iterations, float omega)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (iterations < 0) return fail(CVT_ERR_ARG, "iterations %d", iterations);
    if (!(omega > 0.0f && omega < 2.0f + 1e-6f)) return fail(CVT_ERR_ARG, "omega %g", omega);
    return c->lloyd(iterations, omega);
}

int cvt_lloyd_until(cvt_ctx *ctx, float omega, double rel_tol, int window, double target,
                    int max_iter, int *iterations)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (max_iter < 0) return fail(CVT_ERR_ARG, "max_iter %d", max_iter);
    if (window < 0 || window >= c->hist_cap) return fail(CVT_ERR_ARG, "window %d", window);
    if (!(omega > 0.0f && omega < 2.0f + 1e-6f)) return fail(CVT_ERR_ARG, "omega %g", omega);
    if (!(rel_tol >= 0.0) || target != target)
        return fail(CVT_ERR_ARG, "rel_tol %g, target %g", rel_tol, target);
    if (iterations) *iterations = 0;
    return c->lloyd_until(max_iter, omega, rel_tol, window, target, iterations);
}

int cvt_history_length(cvt_ctx *ctx)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    int st[NUM_STATE];
    HIPCHK(hipMemcpyAsync(st, c->dstate, sizeof st, hipMemcpyDeviceToHost, c->stream), "state");
    HIPCHK(hipStreamSynchronize(c->stream), "state");
    return std::min(st[ST_HIST], c->hist_cap);
}

int cvt_get_history(cvt_ctx *ctx, double *eg, int n)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (n < 0 || n > c->hist_cap) return fail(CVT_ERR_ARG, "n %d", n);
    if (n == 0) return CVT_OK;
    if (!eg) return fail(CVT_ERR_ARG, "null output");
    int rc = c->check_flags();
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(eg, c->hist, sizeof(double) * 2 * n, hipMemcpyDefault, c->stream),
           "history");
    HIPCHK(hipStreamSynchronize(c->stream), "history");
    return CVT_OK;
}

int cvt_clear_history(cvt_ctx *ctx)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    HIPCHK(hipMemsetAsync(c->dstate + ST_HIST, 0, sizeof(int), c->stream), "history clear");
    HIPCHK(hipStreamSynchronize(c->stream), "history clear");
    return CVT_OK;
}

int cvt_energy(cvt_ctx *ctx, double *energy, double *gradnorm)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c->have_density) return fail(CVT_ERR_STATE, "cvt_energy before cvt_set_density");
    int rc = cvt_voronoi(ctx);
    if (rc) return rc;
    if ((rc = launch_moments_on(c, c->labels_map()))) return rc;
    if ((rc = c->enqueue_update(0, 1.0f, nullptr, 0, c->eg))) return rc;
    const long n = (long)c->W * c->H;
    const int nb = 1024;
    hipLaunchKernelGGL(k_energy64, dim3(nb), dim3(256), 0, c->stream, c->labels_map(), c->rho,
                       c->sites, c->S, c->W, c->H, c->blockpart);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, c->stream, c->blockpart, nb, c->dstate,
                       c->dflags, c->hist, c->hist_cap, 0, 1.0 / (double)n, 0.0, c->eg + 2);
    double eg[4];
    HIPCHK(hipMemcpyAsync(eg, c->eg, sizeof eg, hipMemcpyDeviceToHost, c->stream), "energy");
    HIPCHK(hipStreamSynchronize(c->stream), "energy");
    if ((rc = c->check_flags())) return rc;
    if (energy) *energy = eg[2];
    if (gradnorm) *gradnorm = eg[1];
    // the voronoi above consumed the seeds; re-seed for the next iteration
    c->enqueue_seed();
    return CVT_OK;
}

int cvt_profile(cvt_ctx *ctx, int iterations, float omega, double *ms)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c->have_density) return fail(CVT_ERR_STATE, "cvt_profile before cvt_set_density");
    if (iterations < 1 || iterations > 1000 || !ms)
        return fail(CVT_ERR_ARG, "iterations %d (1..1000)", iterations);
    const int np = jfa_num_passes(c->W, c->H);
    // five events per iteration bracket the phases (the first JFA pass, the
    // other passes, moments, update); nothing in between, so the passes run
    // back to back as in a graph replay
    constexpr int NE = 5;
    std::vector<hipEvent_t> ev(NE * (size_t)iterations);
    for (auto &e : ev) HIPCHK(hipEventCreate(&e), "event");
    int rc = CVT_OK;
    for (int it = 0; it < iterations && !rc; ++it) {
        if (!c->seeds_valid) c->enqueue_seed();
        hipEventRecord(ev[NE * it + 0], c->stream);
        rc = c->enqueue_jfa(ev[NE * it + 1]);
        hipEventRecord(ev[NE * it + 2], c->stream);
        if (!rc) rc = c->enqueue_moments();
        hipEventRecord(ev[NE * it + 3], c->stream);
        if (!rc)
            rc = c->enqueue_update(UPD_MOVE | UPD_SEED, omega, nullptr, 1, nullptr, nullptr,
                                   CompatArgs{}, 1);
        hipEventRecord(ev[NE * it + 4], c->stream);
    }
    c->enqueue_finalize();
    double acc[5] = {0, 0, 0, 0, 0};
    if (!rc && hipEventSynchronize(ev.back()) != hipSuccess) rc = fail(CVT_ERR_HIP, "profile");
    for (int it = 0; it < iterations && !rc; ++it) {
        float t;
        for (int k = 0; k < NE - 1; ++k) {
            hipEventElapsedTime(&t, ev[NE * it + k], ev[NE * it + k + 1]);
            acc[k] += t;
        }
        hipEventElapsedTime(&t, ev[NE * it], ev[NE * it + NE - 1]);
        acc[4] += t;
    }
    for (auto &e : ev) hipEventDestroy(e);
    if (rc) return rc;
    c->labels_fresh = true;
    ms[0] = np > 1 ? acc[1] / ((double)iterations * (np - 1)) : acc[0] / iterations;
    ms[2] = np > 1 ? acc[0] / iterations : 0.0;
    ms[1] = acc[2] / iterations;
    ms[3] = acc[3] / iterations;
    ms[4] = (acc[0] + acc[1]) / iterations;
    ms[5] = acc[4] / iterations;
    ms[6] = np;
    ms[7] = c->tile;
    return c->check_flags();
}

int cvt_synchronize(cvt_ctx *ctx)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    HIPCHK(hipStreamSynchronize(c->stream), "synchronize");
    return c->check_flags();
}

int cvt_set_tile(cvt_ctx *ctx, int tile)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    c->tile_auto = (tile == 0);
    return set_tile_internal(c, tile);
}

int cvt_set_jfa_shape(cvt_ctx *ctx, int runs_mask, int blocks_per_cu)
{
    CHECK_CTX(ctx);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (blocks_per_cu < 0 || blocks_per_cu > 16)
        return fail(CVT_ERR_ARG, "blocks_per_cu %d (0 .. 16)", blocks_per_cu);
    if (runs_mask < 0) runs_mask = JFA_AUTO;
    if (c->jplan.runs_mask != runs_mask || c->jplan.blocks_per_cu != blocks_per_cu) {
        c->jplan = plan_jfa(c->W, c->H, c->capS, runs_mask, blocks_per_cu);
        c->graph_valid = false;
    }
    return CVT_OK;
}

int cvt_get_jfa_shape(cvt_ctx *ctx, int *runs_mask, int *blocks_per_cu)
{
    CHECK_CTX(ctx);
    const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
    if (runs_mask) *runs_mask = jfa_runs_mask(c->jplan, c->W, c->H);
    if (blocks_per_cu) *blocks_per_cu = c->jplan.blocks_per_cu;
    return CVT_OK;
}

int cvt_get_tile(cvt_ctx *ctx)
{
    CHECK_CTX(ctx);
    return reinterpret_cast<Ctx *>(ctx)->tile;
}

}  // extern "C"
