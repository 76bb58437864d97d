// batch.hip -- many uneven patches in one set of launches (the batched CVT).
//
// Layout: every per-pixel map (density, seed map, the two JFA label maps) is
// the concatenation of the patches' maps (BatchPatch::pix_off); every per-site
// array (site table, fixed flags, moment slots, overflow chains) is the
// concatenation of the patches' tables (site_off, each table ending with its
// own sentinel at site_off + cap).  Labels are byte offsets into the whole
// site table, so one label space serves all patches.
//
// One Lloyd iteration of the whole batch is Pmax JFA passes (patches run
// their passes right-aligned; a pass is one launch per step class present in
// it), one moment launch over the tiles of every patch, one update launch
// over the site blocks of every patch and one finalize launch (one block per
// patch); it is captured once into a hipGraph.  Each patch's results equal a
// single-patch run on the same inputs bit for bit: labels, moments (integer
// sums over the same 64-pixel runs for every tile of side >= 64) and the
// energy history (the same block partials added in the same order).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/cvt.h"
#include "cvt_engine.h"
#include "cvt_kernels.h"

namespace cvt {

// moments of one tile of one patch (tile side T >= 64): the single-map tile
// routine with 64-pixel segments, so a patch's sums equal its single-map run.
// HC: the hash capacity of the batch's largest tile (256, 512 or 1024 labels
// for tiles of 64, 128 or 256, as in the single-map kernel): the smaller the
// hash, the more tiles each CU keeps in flight.
template <int HC>
static __global__ __launch_bounds__(256) void k_moments_batch(
    const BatchPatch *__restrict__ patches, const int *__restrict__ tile_patch,
    const uint32_t *__restrict__ lab0, const uint32_t *__restrict__ lab1,
    const float *__restrict__ rho, const float2 *__restrict__ sites, Slot *__restrict__ slots,
    long nrows, OvfRec *__restrict__ ovf, int ovf_cap, int *__restrict__ ovf_head, int *__restrict__ dflags,
    int *__restrict__ dstate)
{
    __shared__ TileLds<HC> L;
    const int pid = tile_patch[blockIdx.x];
    const BatchPatch P = patches[pid];
    const int T = P.tile;
    const int tile = blockIdx.x - P.tile_off;
    const int tX = tile % P.tiles_x, tY = tile / P.tiles_x;
    const uint32_t *lab = ((P.passes - 1) & 1 ? lab1 : lab0) + P.pix_off;
    tile_moments<64>(L, lab, rho + P.pix_off, sites + P.site_off, P.site_off, P.S, P.W, P.H,
                     P.rho_scale, tX * T, tY * T, T, nullptr);
    tile_flush(L, sites, P.site_off, T, tX, tY, P.tiles_x, P.tiles_y, dstate[ST_EPOCH], slots, nrows, ovf,
               ovf_cap, ovf_head, dstate, dflags);
}

// site update of one block of one patch (continuous rule, as k_update)
static __global__ __launch_bounds__(256) void k_update_batch(
    const BatchPatch *__restrict__ patches, const int *__restrict__ blk_patch,
    float2 *__restrict__ sites, const uint8_t *__restrict__ fixed, const Slot *__restrict__ slots,
    long nrows, const OvfRec *__restrict__ ovf, int *__restrict__ ovf_head, int *__restrict__ dstate,
    float omega, int mode, double *__restrict__ blockpart, uint32_t *__restrict__ seed)
{
    __shared__ double red_e[256], red_g[256];
    const int pid = blk_patch[blockIdx.x];
    const BatchPatch P = patches[pid];
    const int i = (blockIdx.x - P.blk_off) * 256 + threadIdx.x;
    double e = 0.0, g = 0.0;
    if (i < P.S) {
        const int gi = P.site_off + i;
        // the site and its fixed flag are loaded with the moment slots
        float2 s = sites[gi];
        const int pinned = fixed[gi];
        double w, dx, dy, ee;
        if (cvt_site_totals(gi, dstate[ST_EPOCH], slots, nrows, ovf, ovf_head, w, dx, dy, ee) >= 0)
            ovf_head[gi] = -1;
        e = ee;
        g = dx * dx + dy * dy;
        if ((mode & UPD_MOVE) && w > 0.0 && pinned == 0) {
            s.x = (float)((double)s.x + (double)omega * (dx / w));
            s.y = (float)((double)s.y + (double)omega * (dy / w));
            sites[gi] = s;
        }
        if (mode & UPD_SEED) {
            const int x = cvt_pixel_of(s.x, P.W), y = cvt_pixel_of(s.y, P.H);
            atomicMin(&seed[P.pix_off + (long)y * P.W + x], 8u * (uint32_t)gi);
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
    if (blockIdx.x == 0 && threadIdx.x == 0) dstate[ST_OVFN] = 0;
}

// one block per patch: energy and gradient norm of its iteration into its
// history; block 0 advances the epoch
static __global__ __launch_bounds__(256) void k_finalize_batch(
    const BatchPatch *__restrict__ patches, const double *__restrict__ blockpart,
    const double *__restrict__ e_scale, int *__restrict__ hist_len, double *__restrict__ hist,
    int hist_cap, int *__restrict__ dstate)
{
    __shared__ double re[256], rg[256];
    const BatchPatch P = patches[blockIdx.x];
    double e = 0.0, g = 0.0;
    for (int b = threadIdx.x; b < P.nblk; b += 256) {
        e += blockpart[2 * (P.blk_off + b)];
        g += blockpart[2 * (P.blk_off + b) + 1];
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
        const int h = hist_len[blockIdx.x];
        if (h < hist_cap) {
            double *hp = hist + 2L * ((long)blockIdx.x * hist_cap + h);
            hp[0] = re[0] * e_scale[blockIdx.x];
            hp[1] = 2.0 * sqrt(rg[0]) * e_scale[blockIdx.x];
        }
        hist_len[blockIdx.x] = h + 1;
        if (blockIdx.x == 0) dstate[ST_EPOCH] = dstate[ST_EPOCH] + 1;
    }
}

static __global__ void k_seed_batch(const BatchPatch *__restrict__ patches, int p,
                                    const float2 *__restrict__ sites, uint32_t *__restrict__ seed)
{
    const BatchPatch P = patches[p];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.S) return;
    const int gi = P.site_off + i;
    const float2 s = sites[gi];
    const int x = cvt_pixel_of(s.x, P.W), y = cvt_pixel_of(s.y, P.H);
    atomicMin(&seed[P.pix_off + (long)y * P.W + x], 8u * (uint32_t)gi);
}

static __global__ void k_export_batch(const uint32_t *__restrict__ lab, int32_t *__restrict__ out,
                                      long n, uint32_t base_bytes, uint32_t none)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const uint32_t l = lab[i];
        out[i] = (l >= none) ? CVT_NONE : (int32_t)((l - base_bytes) >> 3);
    }
}

struct Batch {
    int np = 0;
    std::vector<BatchPatch> hp;
    std::vector<BatchUnit> units;
    std::vector<int> tile_patch, blk_patch;
    std::vector<char> have_density;
    std::vector<double> escale;
    long total_pix = 0;
    int total_sites = 0;       // size of the concatenated site table
    int pmax = 0, ntiles = 0, nblk = 0;
    size_t lds_bytes = 0;
    hipStream_t stream = nullptr, own_stream = nullptr;
    BatchPatch *d_patch = nullptr;
    BatchUnit *d_units = nullptr;
    int *d_tile_patch = nullptr, *d_blk_patch = nullptr;
    float *rho = nullptr;
    uint32_t *seed = nullptr, *lab0 = nullptr, *lab1 = nullptr;
    uint32_t *first = nullptr;     // the seed scatter's output, `none` between JFAs
    int *d_site_patch = nullptr;   // patch of each site slot (-1: padding, sentinel)
    float2 *sites = nullptr;
    uint8_t *fixed = nullptr;
    Slot *slots = nullptr;
    OvfRec *ovf = nullptr;
    int ovf_cap = 0;
    int hash_cap = CVT_HASH_CAP;   // moment hash of the largest tile (set with the tiles)
    int *ovf_head = nullptr;
    double *blockpart = nullptr, *hist = nullptr, *d_escale = nullptr;
    int *hist_len = nullptr;
    int hist_cap = 4096;
    int *dstate = nullptr, *dflags = nullptr;
    bool seeds_valid = false, dirty = true, labels_fresh = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    float graph_omega = -1.0f;

    ~Batch()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        void *b[] = {d_patch, d_units, d_tile_patch, d_blk_patch, rho, seed, lab0, lab1, sites,
                     fixed, slots, ovf, ovf_head, blockpart, hist, d_escale, hist_len, dstate,
                     dflags, first, d_site_patch};
        for (void *p : b)
            if (p) (void)hipFree(p);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }

    // work tables, tiles, blocks: rebuilt when site counts or tiles change
    int plan();
    int upload_patches();
    int enqueue_iteration(float omega);
    void enqueue_jfa();
    unsigned pass_modes(int k) const;
};

static int g_cu_count()
{
    static int cu = 0;
    if (!cu) {
        hipDeviceProp_t prop;
        int dev = 0;
        (void)hipGetDevice(&dev);
        cu = (hipGetDeviceProperties(&prop, dev) == hipSuccess) ? prop.multiProcessorCount : 256;
    }
    return cu;
}

int Batch::plan()
{
    // JFA work units: about four per CU over all pixels (a batch of 64
    // patches of 256^2 .. 2048^2: 391 batched iterations/s with two, 426
    // with four, 418 with eight), at least one per patch, at least 16 row
    // segments (one per wave) each
    const long target = std::max<long>(1, total_pix / (4L * g_cu_count()));
    units.clear();
    for (int p = 0; p < np; ++p) {
        const BatchPatch &P = hp[p];
        // a pass holds at least H/4 * nseg runs (four rows of one 256-pixel
        // segment each): at most one part per 16 runs keeps every wave busy
        const int nseg = (P.W + 255) / 256;
        int nb = (int)std::max<long>(1, ((long)P.W * P.H + target / 2) / target);
        nb = std::min(nb, std::max(1, P.H * nseg / 64));
        for (int b = 0; b < nb; ++b)
            units.push_back(BatchUnit{p, (int)((long)P.H * b / nb), (int)((long)P.H * (b + 1) / nb),
                                      b, nb, {0, 0, 0}});
    }
    // moment tiles and update blocks
    ntiles = 0;
    nblk = 0;
    tile_patch.clear();
    blk_patch.clear();
    hash_cap = 0;
    for (int p = 0; p < np; ++p) {
        BatchPatch &P = hp[p];
        hash_cap = std::max(hash_cap, P.tile == 64    ? TileHash<64>::HC
                                      : P.tile == 128 ? TileHash<128>::HC
                                                      : CVT_HASH_CAP);
        P.tiles_x = (P.W + P.tile - 1) / P.tile;
        P.tiles_y = (P.H + P.tile - 1) / P.tile;
        P.tile_off = ntiles;
        ntiles += P.tiles_x * P.tiles_y;
        tile_patch.insert(tile_patch.end(), (size_t)P.tiles_x * P.tiles_y, p);
        P.nblk = std::max(1, (P.S + 255) / 256);
        P.blk_off = nblk;
        nblk += P.nblk;
        blk_patch.insert(blk_patch.end(), (size_t)P.nblk, p);
    }
    if (d_units) (void)hipFree(d_units);
    if (d_tile_patch) (void)hipFree(d_tile_patch);
    if (d_blk_patch) (void)hipFree(d_blk_patch);
    if (blockpart) (void)hipFree(blockpart);
    d_units = nullptr; d_tile_patch = nullptr; d_blk_patch = nullptr; blockpart = nullptr;
    if (hipMalloc(&d_units, sizeof(BatchUnit) * units.size()) != hipSuccess ||
        hipMalloc(&d_tile_patch, sizeof(int) * ntiles) != hipSuccess ||
        hipMalloc(&d_blk_patch, sizeof(int) * nblk) != hipSuccess ||
        hipMalloc(&blockpart, sizeof(double) * 2 * nblk) != hipSuccess)
        return fail(CVT_ERR_NOMEM, "batch work tables");
    (void)hipMemcpy(d_units, units.data(), sizeof(BatchUnit) * units.size(), hipMemcpyHostToDevice);
    (void)hipMemcpy(d_tile_patch, tile_patch.data(), sizeof(int) * ntiles, hipMemcpyHostToDevice);
    (void)hipMemcpy(d_blk_patch, blk_patch.data(), sizeof(int) * nblk, hipMemcpyHostToDevice);
    // overflow record pool: a tile flushes at most CVT_HASH_CAP records
    const int cap = (int)std::min<long>((long)ntiles * CVT_HASH_CAP, 1L << 26);
    if (cap > ovf_cap) {
        if (ovf) (void)hipFree(ovf);
        ovf = nullptr;
        if (hipMalloc(&ovf, sizeof(OvfRec) * (size_t)cap) != hipSuccess)
            return fail(CVT_ERR_NOMEM, "batch overflow pool");
        ovf_cap = cap;
    }
    dirty = false;
    if (gexec) { (void)hipGraphExecDestroy(gexec); gexec = nullptr; }
    if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    return upload_patches();
}

int Batch::upload_patches()
{
    for (int p = 0; p < np; ++p) escale[p] = 0.0;
    if (hipMemcpy(d_patch, hp.data(), sizeof(BatchPatch) * np, hipMemcpyHostToDevice) != hipSuccess)
        return fail(CVT_ERR_HIP, "patch table upload");
    std::vector<double> es(np);
    for (int p = 0; p < np; ++p)
        es[p] = (double)hp[p].rho_scale > 0 ? (1.0 / hp[p].rho_scale) / ((double)hp[p].W * hp[p].H) : 0.0;
    escale = es;
    if (hipMemcpy(d_escale, es.data(), sizeof(double) * np, hipMemcpyHostToDevice) != hipSuccess)
        return fail(CVT_ERR_HIP, "scale upload");
    return CVT_OK;
}

// step classes (bit 0: >= 4, bit 1: 1, bit 2: 2) that batched pass k runs
// densely; bit 3: some patch starts its seed scatter at k
unsigned Batch::pass_modes(int k) const
{
    unsigned m = 0;
    for (const BatchPatch &P : hp) {
        const int local = k - (pmax - P.passes);
        if (local < 0) continue;
        if (local == 0 && P.passes >= 3) {
            m |= 8u;
            continue;
        }
        const int step = (local < P.passes - 1) ? (P.first_step >> local) : 1;
        m |= 1u << (step == 1 ? 1 : step == 2 ? 2 : 0);
    }
    return m;
}

void Batch::enqueue_jfa()
{
    for (int k = 0; k < pmax; ++k) {
        const unsigned modes = pass_modes(k);
        if (modes & 8u)
            launch_jfa_batch_seeds(stream, d_patch, d_site_patch, total_sites, seed, first, sites, k,
                                   pmax);
        if (modes & 7u)
            launch_jfa_batch(stream, d_patch, d_units, (int)units.size(), seed, first, lab0, lab1,
                             sites, k, pmax, lds_bytes, modes & 7u);
    }
}

int Batch::enqueue_iteration(float omega)
{
    enqueue_jfa();
#define CVT_MOMENTS_BATCH(HC)                                                                    \
    hipLaunchKernelGGL(k_moments_batch<HC>, dim3(ntiles), dim3(256), 0, stream, d_patch,          \
                       d_tile_patch, lab0, lab1, rho, sites, slots, (long)total_sites, ovf,       \
                       ovf_cap, ovf_head, dflags, dstate)
    if (hash_cap <= TileHash<64>::HC)
        CVT_MOMENTS_BATCH(TileHash<64>::HC);
    else if (hash_cap <= TileHash<128>::HC)
        CVT_MOMENTS_BATCH(TileHash<128>::HC);
    else
        CVT_MOMENTS_BATCH(CVT_HASH_CAP);
#undef CVT_MOMENTS_BATCH
    hipLaunchKernelGGL(k_update_batch, dim3(nblk), dim3(256), 0, stream, d_patch, d_blk_patch, sites,
                       fixed, slots, (long)total_sites, ovf, ovf_head, dstate, omega,
                       UPD_MOVE | UPD_SEED, blockpart,
                       seed);
    hipLaunchKernelGGL(k_finalize_batch, dim3(np), dim3(256), 0, stream, d_patch, blockpart,
                       d_escale, hist_len, hist, hist_cap, dstate);
    return hipGetLastError() == hipSuccess ? CVT_OK : fail(CVT_ERR_HIP, "batch launch");
}

static int batch_tile_for(int W, int H, int S)
{
    const double cell = sqrt((double)W * H / std::max(S, 1));
    int t = 64;
    while (t < 256 && t < 4.0 * cell) t *= 2;
    return t;
}

}  // namespace cvt

using namespace cvt;

#define BCHK(b) \
    do { if (!(b)) return fail(CVT_ERR_ARG, "null batch"); clear_error(); } while (0)

extern "C" {

int cvt_batch_create(cvt_batch **out, int num_patches, const int *widths, const int *heights,
                     const int *max_sites)
{
    clear_error();
    if (!out || num_patches < 1 || !widths || !heights || !max_sites)
        return fail(CVT_ERR_ARG, "bad arguments");
    *out = nullptr;
    Batch *b = new Batch();
    b->np = num_patches;
    b->hp.resize(num_patches);
    b->have_density.assign(num_patches, 0);
    b->escale.assign(num_patches, 0.0);
    long pix = 0;
    int soff = 0, maxcap = 0;
    for (int p = 0; p < num_patches; ++p) {
        const int W = widths[p], H = heights[p], ms = max_sites[p];
        if (W < 8 || H < 1 || W > 8192 || H > 8192 || (W % 4) != 0 || ms < 0) {
            delete b;
            return fail(CVT_ERR_ARG, "patch %d: %dx%d (width a multiple of 4 in [8, 8192]), %d sites",
                        p, W, H, ms);
        }
        BatchPatch &P = b->hp[p];
        memset(&P, 0, sizeof P);
        P.W = W; P.H = H; P.pix_off = pix;
        P.cap = std::max(ms, 1);
        P.site_off = soff;
        P.S = 0;
        P.passes = jfa_num_passes(W, H);
        P.first_step = jfa_first_step(W, H);
        P.tile = batch_tile_for(W, H, P.cap);
        P.rho_scale = 1.0f;
        pix += (long)W * H;
        soff += ((P.cap + 2 + 63) / 64) * 64;      // table + sentinel, 64-aligned
        b->pmax = std::max(b->pmax, P.passes);
        maxcap = std::max(maxcap, P.cap);
    }
    b->total_pix = pix;
    b->total_sites = soff;
    const size_t tbl = (size_t)(maxcap + 2) / 2 * 16;
    b->lds_bytes = tbl <= 160 * 1024 ? tbl : 0;
    bool ok = hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc(&b->d_patch, sizeof(BatchPatch) * num_patches) == hipSuccess &&
              hipMalloc(&b->rho, sizeof(float) * pix) == hipSuccess &&
              hipMalloc(&b->seed, sizeof(uint32_t) * pix) == hipSuccess &&
              hipMalloc(&b->first, sizeof(uint32_t) * pix) == hipSuccess &&
              hipMalloc(&b->d_site_patch, sizeof(int) * soff) == hipSuccess &&
              hipMalloc(&b->lab0, sizeof(uint32_t) * pix) == hipSuccess &&
              hipMalloc(&b->lab1, sizeof(uint32_t) * pix) == hipSuccess &&
              hipMalloc(&b->sites, sizeof(float2) * soff) == hipSuccess &&
              hipMalloc(&b->fixed, soff) == hipSuccess &&
              hipMalloc(&b->slots, sizeof(Slot) * 9 * (size_t)soff) == hipSuccess &&
              hipMalloc(&b->ovf_head, sizeof(int) * soff) == hipSuccess &&
              hipMalloc(&b->hist, sizeof(double) * 2 * (size_t)b->hist_cap * num_patches) == hipSuccess &&
              hipMalloc(&b->hist_len, sizeof(int) * num_patches) == hipSuccess &&
              hipMalloc(&b->d_escale, sizeof(double) * num_patches) == hipSuccess &&
              hipMalloc(&b->dstate, sizeof(int) * NUM_STATE) == hipSuccess &&
              hipMalloc(&b->dflags, sizeof(int) * NUM_FLAGS) == hipSuccess;
    if (!ok) {
        delete b;
        return fail(CVT_ERR_NOMEM, "batch of %d patches (%ld pixels, %d site slots)", num_patches,
                    pix, soff);
    }
    b->stream = b->own_stream;
    hipStream_t s = b->stream;
    // seed map empty: every pixel holds its patch's `none`
    std::vector<float2> sent(soff, make_float2(0.0f, 0.0f));
    std::vector<int> site_patch(soff, -1);
    for (int p = 0; p < num_patches; ++p) {
        const BatchPatch &P = b->hp[p];
        sent[P.site_off + P.cap] = make_float2(__builtin_inff(), __builtin_inff());
        for (int i = 0; i < P.cap; ++i) site_patch[P.site_off + i] = p;
        hipLaunchKernelGGL(k_fill_u32, dim3(256), dim3(256), 0, s, b->seed + P.pix_off,
                           8u * (uint32_t)(P.site_off + P.cap), (long)P.W * P.H);
        hipLaunchKernelGGL(k_fill_u32, dim3(256), dim3(256), 0, s, b->first + P.pix_off,
                           8u * (uint32_t)(P.site_off + P.cap), (long)P.W * P.H);
    }
    (void)hipMemcpyAsync(b->d_site_patch, site_patch.data(), sizeof(int) * soff,
                         hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(b->sites, sent.data(), sizeof(float2) * soff, hipMemcpyHostToDevice, s);
    (void)hipMemsetAsync(b->rho, 0, sizeof(float) * pix, s);
    (void)hipMemsetAsync(b->fixed, 0, soff, s);
    (void)hipMemsetAsync(b->slots, 0, sizeof(Slot) * 9 * (size_t)soff, s);
    (void)hipMemsetAsync(b->ovf_head, 0xff, sizeof(int) * soff, s);
    (void)hipMemsetAsync(b->hist_len, 0, sizeof(int) * num_patches, s);
    (void)hipMemsetAsync(b->dflags, 0, sizeof(int) * NUM_FLAGS, s);
    int st[NUM_STATE] = {1, 0, 0, 0};
    (void)hipMemcpyAsync(b->dstate, st, sizeof st, hipMemcpyHostToDevice, s);
    if (hipStreamSynchronize(s) != hipSuccess) {
        delete b;
        return fail(CVT_ERR_HIP, "batch initialisation");
    }
    *out = reinterpret_cast<cvt_batch *>(b);
    return CVT_OK;
}

int cvt_batch_destroy(cvt_batch *bh)
{
    clear_error();
    if (!bh) return CVT_OK;
    Batch *b = reinterpret_cast<Batch *>(bh);
    (void)hipStreamSynchronize(b->stream);
    delete b;
    return CVT_OK;
}

int cvt_batch_set_density(cvt_batch *bh, int patch, const float *density)
{
    BCHK(bh);
    Batch *b = reinterpret_cast<Batch *>(bh);
    if (patch < 0 || patch >= b->np || !density) return fail(CVT_ERR_ARG, "patch %d", patch);
    BatchPatch &P = b->hp[patch];
    const long n = (long)P.W * P.H;
    if (hipMemcpyAsync(b->rho + P.pix_off, density, sizeof(float) * n, hipMemcpyDefault, b->stream) !=
        hipSuccess)
        return fail(CVT_ERR_HIP, "density upload");
    // validation and the power-of-two scale, as in cvt_set_density
    std::vector<float> h(n);
    if (hipMemcpyAsync(h.data(), b->rho + P.pix_off, sizeof(float) * n, hipMemcpyDeviceToHost,
                       b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return fail(CVT_ERR_HIP, "density check");
    float mx = 0.0f;
    for (float v : h) {
        if (!(v >= 0.0f) || v == __builtin_inff()) return fail(CVT_ERR_ARG, "bad density value");
        mx = std::max(mx, v);
    }
    int ex = 0;
    if (mx > 0.0f) frexp((double)mx, &ex);
    P.rho_scale = (float)ldexp(1.0, -ex);
    b->have_density[patch] = 1;
    return b->upload_patches();
}

int cvt_batch_set_sites(cvt_batch *bh, int patch, const float *xy, const uint8_t *fixed,
                        int num_sites)
{
    BCHK(bh);
    Batch *b = reinterpret_cast<Batch *>(bh);
    if (patch < 0 || patch >= b->np) return fail(CVT_ERR_ARG, "patch %d", patch);
    BatchPatch &P = b->hp[patch];
    if (num_sites < 0 || num_sites > P.cap || (num_sites > 0 && !xy))
        return fail(CVT_ERR_ARG, "patch %d: %d sites, capacity %d", patch, num_sites, P.cap);
    std::vector<float> h(2 * (size_t)num_sites);
    if (num_sites > 0 && hipMemcpy(h.data(), xy, sizeof(float) * 2 * num_sites, hipMemcpyDefault) !=
                             hipSuccess)
        return fail(CVT_ERR_HIP, "site upload");
    for (int i = 0; i < num_sites; ++i) {
        const float x = h[2 * i], y = h[2 * i + 1];
        if (!(x >= -(float)P.W && x <= 2.0f * P.W && y >= -(float)P.H && y <= 2.0f * P.H))
            return fail(CVT_ERR_ARG, "patch %d: site %d outside [-W, 2W] x [-H, 2H]", patch, i);
    }
    // the seed maps must be empty before seeding: clear the whole batch's
    if (b->seeds_valid) {
        for (int p = 0; p < b->np; ++p) {
            const BatchPatch &Q = b->hp[p];
            hipLaunchKernelGGL(k_fill_u32, dim3(256), dim3(256), 0, b->stream, b->seed + Q.pix_off,
                               8u * (uint32_t)(Q.site_off + Q.cap), (long)Q.W * Q.H);
        }
        b->seeds_valid = false;
    }
    if (num_sites > 0) {
        (void)hipMemcpyAsync(b->sites + P.site_off, h.data(), sizeof(float2) * num_sites,
                             hipMemcpyHostToDevice, b->stream);
        if (fixed)
            (void)hipMemcpyAsync(b->fixed + P.site_off, fixed, num_sites, hipMemcpyDefault, b->stream);
        else
            (void)hipMemsetAsync(b->fixed + P.site_off, 0, num_sites, b->stream);
    }
    if (hipStreamSynchronize(b->stream) != hipSuccess) return fail(CVT_ERR_HIP, "site upload");
    if (P.S != num_sites) b->dirty = true;
    P.S = num_sites;
    P.tile = batch_tile_for(P.W, P.H, std::max(num_sites, 1));
    b->dirty = true;
    b->labels_fresh = false;
    return CVT_OK;
}

static int batch_prepare(Batch *b)
{
    for (int p = 0; p < b->np; ++p)
        if (!b->have_density[p]) return fail(CVT_ERR_STATE, "patch %d has no density", p);
    if (b->dirty) {
        int rc = b->plan();
        if (rc) return rc;
    }
    if (!b->seeds_valid) {
        for (int p = 0; p < b->np; ++p)
            if (b->hp[p].S > 0)
                hipLaunchKernelGGL(k_seed_batch, dim3((b->hp[p].S + 255) / 256), dim3(256), 0,
                                   b->stream, b->d_patch, p, b->sites, b->seed);
        b->seeds_valid = true;
    }
    return CVT_OK;
}

int cvt_batch_lloyd(cvt_batch *bh, int iterations, float omega)
{
    BCHK(bh);
    Batch *b = reinterpret_cast<Batch *>(bh);
    if (iterations < 0 || !(omega > 0.0f && omega <= 2.0f)) return fail(CVT_ERR_ARG, "arguments");
    int rc = batch_prepare(b);
    if (rc || iterations == 0) return rc;
    if (!b->gexec || b->graph_omega != omega) {
        if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
        if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
        if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return fail(CVT_ERR_HIP, "capture");
        rc = b->enqueue_iteration(omega);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess || hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0) != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(CVT_ERR_HIP, "graph");
        }
        b->graph = g;
        b->graph_omega = omega;
    }
    for (int k = 0; k < iterations; ++k)
        if (hipGraphLaunch(b->gexec, b->stream) != hipSuccess) return fail(CVT_ERR_HIP, "launch");
    b->labels_fresh = true;
    return CVT_OK;
}

int cvt_batch_voronoi(cvt_batch *bh)
{
    BCHK(bh);
    Batch *b = reinterpret_cast<Batch *>(bh);
    int rc = batch_prepare(b);
    if (rc) return rc;
    b->enqueue_jfa();
    b->seeds_valid = false;
    b->labels_fresh = true;
    return hipGetLastError() == hipSuccess ? CVT_OK : fail(CVT_ERR_HIP, "voronoi");
}

int cvt_batch_synchronize(cvt_batch *bh)
{
    BCHK(bh);
    Batch *b = reinterpret_cast<Batch *>(bh);
    int f[NUM_FLAGS];
    if (hipMemcpyAsync(f, b->dflags, sizeof f, hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return fail(CVT_ERR_HIP, "synchronize");
    if (f[FLAG_FULL])
        return fail(CVT_ERR_CAPACITY, "a moment tile met more than %d sites", b->hash_cap);
    return CVT_OK;
}

int cvt_batch_get_sites(cvt_batch *bh, int patch, float *xy)
{
    BCHK(bh);
    Batch *b = reinterpret_cast<Batch *>(bh);
    if (patch < 0 || patch >= b->np || !xy) return fail(CVT_ERR_ARG, "patch %d", patch);
    const BatchPatch &P = b->hp[patch];
    if (P.S == 0) return CVT_OK;
    if (hipMemcpyAsync(xy, b->sites + P.site_off, sizeof(float2) * P.S, hipMemcpyDefault, b->stream) !=
            hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return fail(CVT_ERR_HIP, "site download");
    return cvt_batch_synchronize(bh);
}

int cvt_batch_get_labels(cvt_batch *bh, int patch, int32_t *labels)
{
    BCHK(bh);
    Batch *b = reinterpret_cast<Batch *>(bh);
    if (patch < 0 || patch >= b->np || !labels) return fail(CVT_ERR_ARG, "patch %d", patch);
    if (!b->labels_fresh) return fail(CVT_ERR_STATE, "no label map yet");
    const BatchPatch &P = b->hp[patch];
    const long n = (long)P.W * P.H;
    const uint32_t *lab = ((P.passes - 1) & 1 ? b->lab1 : b->lab0) + P.pix_off;
    int32_t *tmp = nullptr;
    if (hipMalloc(&tmp, sizeof(int32_t) * n) != hipSuccess) return fail(CVT_ERR_NOMEM, "labels");
    hipLaunchKernelGGL(k_export_batch, dim3(1024), dim3(256), 0, b->stream, lab, tmp, n,
                       8u * (uint32_t)P.site_off, 8u * (uint32_t)(P.site_off + P.cap));
    int rc = CVT_OK;
    if (hipMemcpyAsync(labels, tmp, sizeof(int32_t) * n, hipMemcpyDefault, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        rc = fail(CVT_ERR_HIP, "label download");
    (void)hipFree(tmp);
    return rc;
}

int cvt_batch_history(cvt_batch *bh, int patch, double *energy_gradnorm, int n)
{
    BCHK(bh);
    Batch *b = reinterpret_cast<Batch *>(bh);
    if (patch < 0 || patch >= b->np) return fail(CVT_ERR_ARG, "patch %d", patch);
    int len = 0;
    if (hipMemcpyAsync(&len, b->hist_len + patch, sizeof(int), hipMemcpyDeviceToHost, b->stream) !=
            hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return fail(CVT_ERR_HIP, "history");
    len = std::min(len, b->hist_cap);
    if (!energy_gradnorm || n <= 0) return len;
    const int k = std::min(n, len);
    if (k > 0 && (hipMemcpyAsync(energy_gradnorm, b->hist + 2L * patch * b->hist_cap,
                                 sizeof(double) * 2 * k, hipMemcpyDefault, b->stream) != hipSuccess ||
                  hipStreamSynchronize(b->stream) != hipSuccess))
        return fail(CVT_ERR_HIP, "history");
    return k;
}

int cvt_batch_set_tile(cvt_batch *bh, int patch, int tile)
{
    BCHK(bh);
    Batch *b = reinterpret_cast<Batch *>(bh);
    if (patch < 0 || patch >= b->np) return fail(CVT_ERR_ARG, "patch %d", patch);
    if (tile != 64 && tile != 128 && tile != 256) return fail(CVT_ERR_ARG, "tile %d", tile);
    b->hp[patch].tile = tile;
    b->dirty = true;
    return CVT_OK;
}

int cvt_batch_get_tile(cvt_batch *bh, int patch)
{
    BCHK(bh);
    Batch *b = reinterpret_cast<Batch *>(bh);
    if (patch < 0 || patch >= b->np) return fail(CVT_ERR_ARG, "patch %d", patch);
    return b->hp[patch].tile;
}

int cvt_batch_set_stream(cvt_batch *bh, void *stream)
{
    BCHK(bh);
    Batch *b = reinterpret_cast<Batch *>(bh);
    (void)hipStreamSynchronize(b->stream);
    // the caller's stream (NULL: the batch's own); the graph is rebuilt on it
    b->stream = stream ? (hipStream_t)stream : b->own_stream;
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    return CVT_OK;
}

}  // extern "C"
