// unet.hip -- the learned warm start: a 4-level convolutional U-Net (32 / 64
// / 96 / 128 channels, two 3x3 convs + ReLU per level, 2x2 average pooling,
// nearest upsampling with skip concatenation, 1x1 head to a 2-channel
// displacement field), BF16 weights and activations with FP32 accumulation on
// MFMA, FP32 master weights and Adam.  Input: an R x R grid (R = 256 for the
// warm start) of channel 0 = density box mean / mean, channel 1 = initial
// site count per cell * R^2 / S averaged over a window of about three sites
// (half-width round(R / sqrt(S)) - 1 cells), channel 2 = 1 where the cell
// has density.
// Output: the displacement of a site (in units of the domain side) sampled
// bilinearly from the field at the site's position.
//
// No reference counterpart: the reference starts Lloyd from random sites
// (source/gcvt.h:76-104, :124-139).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <vector>

#include "../../include/cvt.h"
#include "cvt_engine.h"
#include "nn_ops.h"

namespace cvt {

static const int IN_C = 32;   // input channels padded to the conv's K granularity
static const int HEAD_C = NN_HEAD_C;

struct ConvSpec {
    int cin, cout, level;     // level: resolution R >> level
};
// L0..L13 in execution order
static const ConvSpec SPEC[14] = {
    {IN_C, 32, 0}, {32, 32, 0},        // enc1 (skip1 = L1)
    {32, 64, 1},   {64, 64, 1},        // enc2 (skip2 = L3)
    {64, 96, 2},   {96, 96, 2},        // enc3 (skip3 = L5)
    {96, 128, 3},  {128, 128, 3},      // bottleneck
    {224, 96, 2},  {96, 96, 2},        // dec3: up(L7) ++ skip3
    {160, 64, 1},  {64, 64, 1},        // dec2: up(L9) ++ skip2
    {96, 32, 0},   {32, 32, 0},        // dec1: up(L11) ++ skip1
};

struct UNet {
    int R = 256, maxb = 1;
    hipStream_t stream = nullptr, own = nullptr;
    // parameters: per conv w [9][cout][cin] then b [cout]; head wh [2][32], bh [2]
    long nparam = 0;
    long woff[14], boff[14], whoff = 0, bhoff = 0;
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    uint16_t *wf[14], *wb[14];
    int adam_t = 0;
    // activations (bf16) and their gradients
    uint16_t *x0 = nullptr;                 // input (maxb, R, R, IN_C)
    uint16_t *a[14], *ga[14];
    uint16_t *pool[3], *gpool[3];           // p1 (R/2, 32), p2 (R/4, 64), p3 (R/8, 96)
    uint16_t *up[3], *gup[3];               // u3 (R/4, 224), u2 (R/2, 160), u1 (R, 96)
    float *field = nullptr, *gfield = nullptr;
    float *work = nullptr;
    long work_n = 0;
    float *loss_part = nullptr;
    int loss_blocks = 0;
    float *uv_buf = nullptr, *tgt_buf = nullptr, *count = nullptr, *xy_buf = nullptr;
    int site_cap = 0;
    long site_cap_loss = 0;
    double *mean_part = nullptr;            // NN_SUM_BLOCKS partial sums
    float *inv_mean = nullptr;              // 1 / mean density of the last prepare
    float *rho_buf = nullptr;               // device copy of a host density
    long rho_cap = 0;

    ~UNet();
    int res(int level) const { return R >> level; }
    long npix(int level, int nb) const { return (long)nb * res(level) * res(level); }
    void upload_bf16();
    void forward(int nb);
    void backward(int nb);
};

UNet::~UNet()
{
    std::vector<void *> bufs = {p, g, m1, m2, x0, field, gfield, work, loss_part, uv_buf, tgt_buf,
                                count, xy_buf, mean_part, inv_mean, rho_buf};
    for (int l = 0; l < 14; ++l) {
        bufs.push_back(wf[l]); bufs.push_back(wb[l]); bufs.push_back(a[l]); bufs.push_back(ga[l]);
    }
    for (int k = 0; k < 3; ++k) {
        bufs.push_back(pool[k]); bufs.push_back(gpool[k]); bufs.push_back(up[k]); bufs.push_back(gup[k]);
    }
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (own) (void)hipStreamDestroy(own);
}

void UNet::upload_bf16()
{
    for (int l = 0; l < 14; ++l)
        nn_weights_bf16(stream, p + woff[l], SPEC[l].cin, SPEC[l].cout, wf[l], wb[l]);
}

void UNet::forward(int nb)
{
    const int R0 = res(0), R1 = res(1), R2 = res(2), R3 = res(3);
    auto conv = [&](int l, const uint16_t *in) {
        const int r = res(SPEC[l].level);
        nn_conv3x3(stream, in, wf[l], p + boff[l], a[l], nb, r, r, SPEC[l].cin, SPEC[l].cout, 1);
    };
    conv(0, x0);
    conv(1, a[0]);
    nn_avgpool2(stream, a[1], pool[0], nb, R0, R0, 32);
    conv(2, pool[0]);
    conv(3, a[2]);
    nn_avgpool2(stream, a[3], pool[1], nb, R1, R1, 64);
    conv(4, pool[1]);
    conv(5, a[4]);
    nn_avgpool2(stream, a[5], pool[2], nb, R2, R2, 96);
    conv(6, pool[2]);
    conv(7, a[6]);
    nn_up_concat(stream, a[7], a[5], up[0], nb, R3, R3, 128, 96);
    conv(8, up[0]);
    conv(9, a[8]);
    nn_up_concat(stream, a[9], a[3], up[1], nb, R2, R2, 96, 64);
    conv(10, up[1]);
    conv(11, a[10]);
    nn_up_concat(stream, a[11], a[1], up[2], nb, R1, R1, 64, 32);
    conv(12, up[2]);
    conv(13, a[12]);
    nn_head_fwd(stream, a[13], p + whoff, p + bhoff, field, npix(0, nb));
}

void UNet::backward(int nb)
{
    const int R0 = res(0), R1 = res(1), R2 = res(2), R3 = res(3);
    // head: gfield -> ga[13], dwh, dbh
    nn_head_bwd(stream, a[13], gfield, p + whoff, ga[13], g + whoff, work, npix(0, nb));
    // one conv layer: mask the output gradient by ReLU' and take the bias
    // gradient (one pass), the weight gradient, then the input gradient into
    // gin (if any)
    auto layer = [&](int l, const uint16_t *in, uint16_t *gin) {
        const ConvSpec &c = SPEC[l];
        const int r = res(c.level);
        const long m = npix(c.level, nb);
        nn_colsum(stream, ga[l], m, c.cout, g + boff[l], work, a[l]);
        nn_conv3x3_dw(stream, in, ga[l], g + woff[l], work, nb, r, r, c.cin, c.cout, 0);
        if (gin) nn_conv3x3(stream, ga[l], wb[l], nullptr, gin, nb, r, r, c.cout, c.cin, 0);
    };
    layer(13, a[12], ga[12]);
    layer(12, up[2], gup[2]);
    nn_up_concat_bwd(stream, gup[2], ga[11], ga[1], nb, R1, R1, 64, 32, 0);
    layer(11, a[10], ga[10]);
    layer(10, up[1], gup[1]);
    nn_up_concat_bwd(stream, gup[1], ga[9], ga[3], nb, R2, R2, 96, 64, 0);
    layer(9, a[8], ga[8]);
    layer(8, up[0], gup[0]);
    nn_up_concat_bwd(stream, gup[0], ga[7], ga[5], nb, R3, R3, 128, 96, 0);
    layer(7, a[6], ga[6]);
    layer(6, pool[2], gpool[2]);
    nn_avgpool2_bwd(stream, gpool[2], ga[5], nb, R2, R2, 96, 1);
    layer(5, a[4], ga[4]);
    layer(4, pool[1], gpool[1]);
    nn_avgpool2_bwd(stream, gpool[1], ga[3], nb, R1, R1, 64, 1);
    layer(3, a[2], ga[2]);
    layer(2, pool[0], gpool[0]);
    nn_avgpool2_bwd(stream, gpool[0], ga[1], nb, R0, R0, 32, 1);
    layer(1, a[0], ga[0]);
    layer(0, x0, nullptr);
}

}  // namespace cvt

using namespace cvt;

#define UCHK(n) \
    do { if (!(n)) return fail(CVT_ERR_ARG, "null network"); clear_error(); } while (0)

extern "C" {

int cvt_unet_create(cvt_unet **out, int R, int max_batch, unsigned long long seed)
{
    clear_error();
    if (!out || R < 8 || R > 1024 || (R % 8) != 0 || max_batch < 1 || max_batch > 256)
        return fail(CVT_ERR_ARG, "R = %d (multiple of 8 in [8, 1024]), batch %d", R, max_batch);
    *out = nullptr;
    UNet *u = new UNet();
    memset(u->wf, 0, sizeof u->wf); memset(u->wb, 0, sizeof u->wb);
    memset(u->a, 0, sizeof u->a); memset(u->ga, 0, sizeof u->ga);
    memset(u->pool, 0, sizeof u->pool); memset(u->gpool, 0, sizeof u->gpool);
    memset(u->up, 0, sizeof u->up); memset(u->gup, 0, sizeof u->gup);
    u->R = R;
    u->maxb = max_batch;
    long off = 0;
    for (int l = 0; l < 14; ++l) {
        u->woff[l] = off; off += 9L * SPEC[l].cin * SPEC[l].cout;
        u->boff[l] = off; off += SPEC[l].cout;
    }
    u->whoff = off; off += 2 * HEAD_C;
    u->bhoff = off; off += 2;
    u->nparam = off;
    const int B = max_batch;
    bool ok = hipStreamCreateWithFlags(&u->own, hipStreamNonBlocking) == hipSuccess;
    u->stream = u->own;
    auto alloc = [&](void **p, size_t bytes) {
        if (ok && hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) ok = false;
    };
    alloc((void **)&u->p, sizeof(float) * off);
    alloc((void **)&u->g, sizeof(float) * off);
    alloc((void **)&u->m1, sizeof(float) * off);
    alloc((void **)&u->m2, sizeof(float) * off);
    for (int l = 0; l < 14; ++l) {
        const long wn = 9L * SPEC[l].cin * SPEC[l].cout;
        alloc((void **)&u->wf[l], 2 * wn);
        alloc((void **)&u->wb[l], 2 * wn);
        const long an = u->npix(SPEC[l].level, B) * SPEC[l].cout;
        alloc((void **)&u->a[l], 2 * an);
        alloc((void **)&u->ga[l], 2 * an);
    }
    alloc((void **)&u->x0, 2 * u->npix(0, B) * IN_C);
    const int pc[3] = {32, 64, 96}, uc[3] = {224, 160, 96}, ul[3] = {2, 1, 0};
    for (int k = 0; k < 3; ++k) {
        alloc((void **)&u->pool[k], 2 * u->npix(k + 1, B) * pc[k]);
        alloc((void **)&u->gpool[k], 2 * u->npix(k + 1, B) * pc[k]);
        alloc((void **)&u->up[k], 2 * u->npix(ul[k], B) * uc[k]);
        alloc((void **)&u->gup[k], 2 * u->npix(ul[k], B) * uc[k]);
    }
    alloc((void **)&u->field, sizeof(float) * 2 * u->npix(0, B));
    alloc((void **)&u->gfield, sizeof(float) * 2 * u->npix(0, B));
    // split-K workspace: the largest conv weight gradient and the head partials
    long wn = (long)nn_head_splits(u->npix(0, B)) * (2 * HEAD_C + 2);
    for (int l = 0; l < 14; ++l)
        wn = std::max({wn,
                       (long)nn_dw_splits(u->npix(SPEC[l].level, B), SPEC[l].cin, SPEC[l].cout) * 9L *
                           SPEC[l].cin * SPEC[l].cout,
                       (long)nn_colsum_splits(u->npix(SPEC[l].level, B)) * SPEC[l].cout});
    u->work_n = wn;
    alloc((void **)&u->work, sizeof(float) * wn);
    alloc((void **)&u->count, sizeof(float) * R * R);
    alloc((void **)&u->mean_part, sizeof(double) * NN_SUM_BLOCKS);
    alloc((void **)&u->inv_mean, sizeof(float) * 4);
    if (!ok) {
        delete u;
        return fail(CVT_ERR_NOMEM, "U-Net buffers (R %d, batch %d)", R, max_batch);
    }
    // He-normal initialisation (fan in = 9 * real input channels), zero
    // biases, a small head so the first displacements are near zero
    std::vector<float> h(off, 0.0f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < 14; ++l) {
        const int cin_real = (l == 0) ? 3 : SPEC[l].cin;
        std::normal_distribution<float> nd(0.0f, sqrtf(2.0f / (9.0f * cin_real)));
        for (int tap = 0; tap < 9; ++tap)
            for (int co = 0; co < SPEC[l].cout; ++co)
                for (int ci = 0; ci < SPEC[l].cin; ++ci) {
                    const float v = nd(rng);
                    h[u->woff[l] + ((long)tap * SPEC[l].cout + co) * SPEC[l].cin + ci] =
                        (ci < cin_real) ? v : 0.0f;
                }
    }
    std::normal_distribution<float> hd(0.0f, 1e-3f);
    for (int k = 0; k < 2 * HEAD_C; ++k) h[u->whoff + k] = hd(rng);
    (void)hipMemcpyAsync(u->p, h.data(), sizeof(float) * off, hipMemcpyHostToDevice, u->stream);
    (void)hipMemsetAsync(u->g, 0, sizeof(float) * off, u->stream);
    (void)hipMemsetAsync(u->m1, 0, sizeof(float) * off, u->stream);
    (void)hipMemsetAsync(u->m2, 0, sizeof(float) * off, u->stream);
    (void)hipMemsetAsync(u->x0, 0, 2 * u->npix(0, B) * IN_C, u->stream);
    u->upload_bf16();
    if (hipStreamSynchronize(u->stream) != hipSuccess) {
        delete u;
        return fail(CVT_ERR_HIP, "U-Net initialisation");
    }
    *out = reinterpret_cast<cvt_unet *>(u);
    return CVT_OK;
}

int cvt_unet_destroy(cvt_unet *net)
{
    clear_error();
    if (!net) return CVT_OK;
    UNet *u = reinterpret_cast<UNet *>(net);
    (void)hipStreamSynchronize(u->stream);
    delete u;
    return CVT_OK;
}

int cvt_unet_set_stream(cvt_unet *net, void *stream)
{
    UCHK(net);
    UNet *u = reinterpret_cast<UNet *>(net);
    (void)hipStreamSynchronize(u->stream);
    u->stream = stream ? (hipStream_t)stream : u->own;
    return CVT_OK;
}

long cvt_unet_num_params(cvt_unet *net)
{
    UCHK(net);
    return reinterpret_cast<UNet *>(net)->nparam;
}

int cvt_unet_get_params(cvt_unet *net, float *out)
{
    UCHK(net);
    UNet *u = reinterpret_cast<UNet *>(net);
    if (!out) return fail(CVT_ERR_ARG, "null output");
    if (hipMemcpyAsync(out, u->p, sizeof(float) * u->nparam, hipMemcpyDefault, u->stream) != hipSuccess ||
        hipStreamSynchronize(u->stream) != hipSuccess)
        return fail(CVT_ERR_HIP, "parameter download");
    return CVT_OK;
}

int cvt_unet_set_params(cvt_unet *net, const float *in)
{
    UCHK(net);
    UNet *u = reinterpret_cast<UNet *>(net);
    if (!in) return fail(CVT_ERR_ARG, "null input");
    if (hipMemcpyAsync(u->p, in, sizeof(float) * u->nparam, hipMemcpyDefault, u->stream) != hipSuccess)
        return fail(CVT_ERR_HIP, "parameter upload");
    u->upload_bf16();
    u->adam_t = 0;
    (void)hipMemsetAsync(u->m1, 0, sizeof(float) * u->nparam, u->stream);
    (void)hipMemsetAsync(u->m2, 0, sizeof(float) * u->nparam, u->stream);
    return hipStreamSynchronize(u->stream) == hipSuccess ? CVT_OK : fail(CVT_ERR_HIP, "upload");
}

int cvt_unet_get_grads(cvt_unet *net, float *out)
{
    UCHK(net);
    UNet *u = reinterpret_cast<UNet *>(net);
    if (!out) return fail(CVT_ERR_ARG, "null output");
    if (hipMemcpyAsync(out, u->g, sizeof(float) * u->nparam, hipMemcpyDefault, u->stream) != hipSuccess ||
        hipStreamSynchronize(u->stream) != hipSuccess)
        return fail(CVT_ERR_HIP, "gradient download");
    return CVT_OK;
}

int cvt_unet_set_input(cvt_unet *net, const uint16_t *x, int nb)
{
    UCHK(net);
    UNet *u = reinterpret_cast<UNet *>(net);
    if (!x || nb < 1 || nb > u->maxb) return fail(CVT_ERR_ARG, "batch %d", nb);
    const size_t bytes = 2 * (size_t)u->npix(0, nb) * IN_C;
    if (hipMemcpyAsync(u->x0, x, bytes, hipMemcpyDefault, u->stream) != hipSuccess)
        return fail(CVT_ERR_HIP, "input upload");
    return CVT_OK;
}

static bool is_device_ptr(const void *p)
{
    hipPointerAttribute_t at;
    return hipPointerGetAttributes(&at, p) == hipSuccess && at.type == hipMemoryTypeDevice;
}

// input of `slot` (or into dst) from a device density and device sites, on
// stream st; nothing synchronises
static void prepare_device(UNet *u, hipStream_t st, const float *rho, int n, const float *xy, int S,
                           uint16_t *out)
{
    nn_density_inv_mean(st, rho, n, u->mean_part, u->inv_mean);
    nn_prepare(st, rho, n, u->inv_mean, xy, S, u->R, IN_C, out, u->count);
}

int cvt_unet_prepare(cvt_unet *net, int slot, const float *density, int n, const float *xy, int S,
                     uint16_t *dst)
{
    UCHK(net);
    UNet *u = reinterpret_cast<UNet *>(net);
    if (!density || n < u->R || (n % u->R) != 0 || S < 0 || (S > 0 && !xy))
        return fail(CVT_ERR_ARG, "density %d (a multiple of %d), %d sites", n, u->R, S);
    if (!dst && (slot < 0 || slot >= u->maxb)) return fail(CVT_ERR_ARG, "slot %d", slot);
    const long nn = (long)n * n;
    const float *rho = density;
    if (!is_device_ptr(density)) {
        if (nn > u->rho_cap) {
            if (u->rho_buf) (void)hipFree(u->rho_buf);
            u->rho_buf = nullptr;
            u->rho_cap = 0;
            if (hipMalloc(&u->rho_buf, sizeof(float) * nn) != hipSuccess)
                return fail(CVT_ERR_NOMEM, "density");
            u->rho_cap = nn;
        }
        if (hipMemcpyAsync(u->rho_buf, density, sizeof(float) * nn, hipMemcpyDefault, u->stream) !=
            hipSuccess)
            return fail(CVT_ERR_HIP, "density upload");
        rho = u->rho_buf;
    }
    if (S > u->site_cap) {
        if (u->xy_buf) (void)hipFree(u->xy_buf);
        u->xy_buf = nullptr;
        u->site_cap = 0;
        if (hipMalloc(&u->xy_buf, sizeof(float) * 2 * S) != hipSuccess)
            return fail(CVT_ERR_NOMEM, "sites");
        u->site_cap = S;
    }
    if (S > 0 &&
        hipMemcpyAsync(u->xy_buf, xy, sizeof(float) * 2 * S, hipMemcpyDefault, u->stream) != hipSuccess)
        return fail(CVT_ERR_HIP, "site upload");
    // kernels write device memory only: a caller's dst (host or device) is
    // filled by a copy from a device staging buffer
    const size_t bytes = 2 * (size_t)u->R * u->R * IN_C;
    uint16_t *stage = nullptr;
    if (dst && hipMalloc(&stage, bytes) != hipSuccess) return fail(CVT_ERR_NOMEM, "staging");
    uint16_t *out = dst ? stage : u->x0 + (size_t)slot * u->R * u->R * IN_C;
    prepare_device(u, u->stream, rho, n, u->xy_buf, S, out);
    hipError_t e = hipSuccess;
    if (dst) {
        (void)hipMemcpyAsync(dst, stage, bytes, hipMemcpyDefault, u->stream);
        e = hipStreamSynchronize(u->stream);
        (void)hipFree(stage);
    }
    // a host density or site array is read by the copies above before the
    // call returns only after they complete
    if (!dst && (rho != density || (S > 0 && !is_device_ptr(xy))))
        e = hipStreamSynchronize(u->stream);
    return e == hipSuccess ? CVT_OK : fail(CVT_ERR_HIP, "prepare");
}

int cvt_unet_warm_start(cvt_unet *net, cvt_ctx *ctx)
{
    UCHK(net);
    if (!ctx) return fail(CVT_ERR_ARG, "null context");
    UNet *u = reinterpret_cast<UNet *>(net);
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c->have_density) return fail(CVT_ERR_STATE, "warm start before cvt_set_density");
    if (c->W != c->H || c->W < u->R || c->W % u->R != 0)
        return fail(CVT_ERR_ARG, "warm start needs a square grid, a multiple of %d (got %d x %d)",
                    u->R, c->W, c->H);
    if (c->S == 0) return CVT_OK;
    const hipStream_t st = c->stream;
    // the seed map holds the old sites' seeds: clear them before moving
    if (c->seeds_valid) c->clear_seed_map();
    prepare_device(u, st, c->rho, c->W, reinterpret_cast<const float *>(c->sites), c->S, u->x0);
    const hipStream_t saved = u->stream;
    u->stream = st;                      // the forward runs on the engine's stream
    u->forward(1);
    u->stream = saved;
    nn_apply_disp(st, c->sites, c->fixed, c->S, u->field, u->R, c->rho, c->W);
    c->seeds_valid = false;
    c->labels_fresh = false;
    c->enqueue_seed();
    return hipGetLastError() == hipSuccess ? CVT_OK : fail(CVT_ERR_HIP, "warm start launch");
}

int cvt_unet_forward(cvt_unet *net, int nb)
{
    UCHK(net);
    UNet *u = reinterpret_cast<UNet *>(net);
    if (nb < 1 || nb > u->maxb) return fail(CVT_ERR_ARG, "batch %d", nb);
    u->forward(nb);
    return hipGetLastError() == hipSuccess ? CVT_OK : fail(CVT_ERR_HIP, "forward launch");
}

int cvt_unet_get_field(cvt_unet *net, float *out, int nb)
{
    UCHK(net);
    UNet *u = reinterpret_cast<UNet *>(net);
    if (!out || nb < 1 || nb > u->maxb) return fail(CVT_ERR_ARG, "batch %d", nb);
    if (hipMemcpyAsync(out, u->field, sizeof(float) * 2 * u->npix(0, nb), hipMemcpyDefault, u->stream) !=
            hipSuccess ||
        hipStreamSynchronize(u->stream) != hipSuccess)
        return fail(CVT_ERR_HIP, "field download");
    return CVT_OK;
}

// loss = mean over the nb * S sites of |field(uv) - target|^2, then the full
// backward pass into the parameter gradients.  uv / target: [nb][S][2].
int cvt_unet_loss_backward(cvt_unet *net, int nb, const float *uv, const float *target, int S,
                           float *loss)
{
    UCHK(net);
    UNet *u = reinterpret_cast<UNet *>(net);
    if (nb < 1 || nb > u->maxb || S < 1 || !uv || !target) return fail(CVT_ERR_ARG, "arguments");
    const long ns = (long)nb * S;
    const int blocks = (int)((ns + 255) / 256);
    if (blocks > u->loss_blocks) {
        if (u->loss_part) (void)hipFree(u->loss_part);
        if (hipMalloc(&u->loss_part, sizeof(float) * blocks) != hipSuccess) return fail(CVT_ERR_NOMEM, "loss");
        u->loss_blocks = blocks;
    }
    // uv and target may be host memory: the kernel reads device copies
    if (ns > u->site_cap_loss) {
        if (u->uv_buf) (void)hipFree(u->uv_buf);
        if (u->tgt_buf) (void)hipFree(u->tgt_buf);
        u->uv_buf = u->tgt_buf = nullptr;
        u->site_cap_loss = 0;
        if (hipMalloc(&u->uv_buf, sizeof(float) * 2 * ns) != hipSuccess ||
            hipMalloc(&u->tgt_buf, sizeof(float) * 2 * ns) != hipSuccess)
            return fail(CVT_ERR_NOMEM, "site buffers");
        u->site_cap_loss = ns;
    }
    if (hipMemcpyAsync(u->uv_buf, uv, sizeof(float) * 2 * ns, hipMemcpyDefault, u->stream) != hipSuccess ||
        hipMemcpyAsync(u->tgt_buf, target, sizeof(float) * 2 * ns, hipMemcpyDefault, u->stream) !=
            hipSuccess)
        return fail(CVT_ERR_HIP, "site upload");
    (void)hipMemsetAsync(u->gfield, 0, sizeof(float) * 2 * u->npix(0, nb), u->stream);
    nn_sample_loss(u->stream, u->field, u->R, u->uv_buf, u->tgt_buf, S, nb, 1.0f / (float)ns, nullptr,
                   u->gfield, u->loss_part);
    u->backward(nb);
    if (hipGetLastError() != hipSuccess) return fail(CVT_ERR_HIP, "backward launch");
    if (loss) {
        std::vector<float> lp(blocks);
        if (hipMemcpyAsync(lp.data(), u->loss_part, sizeof(float) * blocks, hipMemcpyDeviceToHost,
                           u->stream) != hipSuccess ||
            hipStreamSynchronize(u->stream) != hipSuccess)
            return fail(CVT_ERR_HIP, "loss readback");
        double s = 0.0;
        for (float v : lp) s += v;
        *loss = (float)s;
    }
    return CVT_OK;
}

int cvt_unet_adam(cvt_unet *net, float lr, float beta1, float beta2, float eps)
{
    UCHK(net);
    UNet *u = reinterpret_cast<UNet *>(net);
    u->adam_t += 1;
    nn_adam(u->stream, u->p, u->g, u->m1, u->m2, u->nparam, lr, beta1, beta2, eps, u->adam_t);
    u->upload_bf16();
    return hipGetLastError() == hipSuccess ? CVT_OK : fail(CVT_ERR_HIP, "adam");
}

// displacement (domain units) of S sites of batch slot `slot` from the last
// forward pass; uv = positions / domain side
int cvt_unet_sample(cvt_unet *net, int slot, const float *uv, int S, float *disp)
{
    UCHK(net);
    UNet *u = reinterpret_cast<UNet *>(net);
    if (slot < 0 || slot >= u->maxb || S < 0 || (S > 0 && (!uv || !disp)))
        return fail(CVT_ERR_ARG, "arguments");
    if (S == 0) return CVT_OK;
    float *d_uv = nullptr, *d_out = nullptr;
    if (hipMalloc(&d_uv, sizeof(float) * 2 * S) != hipSuccess ||
        hipMalloc(&d_out, sizeof(float) * 2 * S) != hipSuccess) {
        if (d_uv) (void)hipFree(d_uv);
        return fail(CVT_ERR_NOMEM, "sample buffers");
    }
    (void)hipMemcpyAsync(d_uv, uv, sizeof(float) * 2 * S, hipMemcpyDefault, u->stream);
    nn_sample_loss(u->stream, u->field + (size_t)slot * u->R * u->R * 2, u->R, d_uv, nullptr, S, 1, 0.0f,
                   d_out, nullptr, nullptr);
    int rc = CVT_OK;
    if (hipMemcpyAsync(disp, d_out, sizeof(float) * 2 * S, hipMemcpyDefault, u->stream) != hipSuccess ||
        hipStreamSynchronize(u->stream) != hipSuccess)
        rc = fail(CVT_ERR_HIP, "sample");
    (void)hipFree(d_uv);
    (void)hipFree(d_out);
    return rc;
}

// raw parameter blob: "CVTUNET1", R, nparam (int64), then nparam bf16 values
int cvt_unet_save(cvt_unet *net, const char *path)
{
    UCHK(net);
    UNet *u = reinterpret_cast<UNet *>(net);
    std::vector<float> h(u->nparam);
    int rc = cvt_unet_get_params(net, h.data());
    if (rc) return rc;
    FILE *f = fopen(path, "wb");
    if (!f) return fail(CVT_ERR_ARG, "cannot write %s", path);
    const long long hdr[2] = {u->R, u->nparam};
    fwrite("CVTUNET1", 1, 8, f);
    fwrite(hdr, sizeof hdr, 1, f);
    std::vector<uint16_t> b(u->nparam);
    for (long i = 0; i < u->nparam; ++i) {
        uint32_t x;
        memcpy(&x, &h[i], 4);
        b[i] = (uint16_t)((x + 0x7fffu + ((x >> 16) & 1u)) >> 16);
    }
    const size_t wr = fwrite(b.data(), 2, b.size(), f);
    fclose(f);
    return wr == b.size() ? CVT_OK : fail(CVT_ERR_ARG, "short write");
}

int cvt_unet_load(cvt_unet *net, const char *path)
{
    UCHK(net);
    UNet *u = reinterpret_cast<UNet *>(net);
    FILE *f = fopen(path, "rb");
    if (!f) return fail(CVT_ERR_ARG, "cannot read %s", path);
    char magic[8];
    long long hdr[2];
    if (fread(magic, 1, 8, f) != 8 || memcmp(magic, "CVTUNET1", 8) != 0 ||
        fread(hdr, sizeof hdr, 1, f) != 1 || hdr[1] != u->nparam) {
        fclose(f);
        return fail(CVT_ERR_ARG, "%s is not a weight file of this network", path);
    }
    std::vector<uint16_t> b(u->nparam);
    const size_t rd = fread(b.data(), 2, b.size(), f);
    fclose(f);
    if (rd != b.size()) return fail(CVT_ERR_ARG, "short read");
    std::vector<float> h(u->nparam);
    for (long i = 0; i < u->nparam; ++i) {
        const uint32_t x = (uint32_t)b[i] << 16;
        memcpy(&h[i], &x, 4);
    }
    return cvt_unet_set_params(net, h.data());
}

int cvt_unet_synchronize(cvt_unet *net)
{
    UCHK(net);
    UNet *u = reinterpret_cast<UNet *>(net);
    return hipStreamSynchronize(u->stream) == hipSuccess ? CVT_OK : fail(CVT_ERR_HIP, "synchronize");
}

}  // extern "C"
