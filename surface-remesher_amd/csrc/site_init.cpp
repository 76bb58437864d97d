// site_init.cpp -- the reference's host-side site initialisation behind
// gCVT (centroidalVoronoi, source/gcvt.h:124-139): putConstrains
// (source/gcvt.h:106-122) marks the constraint pixels as fixed sites, then
// randomPoints (source/gcvt.h:76-104) adds sites by rejection sampling with
// the KISS generator of source/gcvt.h:58-74.
//
// The reference's gcvt.h cannot be compiled on glibc: its `double random()`
// collides with POSIX `long random()`.  It was written for an LLP64 target
// (MSVC), where `unsigned long` is 32 bits, so every generator word wraps at
// 2^32 and random() divides by 2^32.  This file keeps those semantics on any
// host.  The reference never calls randinit, so its generator starts from
// four zero words; with z = w = jsr = 0 the multiply-with-carry and shift
// terms stay 0 and only the congruential term 69069 * j + 1234567 is live.
//
// The rejection loop is sequential: each draw advances the generator and
// each accepted site changes which pixels are free.  It runs on the host, as
// in the reference (O(num / acceptance rate) draws; a 2048^2 patch with
// 16384 sites takes about a millisecond).  Where the reference's loop would
// never end (fewer free pixels of positive density than sites asked for),
// this returns an error instead.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "../../include/cvt.h"
#include "cvt_engine.h"

namespace cvt {
namespace {

constexpr int16_t kMarker = -32768;   // source/gcvt.h:27

// Marsaglia's KISS: multiply-with-carry pair, 3-shift register and the
// congruential generator, combined as (MWC ^ CONG) + SHR3, all in 32 bits.
struct Kiss {
    uint32_t z, w, jsr, jcong;
    uint32_t next()
    {
        z = 36969u * (z & 0xffffu) + (z >> 16);
        w = 18000u * (w & 0xffffu) + (w >> 16);
        const uint32_t mwc = (z << 16) + w;
        jcong = 69069u * jcong + 1234567u;
        jsr ^= jsr << 17;
        jsr ^= jsr >> 13;
        jsr ^= jsr << 5;
        return (mwc ^ jcong) + jsr;
    }
    // [0, 1): the 32-bit draw over 2^32 (ULONG_MAX + 1 on the LLP64 target)
    double uniform() { return (double)next() / 4294967296.0; }
};

// True when p is device memory of the current device; a host pointer, or
// no HIP device at all, gives false (the call then works on the host
// buffers in place).
bool is_device_ptr(const void *p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice;
}

}  // namespace
}  // namespace cvt

using namespace cvt;

extern "C" int cvt_ref_init_sites(int16_t *voronoi, const float *density, const uint8_t *mask,
                                  int size, int num, uint32_t *rng_state, long long *draws)
{
    clear_error();
    if (!voronoi || !density) return fail(CVT_ERR_ARG, "null array");
    if (size < 1 || size > 8192) return fail(CVT_ERR_ARG, "size %d outside [1, 8192]", size);
    if (num < 0) return fail(CVT_ERR_ARG, "num %d < 0", num);
    const long np = (long)size * size;
    // work on host copies of device buffers, in place on host ones
    const bool dv = is_device_ptr(voronoi), dd = is_device_ptr(density),
               dm = mask && is_device_ptr(mask);
    std::vector<int16_t> vbuf;
    std::vector<float> rbuf;
    std::vector<uint8_t> mbuf;
    int16_t *vor = voronoi;
    const float *rho = density;
    const uint8_t *mk = mask;
    if (dd) {
        rbuf.resize(np);
        if (hipMemcpy(rbuf.data(), density, sizeof(float) * np, hipMemcpyDeviceToHost) != hipSuccess)
            return fail(CVT_ERR_HIP, "density download");
        rho = rbuf.data();
    }
    if (dm) {
        mbuf.resize(np);
        if (hipMemcpy(mbuf.data(), mask, np, hipMemcpyDeviceToHost) != hipSuccess)
            return fail(CVT_ERR_HIP, "mask download");
        mk = mbuf.data();
    }
    if (dv) {
        vbuf.resize(2 * np);
        vor = vbuf.data();
    }
    // the density must be a density (the reference's comparisons would accept
    // a NaN pixel unconditionally)
    for (long p = 0; p < np; ++p)
        if (!(rho[p] >= 0.0f) || std::isinf(rho[p]))
            return fail(CVT_ERR_ARG, "density holds a negative, NaN or infinite value");

    // putConstrains (source/gcvt.h:106-122): every pixel is written
    for (long p = 0; p < np; ++p) {
        const bool fixed = mk && mk[p];
        vor[2 * p] = fixed ? (int16_t)(p % size) : kMarker;
        vor[2 * p + 1] = fixed ? (int16_t)(p / size) : kMarker;
    }
    // randomPoints (source/gcvt.h:76-104): the acceptance bound is
    // mx = min(max rho, 100 * mean of the non-zero rho)
    double mx = 0, avg = 0, cnt = 0;
    long avail = 0;
    for (long p = 0; p < np; ++p) {
        if (rho[p] > mx) mx = rho[p];
        if (rho[p] != 0) { cnt += 1; avg += rho[p]; }
        if (rho[p] > 0.0f && vor[2 * p] == kMarker) ++avail;
    }
    if (num > avail)
        return fail(CVT_ERR_ARG,
                    "%d sites asked for, %ld free pixels with density > 0 (the reference's "
                    "rejection loop would not end)", num, avail);
    if (cnt > 0) mx = std::min(mx, avg / cnt * 100.);
    Kiss k{0, 0, 0, 0};
    if (rng_state) k = Kiss{rng_state[0], rng_state[1], rng_state[2], rng_state[3]};
    // every free pixel of positive density is accepted with probability >= its
    // density / mx per hit; the cap only guards against densities so skewed
    // that the loop would run for hours
    const long long cap = 1LL << 36;
    long long nd = 0;
    for (int i = 0; i < num; ++i) {
        int x, y;
        double z;
        do {
            if (nd >= cap) return fail(CVT_ERR_CAPACITY, "rejection sampling: %lld draws", nd);
            x = (int)(k.uniform() * size);
            y = (int)(k.uniform() * size);
            z = k.uniform() * mx;
            nd += 3;
        } while (vor[2 * ((long)y * size + x)] != kMarker || rho[(long)y * size + x] <= z);
        vor[2 * ((long)y * size + x)] = (int16_t)x;
        vor[2 * ((long)y * size + x) + 1] = (int16_t)y;
    }
    if (dv && hipMemcpy(voronoi, vor, sizeof(int16_t) * 2 * np, hipMemcpyHostToDevice) != hipSuccess)
        return fail(CVT_ERR_HIP, "label map upload");
    if (rng_state) {
        rng_state[0] = k.z; rng_state[1] = k.w; rng_state[2] = k.jsr; rng_state[3] = k.jcong;
    }
    if (draws) *draws = nd;
    return CVT_OK;
}
