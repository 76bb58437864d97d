// cvt_types.h -- types shared by the engine's kernels and its host code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define CVT_NONE 0x7fffffff

// Fixed-point moment sums of one (tile, site) pair: W = sum rho', DX = sum
// rho' (px - sx), DY = sum rho' (py - sy), E = sum rho' |p - s|^2, scaled by
// 2^MOM_FW, 2^MOM_FX, 2^MOM_FX, 2^MOM_FE.  rho' = rho * 2^-k <= 1, a tile has
// at most 256^2 = 2^16 pixels and |p - s| < 2^15 (sites stay inside
// [-W, 2W] x [-H, 2H], W, H <= 8192), so each tile sum fits in int64; the
// per-site totals are summed in 128 bits.  Integer sums are exact, hence
// independent of order and grouping.
#define MOM_FW 46
#define MOM_FX 31
#define MOM_FE 15

struct __align__(8) Slot {
    long long w, dx, dy, e;
    int epoch;
    int pad;
};

struct __align__(8) OvfRec {
    long long w, dx, dy, e;
    int next;   // next record of the same site, -1 ends the chain
    int pad;
};

// device flags
enum { FLAG_FULL = 1, FLAG_BAD = 2, NUM_FLAGS = 4 };
// device state words
// (ST_DONE counts the finished blocks of a site-update launch; the last block
// to finish resets it.  ST_STOP, when set, makes every kernel of a Lloyd
// iteration return before its first write: cvt_lloyd_until arms the stop rule
// (ST_UACTIVE, the history index it started at ST_UBASE, the window ST_UWIN)
// and the iteration's finalize raises ST_STOP when the rule holds.  ST_PEND:
// the number of blocks whose energy partials a deferred finalize still has to
// add, 0 when none is pending.)
enum {
    ST_EPOCH = 0, ST_HIST = 1, ST_OVFN = 2, ST_NSITES = 3, ST_DONE = 4,
    ST_STOP = 5, ST_UBASE = 6, ST_UWIN = 7, ST_UACTIVE = 8, ST_PEND = 9, NUM_STATE = 10
};

// what the finalize of a site update does with the energy partials
struct FinalArgs {
    double *hist;          // history (E, |grad E|) [hist_cap][2]
    int hist_cap;
    int record;            // append to the history
    double e_scale, g_scale;
    double *out_eg;        // optional (E, |grad E|) copy
    const double *until;   // stop rule (relative decrease, target energy), read when ST_UACTIVE
    int defer;             // leave the finalize to the next seed pass (or a finalize launch)
};


struct CompatArgs {
    const uint8_t *mask;   // [H][W] fixed-site pixels (may be null)
    const float *rho;      // [H][W]
    int *tgt;              // [S] target pixel index per site
    uint32_t *owner;       // [H][W] lowest site index landing on a pixel, CVT_NONE else
};
