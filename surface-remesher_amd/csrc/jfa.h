// jfa.h -- launcher of the JFA pass kernel (jfa.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cvt_types.h"

namespace cvt {

// Launch shape of a JFA pass for one grid: the quad kernel (4 pixels per
// lane, persistent workgroups over runs or bands of rows) when W % 4 == 0,
// with the site table in a static LDS array of lds_entries float2 (16, 40 or
// 160 KB) when it fits, or, for larger tables, lds_entries = -20480: the
// first 20478 entries and a zero entry in LDS, the rest gathered from global
// memory;
// otherwise one pixel per lane.
// runs_mask < 0 (JFA_AUTO): runs where every wave of the launch gets one,
// bands of rows elsewhere; otherwise bit k set = runs at step 2^k, clear =
// bands (the labels do not depend on it).
enum { JFA_AUTO = -1 };
struct JfaPlan {
    bool quad = false, lds = false;
    int grid = 0, threads = 0, table_f4 = 0, lds_entries = 0;
    size_t lds_bytes = 0;
    int runs_mask = JFA_AUTO, blocks_per_cu = 0;
};
// blocks_per_cu > 0 replaces the built-in grid (four 256-thread workgroups
// per CU for tables of up to 5120 entries, one 1024-thread workgroup per CU
// above); runs_mask as in JfaPlan.
JfaPlan plan_jfa(int W, int H, int capS, int runs_mask = JFA_AUTO, int blocks_per_cu = 0);
// the steps 2^k that take runs under this plan (bit k)
int jfa_runs_mask(const JfaPlan &p, int W, int H);

// One JFA pass over a W x H map of byte-offset labels (8 * site index, `none`
// for no site): out(p) = best of the 9 labels at p + step * {-1, 0, 1}^2.
// If clear is non-null it is reset to `none`: with nclear > 0 on quad plans
// only where it can hold entries, the seed pixels q of sites [0, nclear)
// (rstep == 0) or q + rstep * {-1, 0, 1}^2 (rstep > 0); otherwise pixel by
// pixel.  The site table must hold capS + 1 entries rounded up to an even
// count.
// A non-null `stop` makes the pass write nothing while *stop is set (it is
// read with the pass's first loads).
void launch_jfa_pass(hipStream_t st, const JfaPlan &plan, const uint32_t *in, uint32_t *out,
                     const float2 *sites, int W, int H, int step, uint32_t *clear, uint32_t none,
                     int nclear = 0, int rstep = 0, const int *stop = nullptr);

// The last five passes (steps 8, 4, 2, 1, 1) fused into one launch over
// 32 x 32 tiles with a 16-pixel halo staged in LDS, for plans whose site
// table is staged whole in a 16 or 40 KB array and JFAs of eight passes or
// more: out = those five passes applied to `in` (the step-16 map), bit for
// bit.  `stop` as in launch_jfa_pass.
bool jfa_tail_ok(const JfaPlan &plan, int passes);
void launch_jfa_tail(hipStream_t st, const JfaPlan &plan, const uint32_t *in, uint32_t *out,
                     const float2 *sites, int W, int H, const int *stop = nullptr);

// The first pass from a seed map (labels only at the seed pixels of sites
// [0, S), none elsewhere), computed from the seeds into a map `out` that
// holds none everywhere: the same output as launch_jfa_pass with in = seed.
// Only the receivers q + step * {-1, 0, 1}^2 of the seed pixels q are
// written.  With a non-null fa, one more workgroup runs the previous site
// update's deferred finalize (cvt_finalize.h) if one is pending in dstate.
void launch_jfa_seed_pass(hipStream_t st, const uint32_t *seed, uint32_t *out,
                          const float2 *sites, int S, int W, int H, int step,
                          const int *stop = nullptr, const double *blockpart = nullptr,
                          int *dstate = nullptr, const FinalArgs *fa = nullptr);

// A patch of a batch, as the batched kernels see it.  Maps are concatenated
// (pix_off), site tables too (site_off; entry site_off + cap is the patch's
// sentinel, so `none` = 8 * (site_off + cap)); labels are byte offsets into
// the concatenated table.
struct BatchPatch {
    int W, H;
    long pix_off;
    int site_off, cap, S;      // table base, capacity (sentinel index), live sites
    int passes, first_step;    // JFA schedule
    int tile, tiles_x, tiles_y, tile_off;   // moment tiles
    int blk_off, nblk;         // update blocks [blk_off, blk_off + nblk)
    float rho_scale;           // 2^-k
    float pad;
};
// a workgroup's share of a JFA pass: part b of the nb parts of one patch,
// either its rows [y0, y1) (bands) or the b-th share of the pass's runs
struct BatchUnit {
    int patch, y0, y1, b, nb, pad[3];
};

// Pass `pass` (0 .. pmax-1) of the batched JFA over all units; lds_bytes = 0
// gathers sites from global memory, otherwise every patch's table (at most
// lds_bytes) is staged in LDS.  One launch per step class in `modes` (bit 0:
// step >= 4, bit 1: step 1, bit 2: step 2); a unit runs in the launch of its
// patch's class.
// A patch of three passes or more starts with a scatter of its seeds into
// `first` (launch_jfa_batch_seeds, at the batched pass where its own pass 0
// falls; its units skip that pass), reads `first` in its pass 1, which
// resets the seed map at its seed pixels, and resets `first` at the seeds'
// receivers in pass 2: both maps hold their patch's `none` between JFAs.
void launch_jfa_batch(hipStream_t st, const BatchPatch *patches, const BatchUnit *units, int nunits,
                      uint32_t *seed, uint32_t *first, uint32_t *lab0, uint32_t *lab1,
                      const float2 *sites, int pass, int pmax, size_t lds_bytes, unsigned modes);
// site_patch[i]: the patch owning site slot i of the concatenated table (-1
// for padding and sentinels)
void launch_jfa_batch_seeds(hipStream_t st, const BatchPatch *patches, const int *site_patch,
                            int total_sites, const uint32_t *seed, uint32_t *first,
                            const float2 *sites, int pass, int pmax);

}  // namespace cvt
