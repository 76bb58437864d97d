// jfa.h -- launcher of the JFA pass kernel (jfa.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cvt {

// Launch shape of a JFA pass for one grid: the quad kernel (4 pixels per
// lane, persistent workgroups over runs or bands of rows) when W % 4 == 0,
// with the site table in a static LDS array of lds_entries float2 (16, 40 or
// 160 KB) when it fits, or, for larger tables, lds_entries = -20480: the
// first 20478 entries and a zero entry in LDS, the rest gathered from global
// memory;
// otherwise one pixel per lane.
struct JfaPlan {
    bool quad = false, lds = false;
    int grid = 0, threads = 0, table_f4 = 0, lds_entries = 0;
    size_t lds_bytes = 0;
};
JfaPlan plan_jfa(int W, int H, int capS);

// One JFA pass over a W x H map of byte-offset labels (8 * site index, `none`
// for no site): out(p) = best of the 9 labels at p + step * {-1, 0, 1}^2.
// If clear is non-null it is reset to `none`: with nclear > 0 on quad plans
// only where it can hold entries, the seed pixels q of sites [0, nclear)
// (rstep == 0) or q + rstep * {-1, 0, 1}^2 (rstep > 0); otherwise pixel by
// pixel.  The site table must hold capS + 1 entries rounded up to an even
// count.
// A non-null `stop` makes the pass return at entry while *stop is set.
void launch_jfa_pass(hipStream_t st, const JfaPlan &plan, const uint32_t *in, uint32_t *out,
                     const float2 *sites, int W, int H, int step, uint32_t *clear, uint32_t none,
                     int nclear = 0, int rstep = 0, const int *stop = nullptr);

// The first pass from a seed map (labels only at the seed pixels of sites
// [0, S), none elsewhere), computed from the seeds into a map `out` that
// holds none everywhere: the same output as launch_jfa_pass with in = seed.
// Only the receivers q + step * {-1, 0, 1}^2 of the seed pixels q are
// written.
void launch_jfa_seed_pass(hipStream_t st, const uint32_t *seed, uint32_t *out,
                          const float2 *sites, int S, int W, int H, int step,
                          const int *stop = nullptr);

// ---------------------------------------------------------------------------
// Persistent JFA: the whole JFA(+1) of one map in one launch.  Work items are
// handed out in a fixed order by a ticket counter: the seed scatter, then the
// runs of pass 1, the seed map's reset, the runs of pass 2, the first map's
// reset, then the runs of passes 3 .. np-1.  An item waits only for what it
// reads: a run of pass p for the rows of pass p-1 under its taps (per-row
// completion flags), pass 1 and the seed reset for the whole scatter, the
// first map's reset for the whole of pass 1.  Every item waits only on items
// with smaller tickets, which have been taken by running teams, so the order
// cannot deadlock whatever the number of resident workgroups.  Each pass
// writes its own map (maps[p]), so a map is never rewritten while a launch
// runs and no cache line of it can be stale.
// ---------------------------------------------------------------------------
enum { FLOW_SCATTER = 0, FLOW_PASS = 1, FLOW_SEED_RESET = 2, FLOW_FIRST_RESET = 3 };
// control words (ctl): ticket counter, teams done, launch epoch, scatter
// items done, pass-1 items done
enum { CTL_TICKET = 0, CTL_EXIT = 1, CTL_EPOCH = 2, CTL_SCATTER = 3, CTL_PASS1 = 4, CTL_WORDS = 8 };
#define CVT_FLOW_MAX_PHASES 20
// items of the scatter and reset phases cover this many (site, offset) pairs
// or sites (four per thread of a 256-thread team)
#define CVT_FLOW_CHUNK 1024
struct FlowPhase {
    int start, count;          // tickets [start, start + count)
    int kind, pass, step;
};
struct FlowArgs {
    uint32_t *maps;            // [np][H][W]: map 0 the scatter's, map p pass p's output
    uint32_t *seed;            // seed map [H][W]
    const float2 *sites;
    uint32_t *flags;           // [np][H]: epoch of the launch that completed row y of pass p
    uint32_t *ctl;             // CTL_WORDS control words
    int *dflags;               // FLAG_FLOW on a wait that ran out of spins
    const int *stop;           // non-null: the launch returns at entry while *stop is set
    int W, H, S, np, step0;
    int total, nph;            // tickets, phases
    int upi;                   // run groups (four rows of a chain, full width) per item
    int table_f4;
    uint32_t none;
    FlowPhase ph[CVT_FLOW_MAX_PHASES];
};
// Whether a plan runs as one persistent launch (quad kernel with the site
// table in LDS, whole or split; W a multiple of 32, so that no cache line
// holds pixels of two rows; three passes or more).
bool flow_eligible(const JfaPlan &p, int W, int H);
// The ticket order for S sites (fills a.total, a.nph, a.ph, a.upi).
void flow_phases(FlowArgs &a);
void launch_jfa_flow(hipStream_t st, const JfaPlan &p, const FlowArgs &a);

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
// a workgroup's share of a JFA pass: rows [y0, y1) of one patch
struct BatchUnit {
    int patch, y0, y1, pad;
};

// Pass `pass` (0 .. pmax-1) of the batched JFA over all units; lds_bytes = 0
// gathers sites from global memory, otherwise every patch's table (at most
// lds_bytes) is staged in LDS.  One launch per step class in `modes` (bit 0:
// step >= 4, bit 1: step 1, bit 2: step 2); a unit runs in the launch of its
// patch's class.
void launch_jfa_batch(hipStream_t st, const BatchPatch *patches, const BatchUnit *units, int nunits,
                      uint32_t *seed, uint32_t *lab0, uint32_t *lab1, const float2 *sites, int pass,
                      int pmax, size_t lds_bytes, unsigned modes);

}  // namespace cvt
