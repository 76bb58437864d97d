// jfa.h -- launcher of the JFA pass kernel (jfa.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cvt {

// Launch shape of a JFA pass for one grid: the quad kernel (4 pixels per
// lane, persistent row bands) when W % 4 == 0, with the site table in LDS
// when it fits; otherwise one pixel per lane.
struct JfaPlan {
    bool quad = false, lds = false;
    int grid = 0, threads = 0, table_f4 = 0;
    size_t lds_bytes = 0;
};
JfaPlan plan_jfa(int W, int H, int capS);

// One JFA pass over a W x H map of byte-offset labels (8 * site index, `none`
// for no site): out(p) = best of the 9 labels at p + step * {-1, 0, 1}^2.
// If clear is non-null it is reset to `none` pixel by pixel.  The site table
// must hold capS + 1 entries rounded up to an even count.
void launch_jfa_pass(hipStream_t st, const JfaPlan &plan, const uint32_t *in, uint32_t *out,
                     const float2 *sites, int W, int H, int step, uint32_t *clear, uint32_t none);

}  // namespace cvt
