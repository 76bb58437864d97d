// cvt_finalize.h -- the end of a Lloyd iteration's energy bookkeeping: the
// per-block (energy, squared gradient) partials of the site update summed in
// block order, the history entry, the stop rule of cvt_lloyd_until and the
// epoch step.  Shared by the update kernel (its last block) and by the kernels
// that run it deferred (an extra workgroup of the next seed pass, or a
// one-workgroup launch).
#pragma once

#include <hip/hip_runtime.h>

#include "cvt_types.h"

// The fixed pairwise tree v[t] += v[t + o], o = 32, 16, ..., 1, over the 64
// lanes of a wave (lane 0 ends with the sum).  Shuffles instead of LDS
// steps: the same additions in the same order, so the same bits as the LDS
// tree it replaces.
__device__ __forceinline__ double wave_tree_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double u = __shfl_down(v, o, 64);
        if ((threadIdx.x & 63) < (unsigned)o) v += u;
    }
    return v;
}

// Run by all 256 threads of one workgroup (red_e, red_g: 256 doubles of LDS
// each).  Thread t sums partials t, t + 256, ... of the nblocks blocks, then
// the tree of 256 (steps 128 and 64 in LDS, 32 .. 1 in wave 0).  Then, on
// thread 0: E and |grad E| appended to the history, the stop rule tested,
// the out copy written, the epoch advanced and the update counters cleared.
__device__ __forceinline__ void cvt_finalize_block(const double *__restrict__ blockpart,
                                                   int nblocks, int *__restrict__ dstate,
                                                   const FinalArgs &fa, double *red_e,
                                                   double *red_g)
{
    double te = 0.0, tg = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256) {
        te += __hip_atomic_load(&blockpart[2 * b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        tg += __hip_atomic_load(&blockpart[2 * b + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    red_e[threadIdx.x] = te;
    red_g[threadIdx.x] = tg;
    __syncthreads();
    for (int o = 128; o >= 64; o >>= 1) {
        if (threadIdx.x < o) {
            red_e[threadIdx.x] += red_e[threadIdx.x + o];
            red_g[threadIdx.x] += red_g[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x >= 64) return;
    const double se = wave_tree_sum(red_e[threadIdx.x]);
    const double sg = wave_tree_sum(red_g[threadIdx.x]);
    if (threadIdx.x == 0) {
        const double E = se * fa.e_scale;
        const double G = sqrt(sg) * fa.g_scale;
        if (fa.record) {
            const int h = dstate[ST_HIST];
            if (h < fa.hist_cap) { fa.hist[2 * h] = E; fa.hist[2 * h + 1] = G; }
            dstate[ST_HIST] = h + 1;
            // stop rule of cvt_lloyd_until: this iteration's energy is at or
            // below the target, or fell by less than rel_tol * E over the
            // last `window` iterations of the call
            if (dstate[ST_UACTIVE] && fa.until) {
                const int k = h - dstate[ST_UBASE], win = dstate[ST_UWIN];
                bool stop = E <= fa.until[1];
                if (win > 0 && k >= win && h < fa.hist_cap)
                    stop = stop || (fa.hist[2 * (h - win)] - E) < fa.until[0] * E;
                if (stop) dstate[ST_STOP] = 1;
            }
        }
        if (fa.out_eg) { fa.out_eg[0] = E; fa.out_eg[1] = G; }
        dstate[ST_EPOCH] = dstate[ST_EPOCH] + 1;
        dstate[ST_DONE] = 0;
        dstate[ST_PEND] = 0;
    }
}

