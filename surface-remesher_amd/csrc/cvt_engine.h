// cvt_engine.h -- internal state of a CVT engine context (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/cvt.h"
#include "cvt_types.h"
#include "jfa.h"

namespace cvt {

int fail(int code, const char *fmt, ...);
void clear_error();
const char *last_error();
int jfa_first_step(int W, int H);
int jfa_num_passes(int W, int H);
// exclusive prefix sum of n counts into start[0 .. n] on stream st (three
// launches over tiles of 4096 counts; count must be 16-byte aligned)
void launch_exclusive_scan(hipStream_t st, const int *count, int n, int *start);

struct Ctx {
    int W = 0, H = 0, maxS = 0, S = 0;
    int device = 0;
    hipStream_t stream = nullptr, own_stream = nullptr;
    bool own = false;

    float *rho = nullptr;          // [H][W]
    int rho_exp = 0;               // rho' = rho * 2^-rho_exp <= 1
    bool have_density = false;
    uint32_t *seed = nullptr;      // seed map [H][W]
    bool seeds_valid = false;      // seed map holds the current sites
    uint32_t *lab[2] = {nullptr, nullptr};
    uint32_t *first = nullptr;     // first JFA pass [H][W]: none except at the
                                   // receivers of the current seeds while a JFA runs
    int final_buf = 0;             // lab[final_buf] = labels of the last JFA
    uint32_t *labels_map() const { return lab[final_buf]; }
    // a map the JFA's result does not live in (scratch for exports)
    uint32_t *spare_map() const { return lab[final_buf ^ 1]; }
    bool labels_fresh = false;
    float2 *sites = nullptr;       // [capS + 1], sentinel at capS
    int capS = 1;                  // max(maxS, 1)
    uint32_t none() const { return 8u * (uint32_t)capS; }
    JfaPlan jplan;                 // launch shape of the JFA passes
    uint8_t *fixed = nullptr;      // [maxS]

    int tile = 0, tilesX = 0, tilesY = 0;
    bool tile_auto = true;
    Slot *slots = nullptr;         // [9][capS]: slot k of site i at k * capS + i
    OvfRec *ovf = nullptr;         // [ovf_cap] record pool
    int ovf_cap = 0;
    int *ovf_head = nullptr;       // [maxS] first record of each site, -1 none
    double *blockpart = nullptr;
    double *hist = nullptr;        // [hist_cap][2]
    int hist_cap = 0;
    int *dstate = nullptr;         // ST_*
    int *dflags = nullptr;         // FLAG_*
    float *scratch = nullptr;
    double *eg = nullptr;
    double *until = nullptr;       // stop rule of cvt_lloyd_until (rel_tol, target)
    int *ran = nullptr;            // iterations the last cvt_lloyd_until ran
    int *until_host = nullptr;     // pinned copies of the stop flag, one per chunk in flight
    hipEvent_t until_ev[2] = {nullptr, nullptr};


    bool use_graph = true;
    bool use_tail = true;          // fused last five JFA passes where the plan allows
    bool graph_valid = false;
    float graph_omega = 0.0f;
    hipGraph_t graph = nullptr;            // one iteration
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph_n = nullptr;          // kGraphIters iterations
    hipGraphExec_t graph_exec_n = nullptr;

    ~Ctx();
    int enqueue_seed();
    void clear_seed_map();         // every pixel back to none()
    int enqueue_jfa(hipEvent_t after_first = nullptr);   // event recorded after pass 0
    int enqueue_moments();
    // defer: the update leaves its finalize (energy sums, history, stop rule,
    // epoch) to the next seed pass or enqueue_finalize
    int enqueue_update(int mode, float omega, double *mom_out, int record, double *out_eg,
                       const float *domega = nullptr, CompatArgs ca = CompatArgs{},
                       int defer = 0);
    FinalArgs final_args(int record, double *out_eg) const;
    void enqueue_finalize();       // a deferred finalize, if one is pending
    int enqueue_iteration(float omega);
    int build_graph(float omega);
    int lloyd(int iterations, float omega);
    int lloyd_until(int max_iter, float omega, double rel_tol, int window, double target,
                    int *ran_out);
    int check_flags();
};

}  // namespace cvt
