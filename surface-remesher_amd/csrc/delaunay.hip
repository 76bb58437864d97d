// delaunay.hip -- 2D Delaunay triangulation on the GPU (the unconstrained part
// of gDel2D's GpuDel::compute, source/gDel2D/gDel2D/GPU/GpuDelaunay.cu:46-83).
//
// The triangulation is kept on the sphere: the points plus one vertex at
// infinity (index n), so every edge has two triangles.  A triangle holding
// the infinite vertex stands for the exterior beyond its finite edge; for
// point location that exterior is cut into wedges by rays from an interior
// point O (the centroid of the first triangle, stored as point n): the
// triangle (x, y, inf) owns the wedge between the rays O->x and O->y beyond
// the edge x->y.
//
// Rounds of parallel insertion, each followed by flipping to Delaunay:
//   vote      every uninserted point bids for its triangle (the point closest
//             to the centroid of a finite triangle; beyond the edge of an
//             infinite one, a point far from the edge and projecting near
//             its middle; ties by index)
//   split     each triangle with a bid splits 1 -> 3 around the winner; the
//             two new triangles go to slots after the current ones, placed
//             by a prefix sum over the splitting triangles
//   relocate  the other points of a split triangle move to the child that
//             holds them (orientation tests against the new point)
//   flip      Lawson flips until no edge is illegal: every illegal edge bids
//             its id (3 t + i, t the lower triangle) on both its triangles,
//             an edge that wins both flips, the outer neighbours are re-linked
//             through per-edge forwarding records, and points follow their
//             triangle across the new diagonal.  The lowest illegal edge
//             always wins, so every round makes progress.
// Predicates are exact, and exactly degenerate cases (three collinear, four
// cocircular points) are decided by a symbolic perturbation of the points
// by index (Simulation of Simplicity, dt_predicates.h), as gDel2D does
// (source/gDel2D/gDel2D/GPU/KerPredWrapper.h:69-136, :203-333): the result
// is the Delaunay triangulation of the perturbed points, a Delaunay
// triangulation of the input (every circumcircle empty of points in its
// interior) that includes every distinct point; collinear points on the
// convex hull give zero-area triangles along it.  In general position it is
// the unique Delaunay triangulation.  Exact duplicate points are dropped
// (the lowest index of each set is kept).
//
// Output: the finite triangles, counterclockwise, vertex indices into the
// input; neighbours as gDel2D's TriOpp encoding ((triangle << 4) | edge,
// edge i opposite vertex i; -1 on the convex hull).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "../../include/cvt.h"
#include "cvt_engine.h"
#include "dt_predicates.h"

namespace cvt {

using namespace dt;

enum { C_UNINS = 0, C_DUP = 1, C_ERR = 2, C_FLIPS = 3, C_A = 4, C_B = 5, C_C = 6, C_FTOT = 7,
       C_MISS = 8, C_NCROSS = 9, C_WSLOT = 10, NUM_C = 12 };
static constexpr int kParts = 256;              // workgroups of the first-triangle search
static constexpr unsigned long long NO_BID = ~0ull;
static constexpr int NO_FLIP = 0x7fffffff;

struct Pt {
    double x, y;
};

// orientation of indexed points under the perturbation (never 0 for
// distinct points)
__device__ __forceinline__ int orient_p(const Pt *P, int a, int b, int c, int *err)
{
    return orient2d_sos(P[a].x, P[a].y, P[b].x, P[b].y, P[c].x, P[c].y, a, b, c, err);
}

// the exact, unperturbed orientation
__device__ __forceinline__ int orient_x(const Pt *P, int a, int b, int c, int *err)
{
    return orient2d(P[a].x, P[a].y, P[b].x, P[b].y, P[c].x, P[c].y, err);
}

__device__ __forceinline__ bool same_pt(const Pt &a, const Pt &b) { return a.x == b.x && a.y == b.y; }

// point locations: a triangle (>= 0), inserted (-1), or a dropped duplicate
// of vertex v (-2 - v)
__device__ __forceinline__ int dup_code(int v) { return -2 - v; }

// block-wide count of pred, added to *counter once per block
__device__ __forceinline__ void count_add(int *counter, bool pred)
{
    const int c = __syncthreads_count(pred);
    if (threadIdx.x == 0 && c) atomicAdd(counter, c);
}

template <typename T>
__global__ void k_dt_fill(T *__restrict__ p, T v, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = v;
}

// first triangle: a = lowest x (then y, then index), b = highest, c = the
// point farthest from the line ab (approximate distance, ties by index);
// each search is a per-workgroup pass over the points and a one-workgroup
// pass over the kParts partial results
__device__ __forceinline__ bool pt_before(const Pt *P, int i, int j)   // (x, y, index) order
{
    if (j < 0) return true;
    if (i < 0) return false;
    if (P[i].x != P[j].x) return P[i].x < P[j].x;
    if (P[i].y != P[j].y) return P[i].y < P[j].y;
    return i < j;
}

// block-wide: (a, b) = first and last candidates; reduce pairs in LDS
__device__ void reduce_ab(const Pt *P, int &a, int &b)
{
    __shared__ int sa[256], sb[256];
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const int a2 = sa[threadIdx.x + s], b2 = sb[threadIdx.x + s];
            if (a2 >= 0 && pt_before(P, a2, sa[threadIdx.x])) sa[threadIdx.x] = a2;
            if (b2 >= 0 && (sb[threadIdx.x] < 0 || pt_before(P, sb[threadIdx.x], b2)))
                sb[threadIdx.x] = b2;
        }
        __syncthreads();
    }
    a = sa[0];
    b = sb[0];
}

__global__ __launch_bounds__(256) void k_dt_ab_part(const Pt *__restrict__ P, int n,
                                                    int *__restrict__ part, int *__restrict__ cnt)
{
    int a = -1, b = -1;
    bool bad = false;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (!isfinite(P[i].x) || !isfinite(P[i].y)) {
            bad = true;
            continue;
        }
        if (pt_before(P, i, a)) a = i;
        if (b < 0 || pt_before(P, b, i)) b = i;
    }
    if (bad) atomicOr(cnt + C_ERR, (int)DT_ERR_INPUT);
    reduce_ab(P, a, b);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = b;
    }
}

__global__ __launch_bounds__(256) void k_dt_ab_final(const Pt *__restrict__ P,
                                                     const int *__restrict__ part,
                                                     int *__restrict__ cnt)
{
    int a = part[2 * threadIdx.x], b = part[2 * threadIdx.x + 1];
    reduce_ab(P, a, b);
    if (threadIdx.x == 0) {
        cnt[C_A] = a;
        cnt[C_B] = b;
    }
}

// block-wide: the largest v, ties by the lower index
__device__ void reduce_far(double &v, int &c)
{
    __shared__ double sv[256];
    __shared__ int sc[256];
    sv[threadIdx.x] = v;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double v2 = sv[threadIdx.x + s];
            const int c2 = sc[threadIdx.x + s];
            if (c2 >= 0 && (sc[threadIdx.x] < 0 || v2 > sv[threadIdx.x] ||
                            (v2 == sv[threadIdx.x] && c2 < sc[threadIdx.x]))) {
                sv[threadIdx.x] = v2;
                sc[threadIdx.x] = c2;
            }
        }
        __syncthreads();
    }
    v = sv[0];
    c = sc[0];
}

__global__ __launch_bounds__(256) void k_dt_c_part(const Pt *__restrict__ P, int n,
                                                   const int *__restrict__ cnt,
                                                   double *__restrict__ partv,
                                                   int *__restrict__ partc)
{
    const int a = cnt[C_A], b = cnt[C_B];
    double best = -1.0;
    int c = -1;
    if (a >= 0 && b >= 0) {
        const double ux = P[b].x - P[a].x, uy = P[b].y - P[a].y;
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
            const double v = fabs(ux * (P[i].y - P[a].y) - uy * (P[i].x - P[a].x));
            if (v > best || (v == best && i < c)) { best = v; c = i; }
        }
    }
    reduce_far(best, c);
    if (threadIdx.x == 0) {
        partv[blockIdx.x] = best;
        partc[blockIdx.x] = c;
    }
}

__global__ __launch_bounds__(256) void k_dt_c_final(const double *__restrict__ partv,
                                                    const int *__restrict__ partc,
                                                    int *__restrict__ cnt)
{
    double v = partv[threadIdx.x];
    int c = partc[threadIdx.x];
    reduce_far(v, c);
    if (threadIdx.x == 0) cnt[C_C] = c;
}

// the four triangles of the first triangle on the sphere, and O
__global__ void k_dt_init(Pt *__restrict__ P, int n, int *__restrict__ tv, int *__restrict__ to,
                          int *__restrict__ cnt)
{
    int *err = cnt + C_ERR;
    int a = cnt[C_A], b = cnt[C_B], c = cnt[C_C];
    if (a < 0 || b < 0 || c < 0 || a == b || a == c || b == c) {
        set_flag(err, DT_ERR_DEGENERATE);
        return;
    }
    const int o = orient_x(P, a, b, c, err);
    if (o == 0) {
        set_flag(err, DT_ERR_DEGENERATE);                  // all points collinear
        return;
    }
    if (o < 0) { const int t = b; b = c; c = t; }
    cnt[C_A] = a;
    cnt[C_B] = b;
    cnt[C_C] = c;
    Pt O;
    O.x = (P[a].x + P[b].x + P[c].x) / 3.0;
    O.y = (P[a].y + P[b].y + P[c].y) / 3.0;
    P[n] = O;
    if (orient_x(P, a, b, n, err) <= 0 || orient_x(P, b, c, n, err) <= 0 ||
        orient_x(P, c, a, n, err) <= 0) {
        set_flag(err, DT_ERR_DEGENERATE);                  // a sliver first triangle
        return;
    }
    const int I = n;
    // T0 = (a, b, c), T1 = (b, a, inf), T2 = (c, b, inf), T3 = (a, c, inf)
    const int V[12] = {a, b, c, b, a, I, c, b, I, a, c, I};
    // edge i is opposite vertex i; links (triangle << 4) | edge
    const int L[12] = {(2 << 4) | 2, (3 << 4) | 2, (1 << 4) | 2,
                       (3 << 4) | 1, (2 << 4) | 0, (0 << 4) | 2,
                       (1 << 4) | 1, (3 << 4) | 0, (0 << 4) | 0,
                       (2 << 4) | 1, (1 << 4) | 0, (0 << 4) | 1};
    for (int i = 0; i < 12; ++i) {
        tv[i] = V[i];
        to[i] = L[i];
    }
}

// every other point to T0 or to the wedge of T1..T3 that holds it
__global__ void k_dt_locate0(const Pt *__restrict__ P, int n, int *__restrict__ loc,
                             int *__restrict__ cnt)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    bool dup = false;
    if (p < n) {
        int *err = cnt + C_ERR;
        const int a = cnt[C_A], b = cnt[C_B], c = cnt[C_C], O = n;
        int t;
        if (p == a || p == b || p == c) {
            t = -1;
        } else if (same_pt(P[p], P[a]) || same_pt(P[p], P[b]) || same_pt(P[p], P[c])) {
            t = dup_code(same_pt(P[p], P[a]) ? a : same_pt(P[p], P[b]) ? b : c);
            dup = true;
        } else if (orient_p(P, a, b, p, err) >= 0 && orient_p(P, b, c, p, err) >= 0 &&
                   orient_p(P, c, a, p, err) >= 0) {
            t = 0;
        } else {
            const int sa = orient_p(P, O, a, p, err), sb = orient_p(P, O, b, p, err),
                      sc = orient_p(P, O, c, p, err);
            // wedge (x, y) between the rays O->x and O->y, clockwise
            if (sb <= 0 && sa > 0) t = 1;
            else if (sc <= 0 && sb > 0) t = 2;
            else t = 3;
        }
        loc[p] = t;
    }
    count_add(cnt + C_DUP, dup);
}

__device__ __forceinline__ int inf_pos(const int *v, int I)
{
    return v[0] == I ? 0 : v[1] == I ? 1 : v[2] == I ? 2 : -1;
}

// atomicMin of each lane's bid on vote[t], one atomic per distinct t in the
// wave (in early rounds most of a wave's points share a triangle); called
// by every lane of the wave
__device__ __forceinline__ void wave_min_bid(unsigned long long *vote, int t,
                                             unsigned long long bid, bool active)
{
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(active);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int t0 = __shfl(t, leader);
        const bool mine = active && t == t0;
        unsigned long long v = mine ? bid : NO_BID;
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(v, off);
            v = o < v ? o : v;
        }
        if (lane == leader) atomicMin(&vote[t0], v);
        pending &= ~__ballot(mine);
        active = active && !mine;
    }
}

// the bid of point p on its triangle (NO_BID, and dupv = the vertex, when p
// duplicates one of the triangle's vertices)
__device__ __forceinline__ unsigned long long point_bid(const Pt *P, int n, const int *tv, int p,
                                                        int t, int &dupv)
{
    const int v[3] = {tv[3 * t], tv[3 * t + 1], tv[3 * t + 2]};
    const int k = inf_pos(v, n);
    dupv = -1;
    for (int j = 0; j < 3; ++j)
        if (j != k && same_pt(P[p], P[v[j]])) dupv = v[j];
    if (dupv >= 0) return NO_BID;
    float key;
    if (k < 0) {
        const double cx = (P[v[0]].x + P[v[1]].x + P[v[2]].x) / 3.0 - P[p].x;
        const double cy = (P[v[0]].y + P[v[1]].y + P[v[2]].y) / 3.0 - P[p].y;
        key = (float)(cx * cx + cy * cy);
    } else {
        // beyond a hull edge u -> w: far from the edge (the hull grows fast)
        // and projecting near its middle (points along a line, such as a
        // side of the domain, are halved every round instead of taken one
        // at a time from an end)
        const Pt &u = P[v[(k + 1) % 3]], &w = P[v[(k + 2) % 3]];
        const double ex = w.x - u.x, ey = w.y - u.y;
        const double len2 = ex * ex + ey * ey;
        const double qx = P[p].x - u.x, qy = P[p].y - u.y;
        const double t = len2 > 0.0 ? (qx * ex + qy * ey) / len2 : 0.5;
        const double d = len2 > 0.0 ? fmax(ex * qy - ey * qx, 0.0) / len2 : 0.0;
        key = (float)(1.0 / (1.0 + (d + 1e-6) / (1.0 + fabs(2.0 * t - 1.0))));
    }
    if (!(key >= 0.0f)) key = 3.0e38f;
    return ((unsigned long long)__float_as_uint(key) << 32) | (unsigned)p;
}

__global__ __launch_bounds__(256) void k_dt_vote(const Pt *__restrict__ P, int n,
                                                 const int *__restrict__ tv,
                                                 int *__restrict__ loc,
                                                 unsigned long long *__restrict__ vote,
                                                 int *__restrict__ cnt)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    bool live = false, dup = false;
    int t = -1;
    unsigned long long bid = NO_BID;
    if (p < n && loc[p] >= 0) {
        t = loc[p];
        int dupv;
        bid = point_bid(P, n, tv, p, t, dupv);
        dup = dupv >= 0;
        if (dup) loc[p] = dup_code(dupv);
        else live = true;
    }
    wave_min_bid(vote, t, bid, live);
    count_add(cnt + C_UNINS, live);
    count_add(cnt + C_DUP, dup);
}

// few triangles (T <= kVoteLds): a workgroup takes a grid-stride share of the
// points and keeps the minimum bid per triangle in LDS, then adds its
// non-empty entries to the global minimum (one global atomic per triangle
// per workgroup instead of one per wave)
static constexpr int kVoteLds = 4096;
__global__ __launch_bounds__(1024) void k_dt_vote_lds(const Pt *__restrict__ P, int n, int T,
                                                      const int *__restrict__ tv,
                                                      int *__restrict__ loc,
                                                      unsigned long long *__restrict__ vote,
                                                      int *__restrict__ cnt)
{
    __shared__ unsigned long long sv[kVoteLds];
    __shared__ int scount[2];
    for (int i = threadIdx.x; i < T; i += blockDim.x) sv[i] = NO_BID;
    if (threadIdx.x < 2) scount[threadIdx.x] = 0;
    __syncthreads();
    int live = 0, dups = 0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const int t = loc[p];
        if (t < 0) continue;
        int dupv;
        const unsigned long long bid = point_bid(P, n, tv, p, t, dupv);
        if (dupv >= 0) {
            loc[p] = dup_code(dupv);
            ++dups;
        } else {
            ++live;
            atomicMin(&sv[t], bid);
        }
    }
    if (live) atomicAdd(&scount[0], live);
    if (dups) atomicAdd(&scount[1], dups);
    __syncthreads();
    for (int i = threadIdx.x; i < T; i += blockDim.x)
        if (sv[i] != NO_BID) atomicMin(&vote[i], sv[i]);
    if (threadIdx.x == 0) {
        if (scount[0]) atomicAdd(cnt + C_UNINS, scount[0]);
        if (scount[1]) atomicAdd(cnt + C_DUP, scount[1]);
    }
}

__global__ void k_dt_split_flags(const unsigned long long *__restrict__ vote, int T,
                                 int *__restrict__ flag)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < T) flag[t] = vote[t] != NO_BID;
}

__device__ __forceinline__ int child_slot(int t, int i, int T, const int *base)
{
    return i == 0 ? t : T + 2 * base[t] + (i - 1);
}

// 1 -> 3 split of every triangle with a bid: child i = (v[i+1], v[i+2], p)
// takes slot t (i = 0) or T + 2 rank + i - 1, its edge 2 is the parent's edge i
__global__ void k_dt_split(const Pt *__restrict__ P, int n, int T, int *__restrict__ tv,
                           int *__restrict__ to, const unsigned long long *__restrict__ vote,
                           const int *__restrict__ base, int *__restrict__ loc,
                           int *__restrict__ chk, int next_pass, int *__restrict__ cnt)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T || vote[t] == NO_BID) return;
    int *err = cnt + C_ERR;
    const int p = (int)(vote[t] & 0xffffffffu);
    const int v[3] = {tv[3 * t], tv[3 * t + 1], tv[3 * t + 2]};
    const int o[3] = {to[3 * t], to[3 * t + 1], to[3 * t + 2]};
    for (int i = 0; i < 3; ++i) {
        // p strictly left of every finite edge (v[i+1], v[i+2])
        const int e0 = v[(i + 1) % 3], e1 = v[(i + 2) % 3];
        if (e0 == n || e1 == n) continue;
        if (orient_p(P, e0, e1, p, err) <= 0) set_flag(err, DT_ERR_DEGENERATE);
    }
    int slot[3];
    for (int i = 0; i < 3; ++i) slot[i] = child_slot(t, i, T, base);
    for (int i = 0; i < 3; ++i) {
        const int s = slot[i];
        chk[s] = next_pass;
        tv[3 * s] = v[(i + 1) % 3];
        tv[3 * s + 1] = v[(i + 2) % 3];
        tv[3 * s + 2] = p;
        to[3 * s] = (slot[(i + 1) % 3] << 4) | 1;
        to[3 * s + 1] = (slot[(i + 2) % 3] << 4) | 0;
        const int m = o[i] >> 4, e = o[i] & 3;
        if (vote[m] != NO_BID) {
            to[3 * s + 2] = (child_slot(m, e, T, base) << 4) | 2;
        } else {
            to[3 * s + 2] = o[i];
            to[3 * m + e] = (s << 4) | 2;
        }
    }
    loc[p] = -1;
}

// the other points of a split triangle to the child that holds them
__global__ void k_dt_relocate_split(const Pt *__restrict__ P, int n, int T,
                                    const int *__restrict__ tv,
                                    const unsigned long long *__restrict__ vote,
                                    const int *__restrict__ base, int *__restrict__ loc,
                                    int *__restrict__ cnt)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    bool dup = false;
    if (q < n) {
        const int t = loc[q];
        if (t >= 0 && vote[t] != NO_BID) {
            int *err = cnt + C_ERR;
            const int p = (int)(vote[t] & 0xffffffffu);
            const int c1 = child_slot(t, 1, T, base);
            // child 0 = (v1, v2, p), child 1 = (v2, v0, p)
            const int v[3] = {tv[3 * c1 + 1], tv[3 * t], tv[3 * t + 1]};
            if (same_pt(P[q], P[p])) {
                dup = true;
                loc[q] = dup_code(p);
            } else {
                const int k = inf_pos(v, n);
                int ci = 0;
                if (k < 0) {
                    int s[3];
                    for (int j = 0; j < 3; ++j) s[j] = orient_p(P, p, v[j], q, err);
                    for (int i = 0; i < 3; ++i)
                        if (s[(i + 1) % 3] >= 0 && s[(i + 2) % 3] < 0) ci = i;
                } else {
                    const int u = v[(k + 1) % 3], w = v[(k + 2) % 3];
                    if (orient_p(P, w, p, q, err) >= 0 && orient_p(P, p, u, q, err) >= 0)
                        ci = k;                                 // (u, w, p)
                    else if (orient_p(P, n, p, q, err) > 0)
                        ci = (k + 2) % 3;                       // (inf, u, p)
                    else
                        ci = (k + 1) % 3;                       // (w, inf, p)
                }
                loc[q] = child_slot(t, ci, T, base);
            }
        }
    }
    count_add(cnt + C_DUP, dup);
}

// w strictly inside the circumcircle of t (for t with the infinite vertex:
// strictly beyond its finite edge)
__device__ __forceinline__ bool illegal(const Pt *P, int n, const int *v, int w, int *err)
{
    if (w == n) return false;
    const int k = inf_pos(v, n);
    if (k < 0)
        return incircle_sos(P[v[0]].x, P[v[0]].y, P[v[1]].x, P[v[1]].y, P[v[2]].x, P[v[2]].y,
                            P[w].x, P[w].y, v[0], v[1], v[2], w, err) > 0;
    return orient_p(P, v[(k + 1) % 3], v[(k + 2) % 3], w, err) > 0;
}

// an edge can turn illegal only when one of its triangles changed (split or
// flipped) before this pass: chk[t] holds the pass that must look at t
__global__ void k_dt_flip_mark(const Pt *__restrict__ P, int n, int T, const int *__restrict__ tv,
                               const int *__restrict__ to, const int *__restrict__ chk, int pass,
                               int *__restrict__ fvote, int *__restrict__ cnt)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) cnt[C_FLIPS] = 0;                  // this pass's flips (k_dt_flip counts)
    if (t >= T) return;
    int *err = cnt + C_ERR;
    const bool mine = chk[t] == pass;
    const int v[3] = {tv[3 * t], tv[3 * t + 1], tv[3 * t + 2]};
    for (int i = 0; i < 3; ++i) {
        const int r = to[3 * t + i];
        const int m = r >> 4;
        if (r < 0 || m < t) continue;               // each edge once, by its lower triangle
        if (r & 4) continue;                        // a constraint edge never flips
        if (!mine && chk[m] != pass) continue;
        const int w = tv[3 * m + (r & 3)];
        if (illegal(P, n, v, w, err)) {
            const int id = 3 * t + i;
            atomicMin(&fvote[t], id);
            atomicMin(&fvote[m], id);
        }
    }
}

// flip of the edge i of t shared with edge k of m: t' = (v_i, v_i+1, w) in
// slot t, m' = (w, v_i+2, v_i) in slot m; the outer links are kept as they
// were (fixed by k_dt_flip_fix) and each old outer edge forwards to its new
// place in fwd[3 * slot + old edge]
__global__ void k_dt_flip(int T, int *__restrict__ tv, int *__restrict__ to,
                          const int *__restrict__ fvote, int *__restrict__ fwd,
                          int *__restrict__ fstamp, int4 *__restrict__ fdiag, int *__restrict__ chk,
                          int round, int *__restrict__ cnt)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    bool did = false;
    if (t < T) {
        const int id = fvote[t];
        // a triangle with an illegal edge is looked at again next pass,
        // whether that edge flips now or lost the bid to a lower one
        if (id != NO_FLIP) chk[t] = round + 1;
        if (id != NO_FLIP && id / 3 == t) {
            const int i = id - 3 * t;
            const int r = to[3 * t + i], m = r >> 4, k = r & 3;
            if (fvote[m] == id) {
                did = true;
                const int vi = tv[3 * t + i], vi1 = tv[3 * t + (i + 1) % 3],
                          vi2 = tv[3 * t + (i + 2) % 3];
                const int w = tv[3 * m + k];
                const int ot1 = to[3 * t + (i + 1) % 3], ot2 = to[3 * t + (i + 2) % 3];
                const int om1 = to[3 * m + (k + 1) % 3], om2 = to[3 * m + (k + 2) % 3];
                tv[3 * t] = vi;
                tv[3 * t + 1] = vi1;
                tv[3 * t + 2] = w;
                tv[3 * m] = w;
                tv[3 * m + 1] = vi2;
                tv[3 * m + 2] = vi;
                // t': edge 0 = (vi1, w) was m's edge k+1; edge 1 = diagonal;
                //     edge 2 = (vi, vi1) was t's edge i+2
                // m': edge 0 = (vi2, vi) was t's edge i+1; edge 1 = diagonal;
                //     edge 2 = (w, vi2) was m's edge k+2
                to[3 * t] = om1;
                to[3 * t + 1] = (m << 4) | 1;
                to[3 * t + 2] = ot2;
                to[3 * m] = ot1;
                to[3 * m + 1] = (t << 4) | 1;
                to[3 * m + 2] = om2;
                fwd[3 * t + (i + 2) % 3] = (t << 4) | 2;
                fwd[3 * t + (i + 1) % 3] = (m << 4) | 0;
                fwd[3 * m + (k + 1) % 3] = (t << 4) | 0;
                fwd[3 * m + (k + 2) % 3] = (m << 4) | 2;
                fstamp[t] = round;
                fstamp[m] = round;
                chk[t] = round + 1;
                chk[m] = round + 1;
                const int4 d = make_int4(vi, w, t, m);
                fdiag[t] = d;
                fdiag[m] = d;
            }
        }
    }
    const int c = __syncthreads_count(did);
    if (threadIdx.x == 0 && c) {
        atomicAdd(cnt + C_FLIPS, c);
        atomicAdd(cnt + C_FTOT, c);
    }
}

// after a flip pass, per triangle t < T: re-link the outer edges of flipped
// triangles (through the forwarding records of flipped neighbours) and
// reset the bid for the next pass; per point q < n: follow a flipped
// triangle across its new diagonal
__global__ void k_dt_flip_fix(const Pt *__restrict__ P, int n, int T, int *__restrict__ to,
                              const int *__restrict__ fwd, const int *__restrict__ fstamp,
                              const int4 *__restrict__ fdiag, int *__restrict__ fvote,
                              int *__restrict__ loc, int round, int *__restrict__ cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < T) {
        fvote[i] = NO_FLIP;
        if (fstamp[i] == round) {
            for (int e = 0; e < 3; e += 2) {
                const int r = to[3 * i + e];
                const int m = r >> 4, k = r & 3;
                // bit 2 marks a constraint edge on both of its sides
                if (fstamp[m] == round)
                    to[3 * i + e] = fwd[3 * m + k] | (r & 4);
                else
                    to[3 * m + k] = (i << 4) | e | (r & 4);
            }
        }
    }
    if (i < n) {
        const int t = loc[i];
        if (t >= 0 && fstamp[t] == round) {
            const int4 d = fdiag[t];
            // t' = (vi, vi1, w) lies right of vi -> w
            loc[i] = orient_p(P, d.x, d.y, i, cnt + C_ERR) < 0 ? d.z : d.w;
        }
    }
}

// ---------------------------------------------------------------------------
// Constraint segments (gDel2D's doInsertConstraints, GpuDelaunay.cu:638-747).
// A missing segment is walked through the triangulation: the triangles it
// crosses (its cavity) form a polygon that the segment splits in two.  The
// segment claims its cavity and the cavity's neighbours (the lowest of a
// per-round hashed priority wins each triangle); a segment holding all of them replaces its
// cavity by the segment plus a Delaunay triangulation of each side (the
// apex of each base is the chain vertex whose circle with the base holds no
// other chain vertex, applied recursively: Anglada, "An improved
// incremental algorithm for constructing restricted Delaunay
// triangulations", 1997), in the same triangle slots.  The missing segment
// of lowest priority always proceeds, so every round inserts at least one.  Present
// segments get bit 2 on both sides' links; the Lawson passes that follow
// never flip them.
// ---------------------------------------------------------------------------
static constexpr int kCavMax = 128;            // edges one segment may cross
static constexpr int kWsSlots = 16384;         // cavities re-triangulated per round, at most
// workspace of one cavity re-triangulation (ints, 7.2 KB): cavity slots,
// their saved vertices and links, the two chains, the new triangles and the
// recursion stack; kWsSlots of them, taken by the round's winners
static constexpr int kWsC = 0, kWsOV = kCavMax + 1, kWsOL = 4 * (kCavMax + 1),
                     kWsUL = 7 * (kCavMax + 1), kWsUR = kWsUL + kCavMax + 2,
                     kWsNT = kWsUR + kCavMax + 2, kWsST = kWsNT + 3 * (kCavMax + 1),
                     kWs = kWsST + 2 * (kCavMax + 2);

// vt[v] = a triangle incident to vertex v; deg[v] += its incident triangles
// (deg zeroed before)
__global__ void k_dt_vertex_tri(const int *__restrict__ tv, int T, int n, int *__restrict__ vt,
                                int *__restrict__ deg)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    for (int j = 0; j < 3; ++j) {
        const int v = tv[3 * t + j];
        if (v != n) {
            vt[v] = t;
            atomicAdd(&deg[v], 1);
        }
    }
}

__device__ __forceinline__ int kept_vertex(const int *loc, int v)
{
    const int l = loc[v];
    return l <= -2 ? -2 - l : v;                    // a dropped duplicate's kept twin
}

// one thread per segment: marks it when present (ncr = 0), else lists the
// edges it crosses from a to b (ncr of them; -1 beyond kCavMax).  a is the
// end of lower degree (the turn around it is the shorter), kept in sa.
__global__ void k_dt_walk(const Pt *__restrict__ P, int n, int T, const int *__restrict__ tv,
                          int *__restrict__ to, const int *__restrict__ vt,
                          const int *__restrict__ deg, const int *__restrict__ loc,
                          const int *__restrict__ seg, int nseg, int2 *__restrict__ cross,
                          int *__restrict__ ncr, int *__restrict__ sa, int *__restrict__ cnt)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    bool missing = false;
    if (s < nseg) {
        int *err = cnt + C_ERR;
        int count = 0;
        int a = kept_vertex(loc, seg[2 * s]), b = kept_vertex(loc, seg[2 * s + 1]);
        if (deg[b] < deg[a]) {
            const int x = a;
            a = b;
            b = x;
        }
        sa[s] = a;
        if (a != b) {
            int t = vt[a], ct = -1, ce = -1;
            // around a, counterclockwise, until the edge ab or the wedge holding b
            for (int it = 0; it <= T; ++it) {
                const int i = tv[3 * t] == a ? 0 : tv[3 * t + 1] == a ? 1 : 2;
                const int v1 = tv[3 * t + (i + 1) % 3], v2 = tv[3 * t + (i + 2) % 3];
                if (v1 == b || v2 == b) {
                    const int e = v1 == b ? (i + 2) % 3 : (i + 1) % 3;
                    const int r = to[3 * t + e];
                    atomicOr(&to[3 * t + e], 4);
                    atomicOr(&to[3 * (r >> 4) + (r & 3)], 4);
                    break;
                }
                if (v1 != n && v2 != n && orient_p(P, a, v1, b, err) > 0 &&
                    orient_p(P, a, v2, b, err) < 0) {
                    ct = t;
                    ce = i;
                    break;
                }
                if (it == T) set_flag(err, DT_ERR_DEGENERATE);
                t = to[3 * t + (i + 1) % 3] >> 4;   // across (v2, a): the next one ccw
            }
            if (ct >= 0) {
                missing = true;
                // crossed edge ce of ct, as (R, L): R right of a->b, L left
                for (int it = 0; it <= T; ++it) {
                    if (count < kCavMax) cross[(long)s * kCavMax + count] = make_int2(ct, ce);
                    ++count;
                    const int r = to[3 * ct + ce], m = r >> 4, k = r & 3;
                    if (r & 4) {
                        set_flag(err, DT_ERR_CROSS);        // through a present segment
                        break;
                    }
                    const int w = tv[3 * m + k];
                    if (w == b) break;
                    if (w == n || it == T) {
                        set_flag(err, DT_ERR_DEGENERATE);
                        break;
                    }
                    // m = (w, L, R) from position k: leave through (R, w) or (w, L)
                    ce = orient_p(P, a, b, w, err) > 0 ? (k + 1) % 3 : (k + 2) % 3;
                    ct = m;
                }
                if (count > kCavMax) {
                    set_flag(err, DT_ERR_CAPACITY);
                    count = -1;
                }
            }
        }
        ncr[s] = count;
    }
    count_add(cnt + C_MISS, missing);
}

// the cavity of segment s: the crossed triangles, then the one holding b
__device__ __forceinline__ int cavity_tri(const int2 *cr, int k, int j, const int *to)
{
    if (j < k) return cr[j].x;
    const int2 l = cr[k - 1];
    return to[3 * l.x + l.y] >> 4;
}

// a segment's claim for this round: a hashed priority (so that a chain of
// touching cavities frees about a third of its segments per round, not
// only its lowest), then the segment index
__device__ __forceinline__ unsigned long long claim_key(int s, int round)
{
    unsigned h = (unsigned)s * 0x9E3779B1u ^ (unsigned)round * 0x85EBCA6Bu;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return ((unsigned long long)h << 32) | (unsigned)s;
}

// claims (lowest key) on each cavity triangle and its neighbours
__global__ void k_dt_claim(const int *__restrict__ to, const int2 *__restrict__ cross,
                           const int *__restrict__ ncr, int nseg, int round,
                           unsigned long long *__restrict__ own)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg || ncr[s] <= 0) return;
    const int k = ncr[s];
    const unsigned long long key = claim_key(s, round);
    const int2 *cr = cross + (long)s * kCavMax;
    for (int j = 0; j <= k; ++j) {
        const int t = cavity_tri(cr, k, j, to);
        atomicMin(&own[t], key);
        for (int e = 0; e < 3; ++e) atomicMin(&own[to[3 * t + e] >> 4], key);
    }
}

// apexes for the chain ch[0 .. len) over the base (ch[0], ch[len-1]), the
// chain on the base's left: appends counterclockwise triangles to nt
__device__ int triangulate_chain(const Pt *P, const int *ch, int len, int *nt, int cnt0,
                                 int *stack, int *err)
{
    int sp = 0, nn = cnt0;
    stack[sp++] = 0;
    stack[sp++] = len - 1;
    while (sp > 0) {
        const int i1 = stack[--sp], i0 = stack[--sp];
        if (i1 - i0 < 2) continue;
        int c = i0 + 1;
        for (int i = i0 + 2; i < i1; ++i) {
            const int p = ch[i0], q = ch[i1], r = ch[c], x = ch[i];
            if (incircle_sos(P[p].x, P[p].y, P[q].x, P[q].y, P[r].x, P[r].y, P[x].x, P[x].y, p,
                             q, r, x, err) > 0)
                c = i;
        }
        nt[3 * nn] = ch[i0];
        nt[3 * nn + 1] = ch[i1];
        nt[3 * nn + 2] = ch[c];
        ++nn;
        stack[sp++] = i0;
        stack[sp++] = c;
        stack[sp++] = c;
        stack[sp++] = i1;
    }
    return nn;
}

// one thread per segment that holds its claims: the cavity replaced by the
// segment and the triangulations of its two sides
__global__ void k_dt_cavity(const Pt *__restrict__ P, int n, int *__restrict__ tv,
                            int *__restrict__ to, const int *__restrict__ loc,
                            const int *__restrict__ seg, const int2 *__restrict__ cross,
                            const int *__restrict__ ncr, int nseg, int round,
                            const unsigned long long *__restrict__ own,
                            const int *__restrict__ sa, int *__restrict__ ws_all, int nslots,
                            int *__restrict__ cnt)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg || ncr[s] <= 0) return;
    int *err = cnt + C_ERR;
    const int k = ncr[s];
    const unsigned long long key = claim_key(s, round);
    const int2 *cr = cross + (long)s * kCavMax;
    for (int j = 0; j <= k; ++j) {
        const int t = cavity_tri(cr, k, j, to);
        if (own[t] != key) return;
        for (int e = 0; e < 3; ++e)
            if (own[to[3 * t + e] >> 4] != key) return;
    }
    // a workspace slot; a segment left without one waits for the next round
    const int slot = atomicAdd(cnt + C_WSLOT, 1);
    if (slot >= nslots) return;
    int *ws = ws_all + (long)slot * kWs;
    int *C = ws + kWsC, *OV = ws + kWsOV, *OL = ws + kWsOL, *UL = ws + kWsUL, *UR = ws + kWsUR,
        *NT = ws + kWsNT, *ST = ws + kWsST;
    for (int j = 0; j <= k; ++j) C[j] = cavity_tri(cr, k, j, to);
    // the walk's direction: from sa[s] to the other end
    const int e0 = kept_vertex(loc, seg[2 * s]), e1 = kept_vertex(loc, seg[2 * s + 1]);
    const int a = sa[s], b = a == e0 ? e1 : e0;
    for (int j = 0; j <= k; ++j)
        for (int e = 0; e < 3; ++e) {
            OV[3 * j + e] = tv[3 * C[j] + e];
            OL[3 * j + e] = to[3 * C[j] + e];
        }
    // the chains left and right of a -> b
    int nl = 0, nr = 0;
    UL[nl++] = a;
    UR[nr++] = a;
    for (int j = 0; j < k; ++j) {
        const int R = OV[3 * j + (cr[j].y + 1) % 3], L = OV[3 * j + (cr[j].y + 2) % 3];
        if (UL[nl - 1] != L) UL[nl++] = L;
        if (UR[nr - 1] != R) UR[nr++] = R;
    }
    UL[nl++] = b;
    UR[nr++] = b;
    // the right chain, reversed, lies left of b -> a
    for (int i = 0; i < nr / 2; ++i) {
        const int x = UR[i];
        UR[i] = UR[nr - 1 - i];
        UR[nr - 1 - i] = x;
    }
    int nn = triangulate_chain(P, UL, nl, NT, 0, ST, err);
    nn = triangulate_chain(P, UR, nr, NT, nn, ST, err);
    if (nn != k + 1) {
        set_flag(err, DT_ERR_DEGENERATE);
        return;
    }
    // links: the other new triangle across an edge, else the old outer link
    for (int e = 0; e < nn; ++e) {
        for (int f = 0; f < 3; ++f) {
            const int u = NT[3 * e + (f + 1) % 3], v = NT[3 * e + (f + 2) % 3];
            int link = -1;
            for (int g = 0; g < nn && link < 0; ++g) {
                if (g == e) continue;
                for (int h = 0; h < 3; ++h)
                    if (NT[3 * g + (h + 1) % 3] == v && NT[3 * g + (h + 2) % 3] == u) {
                        const bool ab = (u == a && v == b) || (u == b && v == a);
                        link = (C[g] << 4) | h | (ab ? 4 : 0);
                        break;
                    }
            }
            if (link < 0) {
                for (int j = 0; j <= k && link < 0; ++j)
                    for (int h = 0; h < 3; ++h)
                        if (OV[3 * j + (h + 1) % 3] == u && OV[3 * j + (h + 2) % 3] == v) {
                            link = OL[3 * j + h];
                            break;
                        }
                if (link < 0) {
                    set_flag(err, DT_ERR_DEGENERATE);
                    return;
                }
                // the outer neighbour is claimed by s: nobody else writes it
                to[3 * (link >> 4) + (link & 3)] = (C[e] << 4) | f | (link & 4);
            }
            to[3 * C[e] + f] = link;
        }
    }
    for (int e = 0; e < nn; ++e)
        for (int f = 0; f < 3; ++f) tv[3 * C[e] + f] = NT[3 * e + f];
    atomicAdd(cnt + C_FLIPS, 1);
}

__global__ void k_dt_alive(const int *__restrict__ tv, int T, int n, int *__restrict__ flag)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < T) flag[t] = tv[3 * t] != n && tv[3 * t + 1] != n && tv[3 * t + 2] != n;
}

__global__ void k_dt_output(const int *__restrict__ tv, const int *__restrict__ to, int T, int n,
                            const int *__restrict__ flag, const int *__restrict__ pos,
                            int *__restrict__ tri, int *__restrict__ opp)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T || !flag[t]) return;
    const int j = pos[t];
    for (int e = 0; e < 3; ++e) {
        tri[3 * j + e] = tv[3 * t + e];
        if (opp) {
            const int r = to[3 * t + e], m = r >> 4;
            opp[3 * j + e] = flag[m] ? (pos[m] << 4) | (r & 7) : -1;
        }
    }
}

// predicate check: orient(a, b, c) and incircle(a, b, c, d) per 8 coordinates
__global__ void k_dt_predicates(const double *__restrict__ q, int count, int *__restrict__ o,
                                int *__restrict__ ic, int *__restrict__ err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const double *x = q + 8 * i;
    o[i] = orient2d(x[0], x[1], x[2], x[3], x[4], x[5], err);
    ic[i] = incircle(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], err);
}

struct DtCtx {
    int maxN = 0, Tmax = 0;
    hipStream_t st = nullptr;
    Pt *P = nullptr;                       // [maxN + 1]; point n is O
    int *tv = nullptr, *to = nullptr;      // [Tmax][3]
    unsigned long long *vote = nullptr;    // [Tmax]
    int *flag = nullptr, *pos = nullptr;   // [Tmax], [Tmax + 1]
    int *fvote = nullptr, *fstamp = nullptr, *fwd = nullptr, *chk = nullptr;
    int4 *fdiag = nullptr;
    int *loc = nullptr;                    // [maxN]
    int *otri = nullptr, *oopp = nullptr;  // [Tmax][3] output staging
    int *cnt = nullptr;                    // [NUM_C]
    int *vt = nullptr;                     // [maxN] a triangle per vertex (constraints)
    // constraint buffers, sized by the largest segment count seen:
    // per segment: ends [2], crossings [kCavMax], crossing count; then
    // min(segments, kWsSlots) workspaces [kWs]
    char *cons = nullptr;
    int segCap = 0;
    int *seg = nullptr, *ncr = nullptr, *ws = nullptr, *sa = nullptr, *deg = nullptr;
    unsigned long long *own = nullptr;
    int2 *cross = nullptr;
    int *part = nullptr;                   // [3 kParts] first-triangle search
    double *partv = nullptr;               // [kParts]
    int *hcnt = nullptr;                   // pinned [NUM_C + 1]
    char *arena = nullptr;
    int stats[4] = {0, 0, 0, 0};           // rounds, flip passes, flips, duplicates
};

static void dt_free(DtCtx *c)
{
    if (c->cons) (void)hipFree(c->cons);
    if (c->arena) (void)hipFree(c->arena);
    if (c->hcnt) (void)hipHostFree(c->hcnt);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

static inline unsigned nb(long n) { return (unsigned)std::max<long>(1, (n + 255) / 256); }

}  // namespace cvt

using namespace cvt;

#define DCHK(call, what)                                                                   \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) return fail(CVT_ERR_HIP, "%s: %s", what, hipGetErrorString(e_)); \
    } while (0)

extern "C" int cvt_dt_create(cvt_dt **out, int max_points)
{
    clear_error();
    if (!out) return fail(CVT_ERR_ARG, "null output handle");
    *out = nullptr;
    if (max_points < 3 || max_points > (1 << 26))
        return fail(CVT_ERR_ARG, "max_points %d outside [3, 2^26]", max_points);
    DtCtx *c = new DtCtx;
    c->maxN = max_points;
    c->Tmax = 2 * (max_points + 1) + 8;
    const size_t T = (size_t)c->Tmax, N = (size_t)max_points;
    size_t at = 0;
    auto carve = [&](auto *&p, size_t count, bool set) {
        at = (at + 255) & ~(size_t)255;
        if (set) p = reinterpret_cast<std::remove_reference_t<decltype(p)>>(c->arena + at);
        at += count * sizeof(*p);
    };
    auto layout = [&](bool set) {
        at = 0;
        carve(c->P, N + 1, set); carve(c->tv, 3 * T, set); carve(c->to, 3 * T, set);
        carve(c->vote, T, set); carve(c->flag, T, set); carve(c->pos, T + 1, set);
        carve(c->fvote, T, set); carve(c->fstamp, T, set); carve(c->fwd, 3 * T, set);
        carve(c->chk, T, set);
        carve(c->fdiag, T, set); carve(c->loc, N, set); carve(c->cnt, NUM_C, set);
        carve(c->otri, 3 * T, set); carve(c->oopp, 3 * T, set);
        carve(c->part, 3 * kParts, set); carve(c->partv, kParts, set);
        carve(c->vt, N, set);
    };
    layout(false);
    if (hipMalloc(&c->arena, at) != hipSuccess) {
        c->arena = nullptr;
        dt_free(c);
        return fail(CVT_ERR_NOMEM, "%zu bytes for %d points", at, max_points);
    }
    layout(true);
    if (hipHostMalloc((void **)&c->hcnt, sizeof(int) * (NUM_C + 1), 0) != hipSuccess ||
        hipStreamCreate(&c->st) != hipSuccess) {
        dt_free(c);
        return fail(CVT_ERR_HIP, "stream / pinned counters");
    }
    *out = reinterpret_cast<cvt_dt *>(c);
    return CVT_OK;
}

extern "C" int cvt_dt_destroy(cvt_dt *dt)
{
    if (dt) dt_free(reinterpret_cast<DtCtx *>(dt));
    return CVT_OK;
}

// Lawson flips to Delaunay over T triangles.  Each pass flips at least the
// lowest illegal edge; passes are enqueued kFlipChunk at a time and the
// flips of the chunk's last pass read back (a pass with nothing to check
// costs only its launches).  The bound guards a broken mesh.
static int flip_to_delaunay(DtCtx *c, int n, int T, int &round)
{
    hipStream_t st = c->st;
    int *h = c->hcnt;
    constexpr int kFlipChunk = 4;
    for (int pass = 0;; pass += kFlipChunk) {
        if (pass > 4 * c->Tmax + 64) return fail(CVT_ERR_STATE, "flipping did not converge");
        for (int q = 0; q < kFlipChunk; ++q) {
            ++round;
            hipLaunchKernelGGL(k_dt_flip_mark, dim3(nb(T)), dim3(256), 0, st, c->P, n, T, c->tv,
                               c->to, c->chk, round, c->fvote, c->cnt);
            hipLaunchKernelGGL(k_dt_flip, dim3(nb(T)), dim3(256), 0, st, T, c->tv, c->to,
                               c->fvote, c->fwd, c->fstamp, c->fdiag, c->chk, round, c->cnt);
            hipLaunchKernelGGL(k_dt_flip_fix, dim3(nb(std::max(T, n))), dim3(256), 0, st, c->P, n,
                               T, c->to, c->fwd, c->fstamp, c->fdiag, c->fvote, c->loc, round,
                               c->cnt);
            ++c->stats[1];
        }
        DCHK(hipMemcpyAsync(h, c->cnt, sizeof(int) * NUM_C, hipMemcpyDeviceToHost, st), "counters");
        DCHK(hipStreamSynchronize(st), "flip");
        if (h[C_FLIPS] == 0 || h[C_ERR]) return CVT_OK;
    }
}

// the constraint segments as edges (gDel2D's doInsertConstraints), then the
// constrained Lawson flips
static int insert_constraints(DtCtx *c, int n, int T, int nseg, int &round)
{
    hipStream_t st = c->st;
    int *h = c->hcnt;
    for (int outer = 0;; ++outer) {
        hipLaunchKernelGGL(k_dt_fill<int>, dim3(nb(n)), dim3(256), 0, st, c->deg, 0, (long)n);
        hipLaunchKernelGGL(k_dt_vertex_tri, dim3(nb(T)), dim3(256), 0, st, c->tv, T, n, c->vt,
                           c->deg);
        hipLaunchKernelGGL(k_dt_fill<int>, dim3(1), dim3(64), 0, st, c->cnt + C_MISS, 0, 1L);
        hipLaunchKernelGGL(k_dt_fill<int>, dim3(1), dim3(64), 0, st, c->cnt + C_FLIPS, 0, 1L);
        hipLaunchKernelGGL(k_dt_fill<int>, dim3(1), dim3(64), 0, st, c->cnt + C_WSLOT, 0, 1L);
        hipLaunchKernelGGL(k_dt_walk, dim3(nb(nseg)), dim3(256), 0, st, c->P, n, T, c->tv, c->to,
                           c->vt, c->deg, c->loc, c->seg, nseg, c->cross, c->ncr, c->sa, c->cnt);
        hipLaunchKernelGGL(k_dt_fill<unsigned long long>, dim3(nb(T)), dim3(256), 0, st, c->own,
                           ~0ull, (long)T);
        hipLaunchKernelGGL(k_dt_claim, dim3(nb(nseg)), dim3(256), 0, st, c->to, c->cross, c->ncr,
                           nseg, outer, c->own);
        hipLaunchKernelGGL(k_dt_cavity, dim3((nseg + 63) / 64), dim3(64), 0, st, c->P, n, c->tv,
                           c->to, c->loc, c->seg, c->cross, c->ncr, nseg, outer, c->own, c->sa,
                           c->ws, std::min(nseg, kWsSlots), c->cnt);
        DCHK(hipMemcpyAsync(h, c->cnt, sizeof(int) * NUM_C, hipMemcpyDeviceToHost, st), "counters");
        DCHK(hipStreamSynchronize(st), "constraints");
        ++c->stats[1];
        if (h[C_ERR] & DT_ERR_CAPACITY)
            return fail(CVT_ERR_CAPACITY, "a constraint segment crosses more than %d edges",
                        kCavMax);
        if (h[C_ERR] & DT_ERR_CROSS)
            return fail(CVT_ERR_ARG, "a constraint segment crosses another (round %d)", outer);
        if (h[C_ERR]) return CVT_OK;                 // reported by the caller
        if (h[C_MISS] == 0) break;
        if (h[C_FLIPS] == 0 || outer > nseg)
            return fail(CVT_ERR_ARG,
                        "constraint insertion made no progress (round %d, %d segments missing, "
                        "%d inserted)", outer, h[C_MISS], h[C_FLIPS]);
    }
    // every edge may have to change now: look at all triangles once
    hipLaunchKernelGGL(k_dt_fill<int>, dim3(nb(T)), dim3(256), 0, st, c->chk, round + 1, (long)T);
    return flip_to_delaunay(c, n, T, round);
}

static int dt_run(cvt_dt *dt, const double *points, int num_points, const int *segments,
                  int num_segments, int *tri, int *opp, int capacity, int *num_tri)
{
    clear_error();
    if (!dt) return fail(CVT_ERR_ARG, "null context");
    DtCtx *c = reinterpret_cast<DtCtx *>(dt);
    const int n = num_points;
    if (n < 3 || n > c->maxN) return fail(CVT_ERR_ARG, "num_points %d outside [3, %d]", n, c->maxN);
    if (!points || !num_tri || (capacity > 0 && !tri) || capacity < 0)
        return fail(CVT_ERR_ARG, "null array or negative capacity");
    if (num_segments < 0 || (num_segments > 0 && !segments))
        return fail(CVT_ERR_ARG, "null segment array or negative count");
    if (num_segments > 0) {
        if (num_segments > c->segCap) {
            if (c->cons) (void)hipFree(c->cons);
            c->cons = nullptr;
            c->segCap = 0;
            const size_t S = (size_t)num_segments;
            size_t at = 0;
            auto carve = [&](auto *&p, size_t count, bool set) {
                at = (at + 255) & ~(size_t)255;
                if (set) p = reinterpret_cast<std::remove_reference_t<decltype(p)>>(c->cons + at);
                at += count * sizeof(*p);
            };
            auto layout = [&](bool set) {
                at = 0;
                carve(c->seg, 2 * S, set); carve(c->ncr, S, set); carve(c->sa, S, set);
                carve(c->deg, (size_t)c->maxN, set);
                carve(c->cross, S * kCavMax, set);
                carve(c->ws, std::min<size_t>(S, kWsSlots) * kWs, set);
                carve(c->own, (size_t)c->Tmax, set);
            };
            layout(false);
            if (hipMalloc(&c->cons, at) != hipSuccess) {
                c->cons = nullptr;
                return fail(CVT_ERR_NOMEM, "%zu bytes for %d constraint segments", at,
                            num_segments);
            }
            layout(true);
            c->segCap = num_segments;
        }
        DCHK(hipMemcpyAsync(c->seg, segments, sizeof(int) * 2 * num_segments, hipMemcpyDefault,
                            c->st), "segment upload");
        std::vector<int> hs(2 * (size_t)num_segments);
        DCHK(hipMemcpyAsync(hs.data(), c->seg, sizeof(int) * hs.size(), hipMemcpyDeviceToHost,
                            c->st), "segment check");
        DCHK(hipStreamSynchronize(c->st), "segment check");
        for (int v : hs)
            if (v < 0 || v >= n) return fail(CVT_ERR_ARG, "segment vertex %d outside [0, %d)", v, n);
    }
    hipStream_t st = c->st;
    int *h = c->hcnt;
    for (int i = 0; i < 4; ++i) c->stats[i] = 0;
    DCHK(hipMemcpyAsync(c->P, points, sizeof(Pt) * n, hipMemcpyDefault, st), "point upload");
    hipLaunchKernelGGL(k_dt_fill<int>, dim3(1), dim3(64), 0, st, c->cnt, 0, (long)NUM_C);
    hipLaunchKernelGGL(k_dt_fill<int>, dim3(nb(c->Tmax)), dim3(256), 0, st, c->fstamp, -1,
                       (long)c->Tmax);
    hipLaunchKernelGGL(k_dt_fill<int>, dim3(nb(c->Tmax)), dim3(256), 0, st, c->chk, -1,
                       (long)c->Tmax);
    hipLaunchKernelGGL(k_dt_fill<int>, dim3(nb(c->Tmax)), dim3(256), 0, st, c->fvote, NO_FLIP,
                       (long)c->Tmax);
    hipLaunchKernelGGL(k_dt_ab_part, dim3(kParts), dim3(256), 0, st, c->P, n, c->part, c->cnt);
    hipLaunchKernelGGL(k_dt_ab_final, dim3(1), dim3(kParts), 0, st, c->P, c->part, c->cnt);
    hipLaunchKernelGGL(k_dt_c_part, dim3(kParts), dim3(256), 0, st, c->P, n, c->cnt, c->partv,
                       c->part + 2 * kParts);
    hipLaunchKernelGGL(k_dt_c_final, dim3(1), dim3(kParts), 0, st, c->partv, c->part + 2 * kParts,
                       c->cnt);
    hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, st, c->P, n, c->tv, c->to, c->cnt);
    DCHK(hipMemcpyAsync(h, c->cnt, sizeof(int) * NUM_C, hipMemcpyDeviceToHost, st), "check");
    DCHK(hipStreamSynchronize(st), "first triangle");
    if (h[C_ERR] & DT_ERR_INPUT) return fail(CVT_ERR_ARG, "a point is not finite");
    if (h[C_ERR] & DT_ERR_RANGE)
        return fail(CVT_ERR_ARG, "coordinates span more than 2^71 units of their finest bit");
    if (h[C_ERR]) return fail(CVT_ERR_ARG, "degenerate input: all points collinear or coincident");
    hipLaunchKernelGGL(k_dt_locate0, dim3(nb(n)), dim3(256), 0, st, c->P, n, c->loc, c->cnt);
    int T = 4, round = 0;
    // each round inserts at least one point
    for (int guard = 0; guard <= n; ++guard) {
        hipLaunchKernelGGL(k_dt_fill<unsigned long long>, dim3(nb(T)), dim3(256), 0, st, c->vote,
                           NO_BID, (long)T);
        hipLaunchKernelGGL(k_dt_fill<int>, dim3(1), dim3(64), 0, st, c->cnt + C_UNINS, 0, 1L);
        if (T <= kVoteLds)
            hipLaunchKernelGGL(k_dt_vote_lds, dim3((unsigned)std::min<long>(256, (n + 1023) / 1024)),
                               dim3(1024), 0, st, c->P, n, T, c->tv, c->loc, c->vote, c->cnt);
        else
            hipLaunchKernelGGL(k_dt_vote, dim3(nb(n)), dim3(256), 0, st, c->P, n, c->tv, c->loc,
                               c->vote, c->cnt);
        hipLaunchKernelGGL(k_dt_split_flags, dim3(nb(T)), dim3(256), 0, st, c->vote, T, c->flag);
        launch_exclusive_scan(st, c->flag, T, c->pos);
        DCHK(hipMemcpyAsync(h, c->cnt, sizeof(int) * NUM_C, hipMemcpyDeviceToHost, st), "counters");
        DCHK(hipMemcpyAsync(h + NUM_C, c->pos + T, sizeof(int), hipMemcpyDeviceToHost, st),
             "counters");
        DCHK(hipStreamSynchronize(st), "vote");
        if (h[C_ERR]) break;
        if (h[C_UNINS] == 0) break;
        const int splits = h[NUM_C];
        if (T + 2 * splits > c->Tmax) return fail(CVT_ERR_CAPACITY, "triangle capacity");
        hipLaunchKernelGGL(k_dt_split, dim3(nb(T)), dim3(256), 0, st, c->P, n, T, c->tv, c->to,
                           c->vote, c->pos, c->loc, c->chk, round + 1, c->cnt);
        hipLaunchKernelGGL(k_dt_relocate_split, dim3(nb(n)), dim3(256), 0, st, c->P, n, T, c->tv,
                           c->vote, c->pos, c->loc, c->cnt);
        T += 2 * splits;
        ++c->stats[0];
        const int rc = flip_to_delaunay(c, n, T, round);
        if (rc) return rc;
        if (h[C_ERR]) break;
    }
    if (!h[C_ERR] && num_segments > 0) {
        const int rc = insert_constraints(c, n, T, num_segments, round);
        if (rc) return rc;
    }
    DCHK(hipGetLastError(), "launch");
    c->stats[2] = h[C_FTOT];
    c->stats[3] = h[C_DUP];
    if (h[C_ERR] & DT_ERR_RANGE)
        return fail(CVT_ERR_ARG, "coordinates span more than 2^71 units of their finest bit");
    if (h[C_ERR] & DT_ERR_DEGENERATE)
        return fail(CVT_ERR_ARG, "degenerate input: collinear points where a triangle was needed");
    hipLaunchKernelGGL(k_dt_alive, dim3(nb(T)), dim3(256), 0, st, c->tv, T, n, c->flag);
    launch_exclusive_scan(st, c->flag, T, c->pos);
    DCHK(hipMemcpyAsync(h + NUM_C, c->pos + T, sizeof(int), hipMemcpyDeviceToHost, st), "count");
    DCHK(hipStreamSynchronize(st), "count");
    const int nt = h[NUM_C];
    *num_tri = nt;
    if (nt > capacity) return fail(CVT_ERR_CAPACITY, "%d triangles, capacity %d", nt, capacity);
    // staged, then copied: the caller's arrays may be host or device memory
    hipLaunchKernelGGL(k_dt_output, dim3(nb(T)), dim3(256), 0, st, c->tv, c->to, T, n, c->flag,
                       c->pos, c->otri, opp ? c->oopp : nullptr);
    DCHK(hipGetLastError(), "output launch");
    if (nt > 0) {
        DCHK(hipMemcpyAsync(tri, c->otri, sizeof(int) * 3 * nt, hipMemcpyDefault, st), "download");
        if (opp)
            DCHK(hipMemcpyAsync(opp, c->oopp, sizeof(int) * 3 * nt, hipMemcpyDefault, st), "download");
    }
    DCHK(hipStreamSynchronize(st), "output");
    return CVT_OK;
}

extern "C" int cvt_dt_compute(cvt_dt *dt, const double *points, int num_points, int *tri,
                              int *opp, int capacity, int *num_tri)
{
    return dt_run(dt, points, num_points, nullptr, 0, tri, opp, capacity, num_tri);
}

extern "C" int cvt_dt_compute_constrained(cvt_dt *dt, const double *points, int num_points,
                                          const int *segments, int num_segments, int *tri,
                                          int *opp, int capacity, int *num_tri)
{
    return dt_run(dt, points, num_points, segments, num_segments, tri, opp, capacity, num_tri);
}

extern "C" int cvt_dt_stats(cvt_dt *dt, int *stats)
{
    if (!dt || !stats) return fail(CVT_ERR_ARG, "null argument");
    const DtCtx *c = reinterpret_cast<const DtCtx *>(dt);
    for (int i = 0; i < 4; ++i) stats[i] = c->stats[i];
    return CVT_OK;
}

extern "C" int cvt_dt_predicates(const double *quads, int count, int *orient, int *incircle)
{
    clear_error();
    if (count < 0 || (count > 0 && (!quads || !orient || !incircle)))
        return fail(CVT_ERR_ARG, "null array or negative count");
    if (count == 0) return CVT_OK;
    hipStream_t st = nullptr;
    double *dq = nullptr;
    int *d = nullptr;
    int rc = CVT_OK, err = 0;
    if (hipMalloc(&dq, sizeof(double) * 8 * count) != hipSuccess ||
        hipMalloc(&d, sizeof(int) * (2 * (size_t)count + 1)) != hipSuccess) {
        rc = fail(CVT_ERR_NOMEM, "predicate buffers");
    } else if (hipMemcpyAsync(dq, quads, sizeof(double) * 8 * count, hipMemcpyDefault, st) !=
                   hipSuccess ||
               hipMemsetAsync(d + 2 * count, 0, sizeof(int), st) != hipSuccess) {
        rc = fail(CVT_ERR_HIP, "predicate upload");
    } else {
        hipLaunchKernelGGL(k_dt_predicates, dim3(nb(count)), dim3(256), 0, st, dq, count, d,
                           d + count, d + 2 * count);
        if (hipMemcpyAsync(orient, d, sizeof(int) * count, hipMemcpyDefault, st) != hipSuccess ||
            hipMemcpyAsync(incircle, d + count, sizeof(int) * count, hipMemcpyDefault, st) !=
                hipSuccess ||
            hipMemcpyAsync(&err, d + 2 * count, sizeof(int), hipMemcpyDeviceToHost, st) !=
                hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = fail(CVT_ERR_HIP, "predicates");
        else if (err & DT_ERR_RANGE)
            rc = fail(CVT_ERR_ARG, "coordinates span more than 2^71 units of their finest bit");
    }
    if (dq) (void)hipFree(dq);
    if (d) (void)hipFree(d);
    return rc;
}
