/*
 * cvt_oracle.c -- CPU restatement of the Surface-Remesher CVT hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product (surface-remesher_amd/) never links it.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).  Contraction
 * is off so that every float expression below rounds exactly as written; the
 * distance formula uses an explicit fmaf() so that it rounds identically on
 * the GPU (v_fma_f32) and here.
 *
 * Conventions (shared with the HIP engine):
 *   - grid W x H pixels, row-major, pixel (x, y) at index y*W + x
 *     (reference TOID, source/gcvt.h:26);
 *   - continuous mode: sites are fp32 (x, y) in pixel units, pixel (x, y)
 *     has its centre at (x + 0.5, y + 0.5);
 *   - labels are int32 site indices, ORC_INVALID = 0x7fffffff = no site;
 *   - the nearest-site order is (d, index) lexicographic, with
 *     d = fmaf(dy, dy, dx*dx), dx = sx - (x + 0.5f), dy = sy - (y + 0.5f)
 *     (dx*dx is rounded first, then dy*dy + dx*dx is rounded once).
 *
 * Pinning of each function (what checks it in tests/test_oracle.py):
 *   orc_jfa            -- bit-compared with orc_voronoi_brute on small grids
 *                         (the mismatch is JFA's own error, reported), exact
 *                         on one/two-site closed forms, idempotent;
 *   orc_voronoi_brute  -- the definition (argmin over all sites);
 *   orc_voronoi_exact  -- bit-equal to orc_voronoi_brute (bucketed search);
 *   orc_moments        -- closed forms (uniform square, a x a lattice with
 *                         a^2(a^2-1)/6 per block), mass conservation, and a
 *                         numpy float64 recomputation in the tests;
 *   orc_lloyd          -- energy non-increasing for omega = 1 with the
 *                         exact Voronoi, lattice fixed point;
 *   orc_raster         -- constant / linear vertex weights reproduce the
 *                         constant / linear field, first-hit rule on
 *                         overlapping triangles, degenerate triangles;
 *   orc_ref_rand_*     -- the LCG values 1234567, 3667164066, ... and the
 *                         first uniform-density site (0, 437) at 512;
 *   orc_compat_lloyd   -- update rule on hand-checked cases, fixed sites
 *                         never move, sites never land on rho == 0, merges;
 *   orc_compat_gcvt    -- hand-traced schedules: a lattice fixed point stops
 *                         after exactly 20 iterations; a one-site case whose
 *                         energies, omega = 1.25 after the second decade and
 *                         final position (1, 4) after 30 iterations follow
 *                         from closed forms; a two-level case where omega and
 *                         lastE carried from the coarse level give 30
 *                         iterations and the site (256, 256); depth 2 with
 *                         maxIter 15 runs 16 iterations; the level count.
 *   orc_locate         -- exact weights at the vertices ((1,0,0), (0,1,0),
 *                         (0,0,1)), ~1/3 at centroids, reconstruction
 *                         sum w_j uv_j == q, first face in order on shared
 *                         edges and overlaps, degenerate faces never hit,
 *                         linear 3D fields reproduced by the interpolation.
 *   orc_vertex_weights -- closed forms: a flat copy of the mesh gives 1, a
 *                         mesh scaled by a gives a^2, one tilted by z = k u
 *                         gives sqrt(1 + k^2), a hand-summed two-face fan,
 *                         the > 1e10 rule (largest other weight, 0 when
 *                         there is none), NaN for a vertex in no face.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_INVALID 0x7fffffff

/* ------------------------------------------------------------------------ */
/* distance and order                                                        */
/* ------------------------------------------------------------------------ */

static inline float orc_dist(float sx, float sy, int x, int y)
{
    float dx = sx - ((float)x + 0.5f);
    float dy = sy - ((float)y + 0.5f);
    return fmaf(dy, dy, dx * dx);
}

/* (d, l) < (bd, bl) lexicographically */
static inline int orc_better(float d, int l, float bd, int bl)
{
    return (d < bd) || (d == bd && l < bl);
}

/* ------------------------------------------------------------------------ */
/* JFA(+1)                                                                   */
/* ------------------------------------------------------------------------ */

/* Largest power of two strictly below max(W, H) (at least 1): the first JFA
 * step.  For N = 512 this is 256, the schedule 256, 128, ..., 1. */
int orc_jfa_first_step(int W, int H)
{
    int n = W > H ? W : H;
    int s = 1;
    while (s * 2 < n) s *= 2;
    return s;
}

/* Number of passes: steps s0, s0/2, ..., 1 plus one extra step-1 pass. */
int orc_jfa_num_passes(int W, int H)
{
    int s = orc_jfa_first_step(W, H), p = 1;
    while (s > 1) { s >>= 1; ++p; }
    return p + 1;
}

/* Seed: every site claims the pixel that contains it (coordinates floored
 * and clamped to the grid); on a shared pixel the lowest index wins. */
void orc_seed(int W, int H, const float *xy, int S, int32_t *lab)
{
    for (long i = 0; i < (long)W * H; ++i) lab[i] = ORC_INVALID;
    for (int i = 0; i < S; ++i) {
        int x = (int)floorf(xy[2 * i]), y = (int)floorf(xy[2 * i + 1]);
        x = x < 0 ? 0 : (x >= W ? W - 1 : x);
        y = y < 0 ? 0 : (y >= H ? H - 1 : y);
        long p = (long)y * W + x;
        if (lab[p] > i) lab[p] = i;
    }
}

/* One Jacobi JFA pass: each pixel takes the best of the 9 labels at
 * offsets {-step, 0, step}^2 that fall inside the grid. */
void orc_jfa_pass(int W, int H, const float *xy, int S, const int32_t *in,
                  int32_t *out, int step)
{
    (void)S;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            float bd = INFINITY;
            int bl = ORC_INVALID;
            for (int oy = -1; oy <= 1; ++oy) {
                int yy = y + oy * step;
                if (yy < 0 || yy >= H) continue;
                for (int ox = -1; ox <= 1; ++ox) {
                    int xx = x + ox * step;
                    if (xx < 0 || xx >= W) continue;
                    int l = in[(long)yy * W + xx];
                    if (l == ORC_INVALID) continue;
                    float d = orc_dist(xy[2 * l], xy[2 * l + 1], x, y);
                    if (orc_better(d, l, bd, bl)) { bd = d; bl = l; }
                }
            }
            out[(long)y * W + x] = bl;
        }
    }
}

/* Full JFA(+1): seed, steps s0 .. 1, then one more step-1 pass.  The result
 * is left in lab; tmp is a scratch map of the same size. */
void orc_jfa(int W, int H, const float *xy, int S, int32_t *lab, int32_t *tmp)
{
    orc_seed(W, H, xy, S, lab);
    int32_t *a = lab, *b = tmp;
    int np = orc_jfa_num_passes(W, H);
    int step = orc_jfa_first_step(W, H);
    for (int k = 0; k < np; ++k) {
        orc_jfa_pass(W, H, xy, S, a, b, step);
        int32_t *t = a; a = b; b = t;
        if (step > 1) step >>= 1;
    }
    if (a != lab) memcpy(lab, a, sizeof(int32_t) * (size_t)W * H);
}

/* ------------------------------------------------------------------------ */
/* Exact nearest site (the reference's PBA+ restated by its result,          */
/* source/gcvt.cu:967-978): brute force and a bucketed search that must      */
/* agree with it bit for bit.                                                */
/* ------------------------------------------------------------------------ */

void orc_voronoi_brute(int W, int H, const float *xy, int S, int32_t *lab)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float bd = INFINITY;
            int bl = ORC_INVALID;
            for (int i = 0; i < S; ++i) {
                float d = orc_dist(xy[2 * i], xy[2 * i + 1], x, y);
                if (orc_better(d, i, bd, bl)) { bd = d; bl = i; }
            }
            lab[(long)y * W + x] = bl;
        }
}

void orc_voronoi_exact(int W, int H, const float *xy, int S, int32_t *lab)
{
    if (S <= 0) {
        for (long i = 0; i < (long)W * H; ++i) lab[i] = ORC_INVALID;
        return;
    }
    /* buckets of B x B pixels holding about two sites each */
    double area = (double)W * H;
    int B = (int)ceil(sqrt(2.0 * area / S));
    if (B < 4) B = 4;
    int BW = (W + B - 1) / B, BH = (H + B - 1) / B;
    int nb = BW * BH;
    int *cnt = (int *)calloc((size_t)nb + 1, sizeof(int));
    int *idx = (int *)malloc(sizeof(int) * (size_t)S);
    int *bk = (int *)malloc(sizeof(int) * (size_t)S);
    for (int i = 0; i < S; ++i) {
        int bx = (int)floor(xy[2 * i] / (double)B), by = (int)floor(xy[2 * i + 1] / (double)B);
        bx = bx < 0 ? 0 : (bx >= BW ? BW - 1 : bx);
        by = by < 0 ? 0 : (by >= BH ? BH - 1 : by);
        bk[i] = by * BW + bx;
        cnt[bk[i] + 1]++;
    }
    for (int b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
    int *fill = (int *)malloc(sizeof(int) * (size_t)nb);
    memcpy(fill, cnt, sizeof(int) * (size_t)nb);
    for (int i = 0; i < S; ++i) idx[fill[bk[i]]++] = i;
    free(fill);
    int maxr = BW > BH ? BW : BH;
#pragma omp parallel for schedule(dynamic, 8)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            int bx = x / B, by = y / B;
            float bd = INFINITY;
            int bl = ORC_INVALID;
            for (int r = 0; r <= maxr; ++r) {
                for (int cy = by - r; cy <= by + r; ++cy) {
                    if (cy < 0 || cy >= BH) continue;
                    int ring_row = (cy == by - r) || (cy == by + r);
                    for (int cx = bx - r; cx <= bx + r; ++cx) {
                        if (cx < 0 || cx >= BW) continue;
                        if (!ring_row && cx != bx - r && cx != bx + r) continue;
                        int b = cy * BW + cx;
                        for (int k = cnt[b]; k < cnt[b + 1]; ++k) {
                            int i = idx[k];
                            float d = orc_dist(xy[2 * i], xy[2 * i + 1], x, y);
                            if (orc_better(d, i, bd, bl)) { bd = d; bl = i; }
                        }
                    }
                }
                /* every site in ring r+1 or beyond is at least r*B away
                 * (sites outside the grid sit in the clamped edge buckets,
                 * which only makes them nearer than they are, so the bound
                 * stays a lower bound); the slack covers fp32 rounding of
                 * the formula. */
                if (bl != ORC_INVALID) {
                    double lb = (double)r * B;
                    if (lb * lb > (double)bd * (1.0 + 1e-5) + 1e-3) break;
                }
            }
            lab[(long)y * W + x] = bl;
        }
    free(cnt); free(idx); free(bk);
}

/* ------------------------------------------------------------------------ */
/* Moments, energy, gradient (fp64)                                          */
/* ------------------------------------------------------------------------ */

/* Per-site moments in float64 over the cells of lab:
 *   m[4i+0] = W_i  = sum rho_p
 *   m[4i+1] = DX_i = sum rho_p (px - sx)      (px = x + 0.5)
 *   m[4i+2] = DY_i = sum rho_p (py - sy)
 *   m[4i+3] = E_i  = sum rho_p |p - s|^2
 * The centroid is c_i = s_i + (DX_i, DY_i) / W_i.  This is the reference's
 * (W, X, Y) = (sum rho, sum x rho, sum y rho) of source/gcvt.cu:514-732,
 * taken relative to the site (X - sx W = DX), with the energy term of
 * source/gcvt.cu:788-802 per cell.  Rows are cut into 16 fixed chunks so the
 * float64 sums do not depend on the thread count. */
#define ORC_NCHUNK 16
void orc_moments(int W, int H, const float *rho, const int32_t *lab,
                 const float *xy, int S, double *m)
{
    double *acc = (double *)calloc((size_t)ORC_NCHUNK * S * 4 + 4, sizeof(double));
#pragma omp parallel for schedule(static, 1)
    for (int c = 0; c < ORC_NCHUNK; ++c) {
        int y0 = (int)((long)H * c / ORC_NCHUNK), y1 = (int)((long)H * (c + 1) / ORC_NCHUNK);
        double *a = acc + (size_t)c * S * 4;
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < W; ++x) {
                long p = (long)y * W + x;
                int l = lab[p];
                if (l < 0 || l >= S) continue;
                double r = rho[p];
                if (r == 0.0) continue;
                double dx = ((double)x + 0.5) - (double)xy[2 * l];
                double dy = ((double)y + 0.5) - (double)xy[2 * l + 1];
                a[4 * l + 0] += r;
                a[4 * l + 1] += r * dx;
                a[4 * l + 2] += r * dy;
                a[4 * l + 3] += r * (dx * dx + dy * dy);
            }
    }
    for (long k = 0; k < 4L * S; ++k) {
        double s = 0.0;
        for (int c = 0; c < ORC_NCHUNK; ++c) s += acc[(size_t)c * S * 4 + k];
        m[k] = s;
    }
    free(acc);
}

/* Energy in the reference's normalisation (source/gcvt.cu:788-802 without
 * its reduction bias): E = sum_p rho_p |s(p) - p|^2 / (W H), distances in
 * pixel units.  Gradient norm: dE/ds_i = (2 / (W H)) sum rho (s_i - p), so
 * |grad E| = 2 / (W H) * sqrt(sum_i DX_i^2 + DY_i^2)  (the standard CVT
 * identity dE/ds_i = 2 m_i (s_i - c_i), Du-Faber-Gunzburger 1999). */
void orc_energy_from_moments(int W, int H, const double *m, int S,
                             double *energy, double *gradnorm)
{
    double e = 0.0, g = 0.0;
    for (int i = 0; i < S; ++i) {
        e += m[4 * i + 3];
        g += m[4 * i + 1] * m[4 * i + 1] + m[4 * i + 2] * m[4 * i + 2];
    }
    double a = (double)W * H;
    if (energy) *energy = e / a;
    if (gradnorm) *gradnorm = 2.0 * sqrt(g) / a;
}

/* ------------------------------------------------------------------------ */
/* Lloyd, continuous mode                                                    */
/* ------------------------------------------------------------------------ */

/* Over-relaxed Lloyd update (source/gcvt.cu:753-781 in continuous form):
 * s' = s + omega (c - s) = s + omega DX / W, rounded once to fp32.  Sites
 * with W == 0 (empty or zero-density cells) and fixed sites stay put. */
void orc_update(const double *m, const uint8_t *fixed, int S, float omega,
                const float *xy, float *xy_out)
{
    for (int i = 0; i < S; ++i) {
        double w = m[4 * i + 0];
        if (w > 0.0 && !(fixed && fixed[i])) {
            xy_out[2 * i] = (float)((double)xy[2 * i] + (double)omega * (m[4 * i + 1] / w));
            xy_out[2 * i + 1] = (float)((double)xy[2 * i + 1] + (double)omega * (m[4 * i + 2] / w));
        } else {
            xy_out[2 * i] = xy[2 * i];
            xy_out[2 * i + 1] = xy[2 * i + 1];
        }
    }
}

/* iters Lloyd iterations in place on xy.  Iteration k: labels of the current
 * sites (JFA(+1) if exact == 0, exact nearest site otherwise), energy and
 * gradient norm of the current sites into energy[k] / gradnorm[k], then the
 * update.  lab_out (optional) receives the labels of the last iteration. */
void orc_lloyd(int W, int H, const float *rho, float *xy, const uint8_t *fixed,
               int S, int iters, float omega, int exact, double *energy,
               double *gradnorm, int32_t *lab_out)
{
    size_t np = (size_t)W * H;
    int32_t *lab = (int32_t *)malloc(sizeof(int32_t) * np);
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * np);
    double *m = (double *)malloc(sizeof(double) * 4 * (size_t)(S > 0 ? S : 1));
    float *nxt = (float *)malloc(sizeof(float) * 2 * (size_t)(S > 0 ? S : 1));
    for (int k = 0; k < iters; ++k) {
        if (exact) orc_voronoi_exact(W, H, xy, S, lab);
        else orc_jfa(W, H, xy, S, lab, tmp);
        orc_moments(W, H, rho, lab, xy, S, m);
        orc_energy_from_moments(W, H, m, S, energy ? &energy[k] : NULL,
                                gradnorm ? &gradnorm[k] : NULL);
        orc_update(m, fixed, S, omega, xy, nxt);
        memcpy(xy, nxt, sizeof(float) * 2 * (size_t)S);
    }
    if (lab_out) memcpy(lab_out, lab, sizeof(int32_t) * np);
    free(lab); free(tmp); free(m); free(nxt);
}

/* ------------------------------------------------------------------------ */
/* Rasterizer (source/discretization.cu:28-85, source/discretization.h:69-128)*/
/* ------------------------------------------------------------------------ */

/* Barycentric weights (b1, b2, b3) of (x0, y0) for the triangle
 * (x1, y1), (x2, y2), (x3, y3), by the dot-product formulation of
 * source/discretization.cu:32-55; a zero denominator marks the triangle as
 * degenerate (all weights -1). */
static void orc_bary(double x1, double y1, double x2, double y2, double x3,
                     double y3, double x0, double y0, double *b1, double *b2,
                     double *b3)
{
    double v0x = x2 - x1, v0y = y2 - y1;
    double v1x = x3 - x1, v1y = y3 - y1;
    double v2x = x0 - x1, v2y = y0 - y1;
    double d00 = v0x * v0x + v0y * v0y;
    double d01 = v0x * v1x + v0y * v1y;
    double d11 = v1x * v1x + v1y * v1y;
    double d20 = v2x * v0x + v2y * v0y;
    double d21 = v2x * v1x + v2y * v1y;
    double den = d00 * d11 - d01 * d01;
    if (den == 0) { *b1 = *b2 = *b3 = -1; return; }
    double u = (d11 * d20 - d01 * d21) / den;
    double v = (d00 * d21 - d01 * d20) / den;
    *b2 = u;
    *b3 = v;
    *b1 = 1.0 - u - v;
}

/* density[y*n + x] for the sample point (x*scale, y*scale) in the shifted
 * UV frame: the first triangle (index order) whose three weights are all
 * >= 0 interpolates the vertex weights, else 0. */
void orc_raster(const double *pts, const double *wt, int nv, const int *tri,
                int nt, float *density, double scale, int n)
{
    (void)nv;
#pragma omp parallel for schedule(dynamic, 4)
    for (int ty = 0; ty < n; ++ty)
        for (int tx = 0; tx < n; ++tx) {
            double x = tx * scale, y = ty * scale;
            float res = 0.0f;
            for (int i = 0; i < nt; ++i) {
                int p1 = tri[3 * i], p2 = tri[3 * i + 1], p3 = tri[3 * i + 2];
                double b1, b2, b3;
                orc_bary(pts[2 * p1], pts[2 * p1 + 1], pts[2 * p2], pts[2 * p2 + 1],
                         pts[2 * p3], pts[2 * p3 + 1], x, y, &b1, &b2, &b3);
                if (b1 < 0 || b2 < 0 || b3 < 0) continue;
                res = (float)(b1 * wt[p1] + b2 * wt[p2] + b3 * wt[p3]);
                break;
            }
            density[(long)ty * n + tx] = res;
        }
}

/* ------------------------------------------------------------------------ */
/* Reference host initialisation (source/gcvt.h:58-122), LLP64 semantics     */
/* ------------------------------------------------------------------------ */

/* KISS state; the reference never calls randinit, so all four words start at
 * zero (source/gcvt.h:58) and only the LCG term is live.  unsigned long is 32
 * bits on the reference's MSVC target, so every word wraps at 2^32. */
typedef struct { uint32_t z, w, jsr, jcong; } orc_kiss;

void orc_kiss_init(orc_kiss *k, uint32_t seed)
{
    k->z = k->w = k->jsr = k->jcong = seed;
}

uint32_t orc_kiss_next(orc_kiss *k)
{
    k->z = 36969u * (k->z & 0xffffu) + (k->z >> 16);
    k->w = 18000u * (k->w & 0xffffu) + (k->w >> 16);
    uint32_t mwc = (k->z << 16) + k->w;
    k->jcong = 69069u * k->jcong + 1234567u;
    uint32_t cong = k->jcong;
    k->jsr ^= (k->jsr << 17);
    k->jsr ^= (k->jsr >> 13);
    k->jsr ^= (k->jsr << 5);
    return (mwc ^ cong) + k->jsr;
}

double orc_kiss_uniform(orc_kiss *k)
{
    return (double)orc_kiss_next(k) / 4294967296.0;
}

/* putConstrains (source/gcvt.h:106-122): mask pixels become fixed sites,
 * every other pixel MARKER.  vor holds interleaved (x, y) shorts. */
void orc_ref_put_constraints(int16_t *vor, const uint8_t *mask, int n)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            long p = (long)j * n + i;
            if (mask[p]) { vor[2 * p] = (int16_t)i; vor[2 * p + 1] = (int16_t)j; }
            else vor[2 * p] = vor[2 * p + 1] = (int16_t)-32768;
        }
}

/* randomPoints (source/gcvt.h:76-104): num sites by rejection on free pixels
 * with density > z, z uniform in [0, mx), mx = min(max rho, 100 * mean of
 * the non-zero rho).  Returns the number of RNG draws used. */
long orc_ref_random_points(int16_t *vor, const float *density, int num, int size,
                           orc_kiss *rng)
{
    double mx = 0, avg = 0, cnt = 0;
    for (long i = 0; i < (long)size * size; ++i) {
        if (density[i] > mx) mx = density[i];
        if (density[i] != 0) { cnt += 1; avg += density[i]; }
    }
    double lim = avg / cnt * 100.;
    if (lim < mx) mx = lim;
    long draws = 0;
    for (int i = 0; i < num; ++i) {
        int x, y;
        double z;
        do {
            x = (int)(orc_kiss_uniform(rng) * size);
            y = (int)(orc_kiss_uniform(rng) * size);
            z = orc_kiss_uniform(rng) * mx;
            draws += 3;
        } while (vor[2 * ((long)y * size + x)] != -32768 || density[(long)y * size + x] <= z);
        vor[2 * ((long)y * size + x)] = (int16_t)x;
        vor[2 * ((long)y * size + x) + 1] = (int16_t)y;
    }
    return draws;
}

/* ------------------------------------------------------------------------ */
/* Reference-compatible integer-site mode (source/gcvt.cu:753-781,           */
/* :1087-1156)                                                               */
/* ------------------------------------------------------------------------ */

/* Sites are the pixels whose label equals their own (x, y), in row-major
 * pixel order; a site is fixed when mask is set on its pixel.  Returns the
 * number of sites written to xy (pixel centres, x + 0.5) / fixed. */
int orc_compat_decode(const int16_t *vor, const uint8_t *mask, int n, float *xy,
                      uint8_t *fixed)
{
    int S = 0;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            long p = (long)y * n + x;
            if (vor[2 * p] == x && vor[2 * p + 1] == y) {
                xy[2 * S] = (float)x + 0.5f;
                xy[2 * S + 1] = (float)y + 0.5f;
                fixed[S] = mask ? mask[p] : 0;
                ++S;
            }
        }
    return S;
}

/* One integer-site update (source/gcvt.cu:753-781): a site on a mask pixel
 * is fixed; otherwise its target is t = trunc(s + omega (c - s) + 0.5) per
 * axis (C conversion, toward zero), clamped to [0, n-1], and a target on
 * rho == 0 keeps s.  Sites landing on one pixel merge: the survivor is the
 * lowest old index and survivors keep their relative order (the reference
 * scatters coordinates into a fresh map, so a merge leaves one site).
 * owner is an n*n int32 scratch map.  Returns the new site count. */
int orc_compat_update(int n, const float *rho, const uint8_t *mask,
                      const double *m, int S, float omega, const float *xy,
                      float *xy_out, int32_t *owner)
{
    for (long p = 0; p < (long)n * n; ++p) owner[p] = ORC_INVALID;
    int *tx = (int *)malloc(sizeof(int) * 2 * (size_t)(S > 0 ? S : 1));
    for (int i = 0; i < S; ++i) {
        int sx = (int)floorf(xy[2 * i]), sy = (int)floorf(xy[2 * i + 1]);
        int nx = sx, ny = sy;
        double w = m[4 * i + 0];
        int fixed = mask ? mask[(long)sy * n + sx] : 0;
        if (!fixed && w > 0.0) {
            double vx = (double)sx + (double)omega * (m[4 * i + 1] / w) + 0.5;
            double vy = (double)sy + (double)omega * (m[4 * i + 2] / w) + 0.5;
            int cx = (int)vx, cy = (int)vy;
            cx = cx < 0 ? 0 : (cx > n - 1 ? n - 1 : cx);
            cy = cy < 0 ? 0 : (cy > n - 1 ? n - 1 : cy);
            if (rho[(long)cy * n + cx] != 0.0f) { nx = cx; ny = cy; }
        }
        tx[2 * i] = nx; tx[2 * i + 1] = ny;
        long p = (long)ny * n + nx;
        if (owner[p] > i) owner[p] = i;
    }
    int k = 0;
    for (int i = 0; i < S; ++i) {
        long p = (long)tx[2 * i + 1] * n + tx[2 * i];
        if (owner[p] != i) continue;
        xy_out[2 * k] = (float)tx[2 * i] + 0.5f;
        xy_out[2 * k + 1] = (float)tx[2 * i + 1] + 0.5f;
        ++k;
    }
    free(tx);
    return k;
}

/* Number of pyramid levels gCVT runs for a side n and a requested depth
 * (source/gcvt.cu:1091): the first level whose side falls below 256 and
 * every coarser one are dropped.  The reference takes only powers of two
 * >= 256; for other sides a level is added only while the side halves
 * exactly, and a side below 256 runs one level (the reference would read
 * past the caller's map there). */
int orc_compat_levels(int n, int depth)
{
    int levels = 1;
    while (levels < depth && (n % (1 << levels)) == 0 && (n >> levels) >= 256) ++levels;
    return levels;
}

/* Integer-site CVT with the reference's schedule and pyramid
 * (source/gcvt.cu:1087-1147).  With L levels (orc_compat_levels), level l has
 * side m = n >> l:
 *   - densities: level 0 is rho; level l is the 2x2 mean of level l - 1,
 *     summed in the order (2x, 2y), (2x, 2y+1), (2x+1, 2y), (2x+1, 2y+1) and
 *     divided by 4.0 (kernelDensityScaling, :497-511);
 *   - initial sites: the reference copies the first m_c^2 (x, y) pairs of the
 *     caller's map, m_c = n >> (L-1), into its m_c x m_c map (:1099-1103), so
 *     the initial sites are the self-labelled pixels of that m_c x m_c
 *     row-major prefix (for L = 1 the whole map);
 *   - fixed sites: the update reads mask at the site's pixel index of the
 *     current level, y * m + x (kernelUpdateSites indexes the full-size mask
 *     with the level's TOID, :1033), so level l uses the first m^2 bytes of
 *     mask as an m x m map;
 *   - schedule: one iteration counter `it`, omega, the last sampled energy E
 *     and lastE are shared by all levels (set once at :1105-1108).  On
 *     iterations with it % 10 == 0 the energy of the current map is sampled
 *     before the update and scaled by 4^l (:1082); after every update with
 *     it % 10 == 0: dE = lastE - E, omega = min(2, 1 + dE), stop the level
 *     when dE / 10 < 0.3 (coarse levels) or < 1e-5 (level 0), else
 *     lastE = E.  Each level runs at least one iteration (do ... while
 *     (it < maxIter)), so the total can exceed maxIter by L - 1;
 *   - between levels the site coordinates double (kernelZoomIn, :485-495).
 * The energy is the true float64 value (the reference's float32 reduction
 * is biased, :1070-1078).  One final Voronoi pass labels the result at side n;
 * vor receives the full (x, y) coordinate label map.  exact selects the exact
 * nearest-site map instead of JFA(+1).  Returns the number of iterations;
 * elog (optional, >= maxIter/10 + L + 1 entries) receives the sampled
 * energies (scaled) and *nlog their count. */
int orc_compat_gcvt(int n, const float *rho, int16_t *vor, const uint8_t *mask, int depth,
                    int maxIter, int exact, double *elog, int *nlog, int *num_sites_out)
{
    const int L = orc_compat_levels(n, depth < 1 ? 1 : depth);
    size_t np = (size_t)n * n;
    float *xy = (float *)malloc(sizeof(float) * 2 * np);
    float *xy2 = (float *)malloc(sizeof(float) * 2 * np);
    uint8_t *fx = (uint8_t *)malloc(np);
    int32_t *lab = (int32_t *)malloc(sizeof(int32_t) * np);
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * np);
    double *m = (double *)malloc(sizeof(double) * 4 * np);
    const float *lev[16];
    lev[0] = rho;
    for (int l = 1; l < L; ++l) {
        const int ms = n >> l, w = 2 * ms;
        float *d = (float *)malloc(sizeof(float) * (size_t)ms * ms);
        for (int ty = 0; ty < ms; ++ty)
            for (int tx = 0; tx < ms; ++tx) {
                float s = 0.0f;
                for (int x = 2 * tx; x < 2 * tx + 2; ++x)
                    for (int y = 2 * ty; y < 2 * ty + 2; ++y) s += lev[l - 1][(size_t)y * w + x];
                d[(size_t)ty * ms + tx] = (float)((double)s / 4.0);
            }
        lev[l] = d;
    }
    int S = orc_compat_decode(vor, NULL, n >> (L - 1), xy, fx);
    double E = 0, lastE = 1e18;
    float omega = 2.0f;
    int it = 0, nl = 0;
    for (int l = L - 1; l >= 0; --l) {
        const int ms = n >> l;
        const double escale = ldexp(1.0, 2 * l);
        const double threshold = l ? 3e-1 : 1e-5;
        do {
            if (exact) orc_voronoi_exact(ms, ms, xy, S, lab);
            else orc_jfa(ms, ms, xy, S, lab, tmp);
            orc_moments(ms, ms, lev[l], lab, xy, S, m);
            if (it % 10 == 0) {
                orc_energy_from_moments(ms, ms, m, S, &E, NULL);
                E *= escale;
                if (elog) elog[nl] = E;
                ++nl;
            }
            S = orc_compat_update(ms, lev[l], mask, m, S, omega, xy, xy2, tmp);
            { float *t = xy; xy = xy2; xy2 = t; }
            ++it;
            if (it % 10 == 0) {
                double dE = lastE - E;
                double om = 1.0 + dE;
                omega = (float)(om < 2.0 ? om : 2.0);
                if (dE / 10.0 < threshold) break;
                lastE = E;
            }
        } while (it < maxIter);
        if (l > 0)
            for (int i = 0; i < S; ++i) {
                xy[2 * i] = (float)(2 * (int)floorf(xy[2 * i])) + 0.5f;
                xy[2 * i + 1] = (float)(2 * (int)floorf(xy[2 * i + 1])) + 0.5f;
            }
    }
    if (exact) orc_voronoi_exact(n, n, xy, S, lab);
    else orc_jfa(n, n, xy, S, lab, tmp);
    for (size_t p = 0; p < np; ++p) {
        int l = lab[p];
        if (l == ORC_INVALID) { vor[2 * p] = vor[2 * p + 1] = (int16_t)-32768; continue; }
        vor[2 * p] = (int16_t)floorf(xy[2 * l]);
        vor[2 * p + 1] = (int16_t)floorf(xy[2 * l + 1]);
    }
    if (nlog) *nlog = nl;
    if (num_sites_out) *num_sites_out = S;
    for (int l = 1; l < L; ++l) free((void *)lev[l]);
    free(xy); free(xy2); free(fx); free(lab); free(tmp); free(m);
    return it;
}

/* The single-level form (depth 1) of orc_compat_gcvt. */
int orc_compat_lloyd(int n, const float *rho, int16_t *vor, const uint8_t *mask,
                     int maxIter, int exact, double *elog, int *nlog,
                     int *num_sites_out)
{
    return orc_compat_gcvt(n, rho, vor, mask, 1, maxIter, exact, elog, nlog, num_sites_out);
}

/* ------------------------------------------------------------------------ */
/* CDT input points (source/delaunay.h:48-58)                                */
/* ------------------------------------------------------------------------ */

/* Self-labelled, non-mask pixels in the reference's loop order (x outer,
 * y inner), mapped to (x*scale + l, y*scale + b).  Returns the count. */
long orc_extract_points(const int16_t *vor, const uint8_t *mask, int n,
                        double scale, double l, double b, double *out)
{
    long k = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            long p = (long)j * n + i;
            if (!mask[p] && vor[2 * p] == i && vor[2 * p + 1] == j) {
                out[2 * k] = i * scale + l;
                out[2 * k + 1] = j * scale + b;
                ++k;
            }
        }
    return k;
}

/* ------------------------------------------------------------------------ */
/* Point location (source/recover.h:26-81) and recover()'s vertex step       */
/* (source/recover.h:109-125)                                                 */
/* ------------------------------------------------------------------------ */

/* source/recover.h:30-53: with v0 = p2 - p1, v1 = p3 - p1, v2 = q - p1 and
 * the dot products d00, d01, d11, d20, d21 (x*x + y*y each), den = d00 d11 -
 * d01^2; den == 0 gives -1 weights; w1 = (d11 d20 - d01 d21) / den,
 * w2 = (d00 d21 - d01 d20) / den, and p1's weight is 1 - w1 - w2 (recover()
 * passes its weight array as (w3, w1, w2), :79). */
static void orc_rec_bary(double x1, double y1, double x2, double y2, double x3, double y3,
                         double x0, double y0, double *w0, double *w1, double *w2)
{
    double v0x = x2 - x1, v0y = y2 - y1;
    double v1x = x3 - x1, v1y = y3 - y1;
    double v2x = x0 - x1, v2y = y0 - y1;
    double d00 = v0x * v0x + v0y * v0y;
    double d01 = v0x * v1x + v0y * v1y;
    double d11 = v1x * v1x + v1y * v1y;
    double d20 = v2x * v0x + v2y * v0y;
    double d21 = v2x * v1x + v2y * v1y;
    double den = d00 * d11 - d01 * d01;
    if (den == 0) {
        *w0 = *w1 = *w2 = -1;
        return;
    }
    *w1 = (d11 * d20 - d01 * d21) / den;
    *w2 = (d00 * d21 - d01 * d20) / den;
    *w0 = 1.0 - *w1 - *w2;
}

/* locate() (source/recover.h:63-81): the first face in order whose three
 * weights are all >= 0, by testing every face; -1 (weights -1) for none.  A
 * non-finite point is in no face (the engine's rule; the reference would
 * accept it in the first non-degenerate face).  With p3 and out3, the 3D
 * point sum_j w_j p3[v_j], summed in j order from 0.0 (:117-123), 0 for
 * none. */
void orc_locate(const double *uv, const double *p3, const int *tri, int nf, const double *q,
                int nq, int *face, double *w, double *out3)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int k = 0; k < nq; ++k) {
        double x = q[2 * k], y = q[2 * k + 1];
        int hit = -1;
        double b0 = -1, b1 = -1, b2 = -1;
        if (isfinite(x) && isfinite(y))
            for (int i = 0; i < nf; ++i) {
                const int *t = tri + 3 * i;
                double c0, c1, c2;
                orc_rec_bary(uv[2 * t[0]], uv[2 * t[0] + 1], uv[2 * t[1]], uv[2 * t[1] + 1],
                             uv[2 * t[2]], uv[2 * t[2] + 1], x, y, &c0, &c1, &c2);
                if (c0 < 0 || c1 < 0 || c2 < 0) continue;
                hit = i;
                b0 = c0; b1 = c1; b2 = c2;
                break;
            }
        face[k] = hit;
        if (w) {
            w[3 * k] = b0;
            w[3 * k + 1] = b1;
            w[3 * k + 2] = b2;
        }
        if (p3 && out3) {
            double r[3] = {0.0, 0.0, 0.0};
            if (hit >= 0) {
                const int *t = tri + 3 * hit;
                const double wt[3] = {b0, b1, b2};
                for (int c = 0; c < 3; ++c)
                    for (int j = 0; j < 3; ++j) r[c] += p3[3 * t[j] + c] * wt[j];
            }
            out3[3 * k] = r[0];
            out3[3 * k + 1] = r[1];
            out3[3 * k + 2] = r[2];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Vertex density weights                                                    */
/* ------------------------------------------------------------------------ */

/* The 3D face area as CGAL 5.0.2 computes it (PMP::face_area: the square root
 * of Compute_squared_area_3 on source(h), target(h), target(next(h)) of the
 * face's halfedge h; the squared area is the squared length of the cross
 * product of q - p and r - p, over 4).  A face listed as vertices_around_face
 * (target(h), target(next(h)), source(h)) = (t0, t1, t2) hands it the points
 * (p, q, r) = (t2, t0, t1). */
static double orc_area3(const double *p3, const int *t)
{
    const double *p = p3 + 3 * t[2], *q = p3 + 3 * t[0], *r = p3 + 3 * t[1];
    double ux = q[0] - p[0], uy = q[1] - p[1], uz = q[2] - p[2];
    double vx = r[0] - p[0], vy = r[1] - p[1], vz = r[2] - p[2];
    double cx = uy * vz - uz * vy;
    double cy = uz * vx - ux * vz;
    double cz = ux * vy - uy * vx;
    return sqrt((cx * cx + cy * cy + cz * cz) / 4);
}

/* get_area (source/weighting.h:29-35): |cross(B - A, C - A)| / 2 with
 * cross(P, Q) = P.x Q.y - Q.x P.y, (A, B, C) the face's vertices in order. */
static double orc_area2(const double *uv, const int *t)
{
    const double *a = uv + 2 * t[0], *b = uv + 2 * t[1], *c = uv + 2 * t[2];
    double px = b[0] - a[0], py = b[1] - a[1];
    double qx = c[0] - a[0], qy = c[1] - a[1];
    return fabs(px * qy - qx * py) / 2.;
}

/* weighting_vertices_with_area (source/weighting.h:37-97) on arrays: each
 * face adds its 3D area and its 2D area to each of its corners' sums (in
 * face order, :53-74); a vertex's weight is area3 / area2 (:79); weights
 * above 1e10 are replaced by the largest other weight, which starts at 0
 * (:76-92).  A vertex in no face gets 0 / 0 = NaN, which takes part in
 * neither test.  Both sums run in the order of tri (the reference sums the
 * 3D areas in the seam mesh's face order and the 2D ones in mesh_2d's). */
void orc_vertex_weights(const double *uv, const double *p3, int nv, const int *tri, int nf,
                        double *weight)
{
    double *a3 = (double *)calloc((size_t)(nv > 0 ? nv : 1), sizeof(double));
    double *a2 = (double *)calloc((size_t)(nv > 0 ? nv : 1), sizeof(double));
    for (int f = 0; f < nf; ++f) {
        const int *t = tri + 3 * f;
        const double s3 = orc_area3(p3, t), s2 = orc_area2(uv, t);
        for (int c = 0; c < 3; ++c) a3[t[c]] += s3;
        for (int c = 0; c < 3; ++c) a2[t[c]] += s2;
    }
    double big = 1e10, mx = 0;
    for (int v = 0; v < nv; ++v) {
        double w = a3[v] / a2[v];
        weight[v] = w;
        if (w > big) continue;
        if (w > mx) mx = w;
    }
    for (int v = 0; v < nv; ++v)
        if (weight[v] > big) weight[v] = mx;
    free(a3);
    free(a2);
}

/* ------------------------------------------------------------------------ */
/* Misc                                                                      */
/* ------------------------------------------------------------------------ */

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int t)
{
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}
