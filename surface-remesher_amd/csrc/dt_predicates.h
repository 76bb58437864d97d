// dt_predicates.h -- exact orientation and in-circle signs for float64 points.
//
// Each predicate is evaluated in float64 first, with an error bound that
// decides whether the rounded sign is certain (the filter of Shewchuk's
// adaptive predicates, "Adaptive Precision Floating-Point Arithmetic and Fast
// Robust Geometric Predicates", 1997, with its first-stage bounds doubled).
// When it is not, the sign is computed exactly: every coordinate involved is
// written as m * 2^e with m an odd 53-bit integer, all of them are scaled to
// the smallest e among them, and the determinant is evaluated in fixed-width
// two's-complement integers (int128 differences, 256-bit products, 512-bit
// in-circle terms).  That is exact when the scaled coordinates fit in 124
// bits, i.e. when the coordinates of one predicate span at most 2^71 units
// of their finest bit (any float32 input in pixel units does); otherwise the
// predicate reports a range error through *err and returns 0.
//
// The same source compiles for the host (g++) and for gfx950 (hipcc).  err,
// when not null, must point to global memory on the device.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define DT_HD __host__ __device__ inline
#define DT_HD_COLD __host__ __device__ inline __attribute__((noinline))
#else
#define DT_HD inline
#define DT_HD_COLD inline __attribute__((noinline))
#endif

namespace cvt {
namespace dt {

enum { DT_ERR_DEGENERATE = 1, DT_ERR_RANGE = 2, DT_ERR_CAPACITY = 4, DT_ERR_INPUT = 8,
       DT_ERR_CROSS = 16 };

typedef unsigned __int128 u128;
typedef __int128 i128;

// sets error bits; on the device err is global memory shared by all lanes
DT_HD void set_flag(int *err, int bits)
{
    if (!err) return;
#if defined(__HIP_DEVICE_COMPILE__)
    atomicOr(err, bits);
#else
    *err |= bits;
#endif
}

template <int N> struct Big {
    uint64_t w[N];                                  // two's complement, limb 0 lowest
};

template <int N> DT_HD Big<N> big_from(i128 v)
{
    Big<N> r;
    r.w[0] = (uint64_t)v;
    r.w[1] = (uint64_t)(v >> 64);
    const uint64_t ext = v < 0 ? ~0ull : 0ull;
    for (int i = 2; i < N; ++i) r.w[i] = ext;
    return r;
}

template <int N> DT_HD bool big_neg(const Big<N> &a) { return (a.w[N - 1] >> 63) != 0; }

template <int N> DT_HD Big<N> big_negate(const Big<N> &a)
{
    Big<N> r;
    uint64_t carry = 1;
    for (int i = 0; i < N; ++i) {
        const u128 s = (u128)(~a.w[i]) + carry;
        r.w[i] = (uint64_t)s;
        carry = (uint64_t)(s >> 64);
    }
    return r;
}

template <int N> DT_HD Big<N> big_add(const Big<N> &a, const Big<N> &b)
{
    Big<N> r;
    uint64_t carry = 0;
    for (int i = 0; i < N; ++i) {
        const u128 s = (u128)a.w[i] + b.w[i] + carry;
        r.w[i] = (uint64_t)s;
        carry = (uint64_t)(s >> 64);
    }
    return r;
}

template <int N> DT_HD Big<N> big_sub(const Big<N> &a, const Big<N> &b)
{
    return big_add(a, big_negate(b));
}

template <int N> DT_HD int big_sign(const Big<N> &a)
{
    if (big_neg(a)) return -1;
    for (int i = 0; i < N; ++i)
        if (a.w[i]) return 1;
    return 0;
}

// signed N-limb x signed N-limb -> 2N limbs (the caller keeps |a|, |b| small
// enough that the product fits 2N limbs with its sign)
template <int N> DT_HD Big<2 * N> big_mul(Big<N> a, Big<N> b)
{
    const bool na = big_neg(a), nb = big_neg(b);
    if (na) a = big_negate(a);
    if (nb) b = big_negate(b);
    Big<2 * N> r;
    for (int i = 0; i < 2 * N; ++i) r.w[i] = 0;
    for (int i = 0; i < N; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < N; ++j) {
            const u128 p = (u128)a.w[i] * b.w[j] + r.w[i + j] + carry;
            r.w[i + j] = (uint64_t)p;
            carry = (uint64_t)(p >> 64);
        }
        r.w[i + N] = carry;
    }
    return na != nb ? big_negate(r) : r;
}

// x = m * 2^e with m odd (or m = 0, e = 0)
struct Fx {
    int64_t m;
    int e;
};

DT_HD Fx decompose(double x)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const int be = (int)((u >> 52) & 0x7ff);
    uint64_t f = u & ((1ull << 52) - 1);
    int e;
    if (be == 0) {
        e = -1074;                                  // subnormal or zero
    } else {
        f |= 1ull << 52;
        e = be - 1075;
    }
    Fx r;
    if (f == 0) {
        r.m = 0;
        r.e = 0;
        return r;
    }
    const int tz = __builtin_ctzll(f);
    f >>= tz;
    r.m = (u >> 63) ? -(int64_t)f : (int64_t)f;
    r.e = e + tz;
    return r;
}

// scales k coordinates to integers at their common finest bit; false if one
// would need more than 124 bits
template <int K> DT_HD bool to_integers(const double (&x)[K], i128 (&X)[K])
{
    Fx f[K];
    int emin = 1 << 30;
    for (int i = 0; i < K; ++i) {
        f[i] = decompose(x[i]);
        if (f[i].m != 0 && f[i].e < emin) emin = f[i].e;
    }
    for (int i = 0; i < K; ++i) {
        if (f[i].m == 0) {
            X[i] = 0;
            continue;
        }
        const int s = f[i].e - emin;                // m has at most 53 bits
        if (s > 71) return false;
        X[i] = (i128)f[i].m * ((i128)1 << s);
    }
    return true;
}

// the exact stages are rarely taken: kept out of line so that the callers'
// filters stay small
DT_HD_COLD int orient_exact(double ax, double ay, double bx, double by, double cx, double cy, int *err)
{
    const double x[6] = {ax, ay, bx, by, cx, cy};
    i128 X[6];
    if (!to_integers<6>(x, X)) {
        set_flag(err, DT_ERR_RANGE);
        return 0;
    }
    const i128 acx = X[0] - X[4], acy = X[1] - X[5], bcx = X[2] - X[4], bcy = X[3] - X[5];
    const Big<4> l = big_mul<2>(big_from<2>(acx), big_from<2>(bcy));
    const Big<4> r = big_mul<2>(big_from<2>(acy), big_from<2>(bcx));
    return big_sign(big_sub(l, r));
}

// sign of (a - c) x (b - c): > 0 when a, b, c turn counterclockwise
DT_HD int orient2d(double ax, double ay, double bx, double by, double cx, double cy, int *err)
{
    const double l = (ax - cx) * (by - cy);
    const double r = (ay - cy) * (bx - cx);
    const double det = l - r;
    const double bound = 6.7e-16 * (__builtin_fabs(l) + __builtin_fabs(r));
    // the bound holds away from underflow and overflow; NaN and inf fail the
    // comparisons and go to the exact stage, which decides or reports
    if ((det > bound || -det > bound) && bound > 1e-290 && bound < 1e290)
        return det > 0 ? 1 : -1;
    return orient_exact(ax, ay, bx, by, cx, cy, err);
}

DT_HD_COLD int incircle_exact(const double (&x)[8], int *err)
{
    i128 X[8];
    if (!to_integers<8>(x, X)) {
        set_flag(err, DT_ERR_RANGE);
        return 0;
    }
    const i128 adx = X[0] - X[6], ady = X[1] - X[7];
    const i128 bdx = X[2] - X[6], bdy = X[3] - X[7];
    const i128 cdx = X[4] - X[6], cdy = X[5] - X[7];
    auto sq = [](i128 u, i128 v) {
        return big_add(big_mul<2>(big_from<2>(u), big_from<2>(u)),
                       big_mul<2>(big_from<2>(v), big_from<2>(v)));
    };
    auto cross = [](i128 u1, i128 v1, i128 u2, i128 v2) {   // u1 v2 - u2 v1
        return big_sub(big_mul<2>(big_from<2>(u1), big_from<2>(v2)),
                       big_mul<2>(big_from<2>(u2), big_from<2>(v1)));
    };
    const Big<4> al = sq(adx, ady), bl = sq(bdx, bdy), cl = sq(cdx, cdy);
    const Big<4> bc = cross(bdx, bdy, cdx, cdy), ca = cross(cdx, cdy, adx, ady),
                 ab = cross(adx, ady, bdx, bdy);
    const Big<8> det = big_add(big_add(big_mul<4>(al, bc), big_mul<4>(bl, ca)), big_mul<4>(cl, ab));
    return big_sign(det);
}

// > 0 when d lies inside the circle through a, b, c (counterclockwise), < 0
// outside, 0 on it
DT_HD int incircle(double ax, double ay, double bx, double by, double cx, double cy, double dx,
                   double dy, int *err)
{
    const double adx = ax - dx, ady = ay - dy, bdx = bx - dx, bdy = by - dy;
    const double cdx = cx - dx, cdy = cy - dy;
    const double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const double cdxady = cdx * ady, adxcdy = adx * cdy;
    const double adxbdy = adx * bdy, bdxady = bdx * ady;
    const double alift = adx * adx + ady * ady;
    const double blift = bdx * bdx + bdy * bdy;
    const double clift = cdx * cdx + cdy * cdy;
    const double det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) +
                       clift * (adxbdy - bdxady);
    const double perm = (__builtin_fabs(bdxcdy) + __builtin_fabs(cdxbdy)) * alift +
                        (__builtin_fabs(cdxady) + __builtin_fabs(adxcdy)) * blift +
                        (__builtin_fabs(adxbdy) + __builtin_fabs(bdxady)) * clift;
    const double bound = 2.3e-15 * perm;
    if ((det > bound || -det > bound) && bound > 1e-290 && bound < 1e290)
        return det > 0 ? 1 : -1;
    const double x[8] = {ax, ay, bx, by, cx, cy, dx, dy};
    return incircle_exact(x, err);
}


// ---------------------------------------------------------------------------
// Simulation of Simplicity (Edelsbrunner and Muecke, ACM TOG 9(1), 1990):
// where a predicate is exactly zero, every point i is taken as moved by
// infinitesimals, coordinate c by eps^(2^m(i, c)) with m(i, y) = 3i,
// m(i, x) = 3i + 1 and, for the in-circle lifting z = x^2 + y^2,
// m(i, z) = 3i + 2 (lower indices move more; y before x before z).  The
// perturbed determinant is a polynomial in eps whose terms, largest first,
// have as coefficients +-minors of the unperturbed matrix; the first
// non-zero one gives the sign.  For points sorted by index the order of
// the terms is fixed (the tables below: sign, rows kept, columns kept), so
// no three points are ever collinear and no four cocircular, and
// orientation and in-circle answer for the same perturbed points.
// ---------------------------------------------------------------------------
enum { SOS_ORIENT = 0, SOS_DIFF = 1, SOS_ONE = 2 };
enum { PL_XZ = 0, PL_YZ = 1, PL_XY = 2 };      // planes for SOS_ORIENT
enum { CO_X = 0, CO_Y = 1, CO_Z = 2 };         // coordinates for SOS_DIFF

struct SosTerm {
    signed char sign, kind, r0, r1, r2, what;
};

// orientation of rows (0, 1, 2) of [x y 1] after the determinant itself
static constexpr SosTerm kOrientSos[] = {
    {-1, SOS_DIFF, 1, 2, 0, CO_X},
    {+1, SOS_DIFF, 1, 2, 0, CO_Y},
    {+1, SOS_DIFF, 0, 2, 0, CO_X},
    {+1, SOS_ONE, 0, 0, 0, 0},
};

// in-circle of rows (0, 1, 2, 3) of [x y z 1] after the determinant (terms
// whose minor repeats an earlier one are left out: they are zero when
// reached)
static constexpr SosTerm kIncircleSos[] = {
    {-1, SOS_ORIENT, 1, 2, 3, PL_XZ},
    {+1, SOS_ORIENT, 1, 2, 3, PL_YZ},
    {+1, SOS_ORIENT, 1, 2, 3, PL_XY},
    {+1, SOS_ORIENT, 0, 2, 3, PL_XZ},
    {+1, SOS_DIFF, 2, 3, 0, CO_Z},
    {-1, SOS_DIFF, 2, 3, 0, CO_X},
    {-1, SOS_ORIENT, 0, 2, 3, PL_YZ},
    {+1, SOS_DIFF, 2, 3, 0, CO_Y},
    {-1, SOS_ORIENT, 0, 2, 3, PL_XY},
    {-1, SOS_ORIENT, 0, 1, 3, PL_XZ},
    {-1, SOS_DIFF, 1, 3, 0, CO_Z},
    {+1, SOS_DIFF, 1, 3, 0, CO_X},
    {+1, SOS_DIFF, 0, 3, 0, CO_Z},
    {+1, SOS_ONE, 0, 0, 0, 0},
};

// sorts k indices ascending, permuting the coordinate pairs alongside;
// returns the parity of the permutation (+1 / -1)
template <int K> DT_HD int sort_by_index(int (&idx)[K], double (&x)[2 * K])
{
    int par = 1;
    for (int i = 1; i < K; ++i)
        for (int j = i; j > 0 && idx[j - 1] > idx[j]; --j) {
            const int ti = idx[j]; idx[j] = idx[j - 1]; idx[j - 1] = ti;
            const double tx = x[2 * j], ty = x[2 * j + 1];
            x[2 * j] = x[2 * j - 2]; x[2 * j + 1] = x[2 * j - 1];
            x[2 * j - 2] = tx; x[2 * j - 1] = ty;
            par = -par;
        }
    return par;
}

DT_HD_COLD int orient_sos(const double (&x)[6])
{
    for (const SosTerm &t : kOrientSos) {
        if (t.kind == SOS_ONE) return t.sign;
        const double a = x[2 * t.r0 + t.what], b = x[2 * t.r1 + t.what];
        if (a != b) return a > b ? t.sign : -t.sign;
    }
    return 1;
}

// exact sign of a 2x2 orientation in a plane of (X, Y, Z) among rows r0..r2
DT_HD int lifted_orient(const Big<4> (&U)[4][3], int r0, int r1, int r2, int plane)
{
    const int ca = plane == PL_YZ ? 1 : 0, cb = plane == PL_XY ? 1 : 2;
    const Big<4> a0 = big_sub(U[r0][ca], U[r2][ca]), a1 = big_sub(U[r1][ca], U[r2][ca]);
    const Big<4> b0 = big_sub(U[r0][cb], U[r2][cb]), b1 = big_sub(U[r1][cb], U[r2][cb]);
    return big_sign(big_sub(big_mul<4>(a0, b1), big_mul<4>(b0, a1)));
}

DT_HD_COLD int incircle_sos(const double (&x)[8], int *err)
{
    i128 X[8];
    if (!to_integers<8>(x, X)) {
        set_flag(err, DT_ERR_RANGE);
        return 0;
    }
    // columns x, y, z = x^2 + y^2 at the common scale, as 256-bit integers
    Big<4> U[4][3];
    for (int r = 0; r < 4; ++r) {
        U[r][0] = big_from<4>(X[2 * r]);
        U[r][1] = big_from<4>(X[2 * r + 1]);
        const Big<4> xx = big_mul<2>(big_from<2>(X[2 * r]), big_from<2>(X[2 * r]));
        const Big<4> yy = big_mul<2>(big_from<2>(X[2 * r + 1]), big_from<2>(X[2 * r + 1]));
        U[r][2] = big_add(xx, yy);
    }
    for (const SosTerm &t : kIncircleSos) {
        if (t.kind == SOS_ONE) return t.sign;
        int s;
        if (t.kind == SOS_DIFF)
            s = big_sign(big_sub(U[t.r0][t.what], U[t.r1][t.what]));
        else
            s = lifted_orient(U, t.r0, t.r1, t.r2, t.what);
        if (s) return s * t.sign;
    }
    return 1;
}

// orient2d and incircle of indexed points, never zero for distinct points
// (an exactly degenerate case is decided by the perturbation above)
DT_HD int orient2d_sos(double ax, double ay, double bx, double by, double cx, double cy, int ia,
                       int ib, int ic, int *err)
{
    const int o = orient2d(ax, ay, bx, by, cx, cy, err);
    if (o) return o;
    int idx[3] = {ia, ib, ic};
    double x[6] = {ax, ay, bx, by, cx, cy};
    const int par = sort_by_index<3>(idx, x);
    return par * orient_sos(x);
}

DT_HD int incircle_sos(double ax, double ay, double bx, double by, double cx, double cy,
                       double dx, double dy, int ia, int ib, int ic, int id, int *err)
{
    const int o = incircle(ax, ay, bx, by, cx, cy, dx, dy, err);
    if (o) return o;
    int idx[4] = {ia, ib, ic, id};
    double x[8] = {ax, ay, bx, by, cx, cy, dx, dy};
    const int par = sort_by_index<4>(idx, x);
    return par * incircle_sos(x, err);
}

}  // namespace dt
}  // namespace cvt
