"""Predicate test records and their exact signs (rational arithmetic), shared
by the host and device predicate tests."""
from fractions import Fraction as F

import numpy as np


def exact_orient(r):
    ax, ay, bx, by, cx, cy = (F(float(v)) for v in r[:6])
    d = (ax - cx) * (by - cy) - (ay - cy) * (bx - cx)
    return (d > 0) - (d < 0)


def exact_incircle(r):
    ax, ay, bx, by, cx, cy, dx, dy = (F(float(v)) for v in r)
    adx, ady, bdx, bdy, cdx, cdy = ax - dx, ay - dy, bx - dx, by - dy, cx - dx, cy - dy
    d = ((adx * adx + ady * ady) * (bdx * cdy - cdx * bdy) +
         (bdx * bdx + bdy * bdy) * (cdx * ady - adx * cdy) +
         (cdx * cdx + cdy * cdy) * (adx * bdy - bdx * ady))
    return (d > 0) - (d < 0)


def records(seed=0, n_random=300):
    """random, near-collinear, near-cocircular and exactly degenerate records,
    at magnitudes from 1e-12 to 1e12 (all within the 2^71 span per record)"""
    g = np.random.default_rng(seed)
    out = []
    out.append(g.random((n_random, 8)) * 2048.0)
    out.append((g.random((n_random, 8)) - 0.5).astype(np.float32).astype(np.float64))
    # exactly collinear a, b, c (x = y lines and lattice rows) and 1-ulp nudges
    for k in range(60):
        t = g.random(3) * 1000.0
        r = np.array([t[0], t[0], t[1], t[1], t[2], t[2], 0.0, 0.0])
        out.append(r[None])
        r2 = r.copy()
        r2[5] = np.nextafter(r2[5], np.inf if k % 2 else -np.inf)
        out.append(r2[None])
        base = np.floor(g.random() * 1e6)
        out.append(np.array([[base, 7.0, base + 3.0, 7.0, base + 11.0, 7.0, base, 8.0]]))
    # cocircular: the four corners of an axis-aligned rectangle, then nudged
    for k in range(80):
        x0, y0 = g.random(2) * 1e3
        w, h = g.random(2) * 50.0 + 1.0
        x0, y0, w, h = (float(np.float32(v)) for v in (x0, y0, w, h))
        r = np.array([x0, y0, x0 + w, y0, x0 + w, y0 + h, x0, y0 + h])
        out.append(r[None])
        r2 = r.copy()
        r2[6 + (k % 2)] = np.nextafter(r2[6 + (k % 2)], np.inf if (k // 2) % 2 else -np.inf)
        out.append(r2[None])
    # scale: tiny and large coordinates, and mixed magnitudes
    for s in (1e-12, 1e-3, 1e6, 1e12):
        out.append((g.random((40, 8)) * s))
    m = g.random((40, 8)) * 100.0
    m[:, ::3] *= 1e-9
    out.append(m)
    return np.concatenate(out, axis=0)


def _poly_mul(a, b):
    out = {}
    for ea, ca in a.items():
        for eb, cb in b.items():
            out[ea + eb] = out.get(ea + eb, 0) + ca * cb
    return out


def _poly_det(M):
    """determinant of a matrix of polynomials {exponent: coefficient}"""
    import itertools
    k = len(M)
    total = {}
    for perm in itertools.permutations(range(k)):
        inv = sum(1 for i in range(k) for j in range(i + 1, k) if perm[i] > perm[j])
        term = {0: -1 if inv % 2 else 1}
        for r in range(k):
            term = _poly_mul(term, M[r][perm[r]])
        for e, c in term.items():
            total[e] = total.get(e, 0) + c
    return total


def _leading_sign(poly):
    for e in sorted(poly):
        if poly[e] != 0:
            return 1 if poly[e] > 0 else -1
    return 0


def sos_orient(r, idx):
    """orient2d of points a, b, c (record r) moved by eps^(2^m): m(i, y) = 3i,
    m(i, x) = 3i + 1 with i the rank of the point's index; the sign of the
    lowest-order non-zero term of the perturbed determinant (eps -> 0+)"""
    rank = {v: k for k, v in enumerate(sorted(idx))}
    M = []
    for j in range(3):
        i = rank[idx[j]]
        x, y = F(float(r[2 * j])), F(float(r[2 * j + 1]))
        M.append([{0: x, 2 ** (3 * i + 1): F(1)}, {0: y, 2 ** (3 * i): F(1)}, {0: F(1)}])
    return _leading_sign(_poly_det(M))


def sos_incircle(r, idx):
    """incircle of a, b, c, d (record r) as the 4x4 determinant of rows
    [x y x^2+y^2 1], each coordinate moved by eps^(2^m), m(i, z) = 3i + 2"""
    rank = {v: k for k, v in enumerate(sorted(idx))}
    M = []
    for j in range(4):
        i = rank[idx[j]]
        x, y = F(float(r[2 * j])), F(float(r[2 * j + 1]))
        M.append([{0: x, 2 ** (3 * i + 1): F(1)}, {0: y, 2 ** (3 * i): F(1)},
                  {0: x * x + y * y, 2 ** (3 * i + 2): F(1)}, {0: F(1)}])
    return _leading_sign(_poly_det(M))


def degenerate_records(seed=0):
    """exactly collinear triples and cocircular quadruples (lattice squares,
    rectangles, points of one circle) with distinct points"""
    g = np.random.default_rng(seed)
    out = []
    for _ in range(40):
        x0, y0 = (float(v) for v in np.floor(g.random(2) * 100))
        dx, dy = (float(v) for v in np.floor(g.random(2) * 5) + 1)
        out.append([x0, y0, x0 + dx, y0 + dy, x0 + 2 * dx, y0 + 2 * dy, x0 + dx, y0])
        out.append([x0, y0, x0 + dx, y0, x0 + dx, y0 + dy, x0, y0 + dy])       # rectangle
        out.append([x0, y0, x0 + 3 * dx, y0, x0 + dx, y0, x0, y0 + dy])        # collinear row
    # Pythagorean points on one circle of radius 5 about (c, c)
    ring = [(5, 0), (4, 3), (3, 4), (0, 5), (-3, 4), (-4, 3), (-5, 0), (-4, -3), (0, -5), (3, -4)]
    for _ in range(40):
        c = float(np.floor(g.random() * 50))
        pick = g.choice(len(ring), 4, replace=False)
        q = []
        for k in pick:
            q += [c + ring[k][0], c + ring[k][1]]
        out.append(q)
    for _ in range(40):
        # vertical collinear triples; four points of one line (a degenerate
        # "circle"); mirrored ring points about the origin (equal x and z)
        x0, y0 = (float(v) for v in np.floor(g.random(2) * 100))
        d = float(np.floor(g.random() * 4) + 1)
        out.append([x0, y0, x0, y0 + d, x0, y0 + 3 * d, x0 + d, y0])
        out.append([x0, y0, x0 + d, y0 + 2 * d, x0 + 3 * d, y0 + 6 * d, x0 + 2 * d, y0 + 4 * d])
        out.append([x0, y0, x0, y0 + 2 * d, x0, y0 + 5 * d, x0, y0 + d])
        pick = g.choice(len(ring), 4, replace=False)
        out.append([float(v) for k in pick for v in ring[k]])
    return np.array(out, np.float64)
