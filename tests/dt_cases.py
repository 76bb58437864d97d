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
