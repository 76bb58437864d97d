"""ctypes loader for the CPU oracle (oracle/cvt_oracle.c).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, never by the product package.  Every function
takes and returns numpy arrays; see cvt_oracle.c for the conventions and for
the reference lines each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
INVALID = 0x7FFFFFFF

_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i16p = np.ctypeslib.ndpointer(np.int16, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class Kiss(C.Structure):
    _fields_ = [("z", C.c_uint32), ("w", C.c_uint32), ("jsr", C.c_uint32), ("jcong", C.c_uint32)]


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    L = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    sig = {
        "orc_jfa_first_step": (C.c_int, [C.c_int, C.c_int]),
        "orc_jfa_num_passes": (C.c_int, [C.c_int, C.c_int]),
        "orc_seed": (None, [C.c_int, C.c_int, _f32p, C.c_int, _i32p]),
        "orc_jfa_pass": (None, [C.c_int, C.c_int, _f32p, C.c_int, _i32p, _i32p, C.c_int]),
        "orc_jfa": (None, [C.c_int, C.c_int, _f32p, C.c_int, _i32p, _i32p]),
        "orc_voronoi_brute": (None, [C.c_int, C.c_int, _f32p, C.c_int, _i32p]),
        "orc_voronoi_exact": (None, [C.c_int, C.c_int, _f32p, C.c_int, _i32p]),
        "orc_moments": (None, [C.c_int, C.c_int, _f32p, _i32p, _f32p, C.c_int, _f64p]),
        "orc_energy_from_moments": (None, [C.c_int, C.c_int, _f64p, C.c_int, vp, vp]),
        "orc_update": (None, [_f64p, vp, C.c_int, C.c_float, _f32p, _f32p]),
        "orc_lloyd": (None, [C.c_int, C.c_int, _f32p, _f32p, vp, C.c_int, C.c_int, C.c_float,
                              C.c_int, vp, vp, vp]),
        "orc_raster": (None, [_f64p, _f64p, C.c_int, _i32p, C.c_int, _f32p, C.c_double, C.c_int]),
        "orc_kiss_init": (None, [C.POINTER(Kiss), C.c_uint32]),
        "orc_kiss_next": (C.c_uint32, [C.POINTER(Kiss)]),
        "orc_kiss_uniform": (C.c_double, [C.POINTER(Kiss)]),
        "orc_ref_put_constraints": (None, [_i16p, _u8p, C.c_int]),
        "orc_ref_random_points": (C.c_long, [_i16p, _f32p, C.c_int, C.c_int, C.POINTER(Kiss)]),
        "orc_compat_decode": (C.c_int, [_i16p, vp, C.c_int, _f32p, _u8p]),
        "orc_compat_update": (C.c_int, [C.c_int, _f32p, vp, _f64p, C.c_int, C.c_float, _f32p,
                                        _f32p, _i32p]),
        "orc_compat_lloyd": (C.c_int, [C.c_int, _f32p, _i16p, vp, C.c_int, C.c_int, vp,
                                       C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "orc_compat_levels": (C.c_int, [C.c_int, C.c_int]),
        "orc_compat_gcvt": (C.c_int, [C.c_int, _f32p, _i16p, vp, C.c_int, C.c_int, C.c_int, vp,
                                      C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "orc_extract_points": (C.c_long, [_i16p, _u8p, C.c_int, C.c_double, C.c_double,
                                          C.c_double, _f64p]),
        "orc_locate": (None, [_f64p, vp, _i32p, C.c_int, _f64p, C.c_int, _i32p, vp, vp]),
        "orc_vertex_weights": (None, [_f64p, _f64p, C.c_int, _i32p, C.c_int, _f64p]),
        "orc_max_threads": (C.c_int, []),
        "orc_set_threads": (None, [C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _sites(xy):
    xy = np.ascontiguousarray(xy, dtype=np.float32).reshape(-1, 2)
    return xy, xy.shape[0]


def jfa_first_step(W, H):
    return lib().orc_jfa_first_step(W, H)


def jfa_num_passes(W, H):
    return lib().orc_jfa_num_passes(W, H)


def seed(W, H, xy):
    xy, S = _sites(xy)
    lab = np.empty((H, W), np.int32)
    lib().orc_seed(W, H, xy, S, lab)
    return lab


def jfa_pass(W, H, xy, lab_in, step):
    xy, S = _sites(xy)
    out = np.empty((H, W), np.int32)
    lib().orc_jfa_pass(W, H, xy, S, np.ascontiguousarray(lab_in, np.int32), out, step)
    return out


def jfa(W, H, xy):
    """JFA(+1) label map (int32 [H, W]) of the fp32 sites xy [S, 2]."""
    xy, S = _sites(xy)
    lab = np.empty((H, W), np.int32)
    tmp = np.empty((H, W), np.int32)
    lib().orc_jfa(W, H, xy, S, lab, tmp)
    return lab


def voronoi_brute(W, H, xy):
    xy, S = _sites(xy)
    lab = np.empty((H, W), np.int32)
    lib().orc_voronoi_brute(W, H, xy, S, lab)
    return lab


def voronoi_exact(W, H, xy):
    xy, S = _sites(xy)
    lab = np.empty((H, W), np.int32)
    lib().orc_voronoi_exact(W, H, xy, S, lab)
    return lab


def moments(rho, lab, xy):
    """float64 [S, 4] = (W, DX, DY, E) per site; see orc_moments."""
    rho = np.ascontiguousarray(rho, np.float32)
    H, W = rho.shape
    xy, S = _sites(xy)
    m = np.zeros((max(S, 1), 4), np.float64)
    lib().orc_moments(W, H, rho, np.ascontiguousarray(lab, np.int32), xy, S, m)
    return m[:S]


def energy_from_moments(W, H, m):
    e = C.c_double()
    g = C.c_double()
    m = np.ascontiguousarray(m, np.float64)
    lib().orc_energy_from_moments(W, H, m, m.shape[0], C.byref(e), C.byref(g))
    return e.value, g.value


def update(m, xy, omega=1.0, fixed=None):
    xy, S = _sites(xy)
    out = np.empty_like(xy)
    fx = None if fixed is None else np.ascontiguousarray(fixed, np.uint8)
    lib().orc_update(np.ascontiguousarray(m, np.float64), _ptr(fx), S, omega, xy, out)
    return out


def lloyd(rho, xy, iters, omega=1.0, fixed=None, exact=False, return_labels=False):
    """Continuous-mode Lloyd; returns (sites, energy[iters], gradnorm[iters][, labels])."""
    rho = np.ascontiguousarray(rho, np.float32)
    H, W = rho.shape
    xy = np.array(xy, dtype=np.float32).reshape(-1, 2).copy()
    S = xy.shape[0]
    e = np.zeros(max(iters, 1), np.float64)
    g = np.zeros(max(iters, 1), np.float64)
    lab = np.empty((H, W), np.int32) if return_labels else None
    fx = None if fixed is None else np.ascontiguousarray(fixed, np.uint8)
    lib().orc_lloyd(W, H, rho, xy, _ptr(fx), S, iters, omega, int(exact), _ptr(e), _ptr(g),
                    _ptr(lab))
    out = (xy, e[:iters], g[:iters])
    return out + (lab,) if return_labels else out


def raster(points, weight, tri, scale, n):
    pts = np.ascontiguousarray(points, np.float64).reshape(-1)
    wt = np.ascontiguousarray(weight, np.float64)
    tr = np.ascontiguousarray(tri, np.int32).reshape(-1)
    dens = np.empty((n, n), np.float32)
    lib().orc_raster(pts, wt, wt.shape[0], tr, tr.shape[0] // 3, dens, scale, n)
    return dens


def kiss(seed=0):
    k = Kiss()
    lib().orc_kiss_init(C.byref(k), seed)
    return k


def kiss_next(k):
    return lib().orc_kiss_next(C.byref(k))


def kiss_uniform(k):
    return lib().orc_kiss_uniform(C.byref(k))


def ref_put_constraints(mask):
    mask = np.ascontiguousarray(mask, np.uint8)
    n = mask.shape[0]
    vor = np.empty((n, n, 2), np.int16)
    lib().orc_ref_put_constraints(vor, mask, n)
    return vor


def ref_random_points(vor, density, num, rng=None):
    """Adds num sites to vor in place (reference randomPoints); returns the rng."""
    density = np.ascontiguousarray(density, np.float32)
    n = density.shape[0]
    if rng is None:
        rng = kiss(0)
    assert vor.flags["C_CONTIGUOUS"] and vor.dtype == np.int16
    lib().orc_ref_random_points(vor, density, num, n, C.byref(rng))
    return rng


def compat_decode(vor, mask=None):
    vor = np.ascontiguousarray(vor, np.int16)
    n = vor.shape[0]
    xy = np.empty((n * n, 2), np.float32)
    fx = np.empty(n * n, np.uint8)
    mk = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    S = lib().orc_compat_decode(vor, _ptr(mk), n, xy, fx)
    return xy[:S].copy(), fx[:S].copy()


def compat_update(rho, m, xy, omega, mask=None):
    rho = np.ascontiguousarray(rho, np.float32)
    n = rho.shape[0]
    xy, S = _sites(xy)
    out = np.empty((max(S, 1), 2), np.float32)
    owner = np.empty((n, n), np.int32)
    mk = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    k = lib().orc_compat_update(n, rho, _ptr(mk), np.ascontiguousarray(m, np.float64), S, omega,
                                xy, out, owner)
    return out[:k].copy()


def compat_levels(n, depth):
    return lib().orc_compat_levels(int(n), int(depth))


def compat_gcvt(rho, vor, mask=None, depth=1, max_iter=1000, exact=False):
    """Reference-compatible integer-site CVT with gCVT's pyramid and schedule;
    vor [n, n, 2] int16 is updated in place (for depth > 1 its first
    (n >> (L-1))^2 pairs are read as the coarsest level's map).
    Returns (iterations, energies sampled every 10 iterations, number of sites)."""
    rho = np.ascontiguousarray(rho, np.float32)
    n = rho.shape[0]
    assert vor.flags["C_CONTIGUOUS"] and vor.dtype == np.int16
    elog = np.zeros(max_iter // 10 + 16, np.float64)
    nlog = C.c_int()
    ns = C.c_int()
    mk = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    it = lib().orc_compat_gcvt(n, rho, vor, _ptr(mk), int(depth), int(max_iter), int(exact),
                               _ptr(elog), C.byref(nlog), C.byref(ns))
    return it, elog[: nlog.value].copy(), ns.value


def compat_lloyd(rho, vor, mask=None, max_iter=1000, exact=False):
    """The single-level (depth 1) form of compat_gcvt."""
    return compat_gcvt(rho, vor, mask, 1, max_iter, exact)


def extract_points(vor, mask, scale, l, b):
    vor = np.ascontiguousarray(vor, np.int16)
    n = vor.shape[0]
    out = np.empty((n * n, 2), np.float64)
    k = lib().orc_extract_points(vor, np.ascontiguousarray(mask, np.uint8), n, scale, l, b, out)
    return out[:k].copy()


def locate(uv, tri, q, p3=None):
    """First face (in order) holding each query point, its three weights and,
    with p3 ([nv][3]), the interpolated 3D point (source/recover.h:63-125)."""
    uv = np.ascontiguousarray(uv, np.float64).reshape(-1)
    tr = np.ascontiguousarray(tri, np.int32).reshape(-1)
    q = np.ascontiguousarray(q, np.float64).reshape(-1, 2)
    nq = q.shape[0]
    face = np.empty(nq, np.int32)
    w = np.empty((nq, 3), np.float64)
    out3 = np.empty((nq, 3), np.float64) if p3 is not None else None
    p3c = None if p3 is None else np.ascontiguousarray(p3, np.float64).reshape(-1)
    lib().orc_locate(uv, _ptr(p3c), tr, tr.shape[0] // 3, q.reshape(-1), nq, face, _ptr(w),
                     _ptr(out3))
    return (face, w) if p3 is None else (face, w, out3)


def vertex_weights(uv, p3, tri):
    """Per-vertex density weight, 3D area over 2D area of the incident faces,
    with the > 1e10 rule (source/weighting.h:37-97)."""
    uv = np.ascontiguousarray(uv, np.float64).reshape(-1)
    p3 = np.ascontiguousarray(p3, np.float64).reshape(-1)
    tr = np.ascontiguousarray(tri, np.int32).reshape(-1)
    nv = uv.shape[0] // 2
    out = np.empty(nv, np.float64)
    lib().orc_vertex_weights(uv, p3, nv, tr, tr.shape[0] // 3, out)
    return out


def set_threads(t):
    lib().orc_set_threads(int(t))


def max_threads():
    return lib().orc_max_threads()
