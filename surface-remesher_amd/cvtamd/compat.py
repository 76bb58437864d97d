"""The reference-facing entry points: integer-site gCVT semantics, the density
rasterizer and the CDT point extraction (include/cvt.h), plus ctypes access to
the C++ shims that keep the reference's signatures (include/gcvt_compat.h).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call, lib
from .engine import _out_ptr, _ptr

MARKER = -32768


def gcvt(voronoi: np.ndarray, density, mask=None, depth: int = 1, max_iter: int = 1000):
    """Integer-site CVT on an (n, n, 2) int16 label map, updated in place.
    Returns (iterations, number of sites)."""
    assert voronoi.dtype == np.int16 and voronoi.flags["C_CONTIGUOUS"]
    n = voronoi.shape[0]
    it, ns = C.c_int(), C.c_int()
    pv, k0 = _ptr(voronoi)
    pd, k1 = _ptr(density, np.float32)
    pm, k2 = _ptr(mask, np.uint8)
    call("cvt_compat_gcvt", pv, pd, pm, int(n), int(depth), int(max_iter), C.byref(it), C.byref(ns))
    return it.value, ns.value


def ref_init_sites(density, num: int, mask=None, rng_state=None, out=None):
    """The reference's initial sites for gCVT (putConstrains + randomPoints,
    cvt_ref_init_sites): an (n, n, 2) int16 map with the mask pixels as fixed
    sites and num sites drawn by the reference's rejection rule and LLP64
    KISS generator.  rng_state: None (the reference's zero state) or a
    4-element uint32 array, advanced in place.  out: an (n, n, 2) int16 array
    or tensor to write instead of a new array.  Returns (map, draws)."""
    n = int(density.shape[0])
    if out is None:
        out = np.empty((n, n, 2), np.int16)
    po = _out_ptr(out, np.int16, 2 * n * n)
    pd, kd = _ptr(density, np.float32)
    pm, km = _ptr(mask, np.uint8)
    ps = None if rng_state is None else _out_ptr(rng_state, np.uint32, 4)
    draws = C.c_longlong()
    call("cvt_ref_init_sites", po, pd, pm, n, int(num), ps, C.byref(draws))
    return out, draws.value


def rasterize(points, weight, triangles, scale: float, n: int, out=None):
    """Density grid (n, n) float32 of a 2-D triangle mesh (discretization_d rule).

    points (nv, 2) float64, weight (nv,) float64, triangles (nf, 3) int32:
    numpy arrays or torch tensors in host or device memory.  out: an
    (n, n) float32 array or tensor to write (host or device), or None for a
    new numpy array."""
    pp, kp = _ptr(points, np.float64)
    pw, kw = _ptr(weight, np.float64)
    pt, kt = _ptr(triangles, np.int32)
    nv = int(np.prod(kp.shape)) // 2
    nf = int(np.prod(kt.shape)) // 3
    if int(np.prod(kw.shape)) != nv:
        raise ValueError("weight must hold one value per point")
    if out is None:
        out = np.empty((n, n), np.float32)
    po = _out_ptr(out, np.float32, n * n)
    call("cvt_rasterize", pp, pw, nv, pt, nf, po, float(scale), int(n))
    return out


def extract_points(voronoi, mask, scale: float, l: float, b: float):
    """CDT input points: self-labelled, non-mask pixels in x-major order."""
    v = np.ascontiguousarray(voronoi, np.int16)
    n = v.shape[0]
    cnt = C.c_long()
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    pm = None if m is None else m.ctypes.data_as(C.c_void_p)
    call("cvt_extract_points", v.ctypes.data_as(C.c_void_p), pm, n, scale, l, b, None, 0,
         C.byref(cnt))
    out = np.empty((max(cnt.value, 1), 2), np.float64)
    call("cvt_extract_points", v.ctypes.data_as(C.c_void_p), pm, n, scale, l, b,
         out.ctypes.data_as(C.c_void_p), cnt.value, C.byref(cnt))
    return out[: cnt.value]


def shim_gcvt(voronoi: np.ndarray, density: np.ndarray, mask: np.ndarray, size: int, depth: int,
              max_iter: int):
    """Calls the C++ shim `void gCVT(short*, float*, bool*, int, int, int)` by its
    mangled name, as a reference caller linking libcvt_hip.so would."""
    f = getattr(lib(), "_Z4gCVTPsPfPbiii")
    f.restype = None
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int]
    f(voronoi.ctypes.data_as(C.c_void_p), density.ctypes.data_as(C.c_void_p),
      mask.ctypes.data_as(C.c_void_p), size, depth, max_iter)


def shim_discretization_d(points, weight, num_point, triangle, num_tri, density, scale, n):
    """Calls the C++ shim `void discretization_d(double*, double*, int, int*, int, float*,
    double, int)` by its mangled name."""
    f = getattr(lib(), "_Z16discretization_dPdS_iPiiPfdi")
    f.restype = None
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_double,
                  C.c_int]
    f(points.ctypes.data_as(C.c_void_p), weight.ctypes.data_as(C.c_void_p), num_point,
      triangle.ctypes.data_as(C.c_void_p), num_tri, density.ctypes.data_as(C.c_void_p), scale, n)
