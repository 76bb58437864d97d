"""Delaunay triangulation on the GPU (cvt_dt_* in include/cvt.h): the
unconstrained part of gDel2D's GpuDel::compute
(source/gDel2D/gDel2D/GPU/GpuDelaunay.cu:46-83).

    with Delaunay(max_points) as d:
        tri, opp = d.compute(points)      # float64 (n, 2) -> int32 (t, 3) x 2

`tri` holds counterclockwise vertex triples of the finite triangles; `opp`
their neighbours in gDel2D's TriOpp encoding ((triangle << 4) | edge, edge i
opposite vertex i, -1 on the convex hull).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call
from .engine import _ptr


class Delaunay:
    def __init__(self, max_points: int):
        self._h = C.c_void_p()
        call("cvt_dt_create", C.byref(self._h), int(max_points))
        self.max_points = int(max_points)

    def compute(self, points, want_opp: bool = True, segments=None):
        """float64 points (n, 2) (numpy, or a contiguous torch tensor) ->
        (tri, opp) as int32 numpy arrays (opp None unless want_opp).  With
        segments (int32 (m, 2) point index pairs) the result is the
        constrained triangulation, and opp marks constraint edges with bit 2."""
        if hasattr(points, "numel"):
            n = points.numel() // 2
        else:
            points = np.ascontiguousarray(points, np.float64).reshape(-1, 2)
            n = points.shape[0]
        pp, keep = _ptr(points, np.float64)
        cap = max(2 * n, 1)
        tri = np.zeros((cap, 3), np.int32)
        opp = np.zeros((cap, 3), np.int32) if want_opp else None
        nt = C.c_int(0)
        po = opp.ctypes.data_as(C.c_void_p) if want_opp else None
        if segments is None:
            call("cvt_dt_compute", self._h, pp, int(n), tri.ctypes.data_as(C.c_void_p), po, cap,
                 C.byref(nt))
        else:
            ps, keep_s = _ptr(segments, np.int32)
            m = (segments.numel() if hasattr(segments, "numel") else np.asarray(segments).size) // 2
            call("cvt_dt_compute_constrained", self._h, pp, int(n), ps, int(m),
                 tri.ctypes.data_as(C.c_void_p), po, cap, C.byref(nt))
        k = nt.value
        return tri[:k].copy(), (opp[:k].copy() if want_opp else None)

    def stats(self) -> dict:
        s = np.zeros(4, np.int32)
        call("cvt_dt_stats", self._h, s.ctypes.data_as(C.c_void_p))
        return {"rounds": int(s[0]), "flip_passes": int(s[1]), "flips": int(s[2]),
                "duplicates": int(s[3])}

    def close(self):
        if self._h:
            call("cvt_dt_destroy", self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def delaunay(points, want_opp: bool = True, segments=None):
    """One-shot triangulation (a context sized for these points)."""
    n = (points.numel() // 2) if hasattr(points, "numel") else len(np.asarray(points).reshape(-1, 2))
    with Delaunay(max(n, 3)) as d:
        return d.compute(points, want_opp, segments)


def predicates(quads):
    """(orient, incircle) signs on the device for float64 records
    (ax ay bx by cx cy dx dy): orient2d(a, b, c) and incircle(a, b, c, d)."""
    q = np.ascontiguousarray(quads, np.float64).reshape(-1, 8)
    n = q.shape[0]
    o = np.zeros(n, np.int32)
    ic = np.zeros(n, np.int32)
    call("cvt_dt_predicates", q.ctypes.data_as(C.c_void_p), int(n),
         o.ctypes.data_as(C.c_void_p), ic.ctypes.data_as(C.c_void_p))
    return o, ic


__all__ = ["Delaunay", "delaunay", "predicates"]
