"""cvt_rasterize_noalloc, the rasterizer for callers that cannot allocate: it
gives cvt_rasterize's density bit for bit on two calls in a row (the second
smaller, so the shared scratch holds the first call's leftovers), rejects
pageable host buffers and sizes beyond its limits, and reports a bad vertex
index like cvt_rasterize."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def mesh(seed, npts, extent, thin=False):
    from scipy.spatial import Delaunay
    g = np.random.default_rng(seed)
    pts = g.random((npts, 2)) * extent
    pts = np.vstack([pts, [[0, 0], [extent, 0], [0, extent], [extent, extent]]])
    tri = Delaunay(pts).simplices.astype(np.int32)
    if thin:
        # a sliver (tested everywhere) and a repeated triangle (first hit wins)
        k = len(pts)
        pts = np.vstack([pts, [[1.0, 1.0], [extent - 1.0, 1.0 + 1e-9], [extent / 2, 1.0]]])
        tri = np.vstack([tri, [[k, k + 1, k + 2]], tri[:3]]).astype(np.int32)
    w = 0.1 + g.random(len(pts))
    return pts, w, tri


def raster_both(p, w, t, scale, n):
    import torch
    import cvtamd
    from cvtamd._lib import call
    ref = cvtamd.compat.rasterize(p, w, t, scale, n)
    dp = torch.from_numpy(np.ascontiguousarray(p, np.float64)).cuda()
    dw = torch.from_numpy(np.ascontiguousarray(w, np.float64)).cuda()
    dt = torch.from_numpy(np.ascontiguousarray(t, np.int32)).cuda()
    out = torch.full((n, n), -1.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    call("cvt_rasterize_noalloc", C.c_void_p(dp.data_ptr()), C.c_void_p(dw.data_ptr()),
         len(p), C.c_void_p(dt.data_ptr()), len(t), C.c_void_p(out.data_ptr()),
         float(scale), int(n))
    return ref, out.cpu().numpy()


def test_noalloc_matches_rasterize_twice(gpu):
    # first call: a larger grid and mesh; second call smaller, on the same scratch
    p, w, t = mesh(11, 3000, 1.0, thin=True)
    ref, got = raster_both(p, w, t, 1.0 / 700, 700)
    assert (ref > 0).mean() > 0.9
    assert np.array_equal(ref, got)
    p, w, t = mesh(12, 200, 1.0)
    ref, got = raster_both(p, w, t, 1.0 / 160, 160)
    assert np.array_equal(ref, got)


def test_noalloc_errors(gpu):
    import torch
    import cvtamd
    from cvtamd._lib import call
    p, w, t = mesh(13, 50, 1.0)
    dp = torch.from_numpy(p).cuda()
    dw = torch.from_numpy(w).cuda()
    dt = torch.from_numpy(t).cuda()
    out = torch.zeros((64, 64), dtype=torch.float32, device="cuda")
    args = lambda n, nt, o: (C.c_void_p(dp.data_ptr()), C.c_void_p(dw.data_ptr()), len(p),
                             C.c_void_p(dt.data_ptr()), nt, o, 1.0 / 64, n)
    dev_out = C.c_void_p(out.data_ptr())
    with pytest.raises(cvtamd.CvtError):                 # beyond the grid limit
        call("cvt_rasterize_noalloc", *args(4097, len(t), dev_out))
    with pytest.raises(cvtamd.CvtError):                 # beyond the triangle limit
        call("cvt_rasterize_noalloc", *args(64, (1 << 18) + 1, dev_out))
    host = np.zeros((64, 64), np.float32)                # pageable host output
    with pytest.raises(cvtamd.CvtError):
        call("cvt_rasterize_noalloc", *args(64, len(t), host.ctypes.data_as(C.c_void_p)))
    bad = dt.clone()
    bad[0, 0] = len(p) + 5
    with pytest.raises(cvtamd.CvtError):
        call("cvt_rasterize_noalloc", C.c_void_p(dp.data_ptr()), C.c_void_p(dw.data_ptr()),
             len(p), C.c_void_p(bad.data_ptr()), len(t), dev_out, 1.0 / 64, 64)
    # a good call after the failures
    ref, got = raster_both(p, w, t, 1.0 / 64, 64)
    assert np.array_equal(ref, got)
