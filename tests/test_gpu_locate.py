"""GPU point location (cvt_locate_points, the face search of the reference's
recover()) against the oracle's face-by-face scan, through the C ABI.

Bar: face indices, weights and lifted 3D points bit-exact."""
import numpy as np
import pytest

from test_oracle import jittered_mesh

pytestmark = pytest.mark.gpu


def queries(uv, tri, seed, n):
    """Random points plus the hard cases: vertices, edge midpoints,
    centroids, points just outside the hull and far away."""
    g = np.random.default_rng(seed)
    lo, hi = uv.min(0), uv.max(0)
    q = [g.uniform(lo - 0.5, hi + 0.5, (n, 2)), uv, uv[tri].mean(1),
         0.5 * (uv[tri[:, 0]] + uv[tri[:, 1]]), 0.5 * (uv[tri[:, 1]] + uv[tri[:, 2]]),
         np.array([[lo[0] - 1e-13, lo[1]], [hi[0] + 1e-12, hi[1]], [1e9, -1e9]])]
    return np.concatenate(q)


def check(uv, tri, q, p3=None):
    from cvtamd import recover
    got = recover.locate_points(uv, tri, q, p3=p3)
    from oracle import oracle as orc
    ref = orc.locate(uv, tri, q, p3=p3)
    assert np.array_equal(got[0], ref[0])
    assert np.array_equal(got[1], ref[1])
    if p3 is not None:
        assert np.array_equal(got[2], ref[2])
    return got


@pytest.mark.parametrize("m,seed", [(2, 1), (9, 2), (40, 3), (90, 4)])
def test_locate_matches_oracle(gpu, m, seed):
    uv, tri = jittered_mesh(m, seed)
    p3 = np.stack([np.sin(uv[:, 0]), np.cos(uv[:, 1]), uv[:, 0] * uv[:, 1]], 1)
    face = check(uv, tri, queries(uv, tri, seed, 3000), p3)[0]
    assert (face >= 0).mean() > 0.2


def test_locate_overlaps_thin_and_degenerate(gpu):
    uv, tri = jittered_mesh(12, 5)
    nv = len(uv)
    extra_uv = np.array([[0.0, 0.0], [11.0, 11.0], [5.5, 5.5 + 1e-9], [3.0, 3.0],
                         [6.0, 6.0], [9.0, 9.0]])
    uv2 = np.concatenate([uv, extra_uv])
    extra = np.array([[nv, nv + 1, nv + 2],        # a sliver along the diagonal: the "always" list
                      [nv + 3, nv + 4, nv + 5],    # collinear: never holds a point
                      [nv, nv + 1, 11 * 12]],      # overlaps half the mesh
                     np.int32)
    tri2 = np.concatenate([extra, tri, extra[::-1]])
    q = queries(uv2, tri2, 6, 4000)
    q = np.concatenate([q, np.array([[4.5, 4.5], [2.0, 2.0 + 1e-14], [5.0, 5.0]])])
    check(uv2, tri2, q)


def test_locate_crowded_bin(gpu):
    # a fan of 300 faces around one point: its bin overflows and tests every face
    k = 300
    a = np.linspace(0, 2 * np.pi, k, endpoint=False)
    uv = np.concatenate([[[0.0, 0.0]], np.stack([np.cos(a), np.sin(a)], 1) * 5.0,
                         [[40.0, 40.0], [41.0, 40.0], [40.0, 41.0]]])
    fan = np.array([[0, 1 + i, 1 + (i + 1) % k] for i in range(k)], np.int32)
    tri = np.concatenate([fan, np.array([[k + 1, k + 2, k + 3]], np.int32)])
    g = np.random.default_rng(7)
    q = np.concatenate([g.uniform(-6, 6, (3000, 2)), g.uniform(39.5, 41.5, (300, 2)),
                        np.zeros((1, 2))])
    face = check(uv, tri, q)[0]
    assert face[-1] == 0                          # the centre lies in every fan face: the first


def test_locate_device_pointers_and_nonfinite(gpu):
    import torch
    from cvtamd import recover
    uv, tri = jittered_mesh(20, 8)
    q = queries(uv, tri, 9, 1000)
    q[:3] = [[np.nan, 1.0], [1.0, np.inf], [-np.inf, np.nan]]
    f_host, w_host = recover.locate_points(uv, tri, q)
    f_dev, w_dev = recover.locate_points(torch.from_numpy(uv).cuda(),
                                         torch.from_numpy(tri).cuda(), torch.from_numpy(q).cuda())
    assert np.array_equal(f_host, f_dev) and np.array_equal(w_host, w_dev)
    assert np.all(f_host[:3] == -1)
    check(uv, tri, q)


def test_locate_errors_and_empty(gpu):
    import cvtamd
    from cvtamd import recover
    uv, tri = jittered_mesh(5, 10)
    bad = tri.copy()
    bad[3, 1] = len(uv)
    with pytest.raises(cvtamd.CvtError):
        recover.locate_points(uv, bad, np.zeros((4, 2)))
    f, w = recover.locate_points(uv, tri[:0], np.ones((5, 2)))
    assert np.all(f == -1)
    f, w = recover.locate_points(uv, tri, np.zeros((0, 2)))
    assert f.shape == (0,)


def test_recover_lifts_sites_and_filters_triangles(gpu):
    """recover() on arrays: sites lifted by their face, constraint vertices at
    their own 3D position, CDT triangles kept when their centroid is inside;
    a site outside every face reuses the previous site's location."""
    from cvtamd import recover
    from oracle import oracle as orc
    uv, tri = jittered_mesh(15, 11)
    p3 = np.stack([uv[:, 0], uv[:, 1], np.sin(uv[:, 0]) * np.cos(uv[:, 1])], 1)
    g = np.random.default_rng(12)
    sites = g.uniform(0.2, 13.8, (500, 2))
    sites[100] = [50.0, 50.0]                     # outside every face
    cons = np.array([0, 14, 224], np.int64)       # three mesh vertices as constraint points
    points = np.concatenate([sites, uv[cons]])
    cdt = g.integers(0, len(points), (800, 3)).astype(np.int32)
    verts, faces = recover.recover(uv, p3, tri, points, cdt, len(cons), cons)
    f, w, lifted = orc.locate(uv, tri, sites, p3=p3)
    assert f[100] == -1 and f[99] >= 0
    assert np.array_equal(verts[:100], lifted[:100])
    assert np.array_equal(verts[101:500], lifted[101:])
    assert np.array_equal(verts[100], lifted[99])
    assert np.array_equal(verts[500:], p3[cons])
    cen = recover.centroids(points, cdt)
    fc, _ = orc.locate(uv, tri, cen)
    assert np.array_equal(faces, cdt[fc >= 0])


def test_locate_large_mesh(gpu):
    # 358K faces: the binning is a counting sort, linear in (face, bin) pairs
    uv, tri = jittered_mesh(424, 13)
    g = np.random.default_rng(14)
    q = np.concatenate([g.uniform(-1.0, 424.0, (2000, 2)), uv[g.integers(0, len(uv), 200)]])
    face = check(uv, tri, q)[0]
    assert (face >= 0).mean() > 0.9


@pytest.mark.parametrize("npts,seed", [(300, 31), (5000, 32)])
def test_locate_delaunay_mesh(gpu, npts, seed):
    # irregular triangles of every shape and size, faces shuffled
    from scipy.spatial import Delaunay
    g = np.random.default_rng(seed)
    uv = g.random((npts, 2)) * 10.0
    tri = Delaunay(uv).simplices.astype(np.int32)
    tri = tri[g.permutation(len(tri))]
    p3 = np.stack([uv[:, 0], uv[:, 1], uv[:, 0] * 0.3 - uv[:, 1] * 0.2], 1)
    face = check(uv, tri, queries(uv, tri, seed, 4000), p3)[0]
    assert (face >= 0).mean() > 0.5


def test_locate_writes_device_outputs(gpu):
    import torch
    from cvtamd import recover
    uv, tri = jittered_mesh(25, 15)
    p3 = np.stack([uv[:, 0], uv[:, 1], 0.5 * uv[:, 0]], 1)
    q = queries(uv, tri, 16, 2000)
    host = recover.locate_points(uv, tri, q, p3=p3)
    n = len(q)
    out = (torch.full((n,), 7, dtype=torch.int32, device="cuda"),
           torch.empty((n, 3), dtype=torch.float64, device="cuda"),
           torch.empty((n, 3), dtype=torch.float64, device="cuda"))
    got = recover.locate_points(torch.from_numpy(uv).cuda(), torch.from_numpy(tri).cuda(),
                                torch.from_numpy(q).cuda(), p3=torch.from_numpy(p3).cuda(), out=out)
    assert all(a is b for a, b in zip(got, out))
    for a, b in zip(host, out):
        assert np.array_equal(a, b.cpu().numpy())
    with pytest.raises(ValueError):
        recover.locate_points(uv, tri, q, p3=p3, out=out[:2])
