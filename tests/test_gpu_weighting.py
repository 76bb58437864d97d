"""GPU vertex density weights (cvt_vertex_weights, the reference's
weighting_vertices_with_area) against the oracle, through the C ABI.

Bar: bit-exact (NaN where the oracle has NaN)."""
import numpy as np
import pytest

from test_oracle import jittered_mesh

pytestmark = pytest.mark.gpu


def lift(uv):
    return np.stack([uv[:, 0] + 0.1 * np.sin(uv[:, 1]), uv[:, 1],
                     np.cos(0.3 * uv[:, 0]) * np.sin(0.2 * uv[:, 1]) * 3.0], 1)


def check(uv, p3, tri):
    from cvtamd import weighting
    from oracle import oracle as orc
    got = weighting.vertex_weights(uv, p3, tri)
    ref = orc.vertex_weights(uv, p3, tri)
    assert np.array_equal(got, ref, equal_nan=True)
    return got


@pytest.mark.parametrize("m,seed", [(2, 1), (9, 2), (40, 3), (90, 4)])
def test_weights_match_oracle(gpu, m, seed):
    uv, tri = jittered_mesh(m, seed)
    w = check(uv, lift(uv), tri)
    assert np.all(np.isfinite(w)) and np.all(w > 0)


def test_weights_order_sensitive_fan(gpu):
    # 3000 faces of random 3D areas around vertex 0: its sums depend on the
    # order of addition, and must follow face order
    k = 3000
    g = np.random.default_rng(21)
    a = np.sort(g.uniform(0, 2 * np.pi, k))
    r = g.uniform(0.5, 2.0, k)
    uv = np.concatenate([[[0.0, 0.0]], np.stack([r * np.cos(a), r * np.sin(a)], 1)])
    p3 = np.concatenate([uv, g.uniform(-1, 1, (k + 1, 1))], 1)
    p3[0, 2] = 0.0
    fan = np.array([[0, 1 + i, 1 + (i + 1) % k] for i in range(k)], np.int32)
    fan = fan[g.permutation(k)]
    w = check(uv, p3, fan)
    # the data discriminates the order: a reversed sum rounds differently
    u, v = p3[fan[:, 0]] - p3[fan[:, 2]], p3[fan[:, 1]] - p3[fan[:, 2]]
    area3 = np.sqrt((np.cross(u, v) ** 2).sum(1) / 4)
    fwd = rev = 0.0
    for x in area3:
        fwd += x
    for x in area3[::-1]:
        rev += x
    assert fwd != rev
    assert np.isfinite(w[0])


def test_weights_degenerate_cases(gpu):
    uv, tri = jittered_mesh(20, 5)
    nv = len(uv)
    p3 = lift(uv)
    extra_uv = np.array([[30.0, 30.0], [31.0, 31.0], [32.0, 32.0],
                         [40.0, 40.0], [40.0 + 1e-6, 40.0], [40.0, 40.0 + 1e-6], [50.0, 50.0],
                         [60.0, 60.0], [61.0, 60.0]])
    extra_p3 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0],
                         [0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [9.0, 9.0, 9.0],
                         [0.0, 0.0, 0.0], [1.0, 2.0, 0.0]])
    uv2 = np.concatenate([uv, extra_uv])
    p32 = np.concatenate([p3, extra_p3])
    extra = np.array([[nv, nv + 1, nv + 2],          # collinear in 2D: +inf
                      [nv + 3, nv + 4, nv + 5],      # a sliver: finite, above 1e10
                      [nv + 7, nv + 7, nv + 8],      # a repeated vertex: zero areas
                      [nv + 2, nv + 1, 0]], np.int32)
    tri2 = np.concatenate([extra[:2], tri, extra[2:]])
    w = check(uv2, p32, tri2)
    assert np.isnan(w[nv + 6]) and np.isnan(w[nv + 7])
    assert w[nv + 3] == np.nanmax(w[:nv])
    # every weight above the cap: all become 0
    check(extra_uv[:3], extra_p3[:3], np.array([[0, 1, 2], [2, 1, 0]], np.int32))


def test_weights_large_mesh(gpu):
    uv, tri = jittered_mesh(424, 13)
    check(uv, lift(uv), tri)


def test_weights_device_pointers_and_errors(gpu):
    import torch
    import cvtamd
    from cvtamd import weighting
    uv, tri = jittered_mesh(30, 8)
    p3 = lift(uv)
    host = weighting.vertex_weights(uv, p3, tri)
    dev = weighting.vertex_weights(*(torch.from_numpy(a).cuda() for a in (uv, p3, tri)))
    assert np.array_equal(host, dev)
    with pytest.raises(TypeError):
        weighting.vertex_weights(torch.from_numpy(uv).float().cuda(), p3, tri)
    bad = tri.copy()
    bad[7, 2] = len(uv)
    with pytest.raises(cvtamd.CvtError):
        weighting.vertex_weights(uv, p3, bad)
    bad[7, 2] = -1
    with pytest.raises(cvtamd.CvtError):
        weighting.vertex_weights(uv, p3, bad)
    assert np.all(np.isnan(weighting.vertex_weights(uv, p3, tri[:0])))
    assert weighting.vertex_weights(uv[:0], p3[:0], tri[:0]).shape == (0,)


def test_weights_and_raster_stay_on_the_device(gpu):
    """Weights written into a device tensor feed the rasterizer without a
    host copy; the density lands in a device tensor; both equal the host
    path's results."""
    import torch
    from cvtamd import compat, weighting
    uv, tri = jittered_mesh(18, 14)
    p3 = lift(uv)
    n = 200
    scale = 17.0 / (n - 1)
    d_uv, d_p3, d_tri = (torch.from_numpy(a).cuda() for a in (uv, p3, tri))
    d_w = torch.empty(len(uv), dtype=torch.float64, device="cuda")
    assert weighting.vertex_weights(d_uv, d_p3, d_tri, out=d_w) is d_w
    d_den = torch.empty((n, n), dtype=torch.float32, device="cuda")
    compat.rasterize(d_uv, d_w, d_tri, scale, n, out=d_den)
    w_host = weighting.vertex_weights(uv, p3, tri)
    assert np.array_equal(d_w.cpu().numpy(), w_host)
    assert np.array_equal(d_den.cpu().numpy(), compat.rasterize(uv, w_host, tri, scale, n))
    with pytest.raises(ValueError):
        compat.rasterize(uv, w_host[:-1], tri, scale, n)


def test_weights_feed_the_rasterizer(gpu):
    """The two GPU stages chained: weights, then the density raster, equal
    the oracle's weights rasterized by the oracle."""
    from cvtamd import compat, weighting
    from oracle import oracle as orc
    uv, tri = jittered_mesh(25, 9)
    p3 = lift(uv)
    n = 256
    scale = 24.0 / (n - 1)
    got = compat.rasterize(uv, weighting.vertex_weights(uv, p3, tri), tri, scale, n)
    ref = orc.raster(uv, orc.vertex_weights(uv, p3, tri), tri, scale, n)
    assert np.array_equal(got, ref)


def test_weights_many_vertices(gpu):
    # 4.41M vertices: the prefix sum over the vertices' corner counts spans
    # more than 1024 tiles of 4096
    from cvtamd import workloads as wl
    uv, p3, tri, _ = wl.patch_mesh(m=2100)
    check(uv, p3, tri)


@pytest.mark.parametrize("npts,seed", [(300, 21), (5000, 22)])
def test_weights_delaunay_mesh(gpu, npts, seed):
    # irregular valences: a Delaunay triangulation of random points, faces
    # shuffled, over a curved surface
    from scipy.spatial import Delaunay
    g = np.random.default_rng(seed)
    uv = g.random((npts, 2)) * 10.0
    tri = Delaunay(uv).simplices.astype(np.int32)
    tri = tri[g.permutation(len(tri))]
    p3 = np.stack([uv[:, 0], uv[:, 1], np.sin(uv[:, 0]) * np.cos(0.7 * uv[:, 1])], 1)
    w = check(uv, p3, tri)
    assert np.all(np.isfinite(w)) and np.all(w >= 1.0 - 1e-12)
