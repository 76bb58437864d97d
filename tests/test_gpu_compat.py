"""GPU parity of the reference-facing entry points against the CPU oracle:
the density rasterizer (discretization_d rule, bit-exact), the integer-site
gCVT mode (label map bit-exact, same iteration and site counts), the CDT point
extraction (exact) and the C++ shims with the reference's signatures.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def delaunay_mesh(seed, npts, extent):
    from scipy.spatial import Delaunay
    g = np.random.default_rng(seed)
    pts = g.random((npts, 2)) * extent
    # corners so the hull is the square
    pts = np.vstack([pts, [[0, 0], [extent, 0], [0, extent], [extent, extent]]])
    tri = Delaunay(pts).simplices.astype(np.int32)
    w = 0.1 + g.random(len(pts))
    return pts, w, tri


# -------------------------------------------------------------- rasterizer --
@pytest.mark.parametrize("n,npts,seed", [(64, 200, 1), (100, 500, 2), (257, 50, 3), (512, 2000, 4)])
def test_raster_bit_exact(gpu, orc, n, npts, seed):
    from cvtamd import compat
    pts, w, tri = delaunay_mesh(seed, npts, 37.5)
    scale = 37.5 / (n - 1)                    # discretization.h: max extent / (n - 1)
    got = compat.rasterize(pts, w, tri, scale, n)
    ref = orc.raster(pts, w, tri, scale, n)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert (got > 0).mean() > 0.9


def test_raster_first_hit_thin_and_degenerate(gpu, orc):
    from cvtamd import compat
    n = 48
    pts = np.array([[0, 0], [40, 0], [0, 40], [5, 5], [45, 5], [5, 45],
                    [1, 1], [2, 2], [3, 3],                       # degenerate
                    [0, 10], [47, 10.0000001], [47, 10]], np.float64)  # sliver
    w = np.array([1, 2, 3, 7, 8, 9, 5, 5, 5, 4, 4, 4], np.float64)
    tri = np.array([[6, 7, 8], [9, 10, 11], [0, 1, 2], [3, 4, 5]], np.int32)
    got = compat.rasterize(pts, w, tri, 1.0, n)
    ref = orc.raster(pts, w, tri, 1.0, n)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_raster_many_overlaps_tile_overflow(gpu, orc):
    # > 256 triangles over one 16x16 tile: the tile falls back to all triangles
    from cvtamd import compat
    g = np.random.default_rng(5)
    k = 400
    c = g.random((k, 2)) * 8 + 4
    pts = np.concatenate([c + [-3, -3], c + [3, -2], c + [0, 3]]).astype(np.float64)
    tri = np.stack([np.arange(k), np.arange(k) + k, np.arange(k) + 2 * k], 1).astype(np.int32)
    w = g.random(3 * k)
    got = compat.rasterize(pts, w, tri, 1.0, 32)
    ref = orc.raster(pts, w, tri, 1.0, 32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_raster_bad_index(gpu):
    import cvtamd
    from cvtamd import compat
    pts, w, tri = delaunay_mesh(1, 20, 10.0)
    tri[3, 1] = len(pts) + 5
    with pytest.raises(cvtamd.CvtError):
        compat.rasterize(pts, w, tri, 0.1, 32)


def test_discretization_shim(gpu, orc):
    from cvtamd import compat
    pts, w, tri = delaunay_mesh(7, 100, 20.0)
    n = 80
    pts = np.ascontiguousarray(pts, np.float64)
    tri = np.ascontiguousarray(tri, np.int32)
    dens = np.zeros((n, n), np.float32)
    compat.shim_discretization_d(pts, np.ascontiguousarray(w), len(pts), tri, len(tri), dens,
                                 20.0 / (n - 1), n)
    assert np.array_equal(dens, compat.rasterize(pts, w, tri, 20.0 / (n - 1), n))


# ------------------------------------------------------------- compat mode --
def reference_init(orc, density, mask, num):
    """putConstrains + randomPoints with the reference's zero-seeded RNG."""
    vor = orc.ref_put_constraints(mask)
    orc.ref_random_points(vor, density, num)
    return vor


@pytest.mark.parametrize("n,num,seed", [(64, 40, 1), (256, 300, 2), (512, 1000, 3)])
def test_gcvt_matches_oracle(gpu, orc, n, num, seed):
    from cvtamd import compat
    from cvtamd import workloads as wl
    g = wl.rng(seed)
    rho = wl.gaussian_mixture(n, g, 5, 0.05, wl.disk_mask(n))
    mask = np.zeros((n, n), np.uint8)
    ys, xs = np.nonzero(rho > 0)
    pick = g.integers(0, len(xs), 6)
    mask[ys[pick], xs[pick]] = 1
    vor = reference_init(orc, rho, mask, num)
    ref = vor.copy()
    it_r, elog, ns_r = orc.compat_lloyd(rho, ref, mask, max_iter=200)
    got = vor.copy()
    it, ns = compat.gcvt(got, rho, mask, 1, 200)
    assert (it, ns) == (it_r, ns_r)
    assert np.array_equal(got, ref)
    # fixed sites are still sites, on their pixels
    for y, x in zip(ys[pick], xs[pick]):
        assert got[y, x].tolist() == [x, y]


def test_gcvt_uniform_and_zero_region(gpu, orc):
    from cvtamd import compat
    n = 128
    rho = np.ones((n, n), np.float32)
    rho[:, 90:] = 0.0
    mask = np.zeros((n, n), np.uint8)
    vor = reference_init(orc, rho, mask, 150)
    ref = vor.copy()
    it_r, _, ns_r = orc.compat_lloyd(rho, ref, mask, max_iter=300)
    got = vor.copy()
    it, ns = compat.gcvt(got, rho, mask, 1, 300)
    assert (it, ns) == (it_r, ns_r)
    assert np.array_equal(got, ref)


def test_gcvt_close_to_exact_voronoi_semantics(gpu, orc):
    # the reference's own Voronoi is exact (PBA+): same schedule with the exact
    # map ends at a nearby energy
    from cvtamd import compat
    from cvtamd import workloads as wl
    n = 256
    rho = wl.gaussian_mixture(n, wl.rng(9), 6, 0.05, wl.disk_mask(n))
    mask = np.zeros((n, n), np.uint8)
    vor = reference_init(orc, rho, mask, 400)
    ex = vor.copy()
    _, e_exact, ns_exact = orc.compat_lloyd(rho, ex, mask, max_iter=100, exact=True)
    got = vor.copy()
    _, ns = compat.gcvt(got, rho, mask, 1, 100)
    _, e_jfa, _ = orc.compat_lloyd(rho, vor.copy(), mask, max_iter=100)
    assert abs(e_jfa[-1] - e_exact[-1]) / e_exact[-1] < 2e-2
    assert abs(ns - ns_exact) <= 0.02 * ns_exact


def coarse_start(orc, rho, depth, num, seed):
    """A full-size input map whose first nc^2 pairs hold an nc x nc map of
    reference-initialised sites (nc the coarsest side): the layout gCVT reads
    for its coarsest level."""
    n = rho.shape[0]
    L = orc.compat_levels(n, depth)
    nc = n >> (L - 1)
    f = n // nc
    coarse = rho.reshape(nc, f, nc, f).mean(axis=(1, 3)).astype(np.float32)
    cvor = reference_init(orc, coarse, np.zeros((nc, nc), np.uint8), num)
    vor = np.full((n, n, 2), -32768, np.int16)
    vor.reshape(-1, 2)[: nc * nc] = cvor.reshape(-1, 2)
    return vor


@pytest.mark.parametrize("n,depth,num,seed", [(512, 2, 300, 4), (1024, 3, 500, 5), (512, 4, 200, 6)])
def test_gcvt_multires_matches_oracle(gpu, orc, n, depth, num, seed):
    from cvtamd import compat
    from cvtamd import workloads as wl
    g = wl.rng(seed)
    rho = wl.gaussian_mixture(n, g, 6, 0.05, wl.disk_mask(n))
    vor = coarse_start(orc, rho, depth, num, seed)
    # fixed pixels: every level reads this mask with its own row stride
    mask = (g.random((n, n)) < 0.002).astype(np.uint8)
    ref = vor.copy()
    it_r, elog, ns_r = orc.compat_gcvt(rho, ref, mask, depth, 300)
    got = vor.copy()
    it, ns = compat.gcvt(got, rho, mask, depth, 300)
    print(f"n={n} depth={depth}: {it} iterations, {ns} sites (oracle {it_r}, {ns_r})")
    assert (it, ns) == (it_r, ns_r)
    assert np.array_equal(got, ref)


def test_gcvt_schedule_hand_traced(gpu):
    # one site at (8, 4) on a 9 x 9 grid with rho = 1/64 on columns 0..2: 30
    # iterations, omega = 1.25 after the second decade brings it to (1, 4)
    # (the derivation is in tests/test_oracle.py)
    from cvtamd import compat
    n = 9
    rho = np.zeros((n, n), np.float32)
    rho[:, :3] = 1.0 / 64.0
    vor = np.full((n, n, 2), -32768, np.int16)
    vor[4, 8] = (8, 4)
    it, ns = compat.gcvt(vor, rho, None, 1, 1000)
    assert (it, ns) == (30, 1)
    assert (vor[:, :, 0] == 1).all() and (vor[:, :, 1] == 4).all()


def test_gcvt_multires_carries_schedule(gpu):
    # 512^2, depth 2: a coarse site at (127, 127); omega = 1 and lastE carried
    # from the coarse level give 30 iterations and the site (256, 256); with
    # maxIter 15 the fine level still runs one iteration: 16, site (255, 255)
    from cvtamd import compat
    n = 512
    rho = np.ones((n, n), np.float32)
    vor = np.full((n, n, 2), -32768, np.int16)
    vor.reshape(-1, 2)[127 * 256 + 127] = (127, 127)
    a = vor.copy()
    assert compat.gcvt(a, rho, None, 2, 1000) == (30, 1)
    assert (a == 256).all()
    b = vor.copy()
    assert compat.gcvt(b, rho, None, 2, 15) == (16, 1)
    assert (b == 255).all()


def test_gcvt_shim_and_max_iter(gpu, orc):
    from cvtamd import compat
    n = 256
    rho = np.ones((n, n), np.float32)
    mask = np.zeros((n, n), np.uint8)
    mask[10, 10] = 1
    vor = reference_init(orc, rho, mask, 200)
    a = vor.copy()
    compat.shim_gcvt(a, rho, mask, n, 1, 25)
    b = vor.copy()
    it, _ = compat.gcvt(b, rho, mask, 1, 25)
    assert it == 25
    assert np.array_equal(a, b)


# ------------------------------------------------------------- extraction --
def test_extract_points(gpu, orc):
    from cvtamd import compat
    n = 300
    g = np.random.default_rng(3)
    vor = np.full((n, n, 2), -32768, np.int16)
    ys, xs = g.integers(0, n, 2000), g.integers(0, n, 2000)
    vor[ys, xs, 0], vor[ys, xs, 1] = xs, ys
    vor[5, 7] = [8, 5]               # not self-labelled
    mask = (g.random((n, n)) < 0.1).astype(np.uint8)
    got = compat.extract_points(vor, mask, 0.01, -1.5, 2.25)
    ref = orc.extract_points(vor, mask, 0.01, -1.5, 2.25)
    assert got.shape == ref.shape and np.array_equal(got, ref)
