"""GPU Delaunay triangulation (cvt_dt_*) against scipy.spatial.Delaunay (Qhull)
on inputs in general position, plus the structural checks that pin it
without scipy: counterclockwise triangles, neighbour links that agree both
ways, Euler's count 2n - h - 2, and every interior edge locally Delaunay
under the exact predicates.  The device predicates are checked against
exact rational arithmetic (tests/dt_cases.py).

Reference behaviour: gDel2D's GpuDel::compute
(source/gDel2D/gDel2D/GPU/GpuDelaunay.cu:46-83) returns the finite triangles
(kerMarkInfinityTri, :1494) with TriOpp links."""
import numpy as np
import pytest

from dt_cases import exact_incircle, exact_orient, records

pytestmark = pytest.mark.gpu


def tri_set(t):
    return set(map(tuple, np.sort(np.asarray(t), axis=1).tolist()))


def check_structure(pts, tri, opp):
    pts = np.asarray(pts, np.float64)
    n, nt = len(pts), len(tri)
    a, b, c = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    o = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    assert (o > 0).all(), "every triangle counterclockwise"
    hull_edges = 0
    for t in range(nt):
        for e in range(3):
            r = opp[t, e]
            if r < 0:
                hull_edges += 1
                continue
            m, k = r >> 4, r & 3
            assert opp[m, k] == (t << 4) | e
            assert {tri[t, (e + 1) % 3], tri[t, (e + 2) % 3]} == {tri[m, (k + 1) % 3], tri[m, (k + 2) % 3]}
    used = len(np.unique(tri))
    assert nt == 2 * used - hull_edges - 2, "Euler: 2n - h - 2"
    return hull_edges


def locally_delaunay(pts, tri, opp):
    """every interior edge: the opposite vertex is not inside the circle
    (exact device predicates)"""
    from cvtamd.delaunay import predicates
    q = []
    for t in range(len(tri)):
        for e in range(3):
            r = opp[t, e]
            if r < 0:
                continue
            m, k = r >> 4, r & 3
            v = tri[t]
            w = tri[m, k]
            q.append(np.concatenate([pts[v[0]], pts[v[1]], pts[v[2]], pts[w]]))
    _, ic = predicates(np.array(q))
    return (ic <= 0).all()


def test_device_predicates_exact(gpu):
    from cvtamd.delaunay import predicates
    q = records(seed=1)
    o, ic = predicates(q)
    assert np.array_equal(o, [exact_orient(r) for r in q])
    assert np.array_equal(ic, [exact_incircle(r) for r in q])


@pytest.mark.parametrize("n,kind,seed", [
    (3, "f32", 0), (4, "f32", 1), (10, "f32", 2), (100, "f64", 3), (1000, "f32", 4),
    (5000, "gauss", 5), (20000, "f32", 6), (100000, "f32", 7),
])
def test_equals_scipy(gpu, n, kind, seed):
    from scipy.spatial import Delaunay as QDelaunay
    from cvtamd.delaunay import Delaunay
    g = np.random.default_rng(seed)
    if kind == "f32":
        pts = (g.random((n, 2)) * 2048.0).astype(np.float32).astype(np.float64)
    elif kind == "f64":
        pts = g.random((n, 2))
    else:
        pts = g.normal(size=(n, 2)) * np.array([300.0, 40.0]) + 1000.0
    print(f"\ncase n={n} {kind} seed={seed}", flush=True)
    with Delaunay(n) as d:
        tri, opp = d.compute(pts)
        st = d.stats()
    ref = QDelaunay(pts).simplices
    assert tri_set(tri) == tri_set(ref)
    check_structure(pts, tri, opp)
    if n <= 5000:
        assert locally_delaunay(pts, tri, opp)
    print(st)


def test_cvt_sites(gpu):
    """the sites of a converged Lloyd run (float32 positions of the 1024^2
    disk patch), as the pipeline triangulates them"""
    import cvtamd
    from scipy.spatial import Delaunay as QDelaunay
    from cvtamd import workloads as wl
    from cvtamd.delaunay import delaunay
    w = wl.config2(3, n=1024, num_sites=4096)
    with cvtamd.Engine(w.n, w.n, w.num_sites) as e:
        e.set_density(w.density)
        e.set_sites(w.sites)
        e.lloyd(50)
        xy = e.sites().astype(np.float64)
    tri, opp = delaunay(xy)
    assert tri_set(tri) == tri_set(QDelaunay(xy).simplices)
    check_structure(xy, tri, opp)


def test_duplicates_dropped(gpu):
    """exact duplicates are dropped; the lowest index of each duplicate set
    is kept and the triangulation is that of the distinct points"""
    from scipy.spatial import Delaunay as QDelaunay
    from cvtamd.delaunay import Delaunay
    g = np.random.default_rng(11)
    base = (g.random((500, 2)) * 100.0).astype(np.float32).astype(np.float64)
    pts = np.concatenate([base, base[::7], base[3:9]])
    with Delaunay(len(pts)) as d:
        tri, opp = d.compute(pts)
        assert d.stats()["duplicates"] == len(base[::7]) + 6
    assert tri.max() < len(base)
    assert tri_set(tri) == tri_set(QDelaunay(base).simplices)
    check_structure(pts, tri, opp)


def test_host_and_device_buffers(gpu):
    import torch
    from cvtamd.delaunay import Delaunay
    g = np.random.default_rng(12)
    pts = g.random((3000, 2)) * 512.0
    with Delaunay(4000) as d:
        tri, opp = d.compute(pts)
        tri2, opp2 = d.compute(torch.from_numpy(pts).cuda())
        tri3, _ = d.compute(pts[:100], want_opp=False)
    assert np.array_equal(tri, tri2) and np.array_equal(opp, opp2)
    assert tri_set(tri3) == tri_set(__import__("scipy.spatial", fromlist=["Delaunay"]).Delaunay(pts[:100]).simplices)


def test_errors(gpu):
    import cvtamd
    from cvtamd.delaunay import Delaunay
    with Delaunay(100) as d:
        with pytest.raises(cvtamd.CvtError):
            d.compute(np.array([[0.0, 0.0], [1.0, 1.0], [2.0, 2.0], [3.0, 3.0]]))   # collinear
        with pytest.raises(cvtamd.CvtError):
            d.compute(np.array([[0.0, 0.0], [1.0, 0.0], [np.nan, 1.0]]))
        with pytest.raises(cvtamd.CvtError):
            d.compute(np.zeros((101, 2)))                                          # capacity
        tri, _ = d.compute(np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]))
        assert tri.tolist() == [[0, 1, 2]]
