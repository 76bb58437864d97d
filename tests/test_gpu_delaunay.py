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


def check_structure(pts, tri, opp, flat_ok=False):
    pts = np.asarray(pts, np.float64)
    n, nt = len(pts), len(tri)
    a, b, c = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    o = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    if flat_ok:
        assert (o >= 0).all(), "every triangle counterclockwise or flat"
    else:
        assert (o > 0).all(), "every triangle counterclockwise"
    hull_edges = 0
    for t in range(nt):
        for e in range(3):
            r = opp[t, e]
            if r < 0:
                hull_edges += 1
                continue
            m, k = r >> 4, r & 3
            assert opp[m, k] & ~4 == (t << 4) | e          # bit 2: a constraint edge
            assert {tri[t, (e + 1) % 3], tri[t, (e + 2) % 3]} == {tri[m, (k + 1) % 3], tri[m, (k + 2) % 3]}
    used = len(np.unique(tri))
    assert nt == 2 * used - hull_edges - 2, "Euler: 2n - h - 2"
    return hull_edges


def locally_delaunay(pts, tri, opp):
    """every interior edge between triangles of positive area: the opposite
    vertex is not inside the circle (exact device predicates)"""
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
    q = np.array(q)
    o, ic = predicates(q)
    # a flat triangle (zero orientation) has no finite circle to test
    keep = o != 0
    return (ic[keep] <= 0).all()


def hull_area2(pts):
    """twice the convex hull area, exactly (monotone chain in rationals)"""
    from fractions import Fraction as F
    P = sorted(set((F(float(x)), F(float(y))) for x, y in pts))

    def cross(o, a, b):
        return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])
    lower, upper = [], []
    for p in P:
        while len(lower) >= 2 and cross(lower[-2], lower[-1], p) <= 0:
            lower.pop()
        lower.append(p)
    for p in reversed(P):
        while len(upper) >= 2 and cross(upper[-2], upper[-1], p) <= 0:
            upper.pop()
        upper.append(p)
    H = lower[:-1] + upper[:-1]
    return abs(sum(H[i][0] * H[(i + 1) % len(H)][1] - H[(i + 1) % len(H)][0] * H[i][1]
                   for i in range(len(H))))


def tri_area2(pts, tri):
    from fractions import Fraction as F
    tot = F(0)
    for a, b, c in tri:
        ax, ay = F(float(pts[a][0])), F(float(pts[a][1]))
        bx, by = F(float(pts[b][0])), F(float(pts[b][1]))
        cx, cy = F(float(pts[c][0])), F(float(pts[c][1]))
        tot += (bx - ax) * (cy - ay) - (by - ay) * (cx - ax)
    return tot


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


def circle_lattice_points(R):
    """every integer point of the circle x^2 + y^2 = R^2"""
    pts = []
    for x in range(-R, R + 1):
        y2 = R * R - x * x
        y = int(round(y2 ** 0.5))
        if y * y == y2:
            pts.append((x, y))
            if y:
                pts.append((x, -y))
    return np.array(pts, np.float64)


@pytest.mark.parametrize("case", ["grid", "grid_subset", "circle", "circle_center", "lines"])
def test_degenerate_inputs(gpu, case):
    """collinear and cocircular points everywhere: every distinct point is a
    vertex, triangles are counterclockwise or flat, the links agree, the
    triangles tile the convex hull exactly, and every interior edge between
    triangles of positive area is Delaunay (exact predicates)"""
    from cvtamd.delaunay import Delaunay
    g = np.random.default_rng(7)
    if case == "grid":
        xs, ys = np.meshgrid(np.arange(30.0), np.arange(30.0))
        pts = np.stack([xs.ravel(), ys.ravel()], 1)
    elif case == "grid_subset":
        cells = g.choice(200 * 200, 5000, replace=False)
        pts = np.stack([cells % 200, cells // 200], 1).astype(np.float64) * 0.5 + 3.0
    elif case == "circle":
        pts = circle_lattice_points(5525) + 10000.0           # 108 cocircular points
    elif case == "circle_center":
        pts = np.concatenate([circle_lattice_points(5525), [[0.0, 0.0], [1.0, 2.0]]])
    else:
        t = np.arange(200.0)
        pts = np.concatenate([np.stack([t, 0 * t], 1), np.stack([0 * t, t], 1)[1:],
                              np.stack([t, t], 1)[1:], np.stack([t, 199 - t], 1)[1:199]])
    g.shuffle(pts)
    print(f"\ncase {case} n={len(pts)}", flush=True)
    with Delaunay(len(pts)) as d:
        tri, opp = d.compute(pts)
        st = d.stats()
    print(st, flush=True)
    assert st["duplicates"] == 0
    assert set(np.unique(tri)) == set(range(len(pts)))
    check_structure(pts, tri, opp, flat_ok=True)
    assert tri_area2(pts, tri) == hull_area2(pts)
    assert locally_delaunay(pts, tri, opp)


def test_extracted_sites(gpu, orc):
    """the CDT input points of a compat-mode run (self-labelled pixels of the
    label map scaled to the mesh's coordinates, delaunay.h:48-58), which lie
    on a lattice: triangulated as GpuDel::compute takes them (main.cpp:236,
    unconstrained)"""
    from cvtamd import compat
    from cvtamd.delaunay import delaunay
    from test_gpu_compat import reference_init
    n, num = 256, 600
    dens = np.ones((n, n), np.float32)
    mask = np.zeros((n, n), np.uint8)
    vor = reference_init(orc, dens, mask, num)
    it, ns = compat.gcvt(vor, dens, mask, 1, 200)
    sites = compat.extract_points(vor, mask, 1.0 / 255, -0.5, 0.25)
    assert len(sites) == ns
    tri, opp = delaunay(sites)
    assert set(np.unique(tri)) == set(range(len(sites)))
    check_structure(sites, tri, opp, flat_ok=True)
    assert tri_area2(sites, tri) == hull_area2(sites)
    assert locally_delaunay(sites, tri, opp)


# ---------------------------------------------------------------- constraints --
def edge_set(tri):
    e = set()
    for a, b, c in tri:
        for u, v in ((a, b), (b, c), (c, a)):
            e.add((min(u, v), max(u, v)))
    return e


def constrained_delaunay(pts, tri, opp):
    """every interior edge that is not a constraint (bit 2) between triangles
    of positive area: the opposite vertex is not inside the circle"""
    from cvtamd.delaunay import predicates
    q = []
    for t in range(len(tri)):
        for e in range(3):
            r = opp[t, e]
            if r < 0 or r & 4:
                continue
            m, k = r >> 4, r & 3
            v = tri[t]
            q.append(np.concatenate([pts[v[0]], pts[v[1]], pts[v[2]], pts[tri[m, k]]]))
    if not q:
        return True
    q = np.array(q)
    o, ic = predicates(q)
    return (ic[o != 0] <= 0).all()


def check_constrained(pts, segs, tri, opp):
    check_structure(pts, tri, opp, flat_ok=True)
    assert tri_area2(pts, tri) == hull_area2(pts)
    es = edge_set(tri)
    for a, b in segs:
        assert (min(a, b), max(a, b)) in es, (a, b)
    # bit 2 on both sides of exactly the segment edges
    seg_set = {(min(a, b), max(a, b)) for a, b in segs}
    for t in range(len(tri)):
        for e in range(3):
            r = opp[t, e]
            u, v = tri[t, (e + 1) % 3], tri[t, (e + 2) % 3]
            is_seg = (min(u, v), max(u, v)) in seg_set
            if r >= 0:
                assert bool(r & 4) == is_seg, (t, e)
                m, k = r >> 4, r & 3
                assert opp[m, k] == (t << 4) | e | (r & 4)
    assert constrained_delaunay(pts, tri, opp)


def star_polygon(g, n, c=1000.0, r0=200.0, r1=900.0):
    """a star-shaped polygon about (c, c): its edges never cross"""
    ang = np.sort(g.random(n)) * 2 * np.pi
    rad = r0 + g.random(n) * (r1 - r0)
    pts = np.stack([c + rad * np.cos(ang), c + rad * np.sin(ang)], 1)
    pts = pts.astype(np.float32).astype(np.float64)
    segs = np.stack([np.arange(n), (np.arange(n) + 1) % n], 1)
    return pts, segs


@pytest.mark.parametrize("case", ["star", "star_inside", "lattice_chords", "square_boundary"])
def test_constrained(gpu, case):
    from cvtamd.delaunay import Delaunay
    g = np.random.default_rng(23)
    if case == "star":
        pts, segs = star_polygon(g, 400)
    elif case == "star_inside":
        poly, segs = star_polygon(g, 300)
        inner = 1000.0 + (g.random((4000, 2)) - 0.5) * 2000.0
        pts = np.concatenate([poly, inner.astype(np.float32).astype(np.float64)])
    elif case == "lattice_chords":
        xs, ys = np.meshgrid(np.arange(40.0), np.arange(40.0))
        pts = np.stack([xs.ravel(), ys.ravel()], 1)
        idx = lambda x, y: y * 40 + x
        # parallel chords of slopes 5/13 and -2/5: no lattice point inside
        # them (coprime steps), none crossing
        segs = [(idx(0, y), idx(13, y + 5)) for y in range(0, 30, 4)]
        segs += [(idx(39, y), idx(39 - 5, y + 2)) for y in range(1, 30, 3)]
        segs = np.array(segs)
    else:
        # the domain boundary as constraint points and segments (delaunayInput,
        # delaunay.h:60-79): collinear points along the hull
        side = np.linspace(0.0, 100.0, 21)
        ring = np.concatenate([np.stack([side, 0 * side], 1)[:-1],
                               np.stack([0 * side + 100.0, side], 1)[:-1],
                               np.stack([side[::-1], 0 * side + 100.0], 1)[:-1],
                               np.stack([0 * side, side[::-1]], 1)[:-1]])
        inner = 1.0 + g.random((2000, 2)) * 98.0
        pts = np.concatenate([inner.astype(np.float32).astype(np.float64), ring])
        nr = len(ring)
        segs = np.stack([2000 + np.arange(nr), 2000 + (np.arange(nr) + 1) % nr], 1)
    segs = np.asarray(segs, np.int32)
    print(f"\ncase {case} n={len(pts)} segments={len(segs)}", flush=True)
    with Delaunay(len(pts)) as d:
        tri, opp = d.compute(pts, segments=segs)
        print(d.stats(), flush=True)
        tri0, _ = d.compute(pts)
    missing = [s for s in segs if (min(s), max(s)) not in edge_set(tri0)]
    print(f"{len(missing)} segments were not Delaunay edges", flush=True)
    if case in ("star", "star_inside", "lattice_chords"):
        assert len(missing) > 0
    check_constrained(pts, segs, tri, opp)


def test_constrained_errors(gpu):
    import cvtamd
    from cvtamd.delaunay import Delaunay
    pts = np.array([[0.0, 0.0], [10.0, 0.0], [10.0, 10.0], [0.0, 10.0], [5.0, 1.0], [5.0, 9.0],
                    [1.0, 5.0], [9.0, 5.0]])
    with Delaunay(16) as d:
        with pytest.raises(cvtamd.CvtError):
            d.compute(pts, segments=np.array([[4, 5], [6, 7]], np.int32))       # they cross
        with pytest.raises(cvtamd.CvtError):
            d.compute(pts, segments=np.array([[0, 8]], np.int32))               # out of range
        tri, opp = d.compute(pts, segments=np.array([[4, 5]], np.int32))
        check_constrained(pts, [(4, 5)], tri, opp)
        tri, opp = d.compute(pts, segments=np.zeros((0, 2), np.int32))
        assert tri_set(tri) == tri_set(d.compute(pts)[0])


@pytest.mark.parametrize("n,seed", [(2000, 31), (20000, 32), (100000, 33)])
def test_constrained_by_a_full_triangulation(gpu, n, seed):
    """all edges of another triangulation of the same points (Qhull's Delaunay
    triangulation of the points stretched 3x in y: an affine map keeps it
    planar) as constraints: the constrained triangulation must be that
    triangulation, every edge a constraint.  At 100K points more segments win
    their cavities in one round than there are workspace slots (16384), so
    the rest wait for later rounds."""
    from scipy.spatial import Delaunay as QDelaunay
    from cvtamd.delaunay import Delaunay
    g = np.random.default_rng(seed)
    pts = (g.random((n, 2)) * 1000.0).astype(np.float32).astype(np.float64)
    other = QDelaunay(pts * np.array([1.0, 3.0])).simplices
    segs = np.array(sorted(edge_set(other)), np.int32)
    with Delaunay(n) as d:
        tri0, _ = d.compute(pts)
        tri, opp = d.compute(pts, segments=segs)
        print(f"\n{n} points, {len(segs)} segments, "
              f"{len(edge_set(other) - edge_set(tri0))} not Delaunay edges", flush=True)
        print(d.stats(), flush=True)
    assert tri_set(tri) == tri_set(other)
    assert ((opp < 0) | (opp & 4 != 0)).all()          # every interior edge constrained
    check_structure(pts, tri, opp)
