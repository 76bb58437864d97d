"""Pins the CPU oracle (oracle/cvt_oracle.c) against what the mathematics fixes.

The reference holds no tests, golden vectors or fixtures for this path, so the
oracle is checked against closed forms, brute force, invariants and the
reference RNG's closed-form values; each check is chosen so that a plausible
slip (a dropped term, a wrong sign or index, a swapped operand) fails it.
"""
import numpy as np
import pytest


def rand_sites(seed, n, S, lo=0.0, hi=None):
    g = np.random.default_rng(seed)
    hi = n if hi is None else hi
    return (lo + g.random((S, 2)) * (hi - lo)).astype(np.float32)


# ---------------------------------------------------------------- schedule --
@pytest.mark.parametrize("n,first,passes", [(512, 256, 10), (2048, 1024, 12), (4096, 2048, 13),
                                            (576, 512, 11), (256, 128, 9), (2, 1, 2)])
def test_jfa_schedule(orc, n, first, passes):
    assert orc.jfa_first_step(n, n) == first
    assert orc.jfa_num_passes(n, n) == passes


# ------------------------------------------------------------------ voronoi --
def test_single_site_covers_grid(orc):
    xy = np.array([[10.3, 40.7]], np.float32)
    for lab in (orc.jfa(64, 48, xy), orc.voronoi_brute(64, 48, xy), orc.voronoi_exact(64, 48, xy)):
        assert (lab == 0).all()


def test_two_site_bisector(orc):
    # sites left/right of x = 32: the bisector is the vertical line x = 32 exactly
    xy = np.array([[20.5, 30.25], [43.5, 30.25]], np.float32)
    lab = orc.voronoi_brute(64, 64, xy)
    x = np.arange(64)[None, :] + 0.5
    expect = np.where(x < 32, 0, 1)
    expect = np.broadcast_to(expect, (64, 64))
    assert np.array_equal(lab, expect)
    assert np.array_equal(orc.jfa(64, 64, xy), expect)


def test_tie_goes_to_lower_index(orc):
    # two coincident sites: the lower index owns every pixel
    xy = np.array([[5.5, 5.5], [5.5, 5.5], [50.5, 50.5]], np.float32)
    for lab in (orc.voronoi_brute(64, 64, xy), orc.jfa(64, 64, xy)):
        assert not (lab == 1).any()
    # symmetric pair: pixel centres on the bisector go to index 0
    xy = np.array([[10.5, 20.5], [30.5, 20.5]], np.float32)
    lab = orc.voronoi_brute(41, 41, xy)
    assert (lab[:, 20] == 0).all()


def test_brute_assigns_minimum_distance(orc):
    n, S = 48, 37
    xy = rand_sites(3, n, S)
    lab = orc.voronoi_brute(n, n, xy)
    px = np.arange(n, dtype=np.float64)[None, :, None] + 0.5
    py = np.arange(n, dtype=np.float64)[:, None, None] + 0.5
    d = (xy[None, None, :, 0] - px) ** 2 + (xy[None, None, :, 1] - py) ** 2
    dmin = d.min(axis=2)
    dl = np.take_along_axis(d, lab[..., None].astype(np.int64), axis=2)[..., 0]
    # float64 distances of the chosen site equal the minimum up to fp32 rounding
    assert np.all(dl <= dmin * (1 + 1e-6) + 1e-6)


@pytest.mark.parametrize("n,S,seed", [(64, 1, 0), (64, 7, 1), (100, 50, 2), (128, 300, 3),
                                      (200, 40, 4), (77, 500, 5)])
def test_exact_equals_brute(orc, n, S, seed):
    xy = rand_sites(seed, n, S)
    assert np.array_equal(orc.voronoi_exact(n, n, xy), orc.voronoi_brute(n, n, xy))


def test_exact_equals_brute_awkward_sites(orc):
    n = 96
    g = np.random.default_rng(9)
    xy = rand_sites(9, n, 60, lo=-20, hi=n + 20)       # some outside the grid
    xy[10] = xy[11]                                    # duplicates
    xy[20:30] = np.round(xy[20:30]) + 0.5              # pixel centres: exact ties
    assert np.array_equal(orc.voronoi_exact(n, 80, xy), orc.voronoi_brute(n, 80, xy))


def test_jfa_close_to_exact(orc):
    # JFA(+1) is not exact; its error rate is what the parity report quotes
    n, S = 256, 512
    xy = rand_sites(11, n, S)
    j = orc.jfa(n, n, xy)
    b = orc.voronoi_brute(n, n, xy)
    # almost all of JFA's error comes from sites that share a pixel: only the
    # lowest index is seeded, the others lose their cell for this iteration
    seedmap = orc.seed(n, n, xy)
    seeded = np.zeros(S, bool)
    seeded[seedmap[seedmap != orc.INVALID]] = True
    assert (~seeded).sum() == 4
    ok = seeded[b]
    assert ((j != b) & ok).sum() / ok.sum() < 1e-4
    assert (j != b).mean() < 5e-3
    # a mislabelled pixel's site is never much farther than the true nearest
    px = np.arange(n)[None, :] + 0.5
    py = np.arange(n)[:, None] + 0.5
    dj = (xy[j, 0] - px) ** 2 + (xy[j, 1] - py) ** 2
    db = (xy[b, 0] - px) ** 2 + (xy[b, 1] - py) ** 2
    assert np.all(dj >= db - 1e-3)


def test_jfa_pass_semantics(orc):
    # one pass at step 4 from a seed map equals the 9-tap rule evaluated by hand
    n = 20
    xy = np.array([[1.5, 1.5], [9.5, 2.5], [5.5, 13.5]], np.float32)
    seed = orc.seed(n, n, xy)
    out = orc.jfa_pass(n, n, xy, seed, 4)
    # pixel (5, 5): taps (1..9 step 4)^2 -> holds (1,1)=0 and (9,1)?  seed at (9,2) is not a tap
    assert out[5, 5] == 0
    # pixel (5, 9): taps rows 5, 9, 13 x cols 1, 5, 9 -> (5,13) holds site 2
    assert out[9, 5] == 2
    # pixel (19, 19): taps (15,15),(15,19),(19,15)... and (x-4,y-4): no seeds -> none
    assert out[19, 19] == orc.INVALID


def test_seed_lowest_index_wins(orc):
    xy = np.array([[3.2, 3.9], [3.7, 3.1], [-5, 100]], np.float32)
    lab = orc.seed(8, 8, xy)
    assert lab[3, 3] == 0
    assert lab[7, 0] == 2          # clamped to the grid
    assert (lab != orc.INVALID).sum() == 2


# ------------------------------------------------------------------ moments --
@pytest.mark.parametrize("a", [8, 16, 32])
def test_lattice_closed_form(orc, a):
    n = 4 * a
    c = (np.arange(4) * a + a / 2).astype(np.float32)
    xy = np.stack(np.meshgrid(c, c), -1).reshape(-1, 2).astype(np.float32)
    rho = np.ones((n, n), np.float32)
    lab = orc.voronoi_exact(n, n, xy)
    m = orc.moments(rho, lab, xy)
    assert np.all(m[:, 0] == a * a)
    assert np.all(m[:, 1] == 0) and np.all(m[:, 2] == 0)
    assert np.all(m[:, 3] == a * a * (a * a - 1) / 6)     # 672, 10880, 174592
    # a Lloyd fixed point
    assert np.array_equal(orc.update(m, xy, 1.0), xy)


def test_single_site_uniform_square(orc):
    n = 64
    xy = np.array([[17.25, 40.5]], np.float32)
    rho = np.ones((n, n), np.float32)
    lab = orc.jfa(n, n, xy)
    m = orc.moments(rho, lab, xy)[0]
    assert m[0] == n * n
    # centroid of the square is (n/2, n/2): DX = W (n/2 - sx)
    assert m[1] == pytest.approx(n * n * (n / 2 - 17.25), rel=1e-12)
    assert m[2] == pytest.approx(n * n * (n / 2 - 40.5), rel=1e-12)
    e_centre = n * n * (n * n - 1) / 6
    shift = (n / 2 - 17.25) ** 2 + (n / 2 - 40.5) ** 2
    assert m[3] == pytest.approx(e_centre + n * n * shift, rel=1e-12)
    new = orc.update(orc.moments(rho, lab, xy), xy, 1.0)
    assert np.allclose(new, [[n / 2, n / 2]])


def test_moments_match_numpy_and_conserve_mass(orc):
    n, S = 96, 70
    xy = rand_sites(5, n, S)
    g = np.random.default_rng(6)
    rho = (g.random((n, n)) * (g.random((n, n)) > 0.2)).astype(np.float32)
    lab = orc.jfa(n, n, xy)
    m = orc.moments(rho, lab, xy)
    ys, xs = np.mgrid[0:n, 0:n]
    r = rho.astype(np.float64).ravel()
    L = lab.ravel()
    dx = xs.ravel() + 0.5 - xy[L, 0].astype(np.float64)
    dy = ys.ravel() + 0.5 - xy[L, 1].astype(np.float64)
    ref = np.stack([np.bincount(L, r, S), np.bincount(L, r * dx, S), np.bincount(L, r * dy, S),
                    np.bincount(L, r * (dx * dx + dy * dy), S)], 1)
    assert np.allclose(m, ref, rtol=1e-12, atol=1e-9)
    assert m[:, 0].sum() == pytest.approx(r.sum(), rel=1e-12)


def test_energy_and_gradient_identity(orc):
    n, S = 64, 20
    xy = rand_sites(8, n, S)
    rho = np.ones((n, n), np.float32)
    lab = orc.voronoi_exact(n, n, xy)
    m = orc.moments(rho, lab, xy)
    e, gnorm = orc.energy_from_moments(n, n, m)
    assert e == pytest.approx(m[:, 3].sum() / (n * n), rel=1e-14)
    # finite difference of the energy w.r.t. one site with labels held fixed
    i, h = 3, 1e-3

    def energy_at(sxy):
        mm = orc.moments(rho, lab, sxy.astype(np.float32))
        return mm[:, 3].sum() / (n * n)

    xp, xm = xy.astype(np.float64).copy(), xy.astype(np.float64).copy()
    xp[i, 0] += h
    xm[i, 0] -= h
    fd = (energy_at(xp) - energy_at(xm)) / (2 * h)
    # analytic dE/dsx = -2 DX / (n^2)
    assert fd == pytest.approx(-2 * m[i, 1] / (n * n), rel=1e-3, abs=1e-9)
    assert gnorm == pytest.approx(2 * np.sqrt((m[:, 1] ** 2 + m[:, 2] ** 2).sum()) / (n * n))


def test_lloyd_energy_non_increasing(orc):
    n, S = 128, 64
    xy = rand_sites(12, n, S)
    g = np.random.default_rng(13)
    rho = (0.2 + g.random((n, n))).astype(np.float32)
    _, e, _ = orc.lloyd(rho, xy, 20, omega=1.0, exact=True)
    assert np.all(np.diff(e) <= 1e-6 * e[:-1])
    assert e[-1] < e[0]


def test_lloyd_zero_density_keeps_sites(orc):
    n = 32
    xy = rand_sites(1, n, 5)
    rho = np.zeros((n, n), np.float32)
    out, e, g = orc.lloyd(rho, xy, 3)
    assert np.array_equal(out, xy)
    assert np.all(e == 0)


def test_fixed_sites_do_not_move(orc):
    n, S = 64, 10
    xy = rand_sites(2, n, S)
    fixed = np.zeros(S, np.uint8)
    fixed[[1, 4]] = 1
    out, _, _ = orc.lloyd(np.ones((n, n), np.float32), xy, 5, fixed=fixed)
    assert np.array_equal(out[[1, 4]], xy[[1, 4]])
    assert not np.array_equal(out[0], xy[0])


# --------------------------------------------------------------- rasterizer --
def square_mesh(k, size):
    """k x k grid of vertices over [0, size]^2, two triangles per quad."""
    v = np.linspace(0, size, k)
    X, Y = np.meshgrid(v, v)
    pts = np.stack([X.ravel(), Y.ravel()], 1)
    tris = []
    for j in range(k - 1):
        for i in range(k - 1):
            a, b, c, d = j * k + i, j * k + i + 1, (j + 1) * k + i, (j + 1) * k + i + 1
            tris += [[a, b, d], [a, d, c]]
    return pts, np.array(tris, np.int32)


def test_raster_constant_and_outside(orc):
    n = 64
    pts, tri = square_mesh(5, 40.0)
    scale = 1.0                              # sample (x, y) = pixel index
    dens = orc.raster(pts, np.full(len(pts), 2.5), tri, scale, n)
    x, y = np.arange(n)[None, :], np.arange(n)[:, None]
    inside = (x < 40) & (y < 40)
    hull = (x <= 40) & (y <= 40)
    assert np.all(dens[inside] == 2.5)
    assert np.all(dens[~hull] == 0)
    # on the hull's outer edge the fp64 weights may round to -1e-17: either value
    assert np.all(np.isin(dens[hull & ~inside], [0.0, 2.5]))


def test_raster_linear_field(orc):
    n = 64
    pts, tri = square_mesh(6, 60.0)
    w = 0.5 + 0.01 * pts[:, 0] - 0.02 * pts[:, 1]
    scale = 60.0 / (n - 1)
    dens = orc.raster(pts, w, tri, scale, n)
    x = np.arange(n)[None, :] * scale
    y = np.arange(n)[:, None] * scale
    expect = (0.5 + 0.01 * x - 0.02 * y).astype(np.float32)
    # interior samples (the last row / column lie on the hull's edge)
    assert np.allclose(dens[:-1, :-1], expect[:-1, :-1], rtol=0, atol=1e-6)


def test_raster_first_hit_and_degenerate(orc):
    n = 32
    pts = np.array([[0, 0], [20, 0], [0, 20], [5, 5], [30, 5], [5, 30], [1, 1], [2, 2], [3, 3]],
                   np.float64)
    w = np.array([1, 1, 1, 7, 7, 7, 9, 9, 9], np.float64)
    tri = np.array([[6, 7, 8], [0, 1, 2], [3, 4, 5]], np.int32)    # first one is degenerate
    dens = orc.raster(pts, w, tri, 1.0, n)
    assert dens[6, 6] == 1          # in both triangles: the first listed wins
    assert dens[20, 10] == 7        # only in the second
    assert dens[2, 2] == 1          # on the degenerate triangle: skipped
    assert dens[31, 31] == 0


# ------------------------------------------------------- reference init RNG --
def test_reference_rng_values(orc):
    k = orc.kiss(0)
    got = [orc.kiss_next(k) for _ in range(6)]
    assert got == [1234567, 3667164066, 249762113, 2231956628, 4047385867, 2359287638]
    # only the LCG term is live with a zero seed: x' = 69069 x + 1234567 mod 2^32
    x = 0
    for v in got:
        x = (69069 * x + 1234567) % 2**32
        assert v == x


@pytest.mark.parametrize("n,first", [(512, (0, 437)), (2048, (0, 1748))])
def test_reference_first_site(orc, n, first):
    dens = np.ones((n, n), np.float32)
    vor = orc.ref_put_constraints(np.zeros((n, n), np.uint8))
    orc.ref_random_points(vor, dens, 1)
    ys, xs = np.nonzero(vor[:, :, 0] != -32768)
    assert (xs[0], ys[0]) == first
    assert tuple(vor[ys[0], xs[0]]) == first


def test_reference_random_points_unique_and_on_density(orc):
    n = 64
    dens = np.zeros((n, n), np.float32)
    dens[10:50, 5:60] = 1.0
    mask = np.zeros((n, n), np.uint8)
    mask[12, 7] = 1
    vor = orc.ref_put_constraints(mask)
    orc.ref_random_points(vor, dens, 200)
    xy, fx = orc.compat_decode(vor, mask)
    assert len(xy) == 201 and fx.sum() == 1
    px, py = np.floor(xy[:, 0]).astype(int), np.floor(xy[:, 1]).astype(int)
    assert np.all(dens[py, px] > 0)


# --------------------------------------------------------------- compat mode --
def test_compat_update_rules(orc):
    n = 16
    rho = np.ones((n, n), np.float32)
    rho[0, 15] = 0.0
    mask = np.zeros((n, n), np.uint8)
    mask[8, 8] = 1
    xy = np.array([[2.5, 2.5], [8.5, 8.5], [10.5, 3.5], [14.5, 0.5], [5.5, 5.5]], np.float32)
    m = np.zeros((5, 4))
    m[:, 0] = 1.0
    m[0, 1:3] = [1.3, -0.2]      # s + 2*(1.3, -0.2) + 0.5 = (5.1, 2.1) -> (5, 2)
    m[1, 1:3] = [3.0, 3.0]       # fixed (mask pixel): stays at (8, 8)
    m[2, 1:3] = [-0.75, 0.24]    # (10 - 1.5 + .5, 3 + .48 + .5) = (9.0, 3.98) -> (9, 3)
    m[3, 1:3] = [1.0, -0.1]      # target (15, 0) has rho 0: stays at (14, 0)
    m[4, 1:3] = [-0.1, -1.4]     # (5 - .2 + .5, 5 - 2.8 + .5) = (5.3, 2.7) -> (5, 2): merges with 0
    out = orc.compat_update(rho, m, xy, 2.0, mask)
    assert out.tolist() == [[5.5, 2.5], [8.5, 8.5], [9.5, 3.5], [14.5, 0.5]]


def test_compat_update_truncates_toward_zero_and_clamps(orc):
    n = 8
    rho = np.ones((n, n), np.float32)
    xy = np.array([[0.5, 7.5]], np.float32)
    m = np.array([[1.0, -0.6, 5.0, 0.0]])
    # x: 0 - 0.6 + 0.5 = -0.1 -> trunc 0 (floor would give -1); y: 7 + 5 + .5 -> clamp 7
    out = orc.compat_update(rho, m, xy, 1.0)
    assert out.tolist() == [[0.5, 7.5]]


def test_compat_lloyd_invariants(orc):
    n = 64
    g = np.random.default_rng(4)
    rho = np.ones((n, n), np.float32)
    rho[:, 50:] = 0.0
    mask = np.zeros((n, n), np.uint8)
    mask[[5, 40], [5, 30]] = 1
    vor = orc.ref_put_constraints(mask)
    orc.ref_random_points(vor, rho, 40)
    S0 = len(orc.compat_decode(vor, mask)[0])
    it, elog, S = orc.compat_lloyd(rho, vor, mask, max_iter=60)
    assert 1 <= it <= 60 and S <= S0
    xy, fx = orc.compat_decode(vor, mask)
    assert len(xy) == S
    # fixed sites survive in place; no site on rho == 0
    assert vor[5, 5].tolist() == [5, 5] and vor[40, 30].tolist() == [30, 40]
    assert np.all(rho[np.floor(xy[:, 1]).astype(int), np.floor(xy[:, 0]).astype(int)] > 0)
    # the output is a full coordinate label map of nearest sites
    assert (vor[:, :, 0] >= 0).all()


def sites_map(n, pts, stride=None):
    """A reference input map: MARKER everywhere, (x, y) at each site; with
    stride m < n the sites are written as an m x m map into the first m^2
    pairs (the layout gCVT reads for its coarsest level)."""
    vor = np.full((n, n, 2), -32768, np.int16)
    flat = vor.reshape(-1, 2)
    m = stride or n
    for x, y in pts:
        flat[y * m + x] = (x, y)
    return vor


def test_compat_levels(orc):
    # source/gcvt.cu:1091: levels stop before a side below 256
    assert orc.compat_levels(512, 1) == 1
    assert orc.compat_levels(512, 2) == 2
    assert orc.compat_levels(512, 3) == 2
    assert orc.compat_levels(1024, 3) == 3
    assert orc.compat_levels(2048, 9) == 4
    assert orc.compat_levels(256, 2) == 1
    assert orc.compat_levels(200, 3) == 1


def test_compat_lattice_fixed_point_stops_after_20(orc):
    # 9 x 9 blocks of 7 x 7 pixels, a site at each block centre: every cell is
    # its block and every centroid its site, so nothing moves; E(0) = E(10) =
    # a^2 (a^2 - 1) / 6 per block * 81 / 63^2 = (a^2 - 1) / 6 = 8 exactly.
    # Check 10: dE = 1e18 - 8 (go on); check 20: dE = 0 < 1e-4: stop.
    n, a = 63, 7
    pts = [(a * i + 3, a * j + 3) for j in range(9) for i in range(9)]
    rho = np.ones((n, n), np.float32)
    vor = sites_map(n, pts)
    it, elog, S = orc.compat_gcvt(rho, vor, None, 1, 1000)
    assert it == 20 and S == 81
    assert np.allclose(elog, [8.0, 8.0], rtol=1e-12)
    xy, _ = orc.compat_decode(vor)
    assert sorted(map(tuple, np.floor(xy).astype(int).tolist())) == sorted(pts)
    bx = np.arange(n) // a * a + 3
    assert (vor[:, :, 0] == bx[None, :]).all() and (vor[:, :, 1] == bx[:, None]).all()


def test_compat_omega_schedule_hand_traced(orc):
    # 9 x 9 grid, rho = 1/64 on columns 0..2, one site at (8, 4): its cell is
    # the grid, the centroid is (1, 4).  omega = 2 reflects x about 1:
    # 8 -> trunc(-5.5) = -5 -> clamp 0, then 0 -> 2 -> 0 -> 2 ...  Samples
    # (x at iterations 0, 10, 20 = 8, 2, 2):
    #   E(x) = rho (9 sum_c (x - c)^2 + 3 * 60) / 81, c = 0..2
    #   E(8) = 1521/5184, E(2) = 225/5184.
    # Check 10: omega = min(2, 1 + 1e18) = 2.  Check 20: dE = 1296/5184 =
    # 0.25, omega = 1.25, go on.  Iteration 20: 2 + 1.25 (1 - 2) + .5 = 1.25
    # -> 1, the centroid: fixed.  Check 30: dE = E(2) - E(2) = 0: stop.
    # (With omega still 2 the site would end at 2; an energy sampled after
    # the update, or a reversed dE, would change the count or the samples.)
    n = 9
    rho = np.zeros((n, n), np.float32)
    rho[:, :3] = 1.0 / 64.0
    vor = sites_map(n, [(8, 4)])
    it, elog, S = orc.compat_gcvt(rho, vor, None, 1, 1000)
    assert it == 30 and S == 1
    assert np.allclose(elog, np.array([1521, 225, 225]) / 5184.0, rtol=1e-12, atol=0)
    assert (vor[:, :, 0] == 1).all() and (vor[:, :, 1] == 4).all()


def _sq_sum(c, n):
    return float(sum((c - x) ** 2 for x in range(n)))


def test_compat_multires_carries_omega_and_last_energy(orc):
    # 512^2, depth 2, rho = 1: one site at coarse (127, 127) in the 256 x 256
    # map read from the first 256^2 pairs.  Coarse: centroid 127.5, omega = 2
    # swaps 127 <-> 128, so E(0) = E(10) (= 4 * 2 * 256 * S1 / 256^2, S1 =
    # sum (127 - x)^2 over 0..255); check 20: dE = 0 < 0.3 stops the level
    # with omega = min(2, 1 + 0) = 1 and lastE = E(0) kept.  Zoom: (254, 254).
    # Fine, omega = 1: 254 + 1.5 + .5 -> 256, the centroid's pixel (255.5 -
    # 256 = -0.5 keeps it).  Check 30: dE = E(0) - E_fine(20) < 0: stop.
    # Omega reset to 2 would leave the site swapping 254 <-> 257; lastE reset
    # to 1e18 would run past 30.
    n = 512
    rho = np.ones((n, n), np.float32)
    vor = sites_map(n, [(127, 127)], stride=256)
    it, elog, S = orc.compat_gcvt(rho, vor, None, 2, 1000)
    ec = 4 * 2 * 256 * _sq_sum(127, 256) / 256 ** 2
    ef = 2 * 512 * _sq_sum(254, 512) / 512 ** 2
    assert it == 30 and S == 1
    assert np.allclose(elog, [ec, ec, ef], rtol=1e-12)
    assert ef > ec
    assert (vor == 256).all()


def test_compat_multires_runs_one_iteration_per_level(orc):
    # the same start with maxIter 15: the coarse level runs iterations 0..14
    # (check 10 goes on), the fine level's do-while still runs one: 16 in
    # all.  Coarse iteration 14 leaves the site at 128 -> 256; the fine step
    # with omega = 2: 256 + 2 (-0.5) + .5 = 255.5 -> 255.
    n = 512
    rho = np.ones((n, n), np.float32)
    vor = sites_map(n, [(127, 127)], stride=256)
    it, elog, S = orc.compat_gcvt(rho, vor, None, 2, 15)
    assert it == 16 and S == 1
    assert (vor == 255).all()


def test_compat_multires_coarse_mask_is_strided(orc):
    # a coarse level reads mask with its own row stride: a site at coarse
    # (3, 1) is fixed by mask byte 1 * 256 + 3, which is fine pixel (259, 0)
    n = 512
    rho = np.ones((n, n), np.float32)
    mask = np.zeros((n, n), np.uint8)
    mask.reshape(-1)[1 * 256 + 3] = 1
    vor = sites_map(n, [(3, 1)], stride=256)
    it, elog, S = orc.compat_gcvt(rho, vor, mask, 2, 5)
    # five coarse iterations with the site pinned, then one fine iteration
    # from (6, 2), which is not masked at the fine level: omega = 2 moves it
    assert it == 6 and S == 1
    xy, _ = orc.compat_decode(vor)
    fx, fy = np.floor(xy[0]).astype(int)
    # fine step from (6, 2): trunc(6 + 2 (255.5 - 6) + .5) = 505 -> (505, 509)
    assert (fx, fy) == (505, 509)
    # without the mask the coarse level would move the site
    vor2 = sites_map(n, [(3, 1)], stride=256)
    orc.compat_gcvt(rho, vor2, None, 2, 5)
    assert not (vor2 == vor).all()


def test_extract_points_order_and_mask(orc):
    n = 4
    vor = np.full((n, n, 2), -32768, np.int16)
    vor[0, 3] = [3, 0]
    vor[2, 1] = [1, 2]
    vor[1, 1] = [1, 1]
    vor[3, 0] = [0, 3]     # masked below
    vor[2, 2] = [1, 2]     # not self-labelled
    mask = np.zeros((n, n), np.uint8)
    mask[3, 0] = 1
    pts = orc.extract_points(vor, mask, 0.5, 10.0, 20.0)
    # x-major loop order: (1,1), (1,2), (3,0)
    assert pts.tolist() == [[10.5, 20.5], [10.5, 21.0], [11.5, 20.0]]


# ------------------------------------------------------- point location --
def jittered_mesh(m, seed, jitter=0.3):
    """(m x m) jittered grid of [0, m-1]^2 split into 2 (m-1)^2 triangles,
    faces shuffled and some wound clockwise."""
    g = np.random.default_rng(seed)
    x, y = np.meshgrid(np.arange(m, dtype=np.float64), np.arange(m, dtype=np.float64))
    uv = np.stack([x.ravel(), y.ravel()], 1)
    inner = (uv[:, 0] > 0) & (uv[:, 0] < m - 1) & (uv[:, 1] > 0) & (uv[:, 1] < m - 1)
    uv[inner] += g.uniform(-jitter, jitter, (inner.sum(), 2))
    tri = []
    for j in range(m - 1):
        for i in range(m - 1):
            a, b, c, d = j * m + i, j * m + i + 1, (j + 1) * m + i, (j + 1) * m + i + 1
            tri += [[a, b, d], [a, d, c]]
    tri = np.array(tri, np.int32)[g.permutation(2 * (m - 1) ** 2)]
    flip = g.random(len(tri)) < 0.5
    tri[flip] = tri[flip][:, ::-1]
    return uv, tri


def test_locate_vertices_exact_weights(orc):
    # at a face's own vertices the weights are exactly the unit vectors
    uv = np.array([[0.3, 0.1], [2.7, 0.4], [1.1, 3.3]])
    tri = np.array([[0, 1, 2]], np.int32)
    face, w = orc.locate(uv, tri, uv)
    assert np.all(face == 0)
    assert np.array_equal(w, np.eye(3))


def test_locate_centroid_and_reconstruction(orc):
    uv, tri = jittered_mesh(12, 1)
    g = np.random.default_rng(2)
    q = g.uniform(0.01, 10.99, (2000, 2))
    face, w = orc.locate(uv, tri, q)
    assert np.all(face >= 0)
    assert np.all(w >= 0) and np.allclose(w.sum(1), 1.0, atol=1e-12)
    rec = np.einsum("kj,kjd->kd", w, uv[tri[face]])
    assert np.abs(rec - q).max() < 1e-12
    cen = uv[tri].mean(1)
    fc, wc = orc.locate(uv, tri, cen)
    assert np.array_equal(fc, np.arange(len(tri)))           # each centroid in its own face
    assert np.abs(wc - 1.0 / 3.0).max() < 1e-12


def test_locate_first_face_in_order(orc):
    # the shared diagonal of a unit square lies in both faces: the first listed wins
    uv = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    a = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    q = np.array([[0.5, 0.5], [0.25, 0.25], [0.9, 0.1], [0.1, 0.9], [1.5, 0.5]])
    assert orc.locate(uv, a, q)[0].tolist() == [0, 0, 0, 1, -1]
    assert orc.locate(uv, a[::-1].copy(), q)[0].tolist() == [0, 0, 1, 0, -1]
    # a duplicate face never wins over its earlier copy; a degenerate face never holds a point
    b = np.array([[0, 2, 2], [1, 2, 0], [0, 1, 2]], np.int32)
    f, w = orc.locate(uv, b, np.array([[0.5, 0.2]]))
    assert f.tolist() == [1] and np.allclose(w[0], [0.3, 0.2, 0.5], rtol=0, atol=1e-15)


def test_locate_outside_and_nonfinite(orc):
    uv, tri = jittered_mesh(6, 3)
    q = np.array([[-0.5, 2.0], [2.0, 5.5], [np.nan, 1.0], [1.0, np.inf]])
    face, w = orc.locate(uv, tri, q)
    assert np.all(face == -1) and np.all(w == -1)


def test_locate_lifts_linear_fields(orc):
    uv, tri = jittered_mesh(9, 4)
    p3 = np.stack([uv[:, 0], uv[:, 1], 0.7 - 0.2 * uv[:, 0] + 0.35 * uv[:, 1]], 1)
    q = np.random.default_rng(5).uniform(0.0, 8.0, (500, 2))
    face, w, out3 = orc.locate(uv, tri, q, p3=p3)
    assert np.all(face >= 0)
    assert np.abs(out3[:, :2] - q).max() < 1e-12
    assert np.abs(out3[:, 2] - (0.7 - 0.2 * q[:, 0] + 0.35 * q[:, 1])).max() < 1e-12


# ---------------------------------------------------------------------------
# vertex density weights (source/weighting.h:37-97)

def test_weights_flat_and_scaled_copies_exact(orc):
    # integer grid: every face has 2D area 1/2 and every step is exact, so a
    # flat copy weighs exactly 1 and a copy scaled by a exactly a^2
    uv, tri = jittered_mesh(7, 1, jitter=0.0)
    for a in (1.0, 2.0, 3.0):
        p3 = np.stack([a * uv[:, 0], a * uv[:, 1], np.zeros(len(uv))], 1)
        assert np.array_equal(orc.vertex_weights(uv, p3, tri), np.full(len(uv), a * a))
    # a jittered flat copy: 1 up to rounding; the z axis and the other planes
    uv, tri = jittered_mesh(12, 2)
    for p3 in (np.stack([uv[:, 0], uv[:, 1], np.zeros(len(uv))], 1),
               np.stack([np.zeros(len(uv)), uv[:, 0], uv[:, 1]], 1),
               np.stack([uv[:, 1], np.full(len(uv), 4.0), uv[:, 0]], 1)):
        assert np.abs(orc.vertex_weights(uv, p3, tri) - 1.0).max() < 1e-13


def test_weights_tilted_plane(orc):
    # over the plane z = a u + b v every 3D area is sqrt(1 + a^2 + b^2) times
    # the 2D one: all three cross-product components are non-zero there
    uv, tri = jittered_mesh(15, 3)
    for a, b in ((1.0, 0.0), (0.0, -2.0), (0.6, 1.7), (-3.0, 0.25)):
        p3 = np.stack([uv[:, 0], uv[:, 1], a * uv[:, 0] + b * uv[:, 1] + 0.5], 1)
        w = orc.vertex_weights(uv, p3, tri)
        assert np.abs(w / np.sqrt(1 + a * a + b * b) - 1.0).max() < 1e-13
    # the vertex order inside a face changes the area only by rounding
    p3 = np.stack([uv[:, 0], np.sin(uv[:, 1]), np.cos(uv[:, 0] * uv[:, 1] / 7)], 1)
    w0 = orc.vertex_weights(uv, p3, tri)
    w1 = orc.vertex_weights(uv, p3, np.roll(tri, 1, axis=1))
    assert np.abs(w1 / w0 - 1.0).max() < 1e-13


def test_weights_hand_summed_fan(orc):
    # two faces around vertex 0; vertex 3 lifted to z = 1: face 1's 3D area is
    # |(0,1,0) x (-1,0,1)| / 2 = sqrt(2) / 2, every 2D area is 1/2
    uv = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [-1.0, 0.0]])
    p3 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [-1.0, 0.0, 1.0]])
    tri = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    h = np.sqrt(2.0) / 2
    want = np.array([0.5 + h, 1.0, 0.5 + h, 2 * h])
    assert np.abs(orc.vertex_weights(uv, p3, tri) - want).max() < 1e-15


def test_weights_large_rule_and_isolated_vertex(orc):
    uv, tri = jittered_mesh(5, 4, jitter=0.0)
    nv = len(uv)
    p3 = np.stack([uv[:, 0], uv[:, 1], 0.5 * uv[:, 0]], 1)
    base = orc.vertex_weights(uv, p3, tri)
    # a face collinear in 2D (area 0, weight +inf), a sliver whose weight is
    # finite but above 1e10, and a vertex in no face (0 / 0 = NaN)
    extra_uv = np.array([[10.0, 10.0], [11.0, 11.0], [12.0, 12.0],
                         [20.0, 20.0], [20.0 + 1e-6, 20.0], [20.0, 20.0 + 1e-6], [30.0, 30.0]])
    extra_p3 = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0],
                         [0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [9.0, 9.0, 9.0]])
    uv2 = np.concatenate([uv, extra_uv])
    p32 = np.concatenate([p3, extra_p3])
    tri2 = np.concatenate([tri, np.array([[nv, nv + 1, nv + 2], [nv + 3, nv + 4, nv + 5]], np.int32)])
    w = orc.vertex_weights(uv2, p32, tri2)
    assert np.array_equal(w[:nv], base)
    assert np.all(w[nv:nv + 6] == base.max()) and base.max() > 1.0
    assert np.isnan(w[nv + 6])
    # no weight at or below 1e10: the replacement is the initial maximum, 0
    w = orc.vertex_weights(extra_uv[:3], extra_p3[:3], np.array([[0, 1, 2]], np.int32))
    assert np.array_equal(w, np.zeros(3))
