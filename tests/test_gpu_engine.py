"""GPU parity of the Lloyd hot path (JFA(+1), moments, update, energy) against the
CPU oracle, through the C ABI.

Bar: labels bit-exact; per-site moments within the stated fp32 bound of the
oracle's float64 (centroid offset 1e-4 px, cell energy 1e-4 relative); energy
per iteration 1e-5 relative; the float64 energy check 1e-12 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_engine(n, xy, rho, H=None, tile=0):
    import cvtamd
    H = n if H is None else H
    e = cvtamd.Engine(n, H, max(len(xy), 1))
    e.set_density(rho)
    if tile:
        e.tile = tile
    e.set_sites(xy)
    return e


def rand_sites(seed, W, H, S, lo=0.0):
    g = np.random.default_rng(seed)
    xy = g.random((S, 2))
    xy[:, 0] = lo + xy[:, 0] * (W - 2 * lo)
    xy[:, 1] = lo + xy[:, 1] * (H - 2 * lo)
    return xy.astype(np.float32)


def check_moments(m, ref, W_tol=1e-6):
    w, wr = m[:, 0], ref[:, 0]
    assert np.allclose(w, wr, rtol=W_tol, atol=1e-9)
    nz = wr > 0
    off = m[nz, 1:3] / w[nz, None]
    off_r = ref[nz, 1:3] / wr[nz, None]
    assert np.abs(off - off_r).max() < 1e-4
    assert np.allclose(m[nz, 3], ref[nz, 3], rtol=1e-4, atol=1e-6)
    assert np.all(m[~nz] == 0)


# ----------------------------------------------------------------- labels --
@pytest.mark.parametrize("W,H,S,seed", [
    (512, 512, 1024, 1),      # config 1 shape
    (256, 256, 4096, 2),      # dense: many seed collisions
    (576, 576, 700, 3),       # non power of two (batched patch side)
    (320, 200, 333, 4),       # rectangular, not a multiple of 64 in y
    (100, 70, 1, 5),          # one site
    (64, 64, 64, 6),
    (98, 77, 150, 7),         # width not a multiple of 4: one pixel per lane
    (1024, 1024, 30000, 8),   # site table too large for LDS: split LDS / global
    (700, 600, 20479, 14),    # the largest table that fits in LDS
    (700, 600, 20481, 15),    # just too large: all but the last sites in LDS
    (1000, 600, 25000, 16),   # split table, width not a multiple of 256
    (12, 9, 5, 9),            # tiny grid
    (8, 8, 3, 11),            # narrowest grid of the 4-pixel kernel
    (260, 130, 5000, 12),     # 40 KB site table, width not a multiple of 256
    (512, 384, 8000, 10),     # 160 KB site table on a small grid
    (2560, 1700, 8000, 13),   # runs layout over chains of unequal length
    (704, 608, 20470, 17),    # runs with the largest table that LDS holds whole
    (704, 608, 20471, 18),    # runs with the smallest split table
    (256, 320, 900, 19),      # one 256-pixel segment per row
    (32, 40, 20, 20),         # runs on a grid narrower than one segment
    (8, 200, 30, 21),         # fused tail (last five passes in LDS) on a grid one lane wide
    (200, 8, 30, 22),         # fused tail on a grid eight rows high
    (100, 70, 5000, 23),      # fused tail with the 40 KB table, a site on most pixels
    (68, 40, 50, 24),         # the smallest JFA with a fused tail (eight passes)
])
def test_jfa_labels_bit_exact(gpu, orc, W, H, S, seed):
    xy = rand_sites(seed, W, H, S)
    rho = np.ones((H, W), np.float32)
    with make_engine(W, xy, rho, H) as e:
        e.voronoi()
        lab = e.labels()
    assert np.array_equal(lab, orc.jfa(W, H, xy))


@pytest.mark.parametrize("W,H,S,seed", [
    (512, 512, 1024, 31),     # 4-pixel kernel, table in LDS
    (98, 77, 150, 32),        # one pixel per lane
    (1024, 1024, 30000, 33),  # split LDS / global table
])
def test_jfa_after_lloyd_and_reseeding(gpu, orc, W, H, S, seed):
    """The seed map and the first-pass map are reset sparsely (around the seed
    pixels of the sites): labels of a fresh JFA after Lloyd iterations, after
    moving the sites, and of a repeated JFA must still be bit-exact."""
    xy = rand_sites(seed, W, H, S)
    rho = np.ones((H, W), np.float32)
    with make_engine(W, xy, rho, H) as e:
        e.lloyd(3, 1.0)
        e.voronoi()
        cur = e.sites()
        assert np.array_equal(e.labels(), orc.jfa(W, H, cur))
        moved = rand_sites(seed + 100, W, H, S)
        e.set_sites(moved)
        e.voronoi()
        e.voronoi()
        assert np.array_equal(e.labels(), orc.jfa(W, H, moved))
        e.lloyd(2, 1.0)
        e.voronoi()
        assert np.array_equal(e.labels(), orc.jfa(W, H, e.sites()))


def test_jfa_awkward_sites(gpu, orc):
    n = 200
    xy = rand_sites(7, n, n, 300)
    xy[:20] = xy[20:40]                                   # coincident sites
    xy[40:60] = np.floor(xy[40:60]) + 0.5                 # pixel centres: exact ties
    xy[60] = [-30.0, 50.0]                                # outside the grid
    xy[61] = [n + 10.5, n + 3.0]
    with make_engine(n, xy, np.ones((n, n), np.float32)) as e:
        e.voronoi()
        assert np.array_equal(e.labels(), orc.jfa(n, n, xy))


def test_no_sites(gpu, orc):
    import cvtamd
    n = 64
    with cvtamd.Engine(n, n, 8) as e:
        e.set_density(np.ones((n, n), np.float32))
        e.set_sites(np.zeros((0, 2), np.float32))
        e.voronoi()
        assert (e.labels() == cvtamd.LABEL_NONE).all()
        assert e.moments().shape == (0, 4)
        e.lloyd(2)
        h = e.history()
        assert h.shape == (2, 2) and (h == 0).all()


# ---------------------------------------------------------------- moments --
@pytest.mark.parametrize("tile", [32, 64, 128, 256])
def test_moments_each_tile(gpu, orc, tile):
    from cvtamd import workloads as wl
    w = wl.config2(11, n=512, num_sites=1024)
    with make_engine(w.n, w.sites, w.density, tile=tile) as e:
        assert e.tile == tile
        e.voronoi()
        lab = e.labels()
        m = e.moments()
    assert np.array_equal(lab, orc.jfa(w.n, w.n, w.sites))
    check_moments(m, orc.moments(w.density, lab, w.sites))


def test_moments_independent_of_tile(gpu):
    # tiles >= 64 see the same 64-pixel runs and sum them exactly: identical bits
    from cvtamd import workloads as wl
    w = wl.config2(17, n=512, num_sites=800)
    res = []
    for tile in (64, 128, 256):
        with make_engine(w.n, w.sites, w.density, tile=tile) as e:
            e.voronoi()
            res.append(e.moments())
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[0], res[2])


def test_moments_overflow_path(gpu, orc):
    # 40 sites on a 512 grid: cells reach far beyond a 32-px tile's 3x3 window
    n = 512
    xy = rand_sites(12, n, n, 40)
    g = np.random.default_rng(13)
    rho = (0.1 + g.random((n, n))).astype(np.float32)
    with make_engine(n, xy, rho, tile=32) as e:
        e.voronoi()
        lab = e.labels()
        m = e.moments()
    check_moments(m, orc.moments(rho, lab, xy))


def test_moments_lattice_exact(gpu, orc):
    # a x a lattice at block centres: W = a^2, DX = DY = 0, E = a^2 (a^2 - 1) / 6
    a, k = 16, 8
    n = a * k
    c = (np.arange(k) * a + a / 2).astype(np.float32)
    xy = np.stack(np.meshgrid(c, c), -1).reshape(-1, 2).astype(np.float32)
    with make_engine(n, xy, np.ones((n, n), np.float32)) as e:
        e.voronoi()
        m = e.moments()
        e.lloyd(3)
        out = e.sites()
    assert np.all(m[:, 0] == a * a)
    assert np.abs(m[:, 1:3]).max() < 1e-9
    assert np.allclose(m[:, 3], a * a * (a * a - 1) / 6, rtol=1e-9)
    assert np.array_equal(out, xy)


def test_capacity_error(gpu):
    import cvtamd
    # 2000 sites inside one 64x64 tile exceed its 256-entry table
    n = 256
    g = np.random.default_rng(3)
    xy = (g.random((2000, 2)) * 64).astype(np.float32)
    with make_engine(n, xy, np.ones((n, n), np.float32), tile=64) as e:
        e.voronoi()
        with pytest.raises(cvtamd.CvtError) as err:
            e.moments()
        assert err.value.status == -5


# ------------------------------------------------------------------ lloyd --
def test_lloyd_config1_trajectory(gpu, orc):
    from cvtamd import workloads as wl
    w = wl.config1()
    with make_engine(w.n, w.sites, w.density) as e:
        e.lloyd(w.iterations, 1.0)
        h = e.history()
        out = e.sites()
    sites_r, e_r, g_r = orc.lloyd(w.density, w.sites, w.iterations)
    assert h.shape == (w.iterations, 2)
    assert np.abs(h[:, 0] - e_r).max() / e_r.max() < 1e-5
    assert np.all(np.abs(h[:, 1] - g_r) <= 1e-3 * g_r + 1e-9)
    assert np.abs(out - sites_r).max() < 1e-2
    assert h[-1, 0] < h[0, 0]


def test_lloyd_one_step_matches_update(gpu, orc):
    from cvtamd import workloads as wl
    w = wl.config2(21, n=1024, num_sites=4096)
    with make_engine(w.n, w.sites, w.density) as e:
        e.lloyd(1, 1.0)
        out = e.sites()
    lab = orc.jfa(w.n, w.n, w.sites)
    ref = orc.update(orc.moments(w.density, lab, w.sites), w.sites, 1.0)
    assert np.abs(out - ref).max() < 1e-4


@pytest.mark.parametrize("W,H,S", [(512, 512, 1024), (1024, 768, 5000), (704, 608, 20471)])
def test_jfa_shapes_bit_identical(gpu, orc, W, H, S):
    """Every JFA launch shape (runs/bands per step, workgroups per CU) gives
    the oracle's labels, and the same sites and energies over Lloyd steps."""
    xy = rand_sites(21, W, H, S)
    ref = orc.jfa(W, H, xy)
    rho = np.ones((H, W), np.float32)
    with make_engine(W, xy, rho, H=H) as e:
        auto = e.jfa_shape
        assert auto[1] == 0
        full = (1 << (max(W, H) - 1).bit_length()) - 1
        shapes = [(-1, 0), (0, 0), (full, 0), (0x555 & full, 0), (0xAAA & full, 0),
                  (-1, 1), (full, 2), (0, 8), (auto[0] ^ 1, 4)]
        base = None
        for shape in shapes:
            e.jfa_shape = shape
            e.set_sites(xy)
            e.clear_history()
            e.voronoi()
            assert np.array_equal(e.labels(), ref), shape
            e.lloyd(4)
            got = (e.sites(), e.history())
            if base is None:
                base = got
            assert np.array_equal(got[0], base[0]), shape
            assert np.array_equal(got[1], base[1]), shape
        e.jfa_shape = (-1, 0)
        assert e.jfa_shape == auto


def test_lloyd_deterministic_and_graph_equals_eager(gpu):
    from cvtamd import workloads as wl
    w = wl.config2(5, n=512, num_sites=2000)
    runs = []
    for _ in range(2):
        with make_engine(w.n, w.sites, w.density) as e:
            e.lloyd(7, 1.0)
            runs.append((e.sites(), e.history()))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])


def test_fixed_and_zero_density(gpu, orc):
    n, S = 128, 50
    xy = rand_sites(9, n, n, S)
    fixed = np.zeros(S, np.uint8)
    fixed[::7] = 1
    rho = np.ones((n, n), np.float32)
    rho[:, :40] = 0.0
    import cvtamd
    with cvtamd.Engine(n, n, S) as e:
        e.set_density(rho)
        e.set_sites(xy, fixed)
        e.lloyd(4)
        out = e.sites()
    ref, _, _ = orc.lloyd(rho, xy, 4, fixed=fixed)
    assert np.array_equal(out[fixed == 1], xy[fixed == 1])
    assert np.abs(out - ref).max() < 1e-3


def test_energy_fp64(gpu, orc):
    from cvtamd import workloads as wl
    w = wl.config2(8, n=1024, num_sites=3000)
    with make_engine(w.n, w.sites, w.density) as e:
        E, G = e.energy()
        e.lloyd(1)
        h = e.history()
    lab = orc.jfa(w.n, w.n, w.sites)
    Er, Gr = orc.energy_from_moments(w.n, w.n, orc.moments(w.density, lab, w.sites))
    assert E == pytest.approx(Er, rel=1e-12)
    assert G == pytest.approx(Gr, rel=1e-6)
    assert h[0, 0] == pytest.approx(Er, rel=1e-6)


def test_site_validation(gpu):
    import cvtamd
    n = 32
    with cvtamd.Engine(n, n, 4) as e:
        e.set_density(np.ones((n, n), np.float32))
        for bad in ([[1.0, np.nan]], [[3 * n, 1.0]], [[1.0, -n - 1.0]]):
            with pytest.raises(cvtamd.CvtError):
                e.set_sites(np.array(bad, np.float32))
            assert e.num_sites == 0
        e.set_sites(np.array([[-n + 0.5, 2 * n - 1.0]], np.float32))   # edge of the allowed box
        e.voronoi()
        assert (e.labels() == 0).all()


def test_density_validation(gpu):
    import cvtamd
    n = 32
    with cvtamd.Engine(n, n, 4) as e:
        bad = np.ones((n, n), np.float32)
        bad[3, 3] = np.nan
        with pytest.raises(cvtamd.CvtError):
            e.set_density(bad)
        bad[3, 3] = -1
        with pytest.raises(cvtamd.CvtError):
            e.set_density(bad)
        with pytest.raises(cvtamd.CvtError):
            e.lloyd(1)


def test_device_pointers(gpu, orc):
    import torch
    from cvtamd import workloads as wl
    w = wl.config2(31, n=256, num_sites=256)
    rho = torch.from_numpy(w.density).cuda()
    xy = torch.from_numpy(w.sites).cuda()
    import cvtamd
    with cvtamd.Engine(w.n, w.n, w.num_sites) as e:
        e.set_density(rho)
        e.set_sites(xy)
        e.voronoi()
        lab = torch.empty((w.n, w.n), dtype=torch.int32, device="cuda")
        e.labels(lab)
        out = torch.empty_like(xy)
        e.sites(out)
    assert np.array_equal(lab.cpu().numpy(), orc.jfa(w.n, w.n, w.sites))
    assert torch.equal(out, xy)


# ------------------------------------------------ full-size property checks --
@pytest.mark.parametrize("cfg", ["config2", "config3"])
def test_full_size_labels_and_mass(gpu, orc, cfg):
    from cvtamd import workloads as wl
    w = getattr(wl, cfg)()
    with make_engine(w.n, w.sites, w.density) as e:
        e.voronoi()
        lab = e.labels()
        m = e.moments()
    assert np.array_equal(lab, orc.jfa(w.n, w.n, w.sites))
    # mass conservation and centroid inside the grid
    assert m[:, 0].sum() == pytest.approx(float(w.density.astype(np.float64).sum()), rel=1e-9)
    ref = orc.moments(w.density, lab, w.sites)
    check_moments(m, ref)


# ------------------------------------------------- iteration launch shapes --
@pytest.mark.parametrize("counts", [(1, 7, 8, 9, 17), (16, 3, 8)])
def test_graph_chunks_equal_single_iterations(gpu, counts):
    """cvt_lloyd replays graphs of eight iterations and of one; any split of
    the same iteration count gives the same sites and history bit for bit,
    and so do the event-timed iterations of cvt_profile"""
    n, S = 320, 300
    xy = rand_sites(31, n, n, S)
    rho = (0.5 + np.random.default_rng(32).random((n, n))).astype(np.float32)
    total = sum(counts)
    with make_engine(n, xy, rho) as a:
        for c in counts:
            a.lloyd(c)
        ha, sa = a.history(), a.sites()
    with make_engine(n, xy, rho) as b:
        for _ in range(total):
            b.lloyd(1)
        hb, sb = b.history(), b.sites()
    with make_engine(n, xy, rho) as p:
        p.lloyd(total - 5)
        p.profile(5)
        hp, sp = p.history(), p.sites()
    assert len(ha) == len(hb) == len(hp) == total
    assert np.array_equal(ha, hb) and np.array_equal(sa, sb)
    assert np.array_equal(ha, hp) and np.array_equal(sa, sp)
