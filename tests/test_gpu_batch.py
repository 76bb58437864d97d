"""GPU checks of the batched engine: every patch of a batch ends exactly where a
single-patch context ends on the same inputs (sites, labels and energy
history, bit for bit), and the labels equal the oracle's JFA(+1)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def patches(seed, sides, sites):
    from cvtamd import workloads as wl
    out = []
    for k, (n, s) in enumerate(zip(sides, sites)):
        g = wl.rng(seed * 100 + k)
        rho = wl.gaussian_mixture(n, g, 4 + k % 5, 0.05, wl.disk_mask(n))
        out.append((n, rho, wl.rejection_sites(rho, s, g)))
    return out


def run_single(n, rho, xy, iters, tile):
    import cvtamd
    with cvtamd.Engine(n, n, len(xy)) as e:
        e.set_density(rho)
        e.tile = tile
        e.set_sites(xy)
        e.lloyd(iters)
        out, h = e.sites(), e.history()
        e.voronoi()
        return out, h, e.labels()


def test_batch_equals_single_patch_runs(gpu, orc):
    import cvtamd
    ps = patches(1, [256, 320, 512, 576, 768], [256, 400, 1500, 900, 2500])
    iters = 6
    with cvtamd.Batch([(n, n) for n, _, _ in ps], [len(x) for _, _, x in ps]) as b:
        for k, (n, rho, xy) in enumerate(ps):
            b.set_density(k, rho)
            b.set_sites(k, xy)
        b.voronoi()
        first = [b.labels(k) for k in range(len(ps))]
        b.lloyd(iters)
        b.synchronize()
        res = [(b.sites(k), b.history(k), b.tile(k)) for k in range(len(ps))]
        b.voronoi()
        labs = [b.labels(k) for k in range(len(ps))]
    for k, (n, rho, xy) in enumerate(ps):
        assert np.array_equal(first[k], orc.jfa(n, n, xy))
        s1, h1, l1 = run_single(n, rho, xy, iters, res[k][2])
        assert np.array_equal(res[k][0], s1), k
        assert np.array_equal(res[k][1], h1), k
        assert np.array_equal(labs[k], l1), k


def test_batch_tile_does_not_change_results(gpu):
    import cvtamd
    ps = patches(2, [512, 384], [1200, 600])
    outs = []
    for tile in (64, 128, 256):
        with cvtamd.Batch([(n, n) for n, _, _ in ps], [len(x) for _, _, x in ps]) as b:
            for k, (n, rho, xy) in enumerate(ps):
                b.set_density(k, rho)
                b.set_sites(k, xy)
                b.set_tile(k, tile)
            b.lloyd(4)
            outs.append([b.sites(k) for k in range(len(ps))])
    for k in range(len(ps)):
        assert np.array_equal(outs[0][k], outs[1][k]) and np.array_equal(outs[0][k], outs[2][k])


def test_batch_config4_shape_runs(gpu, orc):
    # a few patches of the batched workload (sides log-uniform in [256, 2048])
    import cvtamd
    from cvtamd import workloads as wl
    ws = wl.config4(4, num_patches=6)
    with cvtamd.Batch([(w.n, w.n) for w in ws], [w.num_sites for w in ws]) as b:
        for k, w in enumerate(ws):
            b.set_density(k, w.density)
            b.set_sites(k, w.sites)
        b.lloyd(3)
        b.synchronize()
        for k, w in enumerate(ws):
            h = b.history(k)
            assert h.shape == (3, 2) and np.all(np.diff(h[:, 0]) < 0)
            _, e_ref, _ = orc.lloyd(w.density, w.sites, 3)
            assert np.allclose(h[:, 0], e_ref, rtol=1e-5)


def test_batch_config4_full(gpu, orc):
    """The whole batched workload: 64 patches (sides log-uniform in
    [256, 2048]) with the tuned moment tiles.  After 3 iterations every patch
    matches a single-patch context run with the same tile bit for bit (sites,
    energy history, labels); energies match the oracle's Lloyd; the labels of
    patches up to 768^2 match the oracle's JFA(+1) of their final sites."""
    import cvtamd
    from cvtamd import autotune
    from cvtamd import workloads as wl
    ws = wl.config4(4)
    assert len(ws) == 64
    iters = 3
    sizes = [(w.n, w.n) for w in ws]
    with cvtamd.Batch(sizes, [w.num_sites for w in ws]) as b:
        for k, w in enumerate(ws):
            b.set_density(k, w.density)
            b.set_sites(k, w.sites)
        tiles = autotune.apply_batch(b, sizes, [w.num_sites for w in ws])
        b.lloyd(iters)
        b.synchronize()
        res = [(b.sites(k), b.history(k)) for k in range(len(ws))]
        b.voronoi()
        labs = [b.labels(k) for k in range(len(ws))]
    for k, w in enumerate(ws):
        s1, h1, l1 = run_single(w.n, w.density, w.sites, iters, tiles[k])
        assert np.array_equal(res[k][0], s1), k
        assert np.array_equal(res[k][1], h1), k
        assert np.array_equal(labs[k], l1), k
        _, e_ref, _ = orc.lloyd(w.density, w.sites, iters)
        assert np.allclose(res[k][1][:, 0], e_ref, rtol=1e-5), k
        if w.n <= 768:
            assert np.array_equal(labs[k], orc.jfa(w.n, w.n, res[k][0])), k
