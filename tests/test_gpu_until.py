"""cvt_lloyd_until: the stop rule tested on the device gives the iteration count
that polling the history from the host after every iteration gives, with the
same energies and sites (the reference tests its rule on the host,
source/gcvt.cu:1123-1140)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def workload(name):
    from cvtamd import workloads as wl
    if name == "config1":
        return wl.config1()
    return wl.config2(n=1024, num_sites=4096)        # the disk patch at a quarter size


def engine(w):
    import cvtamd
    e = cvtamd.Engine(w.n, w.n, w.num_sites)
    e.set_density(w.density)
    e.set_sites(w.sites)
    return e


def host_polled(w, stop, cap):
    """one iteration at a time, the history read after each; (count, history, sites)"""
    with engine(w) as e:
        for k in range(cap):
            e.lloyd(1)
            h = e.history()[:, 0]
            if stop(h):
                break
        return len(h), h, e.sites()


@pytest.mark.parametrize("name,tol", [("config1", 1e-4), ("config2_1024", 1e-3)])
def test_relative_decrease_rule(gpu, name, tol):
    w = workload(name)
    win, cap = 10, 400
    with engine(w) as e:
        n = e.lloyd_until(cap, rel_tol=tol, window=win)
        h = e.history()[:, 0]
        xy = e.sites()
    rule = lambda h: len(h) > win and (h[-1 - win] - h[-1]) < tol * h[-1]
    n2, h2, xy2 = host_polled(w, rule, cap)
    assert 0 < n < cap, n
    assert n == n2 == len(h)
    assert np.array_equal(h, h2)
    assert np.array_equal(xy, xy2)


def test_target_rule_and_repeat(gpu):
    """a target energy; then a second call on the same context continues from
    where the first stopped, and plain cvt_lloyd runs normally afterwards"""
    w = workload("config1")
    with engine(w) as e:
        e.lloyd(30)
        target = float(e.history()[-1, 0]) * (1.0 + 1e-3)
    with engine(w) as e:
        n = e.lloyd_until(100, target=target)
        h = e.history()[:, 0]
        assert h[-1] <= target and (h[:-1] > target).all()
        assert n == len(h) <= 30
        n2 = e.lloyd_until(50, rel_tol=1e-3, window=5)
        assert 5 < n2 < 50
        h2 = e.history()[:, 0]
        assert len(h2) == n + n2
        assert (h2[-6] - h2[-1]) < 1e-3 * h2[-1]
        e.lloyd(3)
        assert len(e.history()) == n + n2 + 3
        assert e.lloyd_until(0, rel_tol=1.0, window=1) == 0


def test_cap_without_stop(gpu):
    w = workload("config1")
    with engine(w) as e:
        assert e.lloyd_until(25) == 25          # no window, no target
        assert len(e.history()) == 25


def test_new_sites_after_a_stop(gpu, orc):
    """a stop raised while the next iteration's seed scatter runs may leave
    part of that scatter behind; new sites set afterwards still get the
    oracle's exact label map, and a Lloyd run from them the same history as
    on a fresh context"""
    import cvtamd
    w = workload("config2_1024")
    g = np.random.default_rng(7)
    other = (g.random((w.num_sites, 2)) * w.n).astype(np.float32)
    with engine(w) as e:
        n = e.lloyd_until(200, rel_tol=1e-3, window=5)
        assert 5 < n < 200
        e.set_sites(other)
        e.voronoi()
        assert np.array_equal(e.labels(), orc.jfa(w.n, w.n, other))
        e.set_sites(other)
        e.lloyd(5)
        h = e.history()[-5:, 0]
    with cvtamd.Engine(w.n, w.n, w.num_sites) as f:
        f.set_density(w.density)
        f.set_sites(other)
        f.lloyd(5)
        assert np.array_equal(h, f.history()[:, 0])
