"""GPU checks of the warm-start network against the numpy restatement
(oracle/unet_ref.py): the forward field and the full parameter gradient,
with bf16 rounding at the same points on both sides."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def setup_net(R, N, seed):
    from cvtamd.warmstart import UNet
    from oracle import unet_ref as ref
    net = UNet(R, N, seed)
    p = net.params().astype(np.float64)
    g = np.random.default_rng(seed)
    x = np.zeros((N, R, R, ref.IN_C))
    x[..., :3] = g.random((N, R, R, 3))
    x = ref.bf16(x)
    xb = (x.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
    net.set_input(np.ascontiguousarray(xb), N)
    return net, p, x


@pytest.mark.parametrize("R,N", [(32, 2), (64, 1)])
def test_forward_matches_reference(gpu, R, N):
    from oracle import unet_ref as ref
    net, p, x = setup_net(R, N, 5)
    # a non-trivial head so the field is not ~0
    wo, bo, who, bho, n = ref.layout(len(p))
    p[who:who + 2 * ref.HEAD_C] = np.random.default_rng(1).normal(0, 0.2, 2 * ref.HEAD_C)
    net.set_params(p.astype(np.float32))
    net.forward(N)
    f = net.field(N)
    fr, _ = ref.forward(p.astype(np.float32).astype(np.float64), x, rb=True)
    err = np.abs(f - fr).max() / np.abs(fr).max()
    assert err < 2e-2, err


def test_gradients_match_reference(gpu):
    from oracle import unet_ref as ref
    R, N, S = 32, 2, 40
    net, p, x = setup_net(R, N, 7)
    wo, bo, who, bho, n = ref.layout(len(p))
    p[who:who + 2 * ref.HEAD_C] = np.random.default_rng(2).normal(0, 0.2, 2 * ref.HEAD_C)
    p32 = p.astype(np.float32)
    net.set_params(p32)
    g = np.random.default_rng(3)
    uv = g.random((N, S, 2)).astype(np.float32)
    tgt = g.normal(0, 0.05, (N, S, 2)).astype(np.float32)
    net.forward(N)
    loss = net.loss_backward(N, uv, tgt, S)
    grad = net.grads().astype(np.float64)
    fr, cache = ref.forward(p32.astype(np.float64), x, rb=True)
    lr_, gf = ref.loss_and_grad_field(fr, uv.astype(np.float64), tgt.astype(np.float64))
    gr = ref.backward(p32.astype(np.float64), cache, gf, rb=True)
    assert loss == pytest.approx(lr_, rel=2e-2)
    cos = grad @ gr / (np.linalg.norm(grad) * np.linalg.norm(gr))
    assert cos > 0.995, cos
    assert np.linalg.norm(grad - gr) / np.linalg.norm(gr) < 0.1
    # every layer's block agrees in direction
    for l in range(14):
        a, b = grad[wo[l]:bo[l]], gr[wo[l]:bo[l]]
        assert a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30) > 0.98, l


def test_training_reduces_loss_and_save_load(gpu, tmp_path):
    from cvtamd.warmstart import UNet
    R, N, S = 32, 4, 64
    net = UNet(R, N, 11)
    g = np.random.default_rng(4)
    x = np.zeros((N, R, R, 32), np.float32)
    x[..., :3] = g.random((N, R, R, 3))
    xb = (x.view(np.uint32) >> 16).astype(np.uint16)
    net.set_input(np.ascontiguousarray(xb), N)
    uv = g.random((N, S, 2)).astype(np.float32)
    tgt = (0.05 * np.sin(6 * uv[..., ::-1])).astype(np.float32)
    losses = []
    for it in range(60):
        net.forward(N)
        losses.append(net.loss_backward(N, uv, tgt, S))
        net.adam(2e-3)
    assert losses[-1] < 0.5 * losses[0]
    net.save(tmp_path / "w.bin")
    other = UNet(R, N, 99)
    other.load(tmp_path / "w.bin")
    other.set_input(np.ascontiguousarray(xb), N)
    net.forward(N)
    other.forward(N)
    f1, f2 = net.field(N), other.field(N)
    # saved as bf16: the reloaded net's field is within bf16 weight rounding
    assert np.abs(f1 - f2).max() <= 0.05 * np.abs(f1).max() + 1e-4


def test_prepare_matches_reference(gpu):
    from cvtamd import workloads as wl
    from cvtamd.warmstart import UNet
    from oracle import unet_ref as ref
    w = wl.config2(3, n=512, num_sites=1024)
    net = UNet(256, 1, 0)
    net.prepare(0, w.density, w.sites)
    dst = np.zeros((256, 256, 32), np.uint16)
    net.prepare(0, w.density, w.sites, dst=dst)
    got = (dst.astype(np.uint32) << 16).view(np.float32)
    exp = ref.bf16(ref.prepare(w.density, w.sites, 256))
    assert np.allclose(got, exp, rtol=1e-2, atol=1e-6)


def test_warm_start_moves_sites_by_the_field(gpu, orc):
    """cvt_unet_warm_start equals prepare + forward + sample through the
    separate entry points, bit for bit; fixed sites and moves onto zero
    density are refused; the engine's next Voronoi map is that of the moved
    sites."""
    import cvtamd
    from cvtamd import workloads as wl
    from cvtamd.warmstart import UNet
    n, S = 512, 2000
    w = wl.config2(5, n=n, num_sites=S)
    net = UNet(256, 1, 3)
    p = net.params()
    p[-66:-2] = np.random.default_rng(8).normal(0, 0.002, 64).astype(np.float32)
    net.set_params(p)
    fixed = np.zeros(S, np.uint8)
    fixed[:50] = 1
    with cvtamd.Engine(n, n, S) as e:
        e.set_density(w.density)
        e.set_sites(w.sites, fixed)
        e.lloyd(1)                       # leaves seeds of the moved sites pending
        start = e.sites()
        net.warm_start(e)
        got = e.sites()
        net.prepare(0, w.density, start)
        net.forward(1)
        d = net.sample(0, start / np.float32(n))
        tgt = start + d * np.float32(n)
        inside = (tgt[:, 0] >= 0) & (tgt[:, 0] < n) & (tgt[:, 1] >= 0) & (tgt[:, 1] < n)
        ix = np.clip(tgt[:, 0].astype(int), 0, n - 1)
        iy = np.clip(tgt[:, 1].astype(int), 0, n - 1)
        ok = inside & (w.density[iy, ix] > 0) & (fixed == 0)
        exp = np.where(ok[:, None], tgt, start)
        assert np.array_equal(got, exp)
        assert ok.sum() > S // 2 and np.abs(d).max() * n > 0.01
        assert np.array_equal(got[:50], start[:50])
        e.voronoi()
        assert np.array_equal(e.labels(), orc.jfa(n, n, got))
