"""Pins the numpy restatement of the warm-start network (oracle/unet_ref.py):
its backward pass against central finite differences of its own forward pass
in float64, and the pieces against closed forms."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def ref():
    from oracle import unet_ref
    return unet_ref


def random_params(ref, seed, scale=1.0):
    wo, bo, who, bho, n = ref.layout()
    g = np.random.default_rng(seed)
    p = np.zeros(n)
    for l, (cin, cout, _) in enumerate(ref.SPEC):
        p[wo[l]:wo[l] + 9 * cin * cout] = g.normal(0, scale * np.sqrt(2 / (9 * cin)), 9 * cin * cout)
        p[bo[l]:bo[l] + cout] = g.normal(0, 0.05, cout)
    p[who:who + 2 * ref.HEAD_C] = g.normal(0, 0.1, 2 * ref.HEAD_C)
    p[bho:bho + 2] = g.normal(0, 0.1, 2)
    return p


def loss_of(ref, p, x, uv, tgt):
    f, _ = ref.forward(p, x)
    return ref.loss_and_grad_field(f, uv, tgt)[0]


def test_backward_matches_finite_differences(ref):
    R, N, S = 8, 2, 7
    g = np.random.default_rng(1)
    p = random_params(ref, 2)
    x = np.zeros((N, R, R, ref.IN_C))
    x[..., :3] = g.random((N, R, R, 3))
    uv = g.random((N, S, 2))
    tgt = g.normal(0, 0.1, (N, S, 2))
    f, cache = ref.forward(p, x)
    loss, gf = ref.loss_and_grad_field(f, uv, tgt)
    grad = ref.backward(p, cache, gf)
    h = 1e-6
    wo, bo, who, bho, n = ref.layout()
    # one random direction per parameter block (perturbing every block at once
    # moves enough pre-activations across the ReLU kink to spoil the
    # difference quotient)
    segs = [(wo[l], bo[l]) for l in range(14)] + \
           [(bo[l], bo[l] + ref.SPEC[l][1]) for l in range(14)] + [(who, bho), (bho, n)]
    for a, b in segs:
        v = np.zeros_like(p)
        v[a:b] = g.normal(size=b - a)
        fd = (loss_of(ref, p + h * v, x, uv, tgt) - loss_of(ref, p - h * v, x, uv, tgt)) / (2 * h)
        assert fd == pytest.approx(grad @ v, rel=1e-4, abs=1e-10)
    # a few single coordinates spread over the layers
    for i in [wo[0] + 5, wo[7] + 100, bo[13] + 3, wo[12] + 77, who + 1, bho]:
        e = np.zeros_like(p)
        e[i] = 1.0
        fd = (loss_of(ref, p + h * e, x, uv, tgt) - loss_of(ref, p - h * e, x, uv, tgt)) / (2 * h)
        assert fd == pytest.approx(grad[i], rel=1e-4, abs=1e-10)


def test_ops_closed_forms(ref):
    g = np.random.default_rng(3)
    x = g.random((1, 4, 6, 2))
    # identity kernel: only the centre tap, identity matrix
    w = np.zeros((9, 2, 2))
    w[4] = np.eye(2)
    assert np.allclose(ref.conv(x, w, np.zeros(2)), x)
    # a one-pixel shift kernel (tap 5 = x + 1)
    w = np.zeros((9, 2, 2))
    w[5] = np.eye(2)
    y = ref.conv(x, w, np.zeros(2))
    assert np.allclose(y[:, :, :-1], x[:, :, 1:]) and np.allclose(y[:, :, -1], 0)
    assert np.allclose(ref.pool(np.ones((1, 4, 4, 3))), 1.0)
    low, skip = g.random((1, 2, 3, 2)), g.random((1, 4, 6, 1))
    u = ref.upcat(low, skip)
    assert u.shape == (1, 4, 6, 3) and np.allclose(u[0, 3, 5, :2], low[0, 1, 2])
    # bilinear sampling reproduces a linear field at interior points
    R = 16
    c = (np.arange(R) + 0.5) / R
    field = np.stack(np.broadcast_arrays(2 * c[None, :] + 0 * c[:, None], -c[:, None] + 0 * c[None, :]), -1)
    uv = g.uniform(0.1, 0.9, (20, 2))
    val, _ = ref.bilinear(field, uv)
    assert np.allclose(val[:, 0], 2 * uv[:, 0]) and np.allclose(val[:, 1], -uv[:, 1])


def test_bf16_rounding(ref):
    v = np.array([1.0, 1.00390625, 1.005859375, -3.0e-3, 65504.0])
    r = ref.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    assert abs(r[3] - v[3]) / abs(v[3]) < 2 ** -8
