"""numpy restatement of the warm-start U-Net (test infrastructure only).

Mirrors surface-remesher_amd/csrc/unet.hip and nn_ops.hip: NHWC tensors,
3x3 convolutions with padding 1 and weights [9][Cout][Cin] (tap = 3 ky + kx),
ReLU after every conv, 2x2 average pooling, nearest 2x upsampling
concatenated with the skip tensor, a 1x1 head to 2 channels, bilinear
sampling at site positions and the mean squared displacement error.
Computed in float64; `bf16=True` rounds weights and every stored activation
and activation gradient to bfloat16 (round to nearest even) at the points
where the GPU stores them, so the two can be compared closely.

There is no reference counterpart (the reference has no learned component);
the backward pass is pinned by central finite differences in the tests.
"""
from __future__ import annotations

import numpy as np

IN_C = 32
HEAD_C = 32
SPEC = [(IN_C, 32, 0), (32, 32, 0), (32, 64, 1), (64, 64, 1), (64, 96, 2), (96, 96, 2),
        (96, 128, 3), (128, 128, 3), (224, 96, 2), (96, 96, 2), (160, 64, 1), (64, 64, 1),
        (96, 32, 0), (32, 32, 0)]


def bf16(x):
    """Round float array to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def layout(nparam_check=None):
    """Offsets of each conv's weights / bias and the head in the flat vector."""
    off, wo, bo = 0, [], []
    for cin, cout, _ in SPEC:
        wo.append(off); off += 9 * cin * cout
        bo.append(off); off += cout
    who = off; off += 2 * HEAD_C
    bho = off; off += 2
    if nparam_check is not None:
        assert off == nparam_check, (off, nparam_check)
    return wo, bo, who, bho, off


def unpack(p):
    wo, bo, who, bho, n = layout(len(p))
    ws = [p[wo[l]:wo[l] + 9 * c * d].reshape(9, d, c) for l, (c, d, _) in enumerate(SPEC)]
    bs = [p[bo[l]:bo[l] + d] for l, (c, d, _) in enumerate(SPEC)]
    return ws, bs, p[who:who + 2 * HEAD_C].reshape(2, HEAD_C), p[bho:bho + 2]


def shift(x, tap):
    """x[n, y + ky - 1, x + kx - 1, :] with zero padding."""
    ky, kx = tap // 3 - 1, tap % 3 - 1
    N, H, W, C = x.shape
    out = np.zeros_like(x)
    ys, ye = max(0, -ky), min(H, H - ky)
    xs, xe = max(0, -kx), min(W, W - kx)
    out[:, ys:ye, xs:xe] = x[:, ys + ky:ye + ky, xs + kx:xe + kx]
    return out


def conv(x, w, b):
    return sum(shift(x, t) @ w[t].T for t in range(9)) + b


def conv_bwd(x, w, g):
    """(dx, dw, db) of conv(x, w, b) for output gradient g."""
    dw = np.stack([np.einsum("nhwc,nhwd->dc", shift(x, t), g) for t in range(9)])
    db = g.sum(axis=(0, 1, 2))
    dx = np.zeros_like(x)
    for t in range(9):
        ky, kx = t // 3 - 1, t % 3 - 1
        contrib = g @ w[t]                       # gradient w.r.t. shift(x, t)
        # shift(x, t)[y] = x[y + k]: scatter back with the opposite shift
        dx += shift(contrib, 8 - t)
    return dx, dw, db


def pool(x):
    N, H, W, C = x.shape
    return 0.25 * (x[:, 0::2, 0::2] + x[:, 1::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 1::2])


def pool_bwd(g):
    return 0.25 * np.repeat(np.repeat(g, 2, axis=1), 2, axis=2)


def upcat(low, skip):
    up = np.repeat(np.repeat(low, 2, axis=1), 2, axis=2)
    return np.concatenate([up, skip], axis=3)


def upcat_bwd(g, c1):
    gl = g[..., :c1]
    N, H, W, _ = gl.shape
    low = gl.reshape(N, H // 2, 2, W // 2, 2, c1).sum(axis=(2, 4))
    return low, g[..., c1:]


def bilinear(field, uv):
    """field (R, R, 2), uv (S, 2) in [0,1): values and (indices, weights)."""
    R = field.shape[0]

    def coords(u):
        p = np.clip(u * R - 0.5, 0, R - 1)
        i0 = np.floor(p).astype(int)
        i1 = np.minimum(i0 + 1, R - 1)
        return i0, i1, p - i0

    x0, x1, tx = coords(uv[:, 0].astype(np.float64))
    y0, y1, ty = coords(uv[:, 1].astype(np.float64))
    idx = [(y0, x0, (1 - tx) * (1 - ty)), (y0, x1, tx * (1 - ty)), (y1, x0, (1 - tx) * ty),
           (y1, x1, tx * ty)]
    val = sum(w[:, None] * field[yy, xx] for yy, xx, w in idx)
    return val, idx


def forward(p, x, rb=False):
    """Returns (field (N, R, R, 2), cache)."""
    r = bf16 if rb else (lambda a: a)
    ws, bs, wh, bh = unpack(p)
    if rb:
        ws = [bf16(w) for w in ws]
    act = [None] * 14

    def c(l, inp):
        act[l] = r(np.maximum(conv(inp, ws[l], bs[l]), 0.0))
        return act[l]

    x = r(x)
    c(0, x); c(1, act[0]); p1 = r(pool(act[1]))
    c(2, p1); c(3, act[2]); p2 = r(pool(act[3]))
    c(4, p2); c(5, act[4]); p3 = r(pool(act[5]))
    c(6, p3); c(7, act[6])
    u3 = upcat(act[7], act[5]); c(8, u3); c(9, act[8])
    u2 = upcat(act[9], act[3]); c(10, u2); c(11, act[10])
    u1 = upcat(act[11], act[1]); c(12, u1); c(13, act[12])
    field = act[13] @ wh.T + bh
    cache = dict(x=x, act=act, p=[p1, p2, p3], u=[u3, u2, u1], ws=ws, wh=wh)
    return field, cache


def loss_and_grad_field(field, uv, target):
    """Mean over all sites of |field(uv) - target|^2; uv, target (N, S, 2)."""
    N, S = uv.shape[:2]
    loss = 0.0
    gf = np.zeros_like(field)
    for n in range(N):
        val, idx = bilinear(field[n], uv[n])
        d = val - target[n]
        loss += (d * d).sum()
        g = 2.0 * d / (N * S)
        for yy, xx, w in idx:
            np.add.at(gf[n], (yy, xx), w[:, None] * g)
    return loss / (N * S), gf


def backward(p, cache, gfield, rb=False):
    """Flat parameter gradient."""
    r = bf16 if rb else (lambda a: a)
    ws, wh = cache["ws"], cache["wh"]
    act, (p1, p2, p3), (u3, u2, u1), x = cache["act"], cache["p"], cache["u"], cache["x"]
    wo, bo, who, bho, n = layout(len(p))
    grad = np.zeros(n)
    a13 = act[13]
    grad[who:who + 2 * HEAD_C] = np.einsum("nhwj,nhwc->jc", gfield, a13).ravel()
    grad[bho:bho + 2] = gfield.sum(axis=(0, 1, 2))
    ga = [None] * 14
    ga[13] = r(gfield @ wh)

    def layer(l, inp, want_dx=True):
        g = ga[l] * (act[l] > 0)
        dx, dw, db = conv_bwd(inp, ws[l], g)
        grad[wo[l]:wo[l] + dw.size] = dw.ravel()
        grad[bo[l]:bo[l] + db.size] = db
        return r(dx) if want_dx else None

    ga[12] = layer(13, act[12])
    gu1 = layer(12, u1)
    ga[11], ga[1] = upcat_bwd(gu1, 64)
    ga[11], ga[1] = r(ga[11]), r(ga[1])
    ga[10] = layer(11, act[10])
    gu2 = layer(10, u2)
    ga[9], ga[3] = upcat_bwd(gu2, 96)
    ga[9], ga[3] = r(ga[9]), r(ga[3])
    ga[8] = layer(9, act[8])
    gu3 = layer(8, u3)
    ga[7], ga[5] = upcat_bwd(gu3, 128)
    ga[7], ga[5] = r(ga[7]), r(ga[5])
    ga[6] = layer(7, act[6])
    gp3 = layer(6, p3)
    ga[5] = r(ga[5] + r(pool_bwd(gp3)))
    ga[4] = layer(5, act[4])
    gp2 = layer(4, p2)
    ga[3] = r(ga[3] + r(pool_bwd(gp2)))
    ga[2] = layer(3, act[2])
    gp1 = layer(2, p1)
    ga[1] = r(ga[1] + r(pool_bwd(gp1)))
    ga[0] = layer(1, act[0])
    layer(0, x, want_dx=False)
    return grad


def prepare(density, xy, R):
    """Network input (R, R, IN_C) from an n x n density and sites (pixel units)."""
    n = density.shape[0]
    f = n // R
    box = density.astype(np.float64).reshape(R, f, R, f).mean(axis=(1, 3)).astype(np.float32)
    mean = density.astype(np.float64).mean()
    x = np.zeros((R, R, IN_C), np.float64)
    x[..., 0] = box * np.float32(1.0 / mean if mean > 0 else 0.0)
    cnt = np.zeros((R, R))
    if len(xy):
        cx = np.clip((xy[:, 0] * R / n).astype(int), 0, R - 1)
        cy = np.clip((xy[:, 1] * R / n).astype(int), 0, R - 1)
        np.add.at(cnt, (cy, cx), R * R / len(xy))
    # site density averaged over a (2h+1)^2 window, zero outside the grid
    h = max(0, int(round(R / np.sqrt(len(xy)))) - 1) if len(xy) else 0
    pad = np.pad(cnt, h)
    win = np.zeros((R, R))
    for dy in range(2 * h + 1):
        for dx in range(2 * h + 1):
            win += pad[dy:dy + R, dx:dx + R]
    x[..., 1] = win / (2 * h + 1) ** 2
    x[..., 2] = (box > 0).astype(np.float64)
    return x
