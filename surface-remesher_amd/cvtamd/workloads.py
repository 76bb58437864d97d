"""Seeded synthetic CVT workloads (densities, domain masks, initial sites).

Every generator takes an integer seed and uses numpy's Philox bit generator,
so the same seed gives the same arrays on any host.  Coordinates: the domain
is the unit square [0, 1)^2 sampled at pixel centres ((i + 0.5) / N,
(j + 0.5) / N); sites are returned in pixel units (x = u * N), fp32, which is
what the engine takes.  Densities are fp32 [N, N] (row j = y).

The reference has no such generators (its densities come from a CGAL
parameterisation of a mesh, source/main.cpp:175-204); these stand in for
that stage with the shapes of the reference's use: a disk-cut patch, a
star-shaped patch, many small patches.

Interpretations fixed here:
  * "unit disk" is the disk inscribed in the unit square: centre (0.5, 0.5),
    radius 0.5;
  * Gaussian-mixture ranges are amplitude U[0.5, 4], sigma_x, sigma_y
    U[0.02, 0.15], rotation U[0, pi), centres uniform in the disk;
  * batched patches draw 4-12 such components and
    S = clamp(round(20000 * integral(rho) * (N / 2048)^2), 256, 16384),
    about one site per 256 pixels;
  * the learned warm start trains on the disk-patch generator at 512^2
    with 1024 sites and evaluates at 2048^2 with 16384 sites on seeds from a
    disjoint range.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(int(seed) & 0xFFFFFFFFFFFFFFFF))


@dataclass
class Workload:
    name: str
    n: int
    density: np.ndarray          # float32 [n, n]
    sites: np.ndarray            # float32 [S, 2], pixel units
    iterations: int
    meta: dict = field(default_factory=dict)

    @property
    def num_sites(self) -> int:
        return int(self.sites.shape[0])


def pixel_centres(n: int):
    c = (np.arange(n, dtype=np.float64) + 0.5) / n
    return c[None, :], c[:, None]          # x [1, n], y [n, 1]


def disk_mask(n: int) -> np.ndarray:
    x, y = pixel_centres(n)
    return ((x - 0.5) ** 2 + (y - 0.5) ** 2) <= 0.25


def gaussian_mixture(n: int, g: np.random.Generator, k: int, base: float = 0.05,
                     mask: np.ndarray | None = None) -> np.ndarray:
    """base + sum_k A exp(-1/2 d^T R diag(1/sx^2, 1/sy^2) R^T d), 0 outside mask."""
    x, y = pixel_centres(n)
    rho = np.full((n, n), base, np.float64)
    for _ in range(k):
        r = 0.5 * np.sqrt(g.random())
        t = 2 * np.pi * g.random()
        cx, cy = 0.5 + r * np.cos(t), 0.5 + r * np.sin(t)
        amp = g.uniform(0.5, 4.0)
        sx, sy = g.uniform(0.02, 0.15), g.uniform(0.02, 0.15)
        ang = g.uniform(0.0, np.pi)
        c, s = np.cos(ang), np.sin(ang)
        dx, dy = x - cx, y - cy
        u = c * dx + s * dy            # R^T d
        v = -s * dx + c * dy
        rho += amp * np.exp(-0.5 * (u * u / (sx * sx) + v * v / (sy * sy)))
    if mask is not None:
        rho[~mask] = 0.0
    return rho.astype(np.float32)


def rejection_sites(density: np.ndarray, num: int, g: np.random.Generator) -> np.ndarray:
    """num continuous sites with probability proportional to density (piecewise
    constant per pixel), in pixel units."""
    n_y, n_x = density.shape
    dmax = float(density.max())
    if dmax <= 0:
        raise ValueError("density is zero everywhere")
    out = np.empty((0, 2), np.float64)
    while out.shape[0] < num:
        m = max(4 * (num - out.shape[0]), 1024)
        u = g.random((m, 2))
        px = np.minimum((u[:, 0] * n_x).astype(np.int64), n_x - 1)
        py = np.minimum((u[:, 1] * n_y).astype(np.int64), n_y - 1)
        z = g.random(m) * dmax
        ok = density[py, px] > z
        pts = np.stack([u[ok, 0] * n_x, u[ok, 1] * n_y], axis=1)
        out = np.concatenate([out, pts])
    return out[:num].astype(np.float32)


def config1(seed: int = 1) -> Workload:
    """512^2, uniform density 1.0 on the unit square, 1024 i.i.d. uniform sites, 50 iterations."""
    n, S = 512, 1024
    g = rng(seed)
    xy = (g.random((S, 2)) * n).astype(np.float32)
    return Workload("c1-uniform-512", n, np.ones((n, n), np.float32), xy, 50, {"seed": seed})


def config2(seed: int = 2, n: int = 2048, num_sites: int = 16384) -> Workload:
    """2048^2 disk-cut patch, 0.05 + 8 anisotropic Gaussians, 16384 sites by rejection, 100 iterations."""
    g = rng(seed)
    rho = gaussian_mixture(n, g, 8, 0.05, disk_mask(n))
    xy = rejection_sites(rho, num_sites, g)
    return Workload(f"c2-disk-gauss8-{n}", n, rho, xy, 100, {"seed": seed})


def star_mask(n: int, g: np.random.Generator, k: int = 64) -> np.ndarray:
    """Star-shaped k-gon about (0.5, 0.5) with vertex radii U[0.6, 1.0] * 0.5."""
    ang = 2 * np.pi * np.arange(k) / k
    rad = g.uniform(0.6, 1.0, k) * 0.5
    vx, vy = 0.5 + rad * np.cos(ang), 0.5 + rad * np.sin(ang)
    x, y = pixel_centres(n)
    theta = np.mod(np.arctan2(y - 0.5, x - 0.5), 2 * np.pi)
    sec = np.minimum((theta / (2 * np.pi / k)).astype(np.int64), k - 1)
    a, b = sec, (sec + 1) % k
    ex, ey = vx[b] - vx[a], vy[b] - vy[a]
    cross = ex * (y - vy[a]) - ey * (x - vx[a])
    return cross >= 0


def config3(seed: int = 3, n: int = 4096, num_sites: int = 65536) -> Workload:
    """4096^2 star 64-gon, 0.1 + |sin 7x cos 5y|^2 + 3 sum_4 exp(-r^2 / 0.01), 65536 sites, 200 iterations."""
    g = rng(seed)
    mask = star_mask(n, g)
    x, y = pixel_centres(n)
    rho = 0.1 + np.abs(np.sin(7 * x) * np.cos(5 * y)) ** 2
    # four bump centres inside the domain
    ys, xs = np.nonzero(mask)
    pick = g.integers(0, xs.shape[0], 4)
    for p in pick:
        cx, cy = (xs[p] + 0.5) / n, (ys[p] + 0.5) / n
        rho = rho + 3.0 * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / 0.01)
    rho = np.where(mask, rho, 0.0).astype(np.float32)
    xy = rejection_sites(rho, num_sites, g)
    return Workload(f"c3-star64-{n}", n, rho, xy, 200, {"seed": seed})


def batch_patch(seed: int) -> Workload:
    """One patch of the batched workload: side log-uniform in [256, 2048] rounded to a
    multiple of 64, 4-12 Gaussians, disk mask, S proportional to integrated density."""
    g = rng(seed)
    side = float(np.exp(g.uniform(np.log(256.0), np.log(2048.0))))
    n = int(min(2048, max(256, 64 * round(side / 64))))
    k = int(g.integers(4, 13))
    rho = gaussian_mixture(n, g, k, 0.05, disk_mask(n))
    integral = float(rho.astype(np.float64).mean())
    S = int(np.clip(round(20000 * integral * (n / 2048) ** 2), 256, 16384))
    xy = rejection_sites(rho, S, g)
    return Workload(f"c4-patch-{n}", n, rho, xy, 100, {"seed": seed, "components": k})


def config4(seed: int = 4, num_patches: int = 64) -> list[Workload]:
    """64 independent patches (batched workload), 100 iterations each."""
    base = rng(seed).integers(0, 2**62, num_patches)
    return [batch_patch(int(s)) for s in base]


def warmstart_train(seed: int, n: int = 512, num_sites: int = 1024) -> Workload:
    """Training sample for the learned warm start: disk patch, 4-12 Gaussians, 512^2, 1024 sites."""
    g = rng(seed)
    k = int(g.integers(4, 13))
    rho = gaussian_mixture(n, g, k, 0.05, disk_mask(n))
    xy = rejection_sites(rho, num_sites, g)
    return Workload(f"c5-train-{n}", n, rho, xy, 200, {"seed": seed, "components": k})


# evaluation seeds for the warm start come from a range the training never uses
WARMSTART_EVAL_SEED_BASE = 1 << 40


def warmstart_eval(index: int, n: int = 2048, num_sites: int = 16384) -> Workload:
    """Held-out evaluation patch at 2048^2 / 16384 sites (config-2 generator)."""
    return config2(WARMSTART_EVAL_SEED_BASE + index, n, num_sites)


def patch_mesh(seed: int = 5, m: int = 78):
    """A parameterized patch as recover() sees it: an m x m jittered grid of
    [0, 1]^2 split into 2 (m-1)^2 triangles (faces shuffled), the 3D surface
    z = 0.3 sin(3u) cos(2v) over it, and the query points of one recover()
    call: 16384 sites sampled in the patch and the centroids of 32768
    triangles over them.  Returns (uv, p3, tri, queries)."""
    g = rng(seed)
    h = 1.0 / (m - 1)
    x, y = np.meshgrid(np.arange(m) * h, np.arange(m) * h)
    uv = np.stack([x.ravel(), y.ravel()], 1)
    inner = (uv[:, 0] > 0) & (uv[:, 0] < 1) & (uv[:, 1] > 0) & (uv[:, 1] < 1)
    uv[inner] += g.uniform(-0.3 * h, 0.3 * h, (int(inner.sum()), 2))
    i, j = np.meshgrid(np.arange(m - 1), np.arange(m - 1))
    a = (j * m + i).ravel()
    tri = np.concatenate([np.stack([a, a + 1, a + m + 1], 1), np.stack([a, a + m + 1, a + m], 1)])
    tri = tri[g.permutation(len(tri))].astype(np.int32)
    p3 = np.stack([uv[:, 0], uv[:, 1], 0.3 * np.sin(3 * uv[:, 0]) * np.cos(2 * uv[:, 1])], 1)
    sites = g.random((16384, 2))
    t = g.integers(0, len(sites), (32768, 3))
    cen = (sites[t[:, 0]] + sites[t[:, 1]] + sites[t[:, 2]]) / 3.0
    return uv, p3, tri, np.concatenate([sites, cen])
