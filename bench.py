#!/usr/bin/env python3
"""Lloyd iterations/s of the CVT engine on one MI355X.

Workload: 2048^2 disk-cut patch, 0.05 + 8 anisotropic Gaussians, 16384 sites
rejection-sampled from the density (the 2048^2 / 16K configuration the
metric is quoted on).  One step = one Lloyd iteration = JFA(+1) (12 passes)
+ moments and energy + site update, replayed from one captured hipGraph.
Inputs are resident in HBM before the timed region.

The JSON line carries:
  roofline      the dominant kernel (the dense JFA pass k_jfa_quad, passes
                1 .. P-1; pass 0 is a scatter of the seeds): algorithmic
                bytes per launch (read + write of the int32 label map,
                8 B/pixel) over its mean duration measured with HIP events
                on the engine's stream, against the 8 TB/s HBM peak;
  cpu_baseline  the oracle's OpenMP Lloyd (JFA(+1) and exact nearest-site
                variants; the faster is reported) on a bounded sample of the
                same workload, on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "surface-remesher_amd"))
sys.path.insert(0, str(ROOT))

PMC_SUMMARY = ROOT / "profiles" / "k_jfa_quad_c2_pmc.txt"


def pmc_traffic(workload: str):
    """HBM bytes per JFA launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE, calibrated on a copy of known size), for the
    workload it was measured on; None otherwise."""
    try:
        for line in PMC_SUMMARY.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch") and workload.startswith("c2"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


METRIC = "Lloyd iter/s (2048², 16K sites); time-to-tol: learned init vs plain Lloyd"
HBM_PEAK_GBS = 8000.0


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, seconds: float):
    """Times the oracle's Lloyd on the host (the 'port' baseline): 3 warm-up
    and up to 20 timed iterations per variant (JFA(+1) and exact nearest
    site), each timed alone; the median per-iteration time of the faster
    variant gives the rate."""
    import numpy as np
    from oracle import oracle as orc
    threads = min(orc.max_threads(), int(os.environ.get("OMP_NUM_THREADS", "64")))
    orc.set_threads(threads)
    res = {}
    for name, exact in (("jfa", False), ("exact", True)):
        xy = w.sites.copy()
        for _ in range(3):
            orc.lloyd(w.density, xy, 1, exact=exact)
        times = []
        t_start = time.perf_counter()
        while len(times) < 20 and time.perf_counter() - t_start < seconds / 2:
            t0 = time.perf_counter()
            orc.lloyd(w.density, xy, 1, exact=exact)
            times.append(time.perf_counter() - t0)
        res[name] = (1.0 / float(np.median(times)), len(times))
    best = max(res, key=lambda n: res[n][0])
    return {
        "value": round(res[best][0], 4),
        "unit": "iter/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{w.name}: oracle Lloyd (gcc -O3 -fopenmp -march=x86-64-v3) on the host, "
                   f"3 warm-up + {res['jfa'][1]} timed iterations of the JFA(+1) variant "
                   f"({res['jfa'][0]:.3f} it/s) and 3 + {res['exact'][1]} of the exact "
                   f"nearest-site variant ({res['exact'][0]:.3f} it/s), median per iteration, "
                   f"OpenMP over {threads} threads of {os.cpu_count()} ({_cpu_model()}); "
                   f"value = the faster ({best})"),
    }


def copy_probe(args):
    """A plain device-to-device copy of known size: its HBM counters per
    dispatch give the factor between counter units and bytes."""
    import torch
    nbytes = args.copy_probe << 20
    src = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda").random_()
    dst = torch.empty_like(src)
    for _ in range(max(args.warmup, 1)):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(args.steps):
        dst.copy_(src)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / args.steps
    print(json.dumps({"probe": "copy", "bytes_read": nbytes, "bytes_written": nbytes,
                      "ms_per_copy": round(ms, 6),
                      "GBps": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}), flush=True)


def time_to_tol_sample(k: int) -> dict:
    """Metric 2 on k held-out patches: plain Lloyd and the learned warm start,
    each timed from the same initial sites to E* (1 + 1e-4)."""
    import torch
    import cvtamd
    from cvtamd import eval_warmstart as ev
    from cvtamd import workloads as wl
    from cvtamd.warmstart import WEIGHTS, UNet
    if not WEIGHTS.exists():
        return {"skipped": f"no trained weights at {WEIGHTS.name}"}
    net = UNet(256, 1, 0)
    net.load(WEIGHTS)
    eng = cvtamd.Engine(2048, 2048, 16384)
    stream = torch.cuda.Stream()
    eng.set_stream(stream.cuda_stream)
    net.set_stream(stream.cuda_stream)
    rows = []
    for idx in range(k):
        w = wl.warmstart_eval(idx)
        rows.append(ev.time_to_tol(eng, net, torch.from_numpy(w.density).cuda(),
                                   torch.from_numpy(w.sites).cuda()))
    eng.close()
    net.close()
    out = ev.summarize(rows)
    out["sample"] = f"held-out patches 0..{k - 1} (2048^2, 16384 sites), tolerance 1e-4"
    return out


# U-Net layers (Cin, Cout, level) as built by csrc/unet.hip
UNET_SPEC = [(32, 32, 0), (32, 32, 0), (32, 64, 1), (64, 64, 1), (64, 96, 2), (96, 96, 2),
             (96, 128, 3), (128, 128, 3), (224, 96, 2), (96, 96, 2), (160, 64, 1), (64, 64, 1),
             (96, 32, 0), (32, 32, 0)]
BF16_PEAK_TFLOPS = 2500.0


def unet_mfma(batch: int = 16, steps: int = 10) -> dict:
    """The warm-start network's convolutions on MFMA: forward and training
    step throughput at 256^2 against the dense bf16 peak.  FLOPs count the
    3x3 convolutions as executed (input channels padded to 32): forward
    2*H*W*9*Cin*Cout per layer; a training step adds the data gradient
    (all layers but the first) and the weight gradient."""
    import numpy as np
    import torch
    from cvtamd.warmstart import UNet
    R = 256
    fwd = sum(2.0 * (R >> l) ** 2 * 9 * ci * co for ci, co, l in UNET_SPEC)
    dgrad = fwd - 2.0 * R * R * 9 * 32 * 32
    train = fwd + dgrad + fwd
    net = UNet(R, batch, 1)
    stream = torch.cuda.Stream()
    net.set_stream(stream.cuda_stream)
    g = np.random.default_rng(0)
    x = np.zeros((batch, R, R, 32), np.float32)
    x[..., :3] = g.random((batch, R, R, 3))
    net.set_input(np.ascontiguousarray((x.view(np.uint32) >> 16).astype(np.uint16)), batch)
    S = 1024
    uv = torch.from_numpy(g.random((batch, S, 2)).astype(np.float32)).cuda()
    tg = torch.zeros_like(uv)
    net.forward(batch)
    net.loss_backward(batch, uv, tg, S)
    net.synchronize()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record(stream)
    for _ in range(steps):
        net.forward(batch)
    e1.record(stream)
    for _ in range(steps):
        net.forward(batch)
        net.loss_backward(batch, uv, tg, S, sync=False)
        net.adam(1e-4)
    e2.record(stream)
    torch.cuda.synchronize()
    f_ms = e0.elapsed_time(e1) / steps
    t_ms = e1.elapsed_time(e2) / steps
    net.close()
    f_tf = fwd * batch / (f_ms * 1e-3) / 1e12
    t_tf = train * batch / (t_ms * 1e-3) / 1e12
    return {"batch": batch, "grid": R, "forward_gflop_per_sample": round(fwd / 1e9, 3),
            "forward_ms": round(f_ms, 4), "forward_tflops": round(f_tf, 2),
            "train_step_ms": round(t_ms, 4), "train_tflops": round(t_tf, 2),
            "samples_per_s_train": round(batch / (t_ms * 1e-3), 1),
            "peak_tflops": BF16_PEAK_TFLOPS, "forward_frac": round(f_tf / BF16_PEAK_TFLOPS, 5),
            "train_frac": round(t_tf / BF16_PEAK_TFLOPS, 5)}


def raster_sample(with_cpu: bool) -> dict:
    """The density rasterizer (discretization_d rule) on the patch mesh of
    locate_sample, scaled to the grid: 11858 triangles at 2048^2 on the GPU
    (host buffers in and out, as the reference's caller passes them, and in
    gpu_ms_device_io with the mesh and the density in device memory); the
    oracle's triangle-by-triangle scan on the host cores at 256^2 (a bounded
    sample of the same mesh), compared per pixel, with a bit-exactness check
    of the GPU at that size."""
    import numpy as np
    from cvtamd import compat, workloads as wl
    uv, p3, tri, _ = wl.patch_mesh()
    wt = 1.0 + p3[:, 2]                          # a smooth positive vertex weight
    n = 2048
    scale = 1.0 / (n - 1)                        # pixel corners span the unit patch
    compat.rasterize(uv, wt, tri, scale, n)     # warm-up
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        dens = compat.rasterize(uv, wt, tri, scale, n)
    gpu_ms = (time.perf_counter() - t0) / reps * 1e3
    # the same calls with the mesh and the density in device memory
    import torch
    d_uv, d_wt, d_tri = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (uv, wt, tri))
    d_den = torch.empty((n, n), dtype=torch.float32, device="cuda")
    compat.rasterize(d_uv, d_wt, d_tri, scale, n, out=d_den)
    t0 = time.perf_counter()
    for _ in range(reps):
        compat.rasterize(d_uv, d_wt, d_tri, scale, n, out=d_den)
    dev_ms = (time.perf_counter() - t0) / reps * 1e3
    res = {"grid": n, "triangles": int(len(tri)), "gpu_ms": round(gpu_ms, 3),
           "gpu_pixels_per_s": round(n * n / gpu_ms * 1e3, 1),
           "gpu_ms_device_io": round(dev_ms, 3),
           "device_io_equal": bool(np.array_equal(d_den.cpu().numpy(), dens)),
           "covered": round(float((dens > 0).mean()), 4)}
    if with_cpu:
        from oracle import oracle as orc
        orc.set_threads(min(orc.max_threads(), int(os.environ.get("OMP_NUM_THREADS", "64"))))
        ns = 256
        ss = 1.0 / (ns - 1)
        t0 = time.perf_counter()
        ref = orc.raster(uv, wt, tri, ss, ns)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        got = compat.rasterize(uv, wt, tri, ss, ns)
        cpu_pps = ns * ns / cpu_ms * 1e3
        res.update({"cpu_sample": f"{ns}^2", "cpu_ms": round(cpu_ms, 1),
                    "cpu_threads": orc.max_threads(), "cpu_pixels_per_s": round(cpu_pps, 1),
                    "bit_exact_at_sample": bool(np.array_equal(got, ref)),
                    "speedup_per_pixel": round(res["gpu_pixels_per_s"] / cpu_pps, 1)})
    return res


def delaunay_sample(n: int, with_cpu: bool, sites=None) -> dict:
    """The GPU Delaunay triangulation (cvt_dt_compute, gDel2D's GpuDel::compute
    without constraints) of n random points (float32 values in [0, 4096)^2,
    held as float64 in device memory; the call includes its host round trips
    and returns when the triangles and links are written to host arrays),
    best of 3; and, when given, of the CVT sites of this run.  Baseline:
    scipy.spatial.Delaunay (Qhull, one host thread) on the same points, whose
    triangle set must be the same."""
    import numpy as np
    import torch
    from cvtamd.delaunay import Delaunay
    g = np.random.default_rng(n)
    pts = (g.random((n, 2)) * 4096.0).astype(np.float32).astype(np.float64)
    dp = torch.from_numpy(pts).cuda()
    with Delaunay(n) as d:
        d.compute(dp)
        ts = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            tri, _ = d.compute(dp)
            ts.append(time.perf_counter() - t0)
        st = d.stats()
    ms = min(ts) * 1e3
    res = {"points": n, "triangles": int(len(tri)), "gpu_ms": round(ms, 3),
           "mpoints_per_s": round(n / ms / 1e3, 2), **st}
    # the same points inside a square boundary of 4 x 1024 collinear constraint
    # points joined by segments (delaunayInput's constraint points and edges)
    side = np.arange(1024) * 4.0
    ring = np.concatenate([np.stack([side, 0 * side - 1.0], 1),
                           np.stack([0 * side + 4097.0, side], 1),
                           np.stack([4096.0 - side, 0 * side + 4097.0], 1),
                           np.stack([0 * side - 1.0, 4096.0 - side], 1)])
    cpts = np.concatenate([pts, ring])
    nr = len(ring)
    segs = np.stack([n + np.arange(nr), n + (np.arange(nr) + 1) % nr], 1).astype(np.int32)
    dc, ds = torch.from_numpy(cpts).cuda(), torch.from_numpy(segs).cuda()
    with Delaunay(len(cpts)) as d:
        d.compute(dc, segments=ds)
        ts = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            tri_c, _ = d.compute(dc, segments=ds)
            ts.append(time.perf_counter() - t0)
    res["constrained"] = {"points": int(len(cpts)), "segments": int(nr),
                          "triangles": int(len(tri_c)), "gpu_ms": round(min(ts) * 1e3, 3)}
    if sites is not None:
        sp = torch.from_numpy(np.ascontiguousarray(sites, np.float64)).cuda()
        with Delaunay(len(sites)) as d:
            d.compute(sp)
            t0 = time.perf_counter()
            tri_s, _ = d.compute(sp)
            res["sites"] = {"points": int(len(sites)), "triangles": int(len(tri_s)),
                            "gpu_ms": round((time.perf_counter() - t0) * 1e3, 3)}
    if with_cpu:
        from scipy.spatial import Delaunay as QDelaunay
        t0 = time.perf_counter()
        ref = QDelaunay(pts).simplices
        cpu_ms = (time.perf_counter() - t0) * 1e3
        same = set(map(tuple, np.sort(tri, 1).tolist())) == set(map(tuple, np.sort(ref, 1).tolist()))
        res.update({"scipy_ms": round(cpu_ms, 1), "scipy_threads": 1,
                    "cpu_kind": "scipy.spatial.Delaunay (Qhull)", "same_triangles": same,
                    "speedup": round(cpu_ms / ms, 1)})
    return res


def locate_sample(with_cpu: bool) -> dict:
    """recover()'s point location on one patch mesh: 49152 queries (16384
    sites + 32768 triangle centroids) in a 11858-face mesh, on the GPU
    (inputs resident in device memory; the call includes its binning and
    returns when the faces, weights and 3D points are written, to host
    arrays and, in gpu_ms_device_out, to device tensors) and, as the
    baseline, the oracle's face-by-face scan on the host cores."""
    import numpy as np
    import torch
    from cvtamd import recover, workloads as wl
    uv, p3, tri, q = wl.patch_mesh()
    d = [torch.from_numpy(a).cuda() for a in (uv, p3, tri, q)]
    recover.locate_points(d[0], d[2], d[3], p3=d[1])          # warm-up
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        face, w, out3 = recover.locate_points(d[0], d[2], d[3], p3=d[1])
    gpu_ms = (time.perf_counter() - t0) / reps * 1e3
    # the same calls with the results kept in device memory (no download)
    dev = (torch.empty(len(q), dtype=torch.int32, device="cuda"),
           torch.empty((len(q), 3), dtype=torch.float64, device="cuda"),
           torch.empty((len(q), 3), dtype=torch.float64, device="cuda"))
    recover.locate_points(d[0], d[2], d[3], p3=d[1], out=dev)
    t0 = time.perf_counter()
    for _ in range(reps):
        recover.locate_points(d[0], d[2], d[3], p3=d[1], out=dev)
    dev_ms = (time.perf_counter() - t0) / reps * 1e3
    res = {"queries": int(len(q)), "faces": int(len(tri)), "located": int((face >= 0).sum()),
           "gpu_ms": round(gpu_ms, 3), "gpu_queries_per_s": round(len(q) / gpu_ms * 1e3, 1),
           "gpu_ms_device_out": round(dev_ms, 3),
           "device_out_equal": bool(np.array_equal(dev[0].cpu().numpy(), face)
                                    and np.array_equal(dev[2].cpu().numpy(), out3))}
    if with_cpu:
        from oracle import oracle as orc
        orc.set_threads(min(orc.max_threads(), int(os.environ.get("OMP_NUM_THREADS", "64"))))
        t0 = time.perf_counter()
        f_cpu, w_cpu, o_cpu = orc.locate(uv, tri, q, p3=p3)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        res.update({"cpu_ms": round(cpu_ms, 1), "cpu_threads": orc.max_threads(),
                    "cpu_kind": "port (face-by-face scan, as source/recover.h:63-81)",
                    "bit_exact": bool(np.array_equal(f_cpu, face) and np.array_equal(w_cpu, w)
                                      and np.array_equal(o_cpu, out3)),
                    "speedup": round(cpu_ms / gpu_ms, 1)})
    return res


def weights_sample(with_cpu: bool) -> dict:
    """Per-vertex density weights (weighting_vertices_with_area rule) on the
    patch mesh of locate_sample (11858 faces) and on a 1024 x 1024 vertex
    patch (2.09M faces), inputs resident in device memory, the call
    returning when the weights are written (to a host array, and in
    gpu_ms_device_out to a device tensor); the baseline is the oracle's
    face loop, single-threaded as the reference's host loop is (its sums
    are sequential in face order), median of as many runs as the GPU's."""
    import numpy as np
    import torch
    from cvtamd import weighting, workloads as wl
    res = {}
    for name, m, reps in (("patch", 78, 20), ("large", 1024, 5)):
        uv, p3, tri, _ = wl.patch_mesh(m=m)
        d = [torch.from_numpy(a).cuda() for a in (uv, p3, tri)]
        weighting.vertex_weights(*d)                          # warm-up
        t0 = time.perf_counter()
        for _ in range(reps):
            w = weighting.vertex_weights(*d)
        gpu_ms = (time.perf_counter() - t0) / reps * 1e3
        d_w = torch.empty(len(uv), dtype=torch.float64, device="cuda")
        weighting.vertex_weights(*d, out=d_w)
        t0 = time.perf_counter()
        for _ in range(reps):
            weighting.vertex_weights(*d, out=d_w)
        dev_ms = (time.perf_counter() - t0) / reps * 1e3
        r = {"faces": int(len(tri)), "vertices": int(len(uv)), "gpu_ms": round(gpu_ms, 3),
             "gpu_faces_per_s": round(len(tri) / gpu_ms * 1e3, 1),
             "gpu_ms_device_out": round(dev_ms, 3),
             "device_out_equal": bool(np.array_equal(d_w.cpu().numpy(), w, equal_nan=True))}
        if with_cpu:
            from oracle import oracle as orc
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                ref = orc.vertex_weights(uv, p3, tri)
                ts.append(time.perf_counter() - t0)
            cpu_ms = float(np.median(ts)) * 1e3
            r.update({"cpu_ms": round(cpu_ms, 3), "cpu_threads": 1,
                      "cpu_kind": "port (face loop, as source/weighting.h:53-92)",
                      "bit_exact": bool(np.array_equal(w, ref, equal_nan=True)),
                      "speedup": round(cpu_ms / gpu_ms, 2)})
        res[name] = r
    return res


def bench_batch(args):
    """The 64-patch batch: one step = one batched Lloyd iteration over every
    patch (one graph launch); value = batched iterations/s, with the
    aggregate pixel rate beside it."""
    import numpy as np
    import torch
    import cvtamd
    from cvtamd import autotune
    from cvtamd import workloads as wl
    ws = wl.config4(args.seed)
    stream = torch.cuda.Stream()
    b = cvtamd.Batch([(w.n, w.n) for w in ws], [w.num_sites for w in ws])
    b.set_stream(stream.cuda_stream)
    for k, w in enumerate(ws):
        b.set_density(k, w.density)
        b.set_sites(k, w.sites)
    tiles = autotune.apply_batch(b, [(w.n, w.n) for w in ws], [w.num_sites for w in ws])
    b.lloyd(args.warmup, 1.0)
    b.synchronize()
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    b.lloyd(args.steps, 1.0)
    t1.record(stream)
    torch.cuda.synchronize()
    b.synchronize()
    ms = t0.elapsed_time(t1) / args.steps
    pix = sum(w.n * w.n for w in ws)
    abytes = sum(w.n * w.n * (8.0 * (int(np.log2(w.n - 1)) + 2) + 8.0) for w in ws)
    for k in range(len(ws)):
        assert np.isfinite(b.history(k)).all()
    b.close()
    print(json.dumps({
        "metric": METRIC, "value": round(1000.0 / ms, 2), "unit": "iter/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 6),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "c4: 64 patches, sides 256..2048, batched", "patches": len(ws),
                   "pixels": pix, "sites": sum(w.num_sites for w in ws), "seed": args.seed,
                   "moment_tiles": {str(t): tiles.count(t) for t in sorted(set(tiles))}},
        "pixels_per_s": round(pix / (ms * 1e-3), 1),
        "iteration_achieved_GBps": round(abytes / (ms * 1e-3) / 1e9, 1),
    }), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--profile-iters", type=int, default=10)
    ap.add_argument("--workload", choices=("c1", "c2", "c3", "c4"), default="c2",
                    help="c2 (2048^2, 16K sites) is the headline line; c1 512^2/1K, "
                         "c3 4096^2/64K star patch, c4 the 64-patch batch")
    ap.add_argument("--ttt", type=int, default=4, metavar="K",
                    help="time-to-tolerance of the learned warm start vs plain Lloyd on K "
                         "held-out 2048^2 / 16K patches (0 = skip; needs trained weights)")
    ap.add_argument("--no-locate", dest="locate", action="store_false",
                    help="skip the rasterizer, vertex-weight and recover() point-location samples")
    ap.add_argument("--unet", type=int, default=16, metavar="B",
                    help="MFMA throughput of the warm-start network at batch B (0 = skip)")
    ap.add_argument("--unet-only", action="store_true",
                    help="only the network measurement (for profiling its kernels)")
    ap.add_argument("--dt", type=int, default=1000000, metavar="N",
                    help="GPU Delaunay triangulation of N random points (0 = skip)")
    ap.add_argument("--dt-only", action="store_true",
                    help="only the triangulation measurement (for profiling its kernels)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="instead of the CVT step, time a plain device copy of MIB MiB "
                         "(calibrates the HBM byte counters against a known size)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the engine runs one grid per device and has "
              "no multi-GPU code path; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import cvtamd
    from cvtamd import autotune
    from cvtamd import workloads as wl

    torch.cuda.set_device(0)
    if args.copy_probe:
        return copy_probe(args)
    if args.unet_only:
        print(json.dumps({"mfma": unet_mfma(max(args.unet, 1))}), flush=True)
        return
    if args.dt_only:
        print(json.dumps({"delaunay": delaunay_sample(max(args.dt, 3), not args.no_cpu)}),
              flush=True)
        return
    if args.workload == "c4":
        return bench_batch(args)
    w = {"c1": wl.config1, "c2": wl.config2, "c3": wl.config3}[args.workload](args.seed)
    n, S = w.n, w.num_sites
    rho = torch.from_numpy(w.density).cuda()
    xy = torch.from_numpy(w.sites).cuda()
    stream = torch.cuda.Stream()
    eng = cvtamd.Engine(n, n, S)
    eng.set_stream(stream.cuda_stream)
    eng.set_density(rho)
    autotune.apply(eng, n, n, S)          # the tuned moment tile (built-in rule if none)
    eng.set_sites(xy)
    torch.cuda.synchronize()

    # warm-up (graph capture happens on the first launch)
    eng.lloyd(args.warmup, 1.0)
    eng.synchronize()
    torch.cuda.synchronize()

    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    eng.lloyd(args.steps, 1.0)
    end.record(stream)
    torch.cuda.synchronize()
    eng.synchronize()
    ms_total = start.elapsed_time(end)
    ms_step = ms_total / args.steps
    value = 1000.0 / ms_step

    hist = eng.history()
    assert np.isfinite(hist).all() and hist.shape[0] == args.warmup + args.steps

    # per-kernel durations (HIP events between launches, same stream)
    prof = eng.profile(args.profile_iters, 1.0)
    P = prof["jfa_passes"]
    kernel, jfa_bytes, launch_ms = "k_jfa_quad", 8.0 * n * n, prof["jfa_pass"]
    jfa_gbs = jfa_bytes / (launch_ms * 1e-3) / 1e9
    iter_bytes = n * n * (8.0 * P + 8.0)           # N^2 (8P + 8)
    iter_gbs = iter_bytes / (ms_step * 1e-3) / 1e9
    final_sites = eng.sites()
    eng.close()

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w.name, "grid": n, "sites": S, "jfa_passes": P,
                   "moment_tile": prof["tile"], "seed": args.seed},
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "achieved": round(jfa_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(jfa_gbs / HBM_PEAK_GBS, 4),
            "traffic": pmc_traffic(w.name),
            "algorithmic_bytes_per_launch": jfa_bytes,
            "launch_ms": round(launch_ms, 6),
            "iteration_achieved": round(iter_gbs, 1),
            "iteration_frac": round(iter_gbs / HBM_PEAK_GBS, 4),
            "kernel_ms": {k: round(v, 6) for k, v in prof.items() if isinstance(v, float)},
        },
        "energy": {"first": hist[0, 0], "last": hist[-1, 0]},
    }
    if args.ttt > 0 and args.workload == "c2":
        out["time_to_tol"] = time_to_tol_sample(args.ttt)
    if args.locate and args.workload == "c2":
        out["raster"] = raster_sample(not args.no_cpu)
        out["locate"] = locate_sample(not args.no_cpu)
        out["weights"] = weights_sample(not args.no_cpu)
    if args.unet > 0 and args.workload == "c2":
        out["mfma"] = unet_mfma(args.unet)
    if args.dt > 0 and args.workload == "c2":
        out["delaunay"] = delaunay_sample(args.dt, not args.no_cpu, final_sites)
    if not args.no_cpu:
        cb = cpu_baseline(w, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["vs_cpu"] = round(value / cb["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
