"""Training of the learned warm start on pairs the plain solver generates.

    python -m cvtamd.train_warmstart --samples 20000 --out weights.bin

1. Pairs: for each training seed, a 512^2 disk patch with a 4-12 component
   Gaussian mixture and 1024 sites rejection-sampled from it
   (workloads.warmstart_train); the batched engine (cvt_batch_*) runs plain
   Lloyd (omega = 1) on 64 patches at a time for --iters iterations.  The
   pair is (density, initial sites) -> converged sites.
2. Data kept on the device: the density box-averaged to 256^2, the initial
   sites, and the target displacement (converged - initial) in units of the
   grid side.
3. Training: the U-Net (cvt_unet_*) with its own forward / backward kernels,
   mean squared displacement error at the initial sites, Adam with a cosine
   learning-rate schedule; a held-out slice reports the fraction of the
   target's variance the network explains.

Densities are generated on the host by a pool of worker processes started
before this process touches the GPU.  torch is used only to hold device
buffers and to gather minibatches.
"""
from __future__ import annotations

import argparse
import json
import math
import multiprocessing as mp
import sys
import time
from pathlib import Path

import numpy as np

N_GRID = 512
N_SITES = 1024
R = 256
PATCHES = 64


def _make(seed: int):
    from . import workloads as wl
    w = wl.warmstart_train(seed, N_GRID, N_SITES)
    small = w.density.reshape(R, N_GRID // R, R, N_GRID // R).mean(axis=(1, 3), dtype=np.float64)
    return seed, w.density, w.sites, small.astype(np.float32)


def generate(pool, seeds, iters, log):
    """Plain-Lloyd pairs for `seeds`: (density 256^2, initial sites, converged sites, final
    relative energy decrease over the last 10 iterations)."""
    import cvtamd
    out_d, out_x, out_c, out_r = [], [], [], []
    it = pool.imap(_make, seeds, chunksize=4)
    t0 = time.time()
    done = 0
    while done < len(seeds):
        batch = []
        for _ in range(min(PATCHES, len(seeds) - done)):
            batch.append(next(it))
        with cvtamd.Batch([(N_GRID, N_GRID)] * len(batch), [N_SITES] * len(batch)) as b:
            for k, (_, rho, xy, _) in enumerate(batch):
                b.set_density(k, rho)
                b.set_sites(k, xy)
            b.lloyd(iters)
            b.synchronize()
            for k, (_, _, xy, small) in enumerate(batch):
                conv = b.sites(k)
                h = b.history(k)[:, 0]
                out_d.append(small)
                out_x.append(xy)
                out_c.append(conv)
                out_r.append(float((h[-11] - h[-1]) / h[-1]) if len(h) > 11 else float("nan"))
        done += len(batch)
        if done % (PATCHES * 16) == 0 or done == len(seeds):
            log(f"pairs {done}/{len(seeds)} in {time.time() - t0:.1f} s; "
                f"median final 10-iteration decrease {np.nanmedian(out_r):.2e}")
    return np.stack(out_d), np.stack(out_x), np.stack(out_c), np.array(out_r)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--samples", type=int, default=20000)
    ap.add_argument("--val", type=int, default=512, help="held-out pairs (from the sample count)")
    ap.add_argument("--iters", type=int, default=400, help="plain Lloyd iterations per pair")
    ap.add_argument("--epochs", type=float, default=6.0)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--workers", type=int, default=14)
    ap.add_argument("--time-limit", type=float, default=1e9, help="stop training after this many s")
    ap.add_argument("--out", type=Path, required=True)
    ap.add_argument("--log", type=Path, default=None)
    args = ap.parse_args(argv)

    records = []

    def log(msg, **kw):
        print(msg, flush=True)
        if kw:
            records.append(kw)

    # the pool forks before this process initialises the GPU
    pool = mp.get_context("fork").Pool(args.workers)
    import torch
    import cvtamd
    from .warmstart import UNet

    seeds = [args.seed * 1_000_003 + i for i in range(args.samples)]
    t0 = time.time()
    dens, init, conv, rel = generate(pool, seeds, args.iters, log)
    pool.close()
    gen_s = time.time() - t0
    log(f"generated {len(seeds)} pairs in {gen_s:.1f} s", stage="generate", seconds=gen_s,
        samples=len(seeds), iters=args.iters, median_rel_decrease=float(np.nanmedian(rel)))

    dev = torch.device("cuda")
    D = torch.from_numpy(dens).to(dev)                                   # (N, 256, 256)
    X = torch.from_numpy(init * np.float32(R / N_GRID)).to(dev)           # sites, 256-px units
    U = torch.from_numpy(init / np.float32(N_GRID)).to(dev)               # uv in [0, 1)
    T = torch.from_numpy((conv - init) / np.float32(N_GRID)).to(dev)     # displacement
    del dens
    n_all = D.shape[0]
    n_val = min(args.val, n_all // 4)
    n_tr = n_all - n_val
    var_val = float((T[n_tr:] ** 2).sum(dim=2).mean())

    B = args.batch
    net = UNet(R, B, args.seed)
    stream = torch.cuda.Stream()
    net.set_stream(stream.cuda_stream)
    steps = int(math.ceil(args.epochs * n_tr / B))
    g = torch.Generator(device="cpu").manual_seed(args.seed)

    def fill(idx):
        for j, i in enumerate(idx.tolist()):
            net.prepare(j, D[i], X[i])
        return U[idx].contiguous(), T[idx].contiguous()

    def evaluate():
        tot, cnt = 0.0, 0
        for s in range(n_tr, n_all, B):
            idx = torch.arange(s, min(s + B, n_all))
            with torch.cuda.stream(stream):
                uv, tg = fill(idx)
            net.forward(len(idx))
            loss = net.loss_backward(len(idx), uv, tg, N_SITES)
            tot += loss * len(idx)
            cnt += len(idx)
        return tot / cnt

    log(f"training: {n_tr} pairs, {n_val} held out, {steps} steps of {B}; "
        f"held-out mean |target|^2 {var_val:.3e}")
    v0 = evaluate()
    log(f"step 0 held-out loss {v0:.4e} (explained {1 - v0 / var_val:.3f})", stage="val", step=0,
        loss=v0, explained=1 - v0 / var_val)
    t1 = time.time()
    perm = torch.randperm(n_tr, generator=g)
    pos = 0
    step = 0
    ema = None
    for step in range(1, steps + 1):
        if pos + B > n_tr:
            perm = torch.randperm(n_tr, generator=g)
            pos = 0
        idx = perm[pos:pos + B]
        pos += B
        with torch.cuda.stream(stream):
            uv, tg = fill(idx)
        net.forward(B)
        want = step % 50 == 0 or step == steps
        loss = net.loss_backward(B, uv, tg, N_SITES, sync=want)
        lr = args.lr * 0.5 * (1 + math.cos(math.pi * step / steps))
        net.adam(lr)
        if want:
            ema = loss if ema is None else 0.9 * ema + 0.1 * loss
            el = time.time() - t1
            if step % 500 == 0 or step == steps:
                log(f"step {step}/{steps} loss {loss:.4e} (ema {ema:.4e}) lr {lr:.2e} "
                    f"{step * B / el:.1f} samples/s")
            if el > args.time_limit:
                log(f"time limit reached at step {step}")
                break
    net.synchronize()
    train_s = time.time() - t1
    v1 = evaluate()
    log(f"held-out loss {v1:.4e} (explained {1 - v1 / var_val:.3f}); trained {step} steps "
        f"in {train_s:.1f} s", stage="val", step=step, loss=v1, explained=1 - v1 / var_val,
        seconds=train_s, samples_per_s=step * B / train_s)
    args.out.parent.mkdir(parents=True, exist_ok=True)
    net.save(args.out)
    if args.log:
        args.log.write_text(json.dumps({"args": {k: str(v) for k, v in vars(args).items()},
                                        "records": records}, indent=1))
    net.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
