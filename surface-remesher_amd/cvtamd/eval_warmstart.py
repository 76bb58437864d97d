"""Time-to-tolerance of the learned warm start against plain Lloyd.

    python -m cvtamd.eval_warmstart --weights W --seeds 200 --out results.json

For each held-out patch (2048^2, 16384 sites, the disk-patch generator on
seeds the training never uses):

1. E* = energy of plain Lloyd (omega = 1) from the initial sites, run until
   the relative decrease over 10 iterations is below 1e-7 (at most 2000
   iterations).  Target: E <= E* (1 + 1e-4).
2. Plain: Lloyd from the initial sites until the target, timed.
3. Learned: from the same initial sites, timed from the start: 10 plain
   Lloyd iterations, then from the initial sites again the warm start (input
   preparation, network, site move: cvt_unet_warm_start) and 10 iterations;
   the lower-energy run continues (plain Lloyd is the arbiter and the
   fallback) until the target.  Both trials are inside the learned time.

Every run stops itself on the device (cvt_lloyd_until: the update kernel
tests the rule after each iteration and the later iterations return at
entry), so plain and learned runs pay the same, small checking cost and the
iteration counts are exact.  Inputs are resident on the device before either
clock starts.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

from .warmstart import TRIAL, warm_start_checked

CHUNK = 10
TOL = 1e-4
CONV = 1e-7
CAP = 2000


def converge(e, cap=CAP):
    """Plain Lloyd until the 10-iteration relative decrease is below CONV
    (at most cap iterations); the energies."""
    e.clear_history()
    e.lloyd_until(cap, rel_tol=CONV, window=CHUNK)
    return e.history()[:, 0]


def run_to(e, target, cap, sync):
    """Lloyd until an energy <= target (at most cap iterations);
    (iterations, reached).  The call synchronises once, at its end."""
    n = e.lloyd_until(cap, target=target)
    h = e.history()[:, 0]
    return n, bool(len(h) and h[-1] <= target)


def time_to_tol(e, net, rho, xy0, cap=CAP, trial=TRIAL):
    """One patch: plain and learned wall times to E* (1 + TOL)."""
    import torch
    e.set_density(rho)
    e.set_sites(xy0)
    h = converge(e, cap)
    e_star = float(h[-1])
    target = e_star * (1.0 + TOL)
    plain_iters_hist = int(np.nonzero(h <= target)[0][0]) + 1 if (h <= target).any() else None

    # plain
    e.set_sites(xy0)
    e.clear_history()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it_p, ok_p = run_to(e, target, cap, e.synchronize)
    t_plain = time.perf_counter() - t0

    # learned: plain Lloyd checks the warm start over the first TRIAL
    # iterations (cvtamd.warmstart.warm_start_checked); both trials are timed
    e.clear_history()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    j = warm_start_checked(net, e, xy0, trial)
    hit = np.nonzero(j["trial_energies"] <= target)[0]
    if len(hit):
        it_l, ok_l = int(hit[0]) + 1, True
    else:
        it, ok_l = run_to(e, target, cap - trial, e.synchronize)
        it_l = trial + it
    t_learned = time.perf_counter() - t0
    fallback = not j["warm"]
    e_init, e_warm = j["e_plain"], j["e_warm"]
    return {
        "e_star": e_star, "e_plain_trial": float(e_init), "e_warm_trial": e_warm,
        "plain_iters": it_p, "plain_reached": ok_p, "plain_s": t_plain,
        "plain_iters_hist": plain_iters_hist,
        "learned_iters": it_l, "learned_reached": ok_l, "learned_s": t_learned,
        "fallback": bool(fallback),
        "ratio": (t_plain / t_learned) if (ok_p and ok_l) else None,
    }


def summarize(rows):
    r = np.array([x["ratio"] for x in rows if x["ratio"] is not None], np.float64)
    out = {
        "patches": len(rows),
        "both_reached": int(len(r)),
        "fallback_rate": float(np.mean([x["fallback"] for x in rows])) if rows else None,
        "median_plain_iters": float(np.median([x["plain_iters"] for x in rows])) if rows else None,
        "median_learned_iters": float(np.median([x["learned_iters"] for x in rows])) if rows else None,
        "median_plain_s": float(np.median([x["plain_s"] for x in rows])) if rows else None,
        "median_learned_s": float(np.median([x["learned_s"] for x in rows])) if rows else None,
    }
    if len(r):
        out.update({"median_ratio": float(np.median(r)), "p10_ratio": float(np.percentile(r, 10)),
                    "p90_ratio": float(np.percentile(r, 90))})
    return out


def _make(idx):
    from . import workloads as wl
    w = wl.warmstart_eval(idx)
    return idx, w.density, w.sites


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--weights", type=Path, required=True)
    ap.add_argument("--seeds", type=int, default=200)
    ap.add_argument("--first", type=int, default=0)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--out", type=Path, default=None)
    ap.add_argument("--trial", type=int, default=TRIAL,
                    help="Lloyd iterations of each checked trial (plain and warm-started)")
    args = ap.parse_args(argv)

    import multiprocessing as mp
    pool = mp.get_context("fork").Pool(args.workers)    # before the GPU is touched
    import torch
    import cvtamd
    from .warmstart import UNet

    net = UNet(256, 1, 0)
    net.load(args.weights)
    rows = []
    eng = cvtamd.Engine(2048, 2048, 16384)
    stream = torch.cuda.Stream()
    eng.set_stream(stream.cuda_stream)
    net.set_stream(stream.cuda_stream)
    for idx, rho, xy in pool.imap(_make, range(args.first, args.first + args.seeds)):
        rho_d = torch.from_numpy(rho).cuda()
        xy_d = torch.from_numpy(xy).cuda()
        res = time_to_tol(eng, net, rho_d, xy_d, trial=args.trial)
        res["index"] = idx
        rows.append(res)
        print(json.dumps(res), flush=True)
    pool.close()
    summ = summarize(rows)
    print(json.dumps({"summary": summ}), flush=True)
    if args.out:
        args.out.write_text(json.dumps({"summary": summ, "patches": rows}, indent=1))
    eng.close()
    net.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
