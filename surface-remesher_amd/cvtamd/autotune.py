"""Autotuner for the Lloyd step's launch shapes, with a persisted table.

    python -m cvtamd.autotune [--workloads c1 c2 c3 c4] [--table PATH]

Two choices are tuned, neither of which changes a result bit (the GPU tests
compare every choice bit for bit):

  * the moment-pass tile (k_moments<T>, one workgroup per T x T tile, T in
    {32, 64, 128, 256}): small tiles give more workgroups and smaller LDS
    hashes, large tiles fewer overflow records for big cells.  The engine's
    built-in rule takes the smallest tile at least four mean cell diameters
    wide.
  * the JFA launch shape (cvt_set_jfa_shape): per step, runs of four rows
    one step apart or bands of rows (a bit mask over the steps 2^k), and the
    workgroups launched per CU.  The built-in rule takes runs wherever every
    wave of the launch gets one, and four 256-thread or one 1024-thread
    workgroup per CU by the size of the site table.

The tuner times whole Lloyd steps on the actual grid and sites: the tile by
HIP events around eager steps (cvt_profile), then, with that tile, the JFA
shape by the graph-launched iterations that cvt_lloyd runs (the runs/bands
mask greedily, one step at a time, from the best of automatic, all runs and
all bands; then the workgroups per CU).  It records the fastest under a key
of grid size and site-density bucket.  `tile_for` / `jfa_shape_for` return
the recorded choice (or None), `apply` sets both on an engine and
`apply_batch` the tile on each patch of a batch (the batched JFA has its own
fixed layout).  The reference's analogue is its fixed launch-shape tables
(source/gcvt.cu:842-847).
"""
from __future__ import annotations

import argparse
import json
import math
import sys
from pathlib import Path

TABLE = Path(__file__).resolve().parent.parent / "tuning" / "launch_shapes.json"
TILES = (32, 64, 128, 256)


def key(W: int, H: int, S: int) -> str:
    """Grid size and site density (pixels per site, half-octave buckets)."""
    b = int(round(2.0 * math.log2(max(W * H / max(S, 1), 1.0))))
    return f"{W}x{H}:{b}"


def load(path: Path = TABLE) -> dict:
    try:
        return json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return {}


def tile_for(W: int, H: int, S: int, table: dict | None = None):
    t = (load() if table is None else table).get(key(W, H, S))
    return int(t["tile"]) if t else None


def jfa_shape_for(W: int, H: int, S: int, table: dict | None = None):
    """(runs_mask, blocks_per_cu) recorded for this grid and site density, or None."""
    t = (load() if table is None else table).get(key(W, H, S))
    if not t or "jfa" not in t:
        return None
    return int(t["jfa"]["runs_mask"]), int(t["jfa"]["blocks_per_cu"])


def apply(engine, W: int, H: int, S: int, table: dict | None = None) -> int:
    """Sets the recorded tile and JFA shape on an Engine (the built-in rules
    where none is recorded); returns the tile."""
    table = load() if table is None else table
    t = tile_for(W, H, S, table)
    engine.tile = t or 0
    engine.jfa_shape = jfa_shape_for(W, H, S, table) or (-1, 0)
    return engine.tile


def apply_batch(batch, sizes, sites, table: dict | None = None):
    """Sets each patch's recorded tile (batched tiles are at least 64: the
    batched moment pass uses 64-pixel segments); returns the tiles in use."""
    table = load() if table is None else table
    out = []
    for p, ((W, H), S) in enumerate(zip(sizes, sites)):
        t = tile_for(W, H, S, table)
        if t:
            batch.set_tile(p, max(t, 64))
        out.append(batch.tile(p))
    return out


def tune(density, sites, iters: int = 20, tiles=TILES) -> dict:
    """Mean step time (ms) per tile on this grid and these sites."""
    import cvtamd
    H, W = density.shape
    S = len(sites)
    res = {}
    with cvtamd.Engine(W, H, S) as e:
        e.set_density(density)
        for t in tiles:
            e.tile = t
            e.set_sites(sites)
            try:
                e.profile(3)
                p = e.profile(iters)
            except cvtamd.CvtError:
                continue                      # the tile cannot hold this site density
            res[t] = p["step"]
    return res


def jfa_steps(W: int, H: int) -> list:
    """log2 of the JFA steps (the first step is the largest power of two
    below max(W, H), halved down to 1)."""
    n, s = max(W, H), 1
    while s * 2 < n:
        s *= 2
    return list(range(s.bit_length()))


def time_lloyd(e, sites, shape, iters: int = 30, repeats: int = 3) -> float:
    """Median ms per graph-launched Lloyd iteration with this JFA shape, each
    repeat from the same initial sites."""
    import time
    e.jfa_shape = shape
    out = []
    for _ in range(repeats + 1):
        e.set_sites(sites)
        e.clear_history()
        e.lloyd(2)
        e.synchronize()
        t0 = time.perf_counter()
        e.lloyd(iters)
        e.synchronize()
        out.append((time.perf_counter() - t0) * 1e3 / iters)
    return sorted(out[1:])[repeats // 2]


def tune_jfa(density, sites, tile: int = 0, iters: int = 30, margin: float = 0.01) -> dict:
    """The fastest JFA shape on this grid and these sites: {"runs_mask",
    "blocks_per_cu", "ms": {shape: ms}}; a change is kept only when it is
    faster by more than `margin` (relative)."""
    import cvtamd
    H, W = density.shape
    bits = jfa_steps(W, H)
    full = sum(1 << k for k in bits)
    ms = {}
    with cvtamd.Engine(W, H, len(sites)) as e:
        e.set_density(density)
        e.tile = tile

        def run(mask, bpc):
            if (mask, bpc) not in ms:
                ms[(mask, bpc)] = time_lloyd(e, sites, (mask, bpc), iters)
            return ms[(mask, bpc)]

        e.jfa_shape = (-1, 0)
        auto = e.jfa_shape[0]                  # the steps the built-in rule gives runs
        mask = min(dict.fromkeys((auto, full, 0)), key=lambda m: run(m, 0))
        best_ms = run(mask, 0)
        for k in reversed(bits):
            cand = mask ^ (1 << k)
            if run(cand, 0) < best_ms * (1.0 - margin):
                mask, best_ms = cand, run(cand, 0)
        bpc = 0
        for b in (1, 2, 4, 8):
            if run(mask, b) < best_ms * (1.0 - margin):
                bpc, best_ms = b, run(mask, b)
    return {"runs_mask": mask, "blocks_per_cu": bpc,
            "ms": {f"{m}/{b}": round(v, 5) for (m, b), v in ms.items()}}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workloads", nargs="+", default=["c1", "c2", "c3", "c4"])
    ap.add_argument("--patches", type=int, default=64, help="c4 patches to tune")
    ap.add_argument("--table", type=Path, default=TABLE)
    ap.add_argument("--force", action="store_true", help="re-time keys already in the table")
    ap.add_argument("--no-jfa", action="store_true", help="tune the moment tile only")
    args = ap.parse_args(argv)
    from . import workloads as wl
    table = load(args.table)
    cases = []
    for name in args.workloads:
        # batched patches run the batched JFA, whose layout is not tuned
        if name == "c4":
            cases += [(w, False) for w in wl.config4(4, num_patches=args.patches)]
        else:
            cases.append(({"c1": wl.config1, "c2": wl.config2, "c3": wl.config3}[name](), True))
    for w, jfa in cases:
        jfa = jfa and not args.no_jfa
        k = key(w.n, w.n, w.num_sites)
        if k in table and not args.force and ("jfa" in table[k] or not jfa):
            continue
        if k in table and not args.force:
            best = int(table[k]["tile"])
        else:
            times = tune(w.density, w.sites)
            if not times:
                continue
            best = min(times, key=times.get)
            table[k] = {"tile": best, "ms": {str(t): round(v, 5) for t, v in times.items()},
                        "workload": w.name}
        if jfa:
            table[k]["jfa"] = tune_jfa(w.density, w.sites, best)
        print(json.dumps({k: table[k]}), flush=True)
    args.table.parent.mkdir(parents=True, exist_ok=True)
    args.table.write_text(json.dumps(dict(sorted(table.items())), indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
