"""Autotuner for the moment-pass tile, with a persisted table.

    python -m cvtamd.autotune [--workloads c1 c2 c3 c4] [--table PATH]

The moment pass (k_moments<T>) runs one workgroup per T x T tile, T in
{32, 64, 128, 256}.  Results do not depend on T (per-site totals are exact
integer sums; the GPU tests compare tiles bit for bit), so T is a pure
speed choice: small tiles give more workgroups and smaller LDS hashes, large
tiles fewer overflow records for big cells.  The engine's built-in rule
takes the smallest tile at least four mean cell diameters wide.

The tuner times whole Lloyd steps (cvt_profile, HIP events) for every tile
that the site density allows, on the actual grid and sites, and records the
fastest under a key of grid size and site-density bucket.  `tile_for`
returns the recorded tile (or None), `apply` / `apply_batch` set it on an
engine or on each patch of a batch.  The reference's analogue is its fixed
launch-shape tables (source/gcvt.cu:842-847).
"""
from __future__ import annotations

import argparse
import json
import math
import sys
from pathlib import Path

TABLE = Path(__file__).resolve().parent.parent / "tuning" / "moment_tiles.json"
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


def apply(engine, W: int, H: int, S: int, table: dict | None = None) -> int:
    """Sets the recorded tile on an Engine (the built-in rule if none); returns it."""
    t = tile_for(W, H, S, table)
    engine.tile = t or 0
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


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workloads", nargs="+", default=["c1", "c2", "c3", "c4"])
    ap.add_argument("--patches", type=int, default=64, help="c4 patches to tune")
    ap.add_argument("--table", type=Path, default=TABLE)
    ap.add_argument("--force", action="store_true", help="re-time keys already in the table")
    args = ap.parse_args(argv)
    from . import workloads as wl
    table = load(args.table)
    cases = []
    for name in args.workloads:
        if name == "c4":
            cases += wl.config4(4, num_patches=args.patches)
        else:
            cases.append({"c1": wl.config1, "c2": wl.config2, "c3": wl.config3}[name]())
    for w in cases:
        k = key(w.n, w.n, w.num_sites)
        if k in table and not args.force:
            continue
        times = tune(w.density, w.sites)
        if not times:
            continue
        best = min(times, key=times.get)
        table[k] = {"tile": best, "ms": {str(t): round(v, 5) for t, v in times.items()},
                    "workload": w.name}
        print(json.dumps({k: table[k]}), flush=True)
    args.table.parent.mkdir(parents=True, exist_ok=True)
    args.table.write_text(json.dumps(dict(sorted(table.items())), indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
