"""Launch-shape tuner for the batched form.

The grouped runner (mcac_batch_run_groups) has three shape parameters: host
worker threads sharing the GPU, boxes stepped in lock step per worker (one
launch per round for all of them) and MC steps of each box per launch.  Their
best values depend on the box sizes and on the host's cores, so they are
measured: every candidate runs the same sample of boxes to completion, the
one with the most runs per hour wins, and every candidate must give each box
the same MC step count (the shape never changes results).

    python3 -m mcac_amd.autotune --boxes 24 --nmax 8000 --out tuning.json

The JSON holds the sample, every candidate's wall time and runs/hour, and the
choice; `best_shape(path)` reads the choice back (bench.py takes it with
--tuned FILE).
"""
from __future__ import annotations

import argparse
import itertools
import json
import sys
import time

import numpy as np


def tune_groups(inis: list[str], workers=(2, 4, 8), boxes_per_worker=(1, 2, 4, 8), depth=(4, 8),
                max_steps: int = 3_000_000, log=None) -> dict:
    """Runs the sample once per candidate shape; returns the table and the fastest shape."""
    from . import ensemble
    rows, steps_ref = [], None
    for w, g, d in itertools.product(workers, boxes_per_worker, depth):
        if g == 1 and d != depth[0]:
            continue  # one box per worker steps one move per launch whatever the depth
        t0 = time.perf_counter()
        r = ensemble.batch_run(inis, workers=w, max_steps=max_steps, boxes_per_worker=g, depth=d)
        wall = time.perf_counter() - t0
        if steps_ref is None:
            steps_ref = r["steps"].copy()
        same = bool(np.array_equal(r["steps"], steps_ref))
        row = {"workers": w, "boxes_per_worker": g, "depth": d, "wall_s": wall,
               "runs_per_hour": 3600.0 * len(inis) / wall, "same_step_counts": same}
        rows.append(row)
        if log:
            print(json.dumps(row), file=log, flush=True)
        if not same:
            raise RuntimeError(f"shape {w}x{g}x{d} changed a box's MC step count")
    best = max(rows, key=lambda r: r["runs_per_hour"])
    return {"candidates": rows, "best": {k: best[k] for k in ("workers", "boxes_per_worker", "depth")}}


def best_shape(path: str) -> dict:
    """The chosen (workers, boxes_per_worker, depth) of a tuning file."""
    with open(path) as f:
        return json.load(f)["best"]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--boxes", type=int, default=24, help="first boxes of the 512-box ensemble list")
    ap.add_argument("--nmax", type=float, default=8000, help="largest box of the sample")
    ap.add_argument("--out", default="tuning.json")
    args = ap.parse_args(argv)
    from . import ensemble
    boxes = [b for b in ensemble.ensemble_configs(args.boxes) if b["n"] <= args.nmax]
    print(f"autotune: {len(boxes)} boxes (N <= {args.nmax:g})", file=sys.stderr, flush=True)
    res = tune_groups([b["ini"] for b in boxes], log=sys.stderr)
    res["sample"] = {"boxes": len(boxes), "of_first": args.boxes, "n_max": args.nmax,
                     "n": [int(b["n"]) for b in boxes]}
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res["best"]), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
