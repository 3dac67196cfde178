"""Digests of whole MC runs: SHA-256 over blocks of step and merge records.

A run's step records (112 B each) and merge records (176 B each) have the same
byte layout on the oracle and on the device driver, so two runs are
bit-identical exactly when every block digest and the final-state digest
agree.  The digests of long oracle runs are committed as fixtures
(tests/golden/digest_*.json, written by tests/golden/make_run_digests.py); a
GPU test recomputes them from the device run.  Each block also keeps its first
record's step, aggregate and pair in clear, which names where two runs part.
"""
import hashlib

import numpy as np


def _blocks(rec, size, fields):
    out = []
    for b in range(0, len(rec), size):
        blk = rec[b:b + size]
        head = {f: (blk[0][f].item() if np.ndim(blk[0][f]) == 0 else blk[0][f].tolist()) for f in fields}
        out.append({"first": b, "count": len(blk), "sha256": hashlib.sha256(blk.tobytes()).hexdigest(),
                    "head": head})
    return out


def final_digest(aggs: dict) -> str:
    """SHA-256 over the live aggregates in label order: stable id, sphere count,
    volume, Rg, rmax, lpm, time step and proper time (the fields both sides export)."""
    h = hashlib.sha256()
    for k, dt in (("uid", np.int64), ("n", np.int64), ("volume", np.float64), ("rg", np.float64),
                  ("rmax", np.float64), ("lpm", np.float64), ("time_step", np.float64),
                  ("proper_time", np.float64)):
        h.update(np.ascontiguousarray(aggs[k], dtype=dt).tobytes())
    return h.hexdigest()


def run_digest(steps, merges, aggs: dict, scalars: dict, step_block: int, merge_block: int) -> dict:
    return {
        "n_steps": int(len(steps)), "n_merges": int(len(merges)),
        "step_block": step_block, "merge_block": merge_block,
        "steps": _blocks(steps, step_block, ("step", "agg", "moving_sphere", "other_sphere", "dist")),
        "merges": _blocks(merges, merge_block, ("step", "keeper", "removed", "n", "volume")),
        "final": final_digest(aggs),
        "scalars": scalars,
    }


def first_difference(a: dict, b: dict) -> str:
    """'' when the digests agree, else a description of the first differing block
    (checked before the counts: where two runs part says more than how long they are)."""
    for kind in ("steps", "merges"):
        for x, y in zip(a[kind], b[kind]):
            if x["sha256"] != y["sha256"] or x["count"] != y["count"]:
                return f"{kind} block at record {x['first']} differs: {x['head']} vs {y['head']}"
    for key in ("n_steps", "n_merges"):
        if a[key] != b[key]:
            return f"{key}: {a[key]} vs {b[key]}"
    if a["final"] != b["final"]:
        return "final aggregate state differs"
    for k, v in a["scalars"].items():
        if b["scalars"].get(k) != v:
            return f"scalar {k}: {v!r} vs {b['scalars'].get(k)!r}"
    return ""
