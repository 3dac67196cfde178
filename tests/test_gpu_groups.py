"""Boxes stepped in lock step, one launch per round for all of a worker's boxes
and up to `depth` steps of each (mcac_batch_run_groups), give the runs each box
gives alone: the same number of MC steps and the same final physical time, bit
for bit, as the oracle; with two launches in flight (the default) and with
one."""
import pathlib
import sys

import numpy as np
import pytest

import oracle_lib as O
from mcac_amd import ensemble

pytestmark = pytest.mark.gpu
CFG = pathlib.Path(__file__).parent / "configs"


@pytest.mark.parametrize("depth,overlap,verify", [(1, 1, 1), (6, 1, 1), (8, 1, 1), (8, 0, 1), (8, 1, 0)])
def test_group_runs_match_oracle_runs(depth, overlap, verify, monkeypatch):
    """verify 0: the cell starts the host counts are used without the read-back
    check the other tests make, as in production runs."""
    monkeypatch.setenv("MCAC_GROUP_OVERLAP", str(overlap))
    monkeypatch.setenv("MCAC_VERIFY_CELLS", str(verify))
    print(f"case: 8 boxes (one dense, hand-over path), 2 workers x 3 boxes, depth {depth}, "
          f"{1 + overlap} launches in flight, cell starts checked {verify}, staggered arrivals", flush=True)
    inis = [ensemble.box_ini(n, fv, t, seed) for n, fv, t, seed in
            [(600, 3e-3, 1500.0, 21), (900, 1e-3, 1200.0, 22), (1500, 4e-3, 1900.0, 23),
             (700, 2e-3, 1000.0, 24), (2500, 5e-3, 1700.0, 25), (1200, 1e-4, 1600.0, 26),
             (3000, 2e-3, 1300.0, 27)]]
    inis.append((CFG / "dense_small.ini").read_text())
    arrivals = np.array([0.0, 0.0, 0.0, 0.02, 0.04, 0.06, 0.08, 0.1])
    res = ensemble.batch_run(inis, arrivals=arrivals, workers=2, boxes_per_worker=3, depth=depth)
    assert np.all(res["status"] == 0)
    assert np.all(res["start"] >= arrivals) and np.all(res["end"] > res["start"])
    for j, ini in enumerate(inis):
        ora = O.OracleSim(ini, fast_init=True, trace=False)
        n = ora.run(10 ** 8)
        assert ora.finished()
        print(f"  job {j}: {n} steps", flush=True)
        assert res["steps"][j] == n, (j, res["steps"][j], n)
        assert res["final_time"][j] == ora.scalar("time"), (j, res["final_time"][j], ora.scalar("time"))
    sys.stdout.flush()


def test_group_runs_with_overlap_corrections():
    """Boxes with each overlap correction (caps, alphas, arvo) stepped in groups
    beside a plain box: every run as the oracle's (the corrections merge through
    their own kernel sequences, alphas with a host phase, on the group stream)."""
    print("case: 4 boxes (none, caps, alphas, arvo), 2 workers x 2 boxes, depth 8", flush=True)
    inis = [ensemble.box_ini(800, 3e-3, 1500.0, 31)]
    inis += [(CFG / f"dense_small_{m}.ini").read_text() for m in ("caps", "alphas", "arvo")]
    res = ensemble.batch_run(inis, workers=2, boxes_per_worker=2, depth=8)
    assert np.all(res["status"] == 0)
    for j, ini in enumerate(inis):
        ora = O.OracleSim(ini, fast_init=True, trace=False)
        n = ora.run(10 ** 8)
        assert ora.finished()
        print(f"  job {j}: {n} steps", flush=True)
        assert res["steps"][j] == n, (j, res["steps"][j], n)
        assert res["final_time"][j] == ora.scalar("time"), (j, res["final_time"][j], ora.scalar("time"))


def test_autotune_shapes_keep_results():
    """The launch-shape tuner runs each candidate on the same boxes; every shape
    gives every box the same MC step count, and the choice is one of them."""
    from mcac_amd import autotune
    print("case: autotune on 3 small boxes, 4 shapes", flush=True)
    inis = [ensemble.box_ini(n, 2e-3, 1500.0, 40 + n) for n in (500, 800, 1200)]
    res = autotune.tune_groups(inis, workers=(1, 2), boxes_per_worker=(1, 2), depth=(8,))
    assert len(res["candidates"]) == 4 and all(r["same_step_counts"] for r in res["candidates"])
    assert res["best"] in [{k: r[k] for k in ("workers", "boxes_per_worker", "depth")} for r in res["candidates"]]


def test_group_enqueue_rejects_interleaved_moves():
    """A box's moves must be contiguous in one group enqueue, boxes must belong
    to the group, and n is bounded by 8 moves per box of the group's capacity."""
    import ctypes
    import mcac_amd
    print("case: group enqueue argument checks", flush=True)
    L = mcac_amd.load_gpu()
    g = ctypes.c_void_p()
    assert L.mcac_group_create(0, 2, ctypes.byref(g)) == 0
    boxes = []
    try:
        rng = np.random.default_rng(2)
        for k in range(3):
            b = mcac_amd.Box(1e-6, 4, 8, 8)
            n = 4
            pos = rng.uniform(0, 1e-6, (n, 3))
            b.upload(dict(x=pos[:, 0], y=pos[:, 1], z=pos[:, 2], r=np.full(n, 1e-8), vol=np.full(n, 4.2e-24),
                          surf=np.full(n, 1.3e-15), rx=np.zeros(n), ry=np.zeros(n), rz=np.zeros(n)),
                     dict(pos=pos, rcm=np.zeros((n, 3)), rmax=np.full(n, 1e-8), volume=np.full(n, 4.2e-24),
                          surface=np.full(n, 1.3e-15), count=np.ones(n, np.int64), members=np.arange(n)))
            boxes.append(b)
        assert L.mcac_group_add(g, boxes[0].h) == 0 and L.mcac_group_add(g, boxes[1].h) == 0
        assert L.mcac_group_add(g, boxes[0].h) == 34  # already in a group

        def enqueue(bs):
            n = len(bs)
            arr = (ctypes.c_void_p * n)(*[b.h.value if hasattr(b.h, "value") else b.h for b in bs])
            aggs = np.zeros(n, np.int64)
            dirs = np.tile([1.0, 0.0, 0.0], n)
            lpm = np.full(n, 1e-9)
            maxr = np.full(n, 1e-8)
            t = (ctypes.c_uint64 * max(n, 1))()
            return L.mcac_move_enqueue_group(g, arr, mcac_amd._ip(aggs), mcac_amd._dp(dirs), mcac_amd._dp(lpm),
                                             mcac_amd._dp(maxr), n, t)

        assert enqueue([boxes[0], boxes[1], boxes[0]]) == 34   # interleaved
        assert enqueue([boxes[2]]) == 34                        # not in the group
        assert enqueue([boxes[0]] * 17) == 34                   # more than 8 per box of capacity 2
    finally:
        for b in boxes:
            b.close()
        L.mcac_group_destroy(g)
