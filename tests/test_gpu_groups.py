"""Boxes stepped in lock step, one launch per round for all of a worker's boxes
and up to `depth` steps of each (mcac_batch_run_groups), give the runs each box
gives alone: the same number of MC steps and the same final physical time, bit
for bit, as the oracle."""
import pathlib
import sys

import numpy as np
import pytest

import oracle_lib as O
from mcac_amd import ensemble

pytestmark = pytest.mark.gpu
CFG = pathlib.Path(__file__).parent / "configs"


@pytest.mark.parametrize("depth", [1, 6])
def test_group_runs_match_oracle_runs(depth):
    print(f"case: 8 boxes (one dense, hand-over path), 2 workers x 3 boxes, depth {depth}, staggered arrivals",
          flush=True)
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
