"""Several boxes sharing the GPU on host worker threads give the runs each box gives alone."""
import sys

import numpy as np
import pytest

import oracle_lib as O
from mcac_amd import ensemble

pytestmark = pytest.mark.gpu


def test_batch_runs_match_oracle_runs():
    print("case: 6 small boxes, 3 workers", flush=True)
    inis = [ensemble.box_ini(n, fv, t, seed) for n, fv, t, seed in
            [(600, 3e-3, 1500.0, 11), (900, 1e-3, 1200.0, 12), (1500, 4e-3, 1900.0, 13),
             (700, 2e-3, 1000.0, 14), (2500, 5e-3, 1700.0, 15), (1200, 1e-4, 1600.0, 16)]]
    res = ensemble.batch_run(inis, workers=3)
    assert np.all(res["status"] == 0)
    for j, ini in enumerate(inis):
        ora = O.OracleSim(ini, fast_init=True, trace=False)
        n = ora.run(10 ** 8)
        assert ora.finished()
        assert res["steps"][j] == n, (j, res["steps"][j], n)
    assert np.all(res["end"] > res["start"])
    sys.stdout.flush()
