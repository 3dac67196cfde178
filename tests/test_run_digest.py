"""The run-digest machinery (tests/run_digest.py) on oracle runs (CPU): equal
runs give equal digests, a one-bit change of one record or a shorter run is
named, and the committed fixtures are well formed."""
import json
import pathlib

import numpy as np

import oracle_lib as O
from run_digest import first_difference, run_digest

CFG = pathlib.Path(__file__).parent / "configs"
GOLD = pathlib.Path(__file__).parent / "golden"


def _digest(sim, steps=None, merges=None):
    _, agg = sim.export()
    sc = {"time": sim.scalar("time"), "n_agg": sim.scalar("n_agg"), "finished": bool(sim.finished())}
    return run_digest(sim.steps() if steps is None else steps, sim.merges() if merges is None else merges, agg, sc,
                      500, 50)


def test_digest_detects_one_bit_and_length():
    text = (CFG / "lognormal_small.ini").read_text()
    a = O.OracleSim(text, fast_init=True)
    a.run(3000)
    b = O.OracleSim(text, fast_init=True)
    b.run(3000)
    da, db = _digest(a), _digest(b)
    assert first_difference(da, db) == ""
    st = b.steps().copy()
    st["dist"][1234] = np.nextafter(st["dist"][1234], np.inf) if np.isfinite(st["dist"][1234]) else 1.0
    diff = first_difference(da, _digest(b, steps=st))
    assert diff.startswith("steps block at record 1000")
    m = b.merges().copy()
    m["rg"][7] *= 1 + 2.0 ** -52
    assert first_difference(da, _digest(b, merges=m)).startswith("merges block at record 0")
    assert first_difference(da, _digest(b, steps=b.steps()[:2999])).startswith("steps block at record 2500")


def test_committed_fixtures_are_complete():
    for name, ini in (("digest_c2_full.json", "c2_plain_gpu.ini"), ("digest_c3_prefix.json", "c3_n1m.ini")):
        d = json.loads((GOLD / name).read_text())
        assert d["ini"] == ini and (CFG / ini).exists()
        assert sum(b["count"] for b in d["steps"]) == d["n_steps"]
        assert sum(b["count"] for b in d["merges"]) == d["n_merges"]
        assert all(len(b["sha256"]) == 64 for b in d["steps"] + d["merges"])
    c2 = json.loads((GOLD / "digest_c2_full.json").read_text())
    assert c2["scalars"]["finished"] and c2["scalars"]["n_agg"] == 1000 and c2["n_merges"] == 99000
