"""Whole-run parity against the oracle's committed run digests
(tests/golden/digest_*.json, written by tests/golden/make_run_digests.py).

* N=1e5 lognormal, FV=1e-3 (tests/configs/c2_plain_gpu.ini) run to its stop
  criterion, N_agg = 1000: every step and merge record and the final state.
* N=1e6 lognormal, FV=3e-3 (tests/configs/c3_n1m.ini), first 3000 MC steps,
  with the FP32 broad phase (precision 1) and with its misses counted
  (precision 2, which must count none): identical to the FP64 oracle, so the
  per-step contact-distance and volume error is zero.
"""
import json
import pathlib
import sys

import pytest

from run_digest import first_difference, run_digest

pytestmark = pytest.mark.gpu
ROOT = pathlib.Path(__file__).resolve().parents[1]
CFG = ROOT / "tests" / "configs"
GOLD = ROOT / "tests" / "golden"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def device_digest(ref: dict):
    import mcac_amd
    text = (CFG / ref["ini"]).read_text()
    sim = mcac_amd.Simulation(text, trace_steps=True, trace_merges=True)
    done = 0
    while done < ref["max_steps"]:
        n = sim.run(min(200000, ref["max_steps"] - done))
        done += n
        say(f"  device: {done} steps, n_agg {sim.scalar('n_agg'):.0f}")
        if n == 0 or sim.finished():
            break
    scalars = {"time": sim.scalar("time"), "n_agg": sim.scalar("n_agg"), "finished": bool(sim.finished())}
    d = run_digest(sim.steps(), sim.merges(), sim.aggregates(), scalars, ref["step_block"], ref["merge_block"])
    return d, sim


def test_c2_whole_run_digest(monkeypatch):
    """N=1e5 -> N_agg=1000: the whole merge sequence, every step and the final state bit-exact."""
    ref = json.loads((GOLD / "digest_c2_full.json").read_text())
    say(f"case: C2 whole run, {ref['n_steps']} oracle steps, {ref['n_merges']} merges")
    monkeypatch.setenv("MCAC_PRECISION", "0")
    got, _ = device_digest(ref)
    diff = first_difference(ref, got)
    assert diff == "", diff
    assert got["scalars"]["finished"] and got["scalars"]["n_agg"] == 1000


@pytest.mark.parametrize("precision", ["1", "2"])
def test_c3_precision_form_prefix_digest(precision, monkeypatch):
    """N=1e6, FP32 broad phase + FP64 solve: the first 3000 steps bit-exact against the FP64 oracle."""
    ref = json.loads((GOLD / "digest_c3_prefix.json").read_text())
    say(f"case: C3 precision {precision}, N=1e6, {ref['n_steps']} steps")
    monkeypatch.setenv("MCAC_PRECISION", precision)
    got, sim = device_digest(ref)
    diff = first_difference(ref, got)
    assert diff == "", diff
    if precision == "2":  # precision 2 counts the FP32 test's work and its misses
        ps = sim.box().prune_stats()
        say("prune stats", ps)
        assert ps["tested"] > 0 and ps["kept"] < ps["tested"] and ps["missed"] == 0
