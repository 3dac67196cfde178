"""Persistent step engine (mcac_box_set_engine, k_step_engine) against one
launch per step and against the oracle: whole runs, queued moves through the
C-ABI, a grid that ends idle and is launched again, and grids of several sizes."""
import pathlib
import sys
import time

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
CFG = pathlib.Path(__file__).parent / "configs"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_sim(text, engine, monkeypatch):
    import mcac_amd
    monkeypatch.setenv("MCAC_ENGINE", str(engine))
    sim = mcac_amd.Simulation(text, trace_steps=True, trace_merges=True)
    n = 0
    while not sim.finished():
        done = sim.run(5000)
        n += done
        if done == 0:
            break
    return sim, n


@pytest.mark.parametrize("cfg,engine", [("lognormal_small.ini", 16), ("lognormal_small.ini", 128),
                                        ("dense_small.ini", 64)])
def test_engine_run_matches_launches_and_oracle(cfg, engine, monkeypatch):
    say(f"case: engine {engine} blocks, whole run {cfg}")
    text = (CFG / cfg).read_text()
    ref, n_ref = run_sim(text, 0, monkeypatch)
    got, n_got = run_sim(text, engine, monkeypatch)
    assert n_got == n_ref
    assert got.steps().tobytes() == ref.steps().tobytes()
    assert got.merges().tobytes() == ref.merges().tobytes()
    st = got.box().engine_stats()
    say("engine stats", st)
    assert st["steps"] > 0.5 * n_got and st["end_reason"] == 3
    ora = O.OracleSim(text, fast_init=True)
    assert ora.run(10 ** 7) == n_got
    assert ora.steps().tobytes() == got.steps().tobytes()


def test_engine_queued_moves_and_idle_relaunch():
    """Queued moves on a box with the engine on give the records and final state
    of one-at-a-time moves; a grid left idle for 1 s ends and is launched again."""
    say("case: engine queued moves with an idle pause, lognormal_small after 3000 steps")
    import mcac_amd
    sim = O.OracleSim((CFG / "lognormal_small.ini").read_text(), fast_init=True, trace=False)
    sim.run(3000)
    sph, agg = sim.export()
    boxes = []
    for _ in range(2):
        b = mcac_amd.Box(sim.scalar("box_length"), 10, len(sph["x"]), len(agg["n"]))
        b.upload(sph, dict(pos=agg["pos"], rcm=agg["rcm"], rmax=agg["rmax"], volume=agg["volume"],
                           surface=agg["surface"], count=agg["n"], members=agg["members"]))
        boxes.append(b)
    a, b = boxes
    b.set_engine(32)
    maxr = sim.scalar("maxradius")
    rng = np.random.default_rng(9)
    live = list(range(len(agg["n"])))
    contacts = 0
    for rnd in range(40):
        if rnd == 20:
            # a short move without contact leaves the grid resident; then it
            # ends after 0.5 s without a step and the next step launches it again
            aid, d, lpm = live[0], np.array([1.0, 0.0, 0.0]), 1e-15
            got, _ = b.move_result(b.move_enqueue(aid, d, lpm, maxr))
            ref = a.next_contact(aid, d, lpm, maxr, move=True)
            assert got["distance"] == ref["distance"] and got["contact"] == 0 == ref["contact"]
            time.sleep(1.0)
        moves = []
        for _ in range(3):
            d = rng.normal(size=3)
            aid = live[int(rng.integers(len(live)))]
            moves.append((aid, d / np.linalg.norm(d), agg["lpm"][aid] * (1 + 6 * rng.random())))
        tickets = [b.move_enqueue(aid, d, lpm, maxr) for aid, d, lpm in moves]
        for t, (aid, d, lpm) in zip(tickets, moves):
            got, skipped = b.move_result(t)
            if skipped == 1:  # behind a contact or a hand-over of this round
                continue
            ref = a.next_contact(aid, d, lpm, maxr, move=True)
            for f in ("distance", "move", "moving_sphere", "other_sphere", "other_agg", "contact"):
                assert got[f] == ref[f], (rnd, f)
            if got["contact"]:
                contacts += 1
                for t2 in tickets[tickets.index(t) + 1:]:
                    assert b.move_result(t2)[1] == 1
                ra, rb = a.merge(ref), b.merge(got)
                assert ra.tobytes() == rb.tobytes()
                b.release()
                if ra["merged"]:
                    live.remove(int(ra["removed"]))
                break
    st = b.engine_stats()
    say("engine stats", st, "contacts", contacts)
    assert st["relaunches"] >= 1 and st["steps"] > 0
    sa, sb = a.spheres(), b.spheres()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    ga, gb = a.aggregates(), b.aggregates()
    for k in ga:
        assert np.array_equal(ga[k], gb[k]), k
