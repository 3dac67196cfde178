"""HIP path (through the C-ABI) against the oracle on the same seeded inputs.

Bar: bit-exact for every distance, index, merge geometry and time value."""
import pathlib
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
CFG = pathlib.Path(__file__).parent / "configs"
INF = float("inf")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mcac():
    import mcac_amd
    return mcac_amd


def test_swept_sphere_batch_bit_exact():
    """a9 on the device vs the oracle on random, touching, grazing and periodic inputs."""
    say("case: swept-sphere batch, 60000 inputs")
    rng = np.random.default_rng(3)
    n = 60000
    L = 2.4e-6
    p1 = rng.uniform(-0.2 * L, 1.2 * L, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r1 = rng.uniform(5e-9, 40e-9, n)
    r2 = rng.uniform(5e-9, 40e-9, n)
    dist = rng.uniform(1e-8, 1.5 * L, n)
    t = rng.uniform(0, 1, n)[:, None] * dist[:, None]
    p2 = p1 + t * d + rng.normal(scale=1, size=(n, 3)) * (r1 + r2)[:, None]
    p2[::7] = rng.uniform(0, L, (len(p2[::7]), 3))            # far ones
    p2[1::11] = p1[1::11] + d[1::11] * (r1 + r2)[1::11, None]  # touching
    p2[2::13] += L * rng.integers(-2, 3, (len(p2[2::13]), 3))  # other periodic images
    got = mcac().swept_sphere_batch(p1, r1, p2, r2, d, dist, L)
    ref = np.array([O.swept_sphere(p1[i], r1[i], p2[i], r2[i], d[i], dist[i], L) for i in range(n)])
    assert np.isfinite(ref).sum() > n // 10
    mism = np.nonzero(~((got == ref) | (np.isnan(got) & np.isnan(ref))))[0]
    assert mism.size == 0, f"{mism.size} mismatches, first {mism[:5]}: {got[mism[:5]]} vs {ref[mism[:5]]}"


def upload_oracle_state(sim):
    sph, agg = sim.export()
    m = mcac()
    box = m.Box(sim.scalar("box_length"), 10, len(sph["x"]), len(agg["n"]))
    box.upload(sph, dict(pos=agg["pos"], rcm=agg["rcm"], rmax=agg["rmax"], volume=agg["volume"],
                         surface=agg["surface"], count=agg["n"], members=agg["members"]))
    return box, sph, agg


@pytest.mark.parametrize("cfg,steps", [("c1_cpu_reference.ini", 4000), ("lognormal_small.ini", 12000)])
def test_next_contact_bit_exact(cfg, steps):
    """a4-a9 on an evolved state: distance, sphere pair and other aggregate."""
    say(f"case: next_contact parity {cfg} after {steps} steps")
    sim = O.OracleSim((CFG / cfg).read_text(), fast_init=True, trace=False)
    sim.run(steps)
    box, sph, agg = upload_oracle_state(sim)
    maxr = sim.scalar("maxradius")
    rng = np.random.default_rng(17)
    hits = 0
    for q in range(400):
        label = int(rng.integers(len(agg["n"])))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        lpm = agg["lpm"][label] * (1 + 4 * rng.random())
        ref = sim.next_contact(label, d, lpm)
        got = box.next_contact(label, d, lpm, maxr)
        assert got["distance"] == ref[0], (q, label)
        assert (got["moving_sphere"], got["other_sphere"], got["other_agg"]) == (ref[1], ref[2], ref[3]), (q, label)
        assert got["n_query"] == ref[4]["query"]
        assert got["n_filtered"] == ref[4]["filtered"]
        hits += ref[0] < INF
    assert hits > 20


def compare_runs(cfg_text, max_steps, fast_init=True):
    ora = O.OracleSim(cfg_text, fast_init=fast_init)
    n_ora = ora.run(max_steps)
    say(f"  oracle: {n_ora} steps, finished={ora.finished()}")
    sim = mcac().Simulation(cfg_text, trace_steps=True, trace_merges=True)
    # the device run is bounded by the oracle's step count, in pieces that report progress
    n_gpu = 0
    while n_gpu < n_ora:
        n_gpu += sim.run(min(20000, n_ora - n_gpu))
        say(f"  device: {n_gpu} steps")
        if sim.finished():
            break
    assert n_gpu == n_ora
    a, b = ora.steps(), sim.steps()
    assert len(a) == len(b)
    for f in a.dtype.names:  # the first differing step names the case
        eq = (a[f] == b[f]) | (np.isnan(a[f]) & np.isnan(b[f])) if a[f].dtype.kind == "f" else a[f] == b[f]
        if eq.ndim > 1:
            eq = eq.all(axis=1)
        bad = np.nonzero(~eq)[0]
        assert bad.size == 0, f"step field {f} differs first at step {bad[0]}: {a[bad[0]]} vs {b[bad[0]]}"
    assert sim.finished() == ora.finished()
    ma, mb = ora.merges(), sim.merges()
    assert len(ma) == len(mb)
    for f in ma.dtype.names:
        eq = ma[f] == mb[f]
        if eq.ndim > 1:
            eq = eq.all(axis=1)
        bad = np.nonzero(~eq)[0]
        assert bad.size == 0, f"merge field {f} differs first at merge {bad[0]}: {ma[bad[0]]} vs {mb[bad[0]]}"
    assert sim.scalar("time") == ora.scalar("time")
    assert sim.scalar("n_agg") == ora.scalar("n_agg")
    return ora, sim


def test_full_run_c1_bit_exact():
    """N=1000 monodisperse to one aggregate: every step and merge identical."""
    say("case: full C1 run, N=1000 -> 1 aggregate")
    text = (CFG / "c1_cpu_reference.ini").read_text()
    ora, sim = compare_runs(text, 10 ** 7, fast_init=False)
    assert sim.finished() and sim.scalar("n_agg") == 1


def test_lognormal_run_bit_exact():
    """N=5000 lognormal to 50 aggregates: every step and merge identical."""
    say("case: lognormal N=5000 -> 50 aggregates")
    compare_runs((CFG / "lognormal_small.ini").read_text(), 10 ** 7)


@pytest.mark.parametrize("rounds,grid", [("0", "1"), ("1", "1"), ("0", "0")])
def test_dense_run_bit_exact(rounds, grid, monkeypatch):
    """N=3000 at FV=0.05 to 60 aggregates: every step and merge identical (large candidate sets),
    with one launch per step and with rounds of up to eight steps per group launch; the large
    movers' searches in one launch with a grid-wide barrier (k_step_grid) and in two
    (k_filter + k_solve)."""
    monkeypatch.setenv("MCAC_ROUNDS", rounds)
    monkeypatch.setenv("MCAC_GRID_SOLVE", grid)
    say(f"case: dense N=3000 FV=0.05 -> 60 aggregates, MCAC_ROUNDS={rounds}, MCAC_GRID_SOLVE={grid}")
    ora, sim = compare_runs((CFG / "dense_small.ini").read_text(), 10 ** 7)
    st = sim.box().stats()
    say("stats", st)
    assert st["two_launch"] > 0  # large movers took the form under test


def test_c2_prefix_bit_exact():
    """N=100,000 lognormal (plain GPU configuration): first 4000 steps identical."""
    say("case: C2 prefix N=1e5, 4000 steps")
    compare_runs((CFG / "c2_plain_gpu.ini").read_text(), 4000)


def test_c3_full_size_invariants():
    """N=1,000,000: sphere count, volume and merge bookkeeping hold over 3000 steps."""
    say("case: C3 N=1e6 invariants, 3000 steps")
    text = (CFG / "c3_n1m.ini").read_text()
    sim = mcac().Simulation(text, trace_steps=True, trace_merges=True)
    v0 = sim.scalar("total_volume")
    n = sim.run(3000)
    assert n == 3000
    st, mg = sim.steps(), sim.merges()
    assert st["merged"].sum() == len(mg) == 1_000_000 - sim.scalar("n_agg")
    aggs = sim.aggregates()
    assert aggs["n"].sum() == 1_000_000
    assert sim.scalar("total_volume") == pytest.approx(v0, rel=1e-12)
    assert np.all(np.diff(st["time"]) > 0)
    assert np.all(st["dist"][st["contact"] == 1] <= st["lpm"][st["contact"] == 1])


@pytest.mark.parametrize("auto_merge", [False, True])
def test_queued_moves_match_synchronous_moves(auto_merge):
    """Gated queued moves (three in flight) give the records and positions of one-at-a-time moves;
    with auto-merge the contact's merge runs in the move's launch and matches a launched merge."""
    say(f"case: queued moves vs synchronous moves, lognormal_small after 3000 steps, auto_merge={auto_merge}")
    sim = O.OracleSim((CFG / "lognormal_small.ini").read_text(), fast_init=True, trace=False)
    sim.run(3000)
    a, sph, agg = upload_oracle_state(sim)
    b, _, _ = upload_oracle_state(sim)
    b.set_auto_merge(auto_merge)
    maxr = sim.scalar("maxradius")
    rng = np.random.default_rng(5)
    live = list(range(len(agg["n"])))
    draws = []

    def draw(k):
        while len(draws) <= k:
            d = rng.normal(size=3)
            draws.append((int(rng.integers(1 << 30)), d / np.linalg.norm(d), 1 + 6 * rng.random()))
        return draws[k]

    def plan(k):
        i, d, f = draw(k)
        agg_id = live[i % len(live)]
        return agg_id, d, agg["lpm"][agg_id] * f

    contacts = 0
    flight = []
    done = 0
    while done < 600:
        while len(flight) < 3:
            aid, d, lpm = plan(len(flight))
            flight.append((b.move_enqueue(aid, d, lpm, maxr), aid, d, lpm))
        ticket, aid, d, lpm = flight.pop(0)
        got, skipped = b.move_result(ticket)
        if skipped == 1:  # behind a step that closed the gate: all queued moves were skipped
            for t, *_ in flight:
                assert b.move_result(t)[1] == 1
            flight.clear()
            continue
        done += 1
        ref = a.next_contact(aid, d, lpm, maxr, move=True)
        for f in ("distance", "move", "moving_sphere", "other_sphere", "other_agg", "contact"):
            assert got[f] == ref[f], (done, f)
        draws.pop(0)
        if skipped == 2:  # second-launch solve: the moves queued behind it were skipped
            for t, *_ in flight:
                assert b.move_result(t)[1] == 1
            flight.clear()
        if got["contact"]:
            contacts += 1
            for t, *_ in flight:
                assert b.move_result(t)[1] == 1, "a move ran past a contact"
            flight.clear()
            ra, rb = a.merge(ref), b.merge(got)
            assert ra.tobytes() == rb.tobytes()
            b.release()
            if ra["merged"]:
                live.remove(int(ra["removed"]))
    while flight:  # moves still queued ran on the device: replay them on the other box
        ticket, aid, d, lpm = flight.pop(0)
        got, skipped = b.move_result(ticket)
        if skipped == 1:
            continue
        ref = a.next_contact(aid, d, lpm, maxr, move=True)
        assert got["distance"] == ref["distance"] and got["move"] == ref["move"]
        if got["contact"]:
            for t, *_ in flight:
                assert b.move_result(t)[1] == 1
            flight.clear()
            assert a.merge(ref).tobytes() == b.merge(got).tobytes()
            b.release()
    assert contacts > 5
    sa, sb = a.spheres(), b.spheres()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    ga, gb = a.aggregates(), b.aggregates()
    for k in ga:
        assert np.array_equal(ga[k], gb[k]), k
    st = b.stats()
    say("stats", st)
    assert (st["inline_merges"] > 0) == auto_merge


def test_one_launch_handover_bit_exact():
    """Searches whose survivors exceed one block's share go to the multi-block solve, same results."""
    say("case: k_step handover, dense_small run to 60 aggregates, small movers, long moves")
    sim = O.OracleSim((CFG / "dense_small.ini").read_text(), fast_init=True, trace=False)
    sim.run(10 ** 7)
    box, sph, agg = upload_oracle_state(sim)
    maxr = sim.scalar("maxradius")
    L = sim.scalar("box_length")
    counts = np.asarray(agg["n"])
    small = [int(i) for i in np.argsort(counts, kind="stable") if counts[i] * counts.max() <= 16384]
    assert small, "no mover small enough for the one-launch search"
    rng = np.random.default_rng(23)
    for q in range(160):
        label = small[int(rng.integers(len(small)))]
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        lpm = L * (0.3 + 0.2 * rng.random())
        ref = sim.next_contact(label, d, lpm)
        got = box.next_contact(label, d, lpm, maxr)
        assert got["distance"] == ref[0], (q, label)
        assert (got["moving_sphere"], got["other_sphere"], got["other_agg"]) == (ref[1], ref[2], ref[3]), (q, label)
        assert got["n_filtered"] == ref[4]["filtered"]
    st = box.stats()
    say("stats", st)
    assert st["handover"] > 0 and st["one_launch"] > st["handover"]


def test_run_in_pieces_bit_exact():
    """Runs split over many mcac_sim_run calls of odd sizes equal one oracle run."""
    say("case: lognormal_small in pieces")
    text = (CFG / "lognormal_small.ini").read_text()
    ora = O.OracleSim(text, fast_init=True)
    ora.run(5000)
    sim = mcac().Simulation(text, trace_steps=True, trace_merges=True)
    done = 0
    for n in (1, 2, 3, 7, 333, 1000, 1, 1654, 1999):
        done += sim.run(n)
    assert done == 5000
    a, b = ora.steps(), sim.steps()
    assert a.tobytes() == b.tobytes()
    assert ora.merges().tobytes() == sim.merges().tobytes()


@pytest.mark.parametrize("cfg", ["lognormal_small.ini", "dense_small.ini"])
def test_fp32_broad_phase_bit_exact(cfg, monkeypatch):
    """FP32 broad phase (precision 2: misses counted): same steps and merges as the oracle, no miss."""
    say(f"case: FP32 broad phase, {cfg}")
    monkeypatch.setenv("MCAC_PRECISION", "2")
    ora, sim = compare_runs((CFG / cfg).read_text(), 10 ** 7)
    ps = sim.box().prune_stats()
    say("prune stats", ps)
    assert ps["tested"] > 0 and ps["missed"] == 0 and ps["kept"] < ps["tested"]  # counted with precision 2


@pytest.mark.parametrize("depth", [1, 2, 5])
def test_pipeline_depths_bit_exact(depth, monkeypatch):
    """Steps kept in flight (1, 2, 5) do not change any step or merge."""
    say(f"case: pipeline depth {depth}, lognormal_small")
    monkeypatch.setenv("MCAC_PIPELINE_DEPTH", str(depth))
    compare_runs((CFG / "lognormal_small.ini").read_text(), 10 ** 7)
