"""Domain duplication in the oracle (aggregat_list.cpp:142-190, calcul.cpp:58,
72-81), pinned by what the duplication must produce: right after it the box
side has doubled, every aggregate has seven copies with the same sphere count,
volume, Rg, rmax, time step, lpm and proper time, whose positions are the
original's shifted by the seven non-zero (i, j, k) * old side and wrapped into
the new box, with each sphere at (position - centre of mass offset) + its
relative position; the copies' spheres follow the original spheres in storage,
copy by copy.  It happens when an event leaves at most N0/8 aggregates (N0
fixed at the start) and again each time that holds anew."""
import itertools
import pathlib

import numpy as np

import oracle_lib as O

CFG = pathlib.Path(__file__).parent / "configs"


def test_duplication_state():
    sim = O.OracleSim((CFG / "duplication_small.ini").read_text(), fast_init=True)
    L0 = sim.scalar("box_length")
    steps = 0
    before = None
    while sim.scalar("box_length") == L0:
        if sim.scalar("n_agg") <= 50:
            before = sim.export()
        assert sim.run(1) == 1
        steps += 1
        assert steps < 200000
    assert before is not None
    sph0, agg0 = before
    sph, agg = sim.export()
    L = sim.scalar("box_length")
    assert L == 2 * L0
    n0, s0 = len(agg0["n"]), len(sph0["x"])
    assert n0 <= 50 and len(sph["x"]) == 8 * s0
    # the step right after the duplication moved one aggregate (and may have merged two):
    # compare the copies, which that step did not touch unless it moved or merged them
    st = sim.steps()[-1]
    touched = {int(st["agg"]), int(st["other_agg"])}
    shifts = [s for s in itertools.product((0, 1), repeat=3) if s != (0, 0, 0)]
    uid = {int(u): l for l, u in enumerate(agg["uid"])}
    checked = 0
    next_uid = 400
    for a in range(n0):
        for (i, j, k) in shifts:
            u = next_uid
            next_uid += 1
            if u in touched:
                continue
            l = uid[u]
            for key in ("n", "volume", "rg", "rmax", "time_step", "lpm", "proper_time"):
                assert agg[key][l] == agg0[key][a], (key, a)
            want = np.mod(agg0["pos"][a] + np.array([i, j, k]) * L0, L)
            assert np.allclose(agg["pos"][l], want, rtol=0, atol=1e-15 * L)
            mem = agg["members"][agg["offset"][l]:agg["offset"][l] + agg["n"][l]]
            orig = agg0["members"][agg0["offset"][a]:agg0["offset"][a] + agg0["n"][a]]
            assert np.all(mem >= s0)
            ref = agg["pos"][l] - agg["rcm"][l]
            for m, o in zip(mem, orig):
                assert sph["r"][m] == sph0["r"][o]
                rel = np.array([sph0["rx"][o], sph0["ry"][o], sph0["rz"][o]])
                assert np.array_equal([sph["x"][m], sph["y"][m], sph["z"][m]], ref + rel)
            checked += 1
    assert checked >= 7 * n0 - 2
    # storage order of the copies' spheres: aggregate by aggregate, copy by copy
    firsts = [agg["members"][agg["offset"][uid[u]]] for u in range(400, next_uid) if u not in touched]
    assert np.all(np.diff(firsts) > 0)


def test_duplication_repeats_and_run_ends_on_mean_size():
    sim = O.OracleSim((CFG / "duplication_small.ini").read_text().replace("mean_monomere_per_aggregate = 60",
                                                                          "mean_monomere_per_aggregate = 300"),
                      fast_init=True)
    sim.run(10 ** 7)
    assert sim.finished()
    sph, agg = sim.export()
    assert len(sph["x"]) == 400 * 64  # two duplications
    assert len(sph["x"]) / len(agg["n"]) >= 300
