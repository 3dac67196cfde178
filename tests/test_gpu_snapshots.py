"""Snapshot files written during a run on the HIP path, against the oracle's run.

Write points follow calcul.cpp:67-71 and physical_model.cpp:338-356 (a write
at the top of a loop iteration when the previous step was an event and the
event count is a multiple of write_events_frequency), plus the final save
(calcul.cpp:282-284).  Every grid is checked for consistency; the last one
must equal the oracle's final state bit for bit."""
import math
import os
import pathlib
import sys

import numpy as np
import pytest

import hdf5_lib as H
import oracle_lib as O

pytestmark = pytest.mark.gpu
CFG = pathlib.Path(__file__).parent / "configs"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def read_grids(out, kind, n):
    width = math.ceil(math.log10(n)) + 4
    grids, k = [], 0
    while os.path.exists(out / f"{kind}_{k:0{width}d}.xmf"):
        info, gl = H.read_xmf(str(out / f"{kind}_{k:0{width}d}.xmf"))
        h5 = str(out / f"{kind}_{k:0{width}d}.h5")
        for g in gl:
            rec = {"time": g["GridTime"], "pos": H.read_dataset(h5, g["Positions"]).reshape(-1, 3)}
            for a, v in g.items():
                if a not in ("Name", "GridTime", "Positions"):
                    rec[a] = H.read_dataset(h5, v[0])
            grids.append(rec)
        k += 1
    return info, grids, k


def test_snapshots_match_oracle(tmp_path):
    import mcac_amd
    every, per_file = 700, 3
    text = (CFG / "lognormal_small.ini").read_text() + (
        f"\n[output]\nwrite_events_frequency = {every}\nwrite_between_event_frequency = 1000000000\n"
        f"n_time_per_file = {per_file}\n")
    say("case: lognormal N=5000 -> 50 aggregates with snapshots every 700 events")
    ora = O.OracleSim(text)
    n_ora = ora.run(10 ** 7)
    assert ora.finished()
    steps = ora.steps()
    sph, agg = ora.export()
    say(f"  oracle: {n_ora} steps")

    out = tmp_path / "run"
    sim = mcac_amd.Simulation(text, output_dir=str(out))
    n_gpu = 0
    while not sim.finished() and n_gpu < n_ora:
        n_gpu += sim.run(min(20000, n_ora - n_gpu))
        say(f"  device: {n_gpu} steps")
    assert sim.finished() and n_gpu == n_ora
    sim.close()

    # expected write points: iteration k writes when k == 0 or step k-1 merged, and the events so far % every == 0
    merged = steps["merged"] != 0
    events = np.concatenate([[0], np.cumsum(merged)])
    at = [k for k in range(n_ora) if (k == 0 or merged[k - 1]) and events[k] % every == 0]
    times = [0.0 if k == 0 else steps["time"][k - 1] for k in at] + [steps["time"][-1]]
    n_aggs = [5000 - events[k] for k in at] + [5000 - events[n_ora]]

    for kind in ("Spheres", "Aggregats"):
        info, grids, n_files = read_grids(out, kind, 5000)
        say(f"  {kind}: {len(grids)} grids in {n_files} files")
        assert len(grids) == len(times) and n_files == math.ceil(len(times) / per_file)
        assert info["N []"] == "5000"
        for g, t, na in zip(grids, times, n_aggs):
            assert g["time"] == t and g["Time"][0] == t
            if kind == "Aggregats":
                assert len(g["Label"]) == na and np.array_equal(g["Label"], np.arange(na))
                assert g["Np"].sum() == 5000
            else:
                assert len(g["Label"]) == 5000 and np.array_equal(g["electric_charge"], np.zeros(5000))
        if kind == "Spheres":
            sgrids = grids
        else:
            agrids = grids
    for s, a in zip(sgrids, agrids):  # each sphere's label names an aggregate with that many spheres
        assert np.array_equal(np.bincount(s["Label"], minlength=len(a["Np"])), a["Np"])

    last_s, last_a = sgrids[-1], agrids[-1]
    assert np.array_equal(last_s["pos"], np.stack([sph["x"], sph["y"], sph["z"]], axis=1))
    assert np.array_equal(last_s["Radius"], sph["r"])
    assert np.array_equal(last_s["Label"], sph["label"])
    assert np.array_equal(last_a["pos"], agg["pos"])
    for name, key in (("Rg", "rg"), ("Np", "n"), ("lpm", "lpm"), ("Deltat", "time_step"), ("Rmax", "rmax"),
                      ("Volume", "volume"), ("Surface", "surface"), ("proper_time", "proper_time")):
        assert np.array_equal(last_a[name], agg[key]), name
