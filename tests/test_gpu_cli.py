"""The command-line driver mcac_amd/bin/mcac_run (the reference's `MCAC params.ini`,
src/main.cpp:26-56) on a small parameter file: exit code, final state and the
advancement.dat rows (calcul.cpp:29-43) against the oracle's run."""
import pathlib
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
ROOT = pathlib.Path(__file__).resolve().parents[1]
CFG = pathlib.Path(__file__).parent / "configs"


def test_mcac_run_small_box(tmp_path):
    print("case: mcac_run lognormal_small.ini", flush=True)
    # a snapshot and an advancement row every 1000 events (the reference default is every event)
    text = (CFG / "lognormal_small.ini").read_text() + (
        "\n[output]\nwrite_events_frequency = 1000\nwrite_between_event_frequency = 1000000\n")
    ini = tmp_path / "params_in.ini"
    ini.write_text(text)
    exe = ROOT / "mcac_amd" / "bin" / "mcac_run"
    assert exe.exists(), "mcac_run is not built"
    out = tmp_path / "run"
    r = subprocess.run([str(exe), str(ini), str(out)], capture_output=True, text=True, timeout=100)
    sys.stdout.write(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    assert "Final number of aggregates : 50" in r.stdout
    ora = O.OracleSim(text, fast_init=True, trace=False)
    ora.run(10 ** 7)
    assert ora.finished() and ora.scalar("n_agg") == 50
    rows = np.loadtxt(out / "advancement.dat", ndmin=2)
    assert rows.shape[1] == 9 and len(rows) == 6  # events 0, 1000, ..., 4000, then the final state
    # first row at t = 0 with every monomer its own aggregate, last row the final state
    assert rows[0, 0] == 0.0 and rows[0, 3] == 1.0
    assert rows[-1, 0] == pytest.approx(ora.scalar("time"), rel=1e-5)
    assert rows[-1, 1] * rows[-1, 5] == pytest.approx(50, rel=1e-4)   # concentration x box volume
    assert rows[-1, 3] == pytest.approx(5000 / 50, rel=1e-5)          # mean monomers per aggregate
    assert np.all(np.diff(rows[:, 0]) >= 0)
    assert (out / "params.ini").exists()


def test_mcac_run_usage_errors(tmp_path):
    exe = ROOT / "mcac_amd" / "bin" / "mcac_run"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=30)
    assert r.returncode == 4 and "usage" in r.stderr
    r = subprocess.run([str(exe), str(tmp_path / "missing.ini")], capture_output=True, text=True, timeout=30)
    assert r.returncode == 2
