"""Surface growth on the device (mcac_box_grow, k_grow_update; calcul.cpp:155-205)
against the oracle: whole runs bit-exact, every step and merge, and the grown
radii, volumes and aggregate state at the end."""
import pathlib
import sys

import numpy as np
import pytest

import oracle_lib as O
from test_oracle_growth import growth_ini

pytestmark = pytest.mark.gpu
CFG = pathlib.Path(__file__).parent / "configs"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("flux,n,stop", [(1e-3, 2000, 200), (300.0, 800, 300), (30.0, 3000, 100)])
def test_growth_run_bit_exact(flux, n, stop):
    import mcac_amd
    say(f"case: surface growth flux={flux} N={n} -> {stop} aggregates")
    text = growth_ini(flux, n=n, stop=stop)
    ora = O.OracleSim(text, fast_init=True)
    n_ora = ora.run(10 ** 7)
    say(f"  oracle: {n_ora} steps")
    sim = mcac_amd.Simulation(text, trace_steps=True, trace_merges=True)
    done = 0
    while done < n_ora and not sim.finished():
        done += sim.run(min(20000, n_ora - done))
        say(f"  device: {done} steps")
    assert done == n_ora and sim.finished() and ora.finished()
    a, b = ora.steps(), sim.steps()
    assert len(a) == len(b)
    for f in a.dtype.names:  # the first differing step names the case
        eq = a[f] == b[f]
        if eq.ndim > 1:
            eq = eq.all(axis=1)
        bad = np.nonzero(~eq)[0]
        assert bad.size == 0, f"step field {f} differs first at step {bad[0]}: {a[bad[0]]} vs {b[bad[0]]}"
    ma, mb = ora.merges(), sim.merges()
    assert len(ma) == len(mb) and ma.tobytes() == mb.tobytes()
    assert sim.scalar("time") == ora.scalar("time")
    sph, agg = ora.export()
    rad = sim.box().radii()
    assert np.array_equal(rad["r"], sph["r"]) and np.array_equal(rad["vol"], sph["vol"])
    got = sim.aggregates()
    for k in ("uid", "n", "volume", "rg", "rmax", "lpm", "time_step", "proper_time"):
        assert np.array_equal(got[k], agg[k]), k
    if flux > 1:  # the growth changed the radii noticeably
        sph0, _ = O.OracleSim(text, fast_init=True, trace=False).export()
        assert (sph["r"] / sph0["r"]).min() > 1.005
