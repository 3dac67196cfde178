"""Domain duplication on the MI355X path (host driver, the reference's default
with_domain_duplication, aggregat_list.cpp:142-190) against the oracle: whole
runs through one and two duplications, every step and merge bit-exact, and the
final state (8x and 64x the spheres) equal."""
import pathlib
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
CFG = pathlib.Path(__file__).parent / "configs"


@pytest.mark.parametrize("npp,dups", [(60, 1), (300, 2)])
def test_duplication_run_bit_exact(npp, dups):
    import mcac_amd
    print(f"case: duplication_small to {npp} monomers per aggregate ({dups} duplications)", flush=True)
    text = (CFG / "duplication_small.ini").read_text().replace("mean_monomere_per_aggregate = 60",
                                                               f"mean_monomere_per_aggregate = {npp}")
    ora = O.OracleSim(text, fast_init=True)
    n_ora = ora.run(10 ** 7)
    sim = mcac_amd.Simulation(text, trace_steps=True, trace_merges=True)
    done = 0
    while done < n_ora and not sim.finished():
        done += sim.run(min(50000, n_ora - done))
        print(f"  device: {done} steps, {sim.scalar('n_sph'):.0f} spheres", flush=True)
    a, b = ora.steps(), sim.steps()
    m = min(len(a), len(b))
    for f in a.dtype.names:
        eq = a[f][:m] == b[f][:m]
        if eq.ndim > 1:
            eq = eq.all(axis=1)
        bad = np.nonzero(~eq)[0]
        assert bad.size == 0, f"step field {f} differs first at step {bad[0]}: {a[bad[0]]} vs {b[bad[0]]}"
    assert done == n_ora and sim.finished() and ora.finished()
    assert sim.scalar("n_sph") == 400 * 8 ** dups
    assert sim.scalar("box_length") == ora.scalar("box_length")
    assert len(a) == len(b)
    assert ora.merges().tobytes() == sim.merges().tobytes()
    assert sim.scalar("time") == ora.scalar("time")
    _, agg = ora.export()
    got = sim.aggregates()
    for k in ("uid", "n", "volume", "rg", "rmax", "lpm", "time_step", "proper_time"):
        assert np.array_equal(got[k], agg[k]), k
