"""Surface reactions beyond uniform growth on the MI355X path (calcul.cpp:155-205)
against the oracle, bit for bit: oxidation with sphere removal (the device
flags aggregates holding a sphere at or below rp_min, the host removes them in
list order and rebuilds the box), splits into connected parts (k_connected
flags them, the host splits them in the reference's component order),
reactions of the moved aggregate alone (individual_surf_reactions), and the
MergeError the reference raises when a contact's aggregate split in the same
step while its spheres still touch."""
import sys

import numpy as np
import pytest

import oracle_lib as O
from test_oracle_reactions import reactions_ini

pytestmark = pytest.mark.gpu
INDIV = "[numerics]\nindividual_surf_reactions = true\n"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def compare_runs(text, max_steps, expect_error=False):
    import mcac_amd
    ora = O.OracleSim(text, fast_init=True)
    ora_error = False
    try:
        n_ora = ora.run(max_steps)
    except RuntimeError:
        ora_error = True
        n_ora = int(ora.scalar("steps"))
    say(f"  oracle: {n_ora} steps, error {ora_error}, {ora.scalar('n_agg'):.0f} aggregates, "
        f"{ora.scalar('n_sph'):.0f} spheres")
    assert ora_error == expect_error
    sim = mcac_amd.Simulation(text, trace_steps=True, trace_merges=True)
    done, sim_error = 0, False
    try:
        while done < n_ora and not sim.finished():
            k = sim.run(min(5000, max_steps - done))
            done += k
            say(f"  device: {done} steps, {sim.scalar('n_agg'):.0f} aggregates, {sim.scalar('n_sph'):.0f} spheres")
            if k == 0:
                break
    except RuntimeError as e:
        sim_error = True
        say(f"  device error: {e}")
    assert sim_error == expect_error
    a, b = ora.steps(), sim.steps()
    assert len(a) == len(b), (len(a), len(b))
    for f in a.dtype.names:  # the first differing step names the case
        eq = a[f] == b[f]
        if eq.ndim > 1:
            eq = eq.all(axis=1)
        bad = np.nonzero(~eq)[0]
        assert bad.size == 0, f"step field {f} differs first at step {bad[0]}: {a[bad[0]]} vs {b[bad[0]]}"
    ma, mb = ora.merges(), sim.merges()
    assert len(ma) == len(mb) and ma.tobytes() == mb.tobytes()
    if expect_error:
        return ora, sim
    assert sim.scalar("time") == ora.scalar("time")
    assert sim.scalar("n_sph") == ora.scalar("n_sph") and sim.scalar("n_agg") == ora.scalar("n_agg")
    sph, agg = ora.export()
    rad = sim.box().radii()
    assert np.array_equal(rad["r"], sph["r"]) and np.array_equal(rad["vol"], sph["vol"])
    got = sim.aggregates()
    for k in ("uid", "n", "volume", "rg", "rmax", "lpm", "time_step", "proper_time"):
        assert np.array_equal(got[k], agg[k]), k
    return ora, sim


def test_oxidation_removes_spheres_bit_exact():
    say("case: oxidation flux -30, N=120 seed 5, whole run (spheres vanish)")
    ora, sim = compare_runs(reactions_ini(-30.0, n=120, seed=5), 10 ** 6)
    assert ora.scalar("n_sph") < 60


@pytest.mark.parametrize("flux,seed,steps", [(-3e-11, 32, 2500), (-3e-10, 32, 1500)])
def test_oxidation_splits_bit_exact(flux, seed, steps):
    say(f"case: oxidation flux {flux}, N=300 seed {seed}, {steps} steps (merges and splits)")
    ora, sim = compare_runs(reactions_ini(flux, n=300, seed=seed), steps)
    assert len(ora.merges()) > 50


def test_merge_error_at_the_same_step():
    say("case: oxidation flux -1e-11, N=300 seed 31: a contact's aggregate splits while its spheres touch")
    compare_runs(reactions_ini(-1e-11, n=300, seed=31), 30000, expect_error=True)


@pytest.mark.parametrize("flux,seed,steps", [(1e-3, 9, 3000), (-3e-10, 32, 1500)])
def test_individual_reactions_bit_exact(flux, seed, steps):
    say(f"case: individual reactions flux {flux}, N=300 seed {seed}, {steps} steps")
    compare_runs(reactions_ini(flux, n=300, seed=seed, extra=INDIV), steps)
