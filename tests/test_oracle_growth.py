"""Surface growth in the oracle (calcul.cpp:155-205, aggregat.cpp:230-243,
sphere.cpp:113-120, physical_model.cpp:587-590) pinned without the oracle's
own code:

* with the reactions on and a zero flux, the run is the run without them
  (sorting the pick table every step and updating unchanged aggregates change
  nothing);
* every sphere's radius is its initial radius plus u_sg times each step's
  time increment, added step by step (u_sg = flux / density);
* every aggregate's volume is the sum of its spheres' 4/3 pi r^3, and every
  remaining monomer's Rg, time step and lpm are the closed forms of a sphere of
  its grown radius (parallel-axis Rg, Stokes-Cunningham drag, 3 m/f,
  sqrt(6 kT/f dt)).
"""
import math
import pathlib

import numpy as np
import pytest

import oracle_lib as O

CFG = pathlib.Path(__file__).parent / "configs"
K_BOLTZMANN = 1.38066e-23


def growth_ini(flux, stop=200, n=2000, seed=31, reactions=True):
    text = (CFG / "growth_small.ini").read_text()
    text = text.replace("flux_surfgrowth = 1e-3", f"flux_surfgrowth = {flux!r}")
    text = text.replace("number_of_aggregates = 200", f"number_of_aggregates = {stop}")
    text = text.replace("number = 2000", f"number = {n}").replace("random_seed = 31", f"random_seed = {seed}")
    if not reactions:
        text = text.replace("with_surface_reactions = true", "with_surface_reactions = false")
    return text


def test_zero_flux_keeps_radii_and_rederives_positions():
    """Zero flux: radii and volumes never change, but every step's update
    re-derives each aggregate's position from its first sphere
    (compute_mass_center, aggregat.cpp:431-452: position = wrap(sphere 0 +
    r_cm)), which a run without reactions does only at merges."""
    a = O.OracleSim(growth_ini(0.0, n=600, stop=60), fast_init=True)
    b = O.OracleSim(growth_ini(0.0, n=600, stop=60, reactions=False), fast_init=True)
    s0, _ = a.export()
    a.run(2000)
    b.run(2000)
    sa, ga = a.export()
    sb, gb = b.export()
    # radii stay; volumes become 4/3 pi r*r*r (sphere.cpp:113-120), which can
    # differ in the last bit from the pow(r, 3) of the initial placement
    assert np.array_equal(sa["r"], s0["r"])
    r = sa["r"]
    assert np.array_equal(sa["vol"], (4 * math.atan(1.0) * 4 / 3) * (r * r * r))
    assert np.abs(sa["vol"] / s0["vol"] - 1).max() < 1e-15
    L = a.scalar("box_length")
    first = sa["x"][ga["members"][ga["offset"]]], sa["y"][ga["members"][ga["offset"]]], sa["z"][ga["members"][ga["offset"]]]
    for k in range(3):
        want = np.mod(first[k] + ga["rcm"][:, k], L)
        assert np.allclose(ga["pos"][:, k], want, rtol=0, atol=1e-12 * L)
    # the same draws pick the same aggregates until the last-bit volume change
    # of the first growth step reaches the time steps of the pick table
    assert np.array_equal(a.steps()["agg"][:10], b.steps()["agg"][:10])


@pytest.mark.parametrize("flux", [1e-3, 300.0])
def test_growth_closed_forms(flux):
    text = growth_ini(flux, n=800, stop=300)
    sim = O.OracleSim(text, fast_init=True)
    ph = O.physics(text)
    u_sg = flux / 1800.0
    assert ph["u_sg"] == u_sg
    sph0, _ = sim.export()
    r = sph0["r"].copy()
    n = sim.run(3000)
    assert n > 500
    for dt in sim.steps()["dt"]:  # each step grows every sphere by u_sg * dt, in turn
        r = r + u_sg * dt
    sph, agg = sim.export()
    assert np.array_equal(sph["r"], r)
    if flux > 1:
        assert (r / sph0["r"]).min() > 1.01  # the growth is not negligible
    vol = 4.0 / 3.0 * math.pi * r ** 3
    off = agg["offset"]
    lam, mu = ph["lambda"], ph["viscosity"]
    monomers = 0
    for l in range(len(agg["n"])):
        mem = agg["members"][off[l]:off[l] + agg["n"][l]]
        assert agg["volume"][l] == pytest.approx(vol[mem].sum(), rel=1e-13)
        if agg["n"][l] == 1:
            monomers += 1
            rr = float(r[mem[0]])
            assert agg["rg"][l] == pytest.approx(math.sqrt(0.6) * rr, rel=1e-14)
            cc = 1.0 + 1.142 * lam / rr + 0.558 * lam / rr * math.exp(-0.999 * rr / lam)
            f = 6.0 * math.pi * mu * rr / cc
            dt = 3.0 * 1800.0 * vol[mem[0]] / f
            assert agg["time_step"][l] == pytest.approx(dt, rel=1e-13)
            assert agg["lpm"][l] == pytest.approx(math.sqrt(6.0 * K_BOLTZMANN * 1700.0 / f * dt), rel=1e-13)
    assert monomers > 50
