"""Pins of the oracle's per-aggregate geometry and physics against closed forms
that do not come from the oracle's own code.

* Centre of mass, rmax and Rg of unequal spheres (aggregat.cpp:431-483): the
  centre of mass of a union of disjoint solid spheres is the volume-weighted
  mean of their centres, and by the parallel-axis theorem a solid sphere of
  radius r contributes v (d^2 + 3/5 r^2) to the second moment about a point at
  distance d (checked here also by sampling points of the solid).
* Monomer physics (physical_model.cpp:550-617, aggregat.cpp:254-276): Stokes
  drag with the Cunningham slip correction of Kim et al. (2005, Aerosol Sci.
  Technol. 39:120, constants 1.142, 0.558, 0.999), the Einstein relation
  D = kT/f, the momentum relaxation time tau = m/f, the reference's time step
  3 tau and mean free path sqrt(6 D dt).
* Aggregates of n equal disjoint spheres: f(n) = f(1) n^(fe/Df) and the
  mobility diameter's power law, on every merge of the monodisperse run.
* The global time step (aggregat_list.cpp:54-58): max/sum(max/dt_i) is the
  harmonic combination 1/sum(1/dt_i); at t=0 in the monodisperse box it is
  dt_mono/N; a contact step scales it by distance/lpm (calcul.cpp:148-152);
  physical time is the running sum of the steps.
"""
import math
import pathlib

import numpy as np
import pytest

import oracle_lib as O

CFG = pathlib.Path(__file__).parent / "configs"
K_BOLTZMANN = 1.38066e-23  # the reference's value (include/constants.hpp)


def _chain(rng, n, rmin=5e-9, rmax=25e-9):
    """n spheres of unequal radii, each touching the previous one, no others
    overlapping; positions relative to sphere 0."""
    r = rng.uniform(rmin, rmax, n)
    pos = np.zeros((n, 3))
    for i in range(1, n):
        for _ in range(1000):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            p = pos[i - 1] + d * (r[i - 1] + r[i])
            if np.all(np.linalg.norm(pos[:i - 1] - p, axis=1) > (r[:i - 1] + r[i]) * (1 + 1e-6)):
                break
        pos[i] = p
    return pos, r


def _independent_geometry(pos, r):
    v = 4.0 / 3.0 * math.pi * r ** 3
    com = (v[:, None] * pos).sum(axis=0) / v.sum()
    d = np.linalg.norm(pos - com, axis=1)
    rg = math.sqrt(float((v * (d ** 2 + 0.6 * r ** 2)).sum() / v.sum()))
    return com, float((r + d).max()), rg, float(v.sum())


@pytest.mark.parametrize("seed", range(40))
def test_unequal_sphere_com_rmax_rg(seed):
    """Merged geometry of random unequal-radius chains (method none) vs the
    volume-weighted centre and the parallel-axis Rg, keeper split at random."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 40))
    pos, r = _chain(rng, n)
    vol = 4.0 / 3.0 * math.pi * r ** 3
    surf = 4.0 * math.pi * r ** 2
    nk = int(rng.integers(1, n))
    g = O.merge_geometry(0, nk, pos, r, vol, surf)
    com, rmax, rg, V = _independent_geometry(pos, r)
    scale = np.abs(pos).max() + r.max()
    assert g["volume"] == pytest.approx(V, rel=1e-13)
    assert np.allclose(g["rcm"], com, rtol=0, atol=1e-13 * scale)
    assert g["rmax"] == pytest.approx(rmax, rel=1e-13)
    assert g["rg"] == pytest.approx(rg, rel=1e-12)
    # a weighting by radius or by count instead of volume moves the centre far beyond that
    wrong = (r[:, None] * pos).sum(axis=0) / r.sum()
    assert np.abs(wrong - com).max() > 1e-6 * scale or n == 2 and r[0] == r[1]


def test_rg_by_sampling_the_solid():
    """Three unequal disjoint spheres: Rg and centre of mass from 4e6 uniform
    points of the solid agree with the oracle's to sampling accuracy."""
    rng = np.random.default_rng(1)
    pos = np.array([[0.0, 0.0, 0.0], [30e-9, 0.0, 0.0], [30e-9, 26e-9, 5e-9]])
    r = np.array([10e-9, 20e-9, 6e-9])
    vol = 4.0 / 3.0 * math.pi * r ** 3
    g = O.merge_geometry(0, 1, pos, r, vol, 4 * math.pi * r ** 2)
    lo, hi = (pos - r[:, None]).min(axis=0), (pos + r[:, None]).max(axis=0)
    pts = rng.uniform(lo, hi, (4_000_000, 3))
    inside = np.zeros(len(pts), bool)
    for p, q in zip(pos, r):
        inside |= ((pts - p) ** 2).sum(axis=1) <= q * q
    pts = pts[inside]
    com = pts.mean(axis=0)
    rg = math.sqrt(float(((pts - com) ** 2).sum(axis=1).mean()))
    assert np.abs(com - g["rcm"]).max() < 3e-3 * r.max()
    assert rg == pytest.approx(g["rg"], rel=2e-3)


def _monodisperse():
    text = (CFG / "c1_cpu_reference.ini").read_text()
    return text, O.physics(text)


def _stokes_cunningham(r, lam, mu):
    cc = 1.0 + 1.142 * lam / r + 0.558 * lam / r * math.exp(-0.999 * r / lam)
    return 6.0 * math.pi * mu * r / cc


def test_monomer_physics_closed_forms():
    """C1 monomer: f = 6 pi mu r / Cc(r), D = kT/f, dt = 3 m/f, lpm = sqrt(6 D dt);
    the values 368.55 nm, 2.7135e-7 s and 8.343e-8 m^2/s."""
    text, ph = _monodisperse()
    sim = O.OracleSim(text, fast_init=True, trace=False)
    sph, agg = sim.export()
    r = float(sph["r"][0])
    assert np.all(sph["r"] == r)
    f = _stokes_cunningham(r, ph["lambda"], ph["viscosity"])
    m = 1800.0 * 4.0 / 3.0 * math.pi * r ** 3
    dt = 3.0 * m / f
    D = K_BOLTZMANN * 1700.0 / f
    lpm = math.sqrt(6.0 * D * dt)
    assert np.allclose(agg["time_step"], dt, rtol=1e-13, atol=0)
    assert np.allclose(agg["lpm"], lpm, rtol=1e-13, atol=0)
    D_oracle = agg["lpm"][0] ** 2 / (6.0 * agg["time_step"][0])
    assert D_oracle == pytest.approx(D, rel=1e-13)
    assert agg["lpm"][0] == pytest.approx(368.55e-9, rel=2e-5)
    assert agg["time_step"][0] == pytest.approx(2.7135e-7, rel=2e-5)
    assert D_oracle == pytest.approx(8.343e-8, rel=1e-4)


def test_equal_sphere_aggregates_power_law():
    """Every merge of the monodisperse run (method none, equal radii r):
    f(n) = f(1) n^(fe/Df), d_m = 2 r n^(fe/Df) Cc(r n^(fe/2Df)) / Cc(r),
    time step 3 rho n v / f(n), lpm sqrt(6 (kT/f) dt), dp = 2 r."""
    text, ph = _monodisperse()
    sim = O.OracleSim(text, fast_init=True, trace=True)
    sim.run(12000)
    sph, _ = sim.export()
    r = float(sph["r"][0])
    lam, mu = ph["lambda"], ph["viscosity"]
    fe = 0.689 * (1.0 + math.erf((lam / r + 4.454) / 10.628))
    f1 = _stokes_cunningham(r, lam, mu)
    v = 4.0 / 3.0 * math.pi * r ** 3

    def cc(x):
        return 1.0 + 1.142 * lam / x + 0.558 * lam / x * math.exp(-0.999 * x / lam)

    m = sim.merges()
    assert len(m) > 200 and m["n"].max() >= 8
    for rec in m:
        n = int(rec["n"])
        fn = f1 * n ** (fe / 1.78)
        assert rec["f_agg"] == pytest.approx(fn, rel=1e-12), n
        assert rec["dp"] == pytest.approx(2 * r, rel=1e-14)
        dm = 2 * r * n ** (fe / 1.78) * cc(r * n ** (fe / 1.78 / 2)) / cc(r)
        assert rec["d_m"] == pytest.approx(dm, rel=1e-12), n
        dt = 3 * 1800.0 * n * v / fn
        assert rec["time_step"] == pytest.approx(dt, rel=1e-12), n
        assert rec["lpm"] == pytest.approx(math.sqrt(6 * K_BOLTZMANN * 1700.0 / fn * dt), rel=1e-12), n


def test_first_step_global_time_step():
    """t=0, N equal monomers: the first step advances time by dt_mono/N times
    (distance/lpm on contact), and time is the running sum of the steps."""
    text, _ = _monodisperse()
    sim = O.OracleSim(text, fast_init=True, trace=True)
    _, agg = sim.export()
    dt_mono = float(agg["time_step"][0])
    sim.run(400)
    st = sim.steps()
    s0 = st[0]
    scale = s0["dist"] / s0["lpm"] if s0["contact"] else 1.0
    assert s0["dt"] == pytest.approx(dt_mono / 1000 * scale, rel=1e-14)
    free = st[:np.argmax(st["merged"] == 1)] if st["merged"].any() else st
    free = free[free["contact"] == 0]
    assert len(free) >= 5 and np.all(free["dt"] == free["dt"][0])
    assert free["dt"][0] == pytest.approx(dt_mono / 1000, rel=1e-14)
    assert np.array_equal(np.cumsum(st["dt"]), st["time"])


@pytest.mark.parametrize("steps", [2000, 6000])
def test_global_time_step_is_harmonic(steps):
    """Lognormal box after some merges: the next step's time increment over
    its move fraction equals 1/sum(1/dt_i) over the live aggregates."""
    text = (CFG / "lognormal_small.ini").read_text()
    sim = O.OracleSim(text, fast_init=True, trace=True)
    sim.run(steps)
    _, agg = sim.export()
    expect = 1.0 / np.sum(1.0 / agg["time_step"])
    sim.run(1)
    s = sim.steps()[-1]
    frac = s["dist"] / s["lpm"] if s["contact"] else 1.0
    assert s["dt"] / frac == pytest.approx(expect, rel=1e-12)
    assert len(np.unique(agg["time_step"])) > 100  # a genuinely mixed set of time steps
