"""Closed forms and brute force for the oracle's geometry (sphere_contact.cpp:47-125,
aggregat_distance.cpp:24-44, aggregat_list.cpp:447-532, aggregat.cpp:431-483)."""
import math
import pathlib

import numpy as np
import pytest

import oracle_lib as O

L = 1e-6
INF = float("inf")


def ss(p1, r1, p2, r2, d, dist, box=L):
    return O.swept_sphere(p1, r1, p2, r2, d, dist, box)


def test_head_on():
    # sphere 2 straight ahead at D: contact after D - (r1 + r2)
    assert ss([0.2e-6, 0.5e-6, 0.5e-6], 10e-9, [0.4e-6, 0.5e-6, 0.5e-6], 15e-9, [1, 0, 0], 0.3e-6) == \
        pytest.approx(0.2e-6 - 25e-9, rel=1e-12)


def test_grazing_and_miss():
    r1, r2 = 10e-9, 15e-9
    off = r1 + r2
    # perpendicular offset exactly r1 + r2: dist_to_axis == dc2 -> contact at proj
    d = ss([0.2e-6, 0.5e-6, 0.5e-6], r1, [0.4e-6, 0.5e-6 + off, 0.5e-6], r2, [1, 0, 0], 0.3e-6)
    assert d == pytest.approx(0.2e-6, rel=1e-9)
    # a little more: no contact
    assert ss([0.2e-6, 0.5e-6, 0.5e-6], r1, [0.4e-6, 0.5e-6 + off * 1.001, 0.5e-6], r2, [1, 0, 0], 0.3e-6) == INF


def test_touching_is_zero_and_behind_is_inf():
    assert ss([0.5e-6] * 3, 10e-9, [0.5e-6 + 19e-9, 0.5e-6, 0.5e-6], 10e-9, [0, 1, 0], 1e-7) == 0.0
    assert ss([0.5e-6] * 3, 10e-9, [0.3e-6, 0.5e-6, 0.5e-6], 10e-9, [1, 0, 0], 1e-7) == INF


def test_too_far_and_end_contact():
    # contact at 0.2e-6 - 25e-9 but travel only 0.1e-6 -> none
    assert ss([0.2e-6, 0.5e-6, 0.5e-6], 10e-9, [0.4e-6, 0.5e-6, 0.5e-6], 15e-9, [1, 0, 0], 0.1e-6) == INF
    # travel ends inside sphere 2 although its centre is beyond the travel: still found
    d = ss([0.2e-6, 0.5e-6, 0.5e-6], 10e-9, [0.4e-6, 0.5e-6, 0.5e-6], 15e-9, [1, 0, 0], 0.19e-6)
    assert d == pytest.approx(0.2e-6 - 25e-9, rel=1e-12)


def test_periodic_image():
    # sphere 2 lies across the periodic boundary ahead of sphere 1
    d = ss([0.95e-6, 0.5e-6, 0.5e-6], 10e-9, [0.10e-6, 0.5e-6, 0.5e-6], 10e-9, [1, 0, 0], 0.3e-6)
    assert d == pytest.approx(0.15e-6 - 20e-9, rel=1e-9)
    # and through two periods when the sweep is longer than the box
    d = ss([0.95e-6, 0.5e-6, 0.5e-6], 10e-9, [0.10e-6, 0.5e-6, 0.5e-6], 10e-9, [-1, 0, 0], 1.5e-6)
    assert d == pytest.approx(0.85e-6 - 20e-9, rel=1e-9)


def test_random_against_ray_march():
    """Brute force: march sphere 1 along dir and find the first touching step."""
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(400):
        p1 = rng.uniform(0, L, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r1, r2 = rng.uniform(5e-9, 20e-9, 2)
        dist = rng.uniform(0.05e-6, 0.3e-6)
        # place sphere 2 near the path so that hits happen
        t = rng.uniform(0, dist)
        p2 = np.mod(p1 + t * d + rng.normal(scale=(r1 + r2) * 0.8, size=3), L)
        got = ss(p1, r1, p2, r2, d, dist)
        steps = np.linspace(0.0, dist, 20001)
        pts = p1[None, :] + steps[:, None] * d[None, :]
        diff = pts - p2[None, :]
        diff -= L * np.round(diff / L)
        hit = np.nonzero(np.sum(diff * diff, axis=1) <= (r1 + r2) ** 2)[0]
        if hit.size == 0:
            assert got == INF or got > dist - dist / 20000 * 2
        else:
            first = steps[hit[0]]
            assert got == pytest.approx(first, abs=dist / 20000 * 1.5)
            checked += 1
    assert checked > 50


def c1():
    return (pathlib.Path(__file__).parent / "configs" / "c1_cpu_reference.ini").read_text()


def brute_force_contact(sph, agg, label, d, lpm, box):
    """Minimum over every other aggregate and every sphere pair, no Verlet grid, no filter."""
    best = (INF, -1, -1)
    me = agg["members"][agg["offset"][label]:agg["offset"][label] + agg["n"][label]]
    for other in range(len(agg["n"])):
        if other == label:
            continue
        ot = agg["members"][agg["offset"][other]:agg["offset"][other] + agg["n"][other]]
        for i in me:
            for j in ot:
                v = O.swept_sphere([sph["x"][i], sph["y"][i], sph["z"][i]], sph["r"][i],
                                   [sph["x"][j], sph["y"][j], sph["z"][j]], sph["r"][j], d, lpm, box)
                if v < best[0]:
                    best = (v, int(i), int(j))
    return best


def test_next_contact_matches_brute_force():
    """Verlet query + bounding filter + ordered scan (a4-a8) loses no contact."""
    sim = O.OracleSim(c1())
    sim.run(3000)
    sph, agg = sim.export()
    box = sim.scalar("box_length")
    rng = np.random.default_rng(11)
    hits = 0
    for _ in range(60):
        label = int(rng.integers(len(agg["n"])))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        lpm = agg["lpm"][label] * 3
        dist, ms, os_, ol, _ = sim.next_contact(label, d, lpm)
        bd, bi, bj = brute_force_contact(sph, agg, label, d, lpm, box)
        assert dist == bd
        if bd < INF:
            assert (ms, os_) == (bi, bj)
            hits += 1
    assert hits > 5


def test_first_merges_closed_forms():
    """Two equal touching spheres: CoM at the midpoint, rmax = 2r, Rg^2 = r^2 + 3/5 r^2, V = 2v."""
    sim = O.OracleSim(c1())
    sim.run(2000)
    m = sim.merges()
    sph, _ = sim.export()
    r = sph["r"][0]
    assert np.all(sph["r"] == r)  # monodisperse
    v = 4 * math.pi / 3 * r ** 3
    pairs = m[m["n"] == 2]
    assert len(pairs) > 5
    for rec in pairs:
        assert rec["volume"] == pytest.approx(2 * v, rel=1e-15)
        assert np.linalg.norm(rec["rcm"]) == pytest.approx(r, rel=1e-12)
        assert rec["rmax"] == pytest.approx(2 * r, rel=1e-12)
        assert rec["rg"] ** 2 == pytest.approx(r ** 2 + 0.6 * r ** 2, rel=1e-12)
        assert rec["coordination"] == pytest.approx(1.0)


def test_full_run_invariants():
    sim = O.OracleSim(c1())
    sim.run(10 ** 7)
    assert sim.finished()
    st, m = sim.steps(), sim.merges()
    assert len(m) == 999 and m[-1]["n"] == 1000
    assert st["merged"].sum() == 999
    sph, agg = sim.export()
    assert agg["n"].tolist() == [1000]
    assert agg["volume"][0] == pytest.approx(sph["vol"].sum(), rel=1e-12)
    assert sim.scalar("total_volume") / sim.scalar("box_length") ** 3 == pytest.approx(1e-3, rel=1e-12)
