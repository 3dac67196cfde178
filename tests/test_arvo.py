"""ARVO (a12, volsurf_method = arvo): the restated method in the oracle against
closed forms and Monte Carlo integration, and the product's per-ball step
(the code the device kernel runs, here on the host) against the oracle.
CPU only."""
import math

import numpy as np
import pytest

import oracle_lib as ora

PI = math.pi


def vs(r):
    return 4 * PI / 3 * r ** 3, 4 * PI * r ** 2


def lens(r1, r2, d):
    return PI * (r1 + r2 - d) ** 2 * (d * d + 2 * d * (r1 + r2) - 3 * (r1 - r2) ** 2) / (12 * d)


def geometry(rel, r, nk=None):
    r = np.asarray(r, float)
    v, s = vs(r)
    return ora.merge_geometry(3, nk or len(r), np.asarray(rel, float), r, v, s)


def test_single_and_disjoint_balls():
    r = np.array([10e-9, 7e-9, 12e-9])
    g = geometry([[0, 0, 0]], r[:1])
    assert g["volume"] == pytest.approx(vs(r[0])[0], rel=1e-14)
    assert g["surface"] == pytest.approx(vs(r[0])[1], rel=1e-14)
    rel = np.array([[0, 0, 0], [40e-9, 0, 0], [0, 50e-9, 0]])
    g = geometry(rel, r, nk=1)
    v, s = vs(r)
    assert g["volume"] == pytest.approx(v.sum(), rel=1e-14)
    assert g["surface"] == pytest.approx(s.sum(), rel=1e-14)


@pytest.mark.parametrize("seed", range(3))
def test_two_balls_lens(seed):
    rng = np.random.default_rng(seed)
    for _ in range(200):
        r1, r2 = rng.uniform(5e-9, 25e-9, 2)
        d = rng.uniform(abs(r1 - r2) * 1.01 + 1e-12, (r1 + r2) * 0.99)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        g = geometry([[0, 0, 0], d * u], [r1, r2], nk=1)
        (v1, s1), (v2, s2) = vs(r1), vs(r2)
        assert g["volume"] == pytest.approx(v1 + v2 - lens(r1, r2, d), rel=1e-9)
        x1 = (d * d + r1 * r1 - r2 * r2) / (2 * d)
        area = s1 + s2 - 2 * PI * r1 * (r1 - x1) - 2 * PI * r2 * (r2 - (d - x1))
        assert g["surface"] == pytest.approx(area, rel=1e-9)


def test_contained_ball():
    rel = [[0, 0, 0], [3e-9, 2e-9, 0]]
    g = geometry(rel, [20e-9, 5e-9], nk=1)
    v, s = vs(20e-9)
    assert g["volume"] == pytest.approx(v, rel=1e-14) and g["surface"] == pytest.approx(s, rel=1e-14)
    g = geometry(rel[::-1], [5e-9, 20e-9], nk=1)  # the small ball first
    assert g["volume"] == pytest.approx(v, rel=1e-14) and g["surface"] == pytest.approx(s, rel=1e-14)


def union_volume_mc(c, r, n, rng):
    lo, hi = (c - r[:, None]).min(0), (c + r[:, None]).max(0)
    p = rng.uniform(lo, hi, size=(n, 3))
    inside = np.zeros(n, bool)
    for i in range(len(r)):
        inside |= ((p - c[i]) ** 2).sum(1) < r[i] ** 2
    return np.prod(hi - lo) * inside.mean()


def exposed_area_mc(c, r, i, n, rng):
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    p = c[i] + r[i] * u
    covered = np.zeros(n, bool)
    for j in range(len(r)):
        if j != i:
            covered |= ((p - c[j]) ** 2).sum(1) < r[j] ** 2
    return 4 * PI * r[i] ** 2 * (1 - covered.mean())


@pytest.mark.parametrize("case", range(5))
def test_overlapping_balls_monte_carlo(case):
    """Three to five mutually overlapping balls: the union's volume and each
    ball's exposed area against Monte Carlo integration."""
    rng = np.random.default_rng(100 + case)
    n = 3 + case % 3
    r = rng.uniform(8, 14, n)
    c = np.zeros((n, 3))
    for i in range(1, n):  # each ball overlaps the first ones
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        c[i] = c[int(rng.integers(0, i))] + u * rng.uniform(0.5, 0.85) * (r[i] + r[0])
    scale = 1e-9
    g = geometry((c - c[0]) * scale, r * scale)
    mc = union_volume_mc(c, r, 2_000_000, rng) * scale ** 3
    assert g["volume"] == pytest.approx(mc, rel=6e-3)
    # per-ball areas of the oracle's split against Monte Carlo on each surface
    balls = np.column_stack([c, r])
    area = 0.0
    for i in range(n):
        d = np.sqrt(((c - c[i]) ** 2).sum(1))
        nb = [j for j in range(n) if j != i and d[j] < r[i] + r[j] and d[j] + r[j] > r[i] and d[j] + r[i] > r[j]]
        part = ora.arvo_part(balls[[i] + nb])
        want = exposed_area_mc(c, r, i, 400_000, rng)
        assert part[1] == pytest.approx(want, rel=2e-2, abs=4 * PI * r[i] ** 2 * 3e-3)
        area += part[1]
    assert g["surface"] == pytest.approx(area * scale ** 2, rel=1e-12)


def test_north_pole_on_a_neighbour():
    """A neighbour whose surface passes through the north pole is shrunk by
    1e-4 (arvo.cpp:690-707): the result stays within that of the lens."""
    r1, r2 = 10.0, 8.0
    c2 = np.array([0.0, 0.0, r1]) + r2 * np.array([0.6, 0.0, 0.8])  # the pole lies on ball 2's surface
    d = np.linalg.norm(c2)
    assert d < r1 + r2 and abs(np.linalg.norm(c2 - [0, 0, r1]) - r2) < 1e-12
    g = geometry([[0, 0, 0], c2 * 1e-9], [r1 * 1e-9, r2 * 1e-9], nk=1)
    v1, v2 = vs(r1 * 1e-9)[0], vs(r2 * 1e-9)[0]
    assert g["volume"] == pytest.approx(v1 + v2 - lens(r1 * 1e-9, r2 * 1e-9, d * 1e-9), rel=2e-4)


def random_neighbourhood(rng, k):
    """Ball 0 at the origin and k neighbours overlapping it, none inside it or containing it."""
    r0 = rng.uniform(8, 15)
    balls = [[0.0, 0.0, 0.0, r0]]
    while len(balls) < k + 1:
        rj = rng.uniform(4, 15)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        d = rng.uniform(abs(r0 - rj) + 0.5, r0 + rj - 0.5)
        balls.append([*(u * d), rj])
    b = np.array(balls)
    b[:, :3] += rng.uniform(-30, 30, 3)  # ball 0 away from the origin (the z-field split uses z)
    return b


@pytest.mark.parametrize("k", [1, 2, 3, 5, 8, 12, 16])
def test_product_ball_step_matches_oracle(k):
    """The product's per-ball step (host build of the device code) against the
    oracle's restatement, within 1e-10 relative of the ball's own volume and area."""
    mcac = pytest.importorskip("mcac_amd")
    rng = np.random.default_rng(k)
    for _ in range(300):
        b = random_neighbourhood(rng, k)
        v0, a0 = ora.arvo_part(b)
        v1, a1 = mcac.arvo_part(b)
        full_v, full_a = vs(b[0, 3])
        assert abs(v1 - v0) <= 1e-10 * max(full_v, abs(v0)), (k, b.tolist(), v0, v1)
        assert abs(a1 - a0) <= 1e-10 * full_a, (k, b.tolist(), a0, a1)


def test_arvo_run_oracle():
    """A dense box with volsurf_method = arvo runs to its stop criterion (the
    contacts only touch, so every part is the ball's own volume and area)."""
    import pathlib
    text = (pathlib.Path(__file__).parent / "configs" / "dense_small_arvo.ini").read_text()
    sim = ora.OracleSim(text, fast_init=True)
    sim.run(10 ** 7)
    assert sim.finished()
