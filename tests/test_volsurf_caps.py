"""Overlap corrections of the merge update (volsurf_method = caps,
aggregat.cpp:354-386, and alphas, :391-426): the oracle against closed forms,
and the order in which the device walks a sphere's neighbour map against
libstdc++'s own std::unordered_map (CPU only)."""
import math

import numpy as np
import pytest

import oracle_lib as ora

PI = math.pi


def lens_volume(r1, r2, d):
    """Volume of the intersection of two balls (textbook lens formula)."""
    return PI * (r1 + r2 - d) ** 2 * (d * d + 2 * d * (r1 + r2) - 3 * (r1 - r2) ** 2) / (12 * d)


def sphere_vs(r):
    return 4 * PI / 3 * r ** 3, 4 * PI * r ** 2


def test_pi_is_the_reference_constant():
    # constants.hpp:113 defines pi as atan(1.0) * 4; the device uses the literal
    assert math.atan(1.0) * 4 == 3.141592653589793


@pytest.mark.parametrize("seed", range(4))
def test_caps_intersection_lens_closed_form(seed):
    rng = np.random.default_rng(seed)
    for _ in range(500):
        r1, r2 = rng.uniform(5e-9, 30e-9, 2)
        d = rng.uniform(abs(r1 - r2) * 1.001, (r1 + r2) * 0.999)
        v1, s1, v2, s2 = ora.caps_intersection(r1, *sphere_vs(r1), r2, *sphere_vs(r2), d)
        assert v1 + v2 == pytest.approx(lens_volume(r1, r2, d), rel=1e-9)
        x1 = (d * d + r1 * r1 - r2 * r2) / (2 * d)  # plane of the intersection circle from centre 1
        assert s1 == pytest.approx(2 * PI * r1 * (r1 - x1), rel=1e-9)
        assert s2 == pytest.approx(2 * PI * r2 * (r2 - (d - x1)), rel=1e-9)
        assert v1 >= 0 and v2 >= 0


def test_caps_intersection_special_cases():
    r1, r2 = 20e-9, 8e-9
    (V1, S1), (V2, S2) = sphere_vs(r1), sphere_vs(r2)
    assert ora.caps_intersection(r1, V1, S1, r2, V2, S2, r1 + r2) == (0.0, 0.0, 0.0, 0.0)  # touching
    assert ora.caps_intersection(r1, V1, S1, r2, V2, S2, 2 * (r1 + r2)) == (0.0, 0.0, 0.0, 0.0)
    assert ora.caps_intersection(r1, V1, S1, r2, V2, S2, 0.0) == (0.0, 0.0, 0.0, 0.0)  # dist <= 0
    # sphere 2 inside sphere 1: the whole of sphere 2 is covered
    assert ora.caps_intersection(r1, V1, S1, r2, V2, S2, 0.5 * (r1 - r2)) == (0.0, 0.0, V2, S2)


def test_two_overlapping_spheres_closed_forms():
    r, d = 15e-9, 24e-9
    v, s = sphere_vs(r)
    rel = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0]])
    g = ora.merge_geometry(1, 1, rel, [r, r], [v, v], [s, s])
    lens = lens_volume(r, r, d)
    h = r - d / 2
    assert g["volume"] == pytest.approx(2 * v - lens, rel=1e-12)
    assert g["surface"] == pytest.approx(2 * s - 2 * (2 * PI * r * h), rel=1e-12)
    assert g["rcm"] == pytest.approx([d / 2, 0, 0], rel=1e-12)
    assert g["rg"] == pytest.approx(math.sqrt(d * d / 4 + 0.6 * r * r), rel=1e-12)
    assert g["rmax"] == pytest.approx(r + d / 2, rel=1e-12)
    assert g["sphere_volume"] == 0.0 + v + v
    assert g["coordination"] == 1.0
    assert g["overlapping"] == pytest.approx((2 * r - d) / (2 * r), rel=1e-12)
    # method none: the sphere sums
    g0 = ora.merge_geometry(0, 1, rel, [r, r], [v, v], [s, s])
    assert g0["volume"] == 0.0 + v + v and g0["surface"] == 0.0 + s + s
    assert g0["rg"] == pytest.approx(math.sqrt(d * d / 4 + 0.6 * r * r), rel=1e-12)


def test_alphas_two_spheres_match_exact_caps():
    """For two equal spheres the alphas fits reduce to the exact cap fractions:
    volume 1 - c²(3 - c)/4 and surface 1 - c/2 of the sphere sums, c the overlap
    coefficient (aggregat.cpp:289-319 with coordination 1 = its minimum)."""
    for r, d in ((15e-9, 24e-9), (10e-9, 19.5e-9), (20e-9, 30e-9)):
        v, s = sphere_vs(r)
        rel = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0]])
        ga = ora.merge_geometry(2, 1, rel, [r, r], [v, v], [s, s])
        gc = ora.merge_geometry(1, 1, rel, [r, r], [v, v], [s, s])
        c = (2 * r - d) / (2 * r)
        assert ga["volume"] == pytest.approx(2 * v * (1 - c * c * (3 - c) / 4), rel=1e-12)
        assert ga["surface"] == pytest.approx(2 * s * (1 - c / 2), rel=1e-12)
        assert ga["volume"] == pytest.approx(gc["volume"], rel=1e-12)
        assert ga["surface"] == pytest.approx(gc["surface"], rel=1e-12)
        assert ga["coordination"] == 1.0 and ga["overlapping"] == pytest.approx(c, rel=1e-12)
        # update_partial weights by the spheres' own volumes but divides by the
        # corrected aggregate volume (aggregat.cpp:431-483 after :417-421)
        assert ga["rcm"] == pytest.approx([d / 2 * (2 * v) / ga["volume"], 0, 0], rel=1e-12)
        rc = ga["rcm"][0]
        rg2 = (v * rc * rc + v * (d - rc) ** 2 + 0.6 * 2 * v * r * r) / ga["volume"]
        assert ga["rg"] == pytest.approx(math.sqrt(rg2), rel=1e-9)


def test_alphas_without_contact_keeps_sphere_values():
    # no contact: no correction, overlap and coordination stay 0 (aggregat.cpp:393, 415)
    r = 10e-9
    v, s = sphere_vs(r)
    g = ora.merge_geometry(2, 1, np.zeros((1, 3)), [r], [v], [s])
    assert g["volume"] == 0.0 + v and g["surface"] == 0.0 + s
    assert g["coordination"] == 0.0 and g["overlapping"] == 0.0


def test_chain_of_three_closed_form():
    # 1-2 and 2-3 overlap, 1-3 apart: the middle sphere loses two caps
    r, d = 10e-9, 16e-9
    v, s = sphere_vs(r)
    rel = np.array([[0.0, 0, 0], [d, 0, 0], [2 * d, 0, 0]])
    g = ora.merge_geometry(1, 2, rel, [r] * 3, [v] * 3, [s] * 3)
    assert g["volume"] == pytest.approx(3 * v - 2 * lens_volume(r, r, d), rel=1e-12)
    assert g["surface"] == pytest.approx(3 * s - 4 * (2 * PI * r * (r - d / 2)), rel=1e-12)
    assert g["rcm"] == pytest.approx([d, 0, 0], rel=1e-12)


def test_contained_sphere_keeps_the_large_one():
    r1, r2 = 20e-9, 5e-9
    (V1, S1), (V2, S2) = sphere_vs(r1), sphere_vs(r2)
    rel = np.array([[0.0, 0, 0], [3e-9, 0, 0]])
    g = ora.merge_geometry(1, 1, rel, [r1, r2], [V1, V2], [S1, S2])
    assert g["volume"] == 0.0 + V1 + 0.0 and g["surface"] == 0.0 + S1 + 0.0


def _random_keys(rng, m, hi):
    return np.sort(rng.choice(hi, size=m, replace=False)).astype(np.int64)


def test_neighbour_order_matches_libstdcxx():
    """The device's replay of the map order equals a real std::unordered_map's
    (same sequence of clear / reserve / inserts as aggregat.cpp:730-751)."""
    mcac = pytest.importorskip("mcac_amd")
    rng = np.random.default_rng(11)
    differs_from_sorted = 0
    for case in range(4000):
        prev = int(rng.integers(0, 24))
        m = int(rng.integers(0, 33))
        keys = _random_keys(rng, m, int(rng.integers(max(m, 1), 3000)))
        want = ora.neighbour_order(prev, keys)
        got = mcac.neighbour_order(prev, keys)
        assert np.array_equal(got, want), (case, prev, keys.tolist(), got.tolist(), want.tolist())
        differs_from_sorted += int(not np.array_equal(want, keys))
    assert differs_from_sorted > 1000  # the order is not the sorted one in general


def test_neighbour_order_argument_checks():
    mcac = pytest.importorskip("mcac_amd")
    with pytest.raises(mcac.MCACError):
        mcac.neighbour_order(0, [3, 1])
    with pytest.raises(mcac.MCACError):
        mcac.neighbour_order(0, list(range(33)))


def overlapping_cluster(rng, n, r_mean=12e-9, overlap=(0.05, 0.35), origin=(0.0, 0.0, 0.0)):
    """n spheres, each attached to a random earlier one with a random overlap;
    positions relative to sphere 0, plus the absolute origin of sphere 0."""
    r = rng.normal(r_mean, 0.2 * r_mean, n).clip(0.4 * r_mean, 2 * r_mean)
    rel = np.zeros((n, 3))
    for i in range(1, n):
        p = int(rng.integers(0, i))
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        rel[i] = rel[p] + u * (r[i] + r[p]) * (1 - rng.uniform(*overlap))
    return rel, r


def test_caps_walk_order_changes_bits():
    """Random overlapping clusters: walking the higher neighbours in the map's
    order and in ascending order give different volumes in some clusters, so
    the bit-exact device comparison checks the order, not only the formulas."""
    rng = np.random.default_rng(5)
    n_diff = 0
    for _ in range(200):
        n = int(rng.integers(6, 40))
        rel, r = overlapping_cluster(rng, n)
        v, s = sphere_vs(r)
        g = ora.merge_geometry(1, n // 2, rel, r, v, s)
        # ascending walk, same formulas
        vols = v.copy()
        nb = [[] for _ in range(n)]
        for i in range(n):
            for j in range(i + 1, n):
                dist = math.sqrt(sum((rel[i, k] - rel[j, k]) * (rel[i, k] - rel[j, k]) for k in range(3)))
                if dist <= (1.0 + 1e-10) * (r[i] + r[j]):
                    nb[i].append((j, dist))
        for i in range(n):
            for j, dist in nb[i]:
                v1, _, v2, _ = ora.caps_intersection(r[i], v[i], s[i], r[j], v[j], s[j], dist)
                vols[i] = vols[i] - v1
                vols[j] = vols[j] - v2
            vols[i] = max(vols[i], 0.0)
        V = 0.0
        for x in vols:
            V = V + x
        assert V == pytest.approx(g["volume"], rel=1e-12)
        n_diff += int(V != g["volume"])
    assert n_diff > 0


@pytest.mark.parametrize("method", ["caps", "alphas"])
def test_oracle_run_with_overlap_correction(method):
    """A dense box with an overlap correction runs to its stop criterion; every
    aggregate's volume is at most the sum of its spheres' volumes."""
    import pathlib
    text = (pathlib.Path(__file__).parent / "configs" / f"dense_small_{method}.ini").read_text()
    sim = ora.OracleSim(text, fast_init=True)
    sim.run(10 ** 7)
    assert sim.finished()
    sph, agg = sim.export()
    for l in range(len(agg["n"])):
        m = agg["members"][agg["offset"][l]:agg["offset"][l] + agg["n"][l]]
        assert agg["volume"][l] <= sph["vol"][m].sum() * (1 + 1e-12)
