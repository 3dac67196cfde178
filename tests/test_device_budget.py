"""Device-memory sizing of boxes (mcac_box_device_bytes, mcac_sim_device_bytes)
and admission of ensemble / serving jobs by an HBM budget.  The sizing is a
host computation (CPU tests); admission and the real allocation run on the GPU."""
import numpy as np
import pytest

import mcac_amd
from mcac_amd import ensemble


def box_bytes(ns, na, n_div=10, volsurf=0):
    p = np.zeros(1, mcac_amd.Box.PARAMS)
    p["box_length"], p["n_div"], p["max_spheres"], p["max_aggregates"], p["volsurf"] = 1e-6, n_div, ns, na, volsurf
    return mcac_amd.load_gpu().mcac_box_device_bytes(p.ctypes.data)


def test_box_bytes_per_sphere_and_method():
    # the overlap corrections add their per-sphere rows: caps and arvo a distance
    # row plus corrected volume and surface, alphas a distance row, r^3, the walk rows and counts
    base = box_bytes(10_000, 10_000)
    assert base > 0
    assert box_bytes(10_000, 10_000, volsurf=1) - base == 10_000 * (16 * 8 + 8 + 8)
    assert box_bytes(10_000, 10_000, volsurf=3) - base == 10_000 * (16 * 8 + 8 + 8)
    assert box_bytes(10_000, 10_000, volsurf=2) - base == 10_000 * (16 * 8 + 8 + 3 * 16 * 8 + 4)
    # growing the sphere capacity alone adds the sphere arrays and the member pool (4 ints per sphere)
    per_sphere = box_bytes(10_001, 10_000) - base
    assert per_sphere == 9 * 8 + 8 + 4 + 16 * 4 + 16 * 8 + 4 * 4
    assert box_bytes(0, 10) == -1 and box_bytes(10, 10, volsurf=4) == -1


def test_sim_bytes_follow_the_parameter_file():
    small = ensemble.box_ini(1000, 1e-3, 1500.0, 1)
    big = ensemble.box_ini(50_000, 1e-3, 1500.0, 1)
    b_small, b_big = ensemble.device_bytes(small), ensemble.device_bytes(big)
    assert 0 < b_small < b_big
    assert b_big == box_bytes(50_000, 50_000)
    assert ensemble.device_bytes("[numerics]\nrandom_seed = -1\n") == -1  # unsupported (clock-seeded) run


def test_sim_bytes_cover_the_duplications_the_mean_size_limit_allows():
    # k-th duplication needs at least 8^k monomers per aggregate: a limit of
    # 60 allows one (8 < 60 <= 64), 300 two, 8 none
    base = ensemble.box_ini(1000, 1e-3, 1500.0, 1)
    def dup(npp):
        return (base.replace("with_domain_duplication = false", "with_domain_duplication = true")
                .replace("number_of_aggregates = 100", f"number_of_aggregates = 1\nmean_monomere_per_aggregate = {npp}"))
    assert "with_domain_duplication = true" in dup(60) and "mean_monomere_per_aggregate = 60" in dup(60)
    assert ensemble.device_bytes(dup(8)) == box_bytes(1000, 1000)
    assert ensemble.device_bytes(dup(60)) == box_bytes(8000, 1000 + 7 * 125)
    assert ensemble.device_bytes(dup(300)) == box_bytes(64000, 1000 + 14 * 125)


@pytest.mark.gpu
def test_box_allocation_matches_the_estimate():
    n = 400_000
    warm = mcac_amd.Box(1e-6, 10, 1000, 1000)  # runtime and module set-up are not the box's
    free0, _ = mcac_amd.device_memory()
    box = mcac_amd.Box(1e-6, 10, n, n)
    free1, _ = mcac_amd.device_memory()
    est = box_bytes(n, n)
    used = free0 - free1
    print(f"case: box of {n} spheres: estimate {est} B, measured {used} B", flush=True)
    assert 0.9 * est <= used <= 1.1 * est + (64 << 20)  # allocation granularity
    box.close()
    warm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("groups", [0, 2])
def test_budget_admits_one_box_at_a_time(groups):
    """With a budget below any two boxes, jobs run one after the other (their
    run intervals do not overlap) and each run is the one it is without a budget."""
    print(f"case: HBM budget admission, boxes per worker {groups}", flush=True)
    inis = [ensemble.box_ini(n, 2e-3, 1500.0, 60 + n) for n in (3000, 3200, 3400, 3600)]
    sizes = [ensemble.device_bytes(s) for s in inis]
    budget = int(1.5 * max(sizes))
    assert 2 * min(sizes) > budget
    free = ensemble.batch_run(inis, workers=4, boxes_per_worker=groups, depth=8)
    capped = ensemble.batch_run(inis, workers=4, boxes_per_worker=groups, depth=8, hbm_budget=budget)
    assert np.array_equal(free["steps"], capped["steps"])
    if groups:
        assert np.array_equal(free["final_time"], capped["final_time"])
    iv = sorted(zip(capped["start"], capped["end"]))
    assert all(iv[k][1] <= iv[k + 1][0] for k in range(len(iv) - 1)), iv
    assert max(sizes) <= capped["peak_bytes"] <= budget
    assert free["peak_bytes"] == 0
