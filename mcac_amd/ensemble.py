"""Ensemble and serving workloads over many independent boxes on one GPU.

Batched form: 512 boxes, N log-uniform in [500, 5e4], normal diameters (mean
30 nm, sd 3 nm), FV uniform in [1e-4, 5e-3], T uniform in [1000, 2000] K, box b
seeded s0 + b, each run until N_agg <= N/10.  Serving form: requests with N
log-uniform in [1e3, 2e5], FV in [1e-4, 5e-3], T in [1200, 2000] K arriving as
a Poisson process; time-to-result = completion - arrival.

Parameter files use the reference's INI keys (src/physical_model/physical_model.cpp).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _check, _dp, _ip, load_sim


def box_ini(n: int, fv: float, temperature: float, seed: int, mean_d: float = 30.0, sd_d: float = 3.0,
            mode: str = "normal", stop_fraction: float = 0.1, growth_flux: float = 0.0) -> str:
    """One parameter file; the run stops at N_agg <= stop_fraction * N.  With
    growth_flux > 0 the spheres grow by surface reactions at that flux
    (kg/m^2/s; [surface_growth] flux_surfgrowth, u_sg = flux / density)."""
    growth = (f"""
[surface_growth]
with_surface_reactions = true
flux_surfgrowth = {growth_flux!r}
""" if growth_flux > 0 else "")
    return f"""[monomers]
number = {n}
density = 1800
mean_diameter = {mean_d}
dispersion_diameter = {sd_d}
initialisation_mode = {mode}

[environment]
volume_fraction = {fv!r}
temperature = {temperature!r}
pressure = 101300
fractal_prefactor = 1.30
fractal_dimension = 1.78

[limits]
number_of_aggregates = {max(1, int(n * stop_fraction))}

[numerics]
random_seed = {seed}
with_domain_duplication = false
n_verlet_divisions = 10
{growth}"""


def ensemble_configs(n_boxes: int = 512, s0: int = 1000, rng_seed: int = 4) -> list[dict]:
    """The batched form's box list (the first n_boxes of the 512-box list)."""
    rng = np.random.default_rng(rng_seed)
    out = []
    for b in range(512):
        n = int(round(np.exp(rng.uniform(np.log(500), np.log(5e4)))))
        fv = float(rng.uniform(1e-4, 5e-3))
        t = float(rng.uniform(1000, 2000))
        out.append(dict(n=n, fv=fv, temperature=t, seed=s0 + b, ini=box_ini(n, fv, t, s0 + b)))
    return out[:n_boxes]


def serving_requests(n_requests: int, rate_per_s: float, s0: int = 5000, rng_seed: int = 5,
                     n_max: float = 2e5, growth_flux: float = 1e-3, growth_n_max: float = 2e5) -> list[dict]:
    """Serving form: Poisson arrivals at rate_per_s (first arrival at 0); every
    other request (the odd ones) with surface growth at growth_flux, its N
    drawn up to growth_n_max (growth_flux 0: none with growth)."""
    rng = np.random.default_rng(rng_seed)
    t = 0.0
    out = []
    for r in range(n_requests):
        grow = growth_flux > 0 and r % 2 == 1
        u = rng.uniform(0.0, 1.0)
        hi = min(n_max, growth_n_max) if grow else n_max
        n = int(round(np.exp(np.log(1e3) + u * (np.log(hi) - np.log(1e3)))))
        fv = float(rng.uniform(1e-4, 5e-3))
        temp = float(rng.uniform(1200, 2000))
        out.append(dict(n=n, fv=fv, temperature=temp, seed=s0 + r, arrival=t, growth=grow,
                        ini=box_ini(n, fv, temp, s0 + r, mode="lognormal", sd_d=1.25,
                                    growth_flux=growth_flux if grow else 0.0)))
        t += float(rng.exponential(1.0 / rate_per_s))
    return out


def device_bytes(ini: str) -> int:
    """Device memory the box of a parameter file takes (mcac_sim_device_bytes)."""
    return int(load_sim().mcac_sim_device_bytes(ini.encode()))


def batch_run(inis: list[str], arrivals=None, workers: int = 8, device: int = 0, max_steps: int = 2 ** 62,
              boxes_per_worker: int = 0, depth: int = 8, hbm_budget: int = 0) -> dict:
    """Runs the jobs on `workers` host threads sharing the GPU (include/mcac_sim.h):
    one box per worker at a time (mcac_batch_run), or with boxes_per_worker >= 1
    up to that many boxes per worker stepped in lock step, up to `depth` steps
    of each box per launch (mcac_batch_run_groups).  With hbm_budget > 0 (bytes)
    a job starts only when its box fits in what the running boxes leave of it;
    out["peak_bytes"] is the most the admitted boxes held at once."""
    n = len(inis)
    texts = (ctypes.c_char_p * n)(*[s.encode() for s in inis])
    arr = None if arrivals is None else np.ascontiguousarray(arrivals, np.float64)
    start, end = np.zeros(n), np.zeros(n)
    steps = np.zeros(n, np.int64)
    status = np.zeros(n, np.int32)
    st = status.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    L = load_sim()
    peak = np.zeros(1, np.int64)
    out = dict(start=start, end=end, steps=steps, status=status)
    if boxes_per_worker >= 1:
        final_time = np.zeros(n)
        rc = L.mcac_batch_run_groups(texts, None if arr is None else _dp(arr), n, workers, boxes_per_worker, depth,
                                     device, max_steps, int(hbm_budget), _dp(start), _dp(end), _ip(steps), st,
                                     _dp(final_time), _ip(peak))
        _check(rc, "mcac_batch_run_groups")
        out["final_time"] = final_time
    else:
        rc = L.mcac_batch_run(texts, None if arr is None else _dp(arr), n, workers, device, max_steps,
                              int(hbm_budget), _dp(start), _dp(end), _ip(steps), st, _ip(peak))
        _check(rc, "mcac_batch_run")
    out["peak_bytes"] = int(peak[0])
    return out
