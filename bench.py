"""Benchmark of the MI355X MCAC collision path: MC steps/s at N = 1,000,000 spheres.

Workload: N=1e6 lognormal monomers (geometric mean 30 nm, sigma_g 1.25),
FV=3e-3, T=1700 K, P=101300 Pa, FP64 throughout, the reference MC loop from
t=0 (host pick + device contact search / translate / merge).  One "step" is
one iteration of the reference loop (src/calcul.cpp:66-281).  The box state is
resident in HBM before the timed region; W warm-up steps run first.

Prints one JSON line.  Also reports the dominant kernel's achieved HBM
bandwidth (algorithmic bytes / average launch duration, HIP events on the box
stream) and the CPU path (the oracle, 1 thread) timed on a bounded sample of
the same run.
"""
import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = ROOT / "profiles" / "k_step_n1m_hbm_traffic.json"
WORKLOAD_INI = ROOT / "mcac_amd" / "configs" / "n1m_lognormal_fv3e-3.ini"
METRIC = "MC steps/s at N=1M; ensemble runs/hour + p99 latency vs plain GPU & CPU"
MIN_TIMED_STEPS = 20000   # floor of the timed region (about 0.4 s of MC steps at N=1e6)
MIN_WARMUP_STEPS = 2000


def cpu_baseline(text: str, steps: int):
    """Oracle (reference algorithm, g++ -O2, one thread) on the first `steps` MC steps."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib
    t0 = time.perf_counter()
    sim = oracle_lib.OracleSim(text, fast_init=True, trace=False)
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    n = sim.run(steps)
    dt = time.perf_counter() - t0
    return dict(value=n / dt, unit="MC steps/s", cores=1, kind="port",
                sample=f"first {n} MC steps from t=0 of the same N=1e6 run (oracle, 1 thread; "
                       f"placement {t_init:.1f} s not timed); {dt:.1f} s of CPU work")


def c3_windows(args) -> int:
    """The N=1e6 box run on for a wall-time budget in 20,000-step chunks; one
    JSON line per chunk on stderr, the summary on stdout (steps/s by N_agg)."""
    import mcac_amd
    sim = mcac_amd.Simulation(WORKLOAD_INI.read_text(), device=0)
    t_end = time.perf_counter() + args.c3_windows
    rows = []
    while time.perf_counter() < t_end and not sim.finished():
        n0 = sim.scalar("n_agg")
        t0 = time.perf_counter()
        done = sim.run(20000)
        dt = time.perf_counter() - t0
        rows.append({"n_agg_start": n0, "n_agg_end": sim.scalar("n_agg"), "steps": done, "seconds": dt,
                     "steps_per_s": done / dt, "time": sim.scalar("time"), "merges": sim.scalar("merges")})
        print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
        if done == 0:
            break
    print(json.dumps({"c3_windows": rows}), flush=True)
    return 0


OVERLAP_INI = """[monomers]
number = {n}
density = 1800
mean_diameter = 30
dispersion_diameter = 1.25
initialisation_mode = lognormal
[environment]
volume_fraction = {fv}
temperature = 1700
pressure = 101300
fractal_prefactor = 1.30
fractal_dimension = 1.78
[limits]
number_of_aggregates = {stop}
[numerics]
random_seed = 11
with_domain_duplication = false
[surface_growth]
volsurf_method = {method}
"""


def overlap_runs() -> int:
    """Whole runs with each volsurf_method, overlap-correction merges in one
    workgroup (default) and as the kernel sequence (MCAC_FUSED_MERGE=0, read
    at box creation); both forms must give the same steps, time and merges."""
    import mcac_amd
    rows = []
    for n, fv, stop in ((20000, 0.01, 2000), (3000, 0.05, 60)):
        for method in ("none", "caps", "alphas", "arvo"):
            ref = None
            for fused in ((1,) if method == "none" else (1, 0)):
                os.environ["MCAC_FUSED_MERGE"] = str(fused)
                sim = mcac_amd.Simulation(OVERLAP_INI.format(n=n, fv=fv, stop=stop, method=method), device=0)
                t0 = time.perf_counter()
                steps = sim.run()
                wall = time.perf_counter() - t0
                key = (steps, sim.scalar("time"), sim.scalar("merges"))
                ref = ref or key
                rows.append(dict(n=n, fv=fv, stop=stop, method=method, fused=fused, wall_s=wall, steps=steps,
                                 merges=key[2], steps_per_s=steps / wall, same_run=key == ref))
                print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
                sim.close()
                if key != ref:
                    print("bench: the two merge forms gave different runs", file=sys.stderr)
                    return 1
    os.environ.pop("MCAC_FUSED_MERGE", None)
    print(json.dumps({"overlap_runs": rows}), flush=True)
    return 0


def _progress_every(k: int):
    """The batch runner prints a line to stderr every k-th finished job."""
    os.environ["MCAC_BATCH_PROGRESS"] = str(k)


def _pct(x, q):
    import numpy as np
    return float(np.percentile(np.asarray(x, float), q)) if len(x) else None


def run_ensemble(args) -> dict:
    """Batched form on the 512-box list (the first --ensemble-boxes of it): every
    box run to N_agg <= N/10 from t=0, --workers host threads each stepping
    --group-boxes boxes in lock step.  The plain GPU path one box at a time is
    measured on every --plain-stride-th box of the list and scaled to the whole
    list by MC steps (its MC steps per second of service, times the list's MC
    steps); the sampled boxes must take the same step counts in both forms."""
    import numpy as np
    from mcac_amd import ensemble
    boxes = ensemble.ensemble_configs(args.ensemble_boxes)
    print(f"bench: ensemble of {len(boxes)} boxes, {args.workers} workers x {args.group_boxes} boxes",
          file=sys.stderr, flush=True)
    _progress_every(32)
    t0 = time.perf_counter()
    r = ensemble.batch_run([b["ini"] for b in boxes], workers=args.workers, max_steps=args.max_job_steps,
                           boxes_per_worker=args.group_boxes, depth=args.group_depth)
    wall = time.perf_counter() - t0
    done = [j for j in range(len(boxes)) if r["steps"][j] < args.max_job_steps]
    dur = (r["end"] - r["start"])[done]
    out = {"boxes": len(boxes), "of": 512, "completed": len(done), "workers": args.workers,
           "boxes_per_worker": args.group_boxes, "steps_per_launch": args.group_depth, "wall_s": wall,
           "runs_per_hour": 3600.0 * len(done) / wall, "mc_steps": int(r["steps"].sum()),
           "mc_steps_per_s": float(r["steps"].sum()) / wall,
           "box_time_p50_s": _pct(dur, 50), "box_time_p99_s": _pct(dur, 99)}
    if args.plain_ensemble:
        idx = list(range(0, len(boxes), max(1, args.plain_stride)))
        print(f"bench: {len(idx)} of the boxes one at a time (plain path)", file=sys.stderr, flush=True)
        _progress_every(1)
        t0 = time.perf_counter()
        p = ensemble.batch_run([boxes[i]["ini"] for i in idx], workers=1, max_steps=args.max_job_steps)
        pw = time.perf_counter() - t0
        service = float((p["end"] - p["start"]).sum())
        rate = float(p["steps"].sum()) / service
        est = float(r["steps"].sum()) / rate
        out["plain_one_at_a_time"] = {
            "sample_boxes": len(idx), "sample_mc_steps": int(p["steps"].sum()), "sample_service_s": service,
            "sample_wall_s": pw, "mc_steps_per_s": rate, "estimated_total_s": est,
            "runs_per_hour": 3600.0 * len(done) / est,
            "same_step_counts": bool(np.array_equal(p["steps"], r["steps"][idx])),
            "method": f"every {args.plain_stride}th box of the list run one at a time on the single-box "
                      "pipeline; the list's MC steps divided by the sample's MC steps per second of service"}
        out["vs_plain_one_at_a_time"] = est / wall
    return out | {
        "sample": f"boxes 0..{len(boxes) - 1} of the 512-box list (N log-uniform 500..5e4, normal 30/3 nm, "
                  f"FV U[1e-4,5e-3], T U[1000,2000] K, until N_agg <= N/10), all started at t=0"}


def run_ensemble_compare(args) -> dict:
    """Runs/hour of the same boxes on the GPU (batch runner) and on the CPU
    (the oracle, one thread per run, P = min(cores, 16, jobs) runs at a time):
    the boxes of the ensemble sample with N <= --compare-nmax, so that the CPU
    side stays a bounded sample."""
    from concurrent.futures import ThreadPoolExecutor
    from mcac_amd import ensemble
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib
    boxes = [b for b in ensemble.ensemble_configs(args.ensemble_boxes)[::max(1, args.compare_stride)]
             if b["n"] <= args.compare_nmax]
    inis = [b["ini"] for b in boxes]
    print(f"bench: ensemble comparison on {len(boxes)} boxes (N <= {args.compare_nmax:g})", file=sys.stderr,
          flush=True)
    _progress_every(4)
    t0 = time.perf_counter()
    r = ensemble.batch_run(inis, workers=args.workers, max_steps=args.max_job_steps, boxes_per_worker=args.group_boxes,
                           depth=args.group_depth)
    gpu_wall = time.perf_counter() - t0

    def one(ini):  # ctypes drops the GIL: the oracle runs proceed in parallel
        sim = oracle_lib.OracleSim(ini, fast_init=True, trace=False)
        n = sim.run(args.max_job_steps)
        return n, sim.finished()

    procs = max(1, min(len(os.sched_getaffinity(0)), 16, len(inis)))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(procs) as ex:
        cpu = list(ex.map(one, inis))
    cpu_wall = time.perf_counter() - t0
    same = all(int(c[0]) == int(g) for c, g in zip(cpu, r["steps"]))
    return {"boxes": len(boxes), "gpu_runs_per_hour": 3600.0 * len(boxes) / gpu_wall, "gpu_wall_s": gpu_wall,
            "gpu_workers": args.workers, "gpu_boxes_per_worker": args.group_boxes, "cpu_runs_per_hour": 3600.0 * len(boxes) / cpu_wall, "cpu_wall_s": cpu_wall,
            "cpu_threads": procs, "ratio": cpu_wall / gpu_wall, "same_step_counts": same,
            "sample": f"boxes with N <= {args.compare_nmax:g} among every {args.compare_stride}th of boxes "
                      f"0..{args.ensemble_boxes - 1} of the 512-box list, run to N_agg <= N/10 on both sides; "
                      f"CPU = the oracle, one thread per run"}


def run_serving(args) -> dict:
    """Serving form: Poisson arrivals, time-to-result percentiles."""
    import numpy as np
    from mcac_amd import ensemble
    reqs = ensemble.serving_requests(args.serving_requests, args.serving_rate, n_max=args.serving_nmax,
                                     growth_flux=args.serving_growth_flux, growth_n_max=args.serving_growth_nmax)
    print(f"bench: serving {len(reqs)} requests at {args.serving_rate}/s", file=sys.stderr, flush=True)
    _progress_every(4)
    t0 = time.perf_counter()
    budget = int(args.serving_hbm_mib) << 20
    r = ensemble.batch_run([q["ini"] for q in reqs], arrivals=[q["arrival"] for q in reqs], workers=args.workers,
                           max_steps=args.max_job_steps, boxes_per_worker=args.group_boxes, depth=args.group_depth,
                           hbm_budget=budget)
    wall = time.perf_counter() - t0
    arrival = np.array([q["arrival"] for q in reqs])
    ttr = r["end"] - arrival
    out = {"requests": len(reqs), "rate_per_s": args.serving_rate, "workers": args.workers,
           "boxes_per_worker": args.group_boxes, "steps_per_launch": args.group_depth, "wall_s": wall,
           "time_to_result_p50_s": _pct(ttr, 50), "time_to_result_p99_s": _pct(ttr, 99),
           "runs_per_hour": 3600.0 * len(reqs) / wall, "hbm_budget_bytes": budget,
           "hbm_peak_bytes": r["peak_bytes"]}
    if args.plain_ensemble:
        # the plain GPU path as a queue: each request's own run time (measured one at a time),
        # served first come first served on the same arrival times
        t1 = time.perf_counter()
        p = ensemble.batch_run([q["ini"] for q in reqs], workers=1, max_steps=args.max_job_steps)
        pw = time.perf_counter() - t1
        service = p["end"] - p["start"]
        finish, free = np.zeros(len(reqs)), 0.0
        for j in np.argsort(arrival, kind="stable"):
            free = max(free, arrival[j]) + service[j]
            finish[j] = free
        pt = finish - arrival
        out["plain_queue"] = {"service_s_total": float(service.sum()), "wall_s": pw,
                              "time_to_result_p50_s": _pct(pt, 50), "time_to_result_p99_s": _pct(pt, 99),
                              "runs_per_hour": 3600.0 * len(reqs) / max(float(finish.max()), 1e-9),
                              "same_step_counts": bool(np.array_equal(p["steps"], r["steps"])),
                              "note": "first-come-first-served over the measured one-at-a-time run times"}
    return out | {
        "sample": f"N log-uniform 1e3..{args.serving_nmax:g} (full form 2e5; odd requests with surface growth at "
                  f"flux {args.serving_growth_flux:g}, N up to {args.serving_growth_nmax:g}), lognormal 30/1.25, "
                  f"FV U[1e-4,5e-3], "
                  f"T U[1200,2000] K, until N_agg <= N/10"}


LEGS = {"ensemble": run_ensemble, "serving": run_serving, "compare": run_ensemble_compare}


def run_leg(name: str) -> dict:
    """One of the batch legs in a process of its own (this command with --leg):
    each starts from a fresh HIP runtime, as a separate ensemble or serving job
    would; run in the process that has just run the N=1e6 box, the ensemble leg
    took 80-87 s instead of 63 s."""
    import subprocess
    cmd = [sys.executable, "-u", str(pathlib.Path(__file__).resolve())] + sys.argv[1:] + ["--leg", name]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"bench leg {name} ended with status {p.returncode}")
    return json.loads(p.stdout.strip().splitlines()[-1])


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy 20 times (HBM counter calibration)")
    ap.add_argument("--ensemble-boxes", type=int, default=512,
                    help="boxes of the 512-box ensemble list run to completion (0: skip)")
    ap.add_argument("--plain-stride", type=int, default=32,
                    help="every n-th box of the ensemble list is also run one at a time on the plain path")
    ap.add_argument("--compare-stride", type=int, default=4,
                    help="every n-th box of the ensemble list (with N <= --compare-nmax) is run on the CPU too")
    ap.add_argument("--no-plain-ensemble", dest="plain_ensemble", action="store_false",
                    help="skip running the ensemble sample again one box at a time")
    ap.add_argument("--compare-nmax", type=float, default=8000,
                    help="largest box of the ensemble sample also run on the CPU for the runs/hour comparison")
    ap.add_argument("--serving-requests", type=int, default=16, help="serving requests (0: skip)")
    ap.add_argument("--serving-rate", type=float, default=4.0, help="serving arrival rate, requests/s")
    ap.add_argument("--serving-hbm-mib", type=int, default=4096,
                    help="device memory the admitted serving boxes may hold at once (0: no limit)")
    ap.add_argument("--serving-growth-flux", type=float, default=1e-3,
                    help="surface growth flux of the odd serving requests (0: none)")
    ap.add_argument("--serving-growth-nmax", type=float, default=5e3,
                    help="largest N of a serving request with surface growth (its steps update every aggregate)")
    ap.add_argument("--serving-nmax", type=float, default=2e4,
                    help="largest request size in the serving sample (the full form allows 2e5)")
    ap.add_argument("--tuned", default="",
                    help="shape file of python3 -m mcac_amd.autotune: its workers, boxes per worker and depth")
    ap.add_argument("--workers", type=int, default=8, help="host threads sharing the GPU (ensemble, serving)")
    ap.add_argument("--group-boxes", type=int, default=32,
                    help="boxes per worker stepped in lock step, one launch per round (0: one box per worker)")
    ap.add_argument("--group-depth", type=int, default=8, help="steps of each box per grouped launch (1..8)")
    ap.add_argument("--max-job-steps", type=int, default=3_000_000, help="MC step cap per ensemble/serving job")
    ap.add_argument("--c3-windows", type=float, default=0.0,
                    help="seconds: run the N=1e6 box on and report MC steps/s per window of N_agg (0: skip)")
    ap.add_argument("--overlap-runs", action="store_true",
                    help="time whole runs per volsurf_method in both merge launch forms, then exit")
    ap.add_argument("--precision", type=int, default=0, choices=(0, 1, 2),
                    help="contact search: 0 FP64, 1 FP32 broad phase + FP64 solve, 2 as 1 with miss counting")
    ap.add_argument("--timing-stride", type=int, default=64,
                    help="HIP events around every n-th contact search of the timed region")
    ap.add_argument("--leg", choices=sorted(LEGS), default="",
                    help="run only this batch leg and print its JSON (the bench runs each leg so, in a child process)")
    ap.add_argument("--hw-queues", type=int, default=8,
                    help="HIP hardware queues of this process (GPU_MAX_HW_QUEUES, 1..32; 0: leave as set)")
    args = ap.parse_args()
    if 1 <= args.hw_queues <= 32:
        # the host workers' streams each on a queue of their own: HIP's default
        # of four puts two workers' group launches on one queue, one behind the
        # other (set before anything starts the HIP runtime in this process)
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    if args.tuned:
        from mcac_amd.autotune import best_shape
        shape = best_shape(args.tuned)
        args.workers, args.group_boxes, args.group_depth = shape["workers"], shape["boxes_per_worker"], shape["depth"]
    if args.gpus != 1:
        print("bench.py measures one GPU: this path runs one aggregation box per GPU "
              "(independent boxes shard across GPUs with no exchange); use --gpus 1", file=sys.stderr)
        return 2
    os.environ["MCAC_PRECISION"] = str(args.precision)
    if args.leg:
        print(json.dumps(LEGS[args.leg](args)), flush=True)
        return 0
    print("bench: importing torch", file=sys.stderr, flush=True)
    import torch
    torch.cuda.synchronize()  # torch's HIP runtime starts before the library's
    if args.copy_calibration:
        n = 256 << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(20):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n, "copy_calibration_bytes_written": n, "launches": 20}))
        return 0
    import mcac_amd
    if args.c3_windows > 0:
        return c3_windows(args)
    if args.overlap_runs:
        return overlap_runs()

    text = WORKLOAD_INI.read_text()
    # --steps and --warmup are floors: a timed region of a few MC steps (tens of
    # microseconds each) would measure launch jitter, not the path
    steps = max(args.steps, MIN_TIMED_STEPS)
    warmup = max(args.warmup, MIN_WARMUP_STEPS)
    stride = max(1, min(args.timing_stride, steps // 16))
    print("bench: placing N=1e6 monomers", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    sim = mcac_amd.Simulation(text, device=0)
    t_init = time.perf_counter() - t0
    print(f"bench: warm-up, {warmup} MC steps", file=sys.stderr, flush=True)
    done = sim.run(warmup)
    if done != warmup:
        raise RuntimeError("warm-up ended early")
    box = sim.box()
    box.set_timing(stride)  # the first search of the timed region is timed, then every stride-th
    s0 = {k: sim.scalar(k) for k in ("query", "filtered", "pairs", "merges", "t_host_pick_ms", "t_host_merge_ms",
                                     "t_device_ms", "pick_rescales", "ties", "t_tie_ms", "t_enqueue_ms", "t_wait_ms",
                                     "t_merge_call_ms", "t_drain_ms")}
    print(f"bench: timed region, {steps} MC steps", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = sim.run(steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if done != steps:
        raise RuntimeError("timed run ended early")
    s1 = {k: sim.scalar(k) for k in s0}
    tm = box.timing()
    box.set_timing(0)
    st = box.stats()
    prune = box.prune_stats() if args.precision > 0 else None
    d = {k: s1[k] - s0[k] for k in s0}
    steps_per_s = steps / elapsed
    # dominant kernel: the contact search (k_step, or k_filter + k_solve for
    # large movers), HBM-bound.  Algorithmic bytes per launch: 36 B per
    # aggregate of the visited Verlet cells (CoM x,y,z + rmax 8 B each, 4 B id).
    one_launch = st["one_launch"] >= st["two_launch"]
    traffic = None
    if TRAFFIC_FILE.exists():  # rocprofv3 FETCH_SIZE + WRITE_SIZE per k_step dispatch, copy-calibrated
        traffic = json.loads(TRAFFIC_FILE.read_text()).get("k_step_bytes_per_dispatch")
    n_search = max(tm["n_search"], 1.0)
    filt_ms = tm["filter_ms"] / n_search if tm["n_search"] >= 1 else None
    solve_ms = tm["solve_ms"] / n_search if tm["n_search"] >= 1 else None
    bytes_per_launch = 36.0 * d["query"] / steps
    achieved = bytes_per_launch / (filt_ms * 1e-3) / 1e9 if filt_ms else None
    pair_tests = d["query"] + d["pairs"]
    out = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "MC steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "steps_requested": args.steps,
        "warmup_requested": args.warmup,
        "ms_per_step": 1e3 * elapsed / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lognormal monomers placed by the reference algorithm)",
        "config": {"workload": "N=1e6 spheres lognormal 30nm/1.25, FV=3e-3, 1700 K, "
                               + ("plain FP64 path, " if args.precision == 0 else "FP32 broad phase + FP64 solve, ")
                               + f"MC steps {warmup}..{warmup + steps} from t=0",
                   "n_spheres": 1_000_000, "parallelism": "single box, 1 GPU", "precision": args.precision,
                   "hip_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "HIP default")},
        "roofline": {"bound": "hbm", "kernel": "k_step" if one_launch else "k_filter", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": filt_ms,
                     "timed_launches": int(tm["n_search"]),
                     "note": None if achieved else "no search was timed (events on every n-th search)"},
        "candidate_pair_tests_per_s": pair_tests / elapsed,
        "breakdown": {"init_s": t_init, "host_pick_ms_per_step": d["t_host_pick_ms"] / steps,
                      "host_merge_ms_per_step": d["t_host_merge_ms"] / steps,
                      "pick_rescales": d["pick_rescales"], "pick_ties": d["ties"],
                      "host_us_per_step": {k[2:-3]: 1e3 * d[k] / steps for k in
                                           ("t_tie_ms", "t_enqueue_ms", "t_wait_ms", "t_merge_call_ms", "t_drain_ms")},
                      "device_call_ms_per_step": d["t_device_ms"] / steps,
                      "search_launch1_ms": filt_ms, "search_launch2_ms": solve_ms,
                      "k_translate_ms": tm["translate_ms"] / max(tm["n_translate"], 1.0),
                      "merge_ms": tm["merge_ms"] / max(tm["n_merge"], 1.0),
                      "merges": d["merges"], "avg_candidates_visited": d["query"] / steps,
                      "avg_candidates_filtered": d["filtered"] / steps,
                      "timed_searches": tm["n_search"], "launch_stats": st, "fp32_prune": prune},
    }
    # the N=1e6 box gives its device memory and stream back before the other legs
    del box
    sim.close()
    if args.ensemble_boxes > 0:
        out["ensemble"] = run_leg("ensemble")
    if args.serving_requests > 0:
        out["serving"] = run_leg("serving")
    if args.ensemble_boxes > 0 and not args.no_cpu:
        out["ensemble_vs_cpu"] = run_leg("compare")
    if not args.no_cpu:
        cb = cpu_baseline(text, args.cpu_steps)
        cb["ratio"] = steps_per_s / cb["value"]
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
