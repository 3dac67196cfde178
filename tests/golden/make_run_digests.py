"""Write the oracle's run digests (tests/golden/digest_<case>.json).

The oracle (oracle/libmcac_oracle.so, the reference loop restated) runs each
case from its parameter file with step and merge tracing on; tests/run_digest.py
turns the traces into block digests.  The GPU tests run the same case on the
device and compare digests, so the whole run, not only a prefix, is checked bit
for bit without committing megabytes of records.

Cases:
  c2_full     tests/configs/c2_plain_gpu.ini (N=1e5 lognormal, FV=1e-3) run to
              its stop criterion, N_agg = 1000 (about half an hour of oracle time)
  c3_prefix   tests/configs/c3_n1m.ini (N=1e6 lognormal, FV=3e-3), the first
              3000 MC steps (a few minutes)

Run in the build container:  python tests/golden/make_run_digests.py c3_prefix [c2_full]
"""
import json
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "tests"))

import oracle_lib as O  # noqa: E402
from run_digest import run_digest  # noqa: E402

CASES = {
    "c2_full": dict(ini="c2_plain_gpu.ini", steps=10 ** 8, step_block=20000, merge_block=2000),
    "c3_prefix": dict(ini="c3_n1m.ini", steps=3000, step_block=100, merge_block=100),
}


def oracle_digest(case: str) -> dict:
    c = CASES[case]
    text = (ROOT / "tests" / "configs" / c["ini"]).read_text()
    t0 = time.perf_counter()
    sim = O.OracleSim(text, fast_init=True, trace=True)
    done = 0
    while done < c["steps"]:
        n = sim.run(min(50000, c["steps"] - done))
        done += n
        print(f"{case}: {done} steps, n_agg {sim.scalar('n_agg'):.0f}, {time.perf_counter() - t0:.0f} s", flush=True)
        if n == 0 or sim.finished():
            break
    _, agg = sim.export()
    scalars = {"time": sim.scalar("time"), "n_agg": sim.scalar("n_agg"), "finished": bool(sim.finished())}
    d = run_digest(sim.steps(), sim.merges(), agg, scalars, c["step_block"], c["merge_block"])
    d["case"] = case
    d["ini"] = c["ini"]
    d["max_steps"] = c["steps"]
    return d


def main() -> int:
    for case in sys.argv[1:] or ["c3_prefix"]:
        d = oracle_digest(case)
        dst = ROOT / "tests" / "golden" / f"digest_{case}.json"
        dst.write_text(json.dumps(d, indent=1) + "\n")
        print(f"wrote {dst}: {d['n_steps']} steps, {d['n_merges']} merges")
    return 0


if __name__ == "__main__":
    sys.exit(main())
