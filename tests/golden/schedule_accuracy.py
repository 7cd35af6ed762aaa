"""Per-stage mean Cauchy cost of whole images under the default Jacobi
schedule against one pixel per batch (the reference's max_thread_count(1)
semantics, lib/src/ms.rs:887-1011 with the serial rule of ms.rs:815), on
the CPU oracle, for the two 1024^2 workloads:

* plain: 512^2 sinusoid+noise example (seed 512) -> 1024^2, defaults;
* guided: 4 x 300^2 examples (seeds 1000+i) with 300^2 blob guides (seeds
  2000+i), a 1024^2 blob target guide (seed 3000), guide_alpha 0.8,
  random_init 10, k = m = 50.

The per-stage mean is chaotic in the synthesis seed, so each workload runs
under several seeds and the seed-averaged means are compared. Writes
schedule_accuracy.json beside this script: every run's per-stage means
(stage P first) and the relative deviation of the default schedule's
averages from the sequential ones. tests/test_schedule_quality.py checks the
recorded deviations against the 2% bar.

    python tests/golden/schedule_accuracy.py [--seeds 4] [--procs 8]

Each run is single-threaded (the oracle's results do not depend on its
thread count); runs go to a process pool. A sequential 1024^2 run takes
about 40 s (plain) to 90 s (guided) on one core.
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "texture-synthesis_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

SEQUENTIAL = dict(px_per_batch=1, redo_shift=0, new_shift=0, max_batch=0, first_new_shift=0)
# the device's default schedule (tsg_schedule_default): redo >> 5, new >> 6,
# first-stage new >> 7
DEFAULT = dict(px_per_batch=0, redo_shift=5, new_shift=6, max_batch=0, first_new_shift=7)

WORKLOADS = {
    "plain_512_to_1024": dict(seed0=0),
    "guided_4x300_to_1024": dict(seed0=211),
}


def _job(workload, seed):
    from oracle import refcpu as R
    from tsgpu import synth
    if workload == "plain_512_to_1024":
        ex = synth.example(512, 512, 512)
        return R.Job(R.default_params(seed=seed), [ex], 1024, 1024)
    exs = [synth.example(1000 + i, 300, 300) for i in range(4)]
    gds = [synth.blobs(2000 + i, 300, 300) for i in range(4)]
    tg = synth.blobs(3000, 1024, 1024)
    prm = R.default_params(seed=seed, guide_alpha=0.8, random_init=10, nearest_neighbors=50,
                           random_sample_locations=50)
    return R.Job(prm, exs, 1024, 1024, example_guides=gds, target_guide=tg)


def _run(task):
    workload, seed, name = task
    sched = SEQUENTIAL if name == "sequential" else DEFAULT
    t0 = time.perf_counter()
    job = _job(workload, seed)
    job.run(threads=1, **sched)
    means = [float(x) for x in job.stage_mean_costs()]
    return workload, seed, name, means, time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=4)
    ap.add_argument("--procs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--out", default=os.path.join(HERE, "schedule_accuracy.json"))
    args = ap.parse_args()
    tasks = [(w, cfg["seed0"] + s, name) for w, cfg in WORKLOADS.items() for s in range(args.seeds)
             for name in ("sequential", "default")]
    # longest first: guided sequential runs
    tasks.sort(key=lambda t: (t[2] != "sequential", not t[0].startswith("guided")))
    runs = {}
    with ProcessPoolExecutor(args.procs) as pool:
        for workload, seed, name, means, dt in pool.map(_run, tasks):
            print(f"{workload} seed {seed} {name}: {dt:.1f} s", flush=True)
            runs.setdefault(workload, {}).setdefault(name, {})[str(seed)] = dict(stage_means=means,
                                                                                 seconds=round(dt, 1))
    result = {"bar": 0.02, "schedule_default": DEFAULT, "schedule_sequential": SEQUENTIAL, "workloads": {}}
    for w, r in runs.items():
        seeds = sorted(r["sequential"], key=int)
        n_st = len(r["sequential"][seeds[0]]["stage_means"])
        avg = {name: [sum(r[name][s]["stage_means"][i] for s in seeds) / len(seeds) for i in range(n_st)]
               for name in ("sequential", "default")}
        dev = [avg["default"][i] / avg["sequential"][i] - 1.0 for i in range(n_st)]
        result["workloads"][w] = dict(seeds=[int(s) for s in seeds], runs=r, mean_sequential=avg["sequential"],
                                      mean_default=avg["default"], deviation=dev,
                                      max_abs_deviation=max(abs(d) for d in dev))
        print(w, "deviation per stage (P..0, %):", [round(100 * d, 2) for d in dev], flush=True)
    with open(args.out, "w") as f:
        json.dump(result, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
