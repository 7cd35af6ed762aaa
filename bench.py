"""Benchmark of the plain device path on the 512^2 -> 1024^2 workload.

One step = one Session::run (random init + every backtracking stage) of a
1024x1024 RGBA8 synthesis from a seeded 512x512 sinusoid+noise example,
default parameters (k=50, m=50, P=5, p=0.5, cauchy 1.0, seed 0). Sessions
(device pyramids, initial state) are built before the timed region, as the
reference's criterion benches build outside the timer
(lib/benches/all-the-things.rs:20-29), so inputs are resident in HBM.

Prints one JSON line: Mpix/s of output, per-step latency p50/p99, the
roofline of the dominant kernel (k_resolve) from HIP events on the
library's stream, and the CPU oracle timed on the host's cores.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "texture-synthesis_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
K, M = 50, 50
# algorithmic bytes per pixel-resolve: RGBA8 example gathers of every
# candidate neighbourhood 4k(k+m), neighbour state 24k, commit 28
BYTES_PER_RESOLVE = 4 * K * (K + M) + 24 * K + 28
WORKLOAD = "plain: 512x512 RGBA8 sinusoid+noise example (seed 512) -> 1024x1024, defaults (k=50, m=50, P=5, p=0.5)"


def build_session(ts, ex, profile=False):
    sess = ts.Session.builder().add_example(ex).seed(0).output_size(ts.Dims.square(1024)).build()
    if profile:
        sess.set_profiling(True)
    return sess


def cpu_baseline(ex, threads):
    from oracle import refcpu as R
    job = R.Job(R.default_params(seed=0), [ex], 1024, 1024)
    t0 = time.perf_counter()
    job.run(threads=threads, px_per_batch=0, redo_shift=5, new_shift=6, max_batch=0)
    dt = time.perf_counter() - t0
    return dict(value=round(1024 * 1024 / dt / 1e6, 4), unit="Mpix/s", cores=threads, kind="port",
                sample=f"one full 1024^2 synthesis of the same workload on {threads} host threads "
                       f"(oracle, same Jacobi schedule), {dt:.1f} s")


def traffic_from_profiles():
    path = os.path.join(ROOT, "profiles", "resolve_c2_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import tsgpu as ts
    from tsgpu import synth

    ex = synth.example(512, 512, 512)
    ctx = ts.Context.default()
    for _ in range(args.warmup):
        build_session(ts, ex).run()
    sessions = [build_session(ts, ex) for _ in range(args.steps)]
    prof = build_session(ts, ex, profile=True)
    lat = []
    t_all0 = time.perf_counter()
    for s in sessions:
        t0 = time.perf_counter()
        s.run()
        lat.append((time.perf_counter() - t0) * 1e3)
    t_all = time.perf_counter() - t_all0
    # profiled run (separate, after the timed region): per-kernel device time
    g = prof.run()
    ks = g.kernel_stats
    res = ks["resolve"]
    resolves = res["items"]
    kern_s = res["ms"] / 1e3
    achieved = resolves * BYTES_PER_RESOLVE / kern_s / 1e9 if kern_s > 0 else 0.0
    ms_per_step = t_all * 1e3 / args.steps
    value = 1024 * 1024 * args.steps / t_all / 1e6
    lat_sorted = sorted(lat)
    p99 = lat_sorted[min(len(lat_sorted) - 1, int(round(0.99 * (len(lat_sorted) - 1))))]
    out = {
        "metric": "output Mpix/s at 1024^2 (plain)",
        "value": round(value, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "output_px": 1024 * 1024, "resolves_per_step": resolves,
                   "schedule": "jacobi redo>>5 new>>6", "batches_per_step": g.n_batches},
        "latency_ms": {"p50": round(statistics.median(lat), 3), "p99": round(p99, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(),
                     "kernel": "k_resolve", "bytes_per_resolve": BYTES_PER_RESOLVE,
                     "avg_launch_ms": round(res["ms"] / max(1, res["launches"]), 4),
                     "launches": res["launches"]},
        "kernels_ms": {k: round(v["ms"], 3) for k, v in ks.items()},
        "device_ms_per_step": round(g.run_ms, 3),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ex, args.cpu_threads)
        out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
