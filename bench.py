"""Benchmark of the plain device path on the 512^2 -> 1024^2 workload.

One step = one Session::run (random init + every backtracking stage) of a
1024x1024 RGBA8 synthesis from a seeded 512x512 sinusoid+noise example,
default parameters (k=50, m=50, P=5, p=0.5, cauchy 1.0, seed 0). Sessions
(device pyramids, initial state) are built before the timed region, as the
reference's criterion benches build outside the timer
(lib/benches/all-the-things.rs:20-29), so inputs are resident in HBM.

Prints one JSON line: Mpix/s of output, per-step latency p50/p99, the
roofline of the dominant kernel (the throughput-bound batch launches:
k_resolve_persist for single-wave items, k_resolve for teams) from the
launches' device-clock spans in the timed steps, and the CPU oracle timed on
the host's cores.

Beside the main line: "guided" (4 x 300^2 examples with guides -> 1024^2)
and "plain_2048" (the plain path at 2048^2, with the job's device bytes).
Two more legs report in the same line:
* "batched": 256 independent syntheses (examples uniform in {64..512}^2,
  outputs uniform in {128..1024}^2, k in {20, 50}, m = 50, seeds 4000+j)
  run together through the work-queue engine, against the same 256 run one
  after another; aggregate output Mpix/s of both with the outputs resident
  (as the main line), the download-inclusive rate apart, and whether every
  batched result equals its sequential twin bit for bit.
* "serving": a Poisson request loop at a fixed rate; outputs log-uniform
  over 128^2..2048^2 from seeded 256^2 examples, 30% inpainting with 1-3
  elliptical holes covering 20-40% (the example resized to the output with
  CatmullRom); admission by device-memory footprint; p50/p99 latency
  (arrival to result on the host) and output Mpix/s. Requests starting
  while fewer than 32 are in flight run alone with the whole device; as they
  pile up they share the work-queue engine (the line says how many took
  which path). A third run at an overloading rate ("serving_loaded", 40/s)
  shows the engine taking over, and a second at 30/s under a tight budget
  shows admission held back by memory. The requests' images are generated before the
  clock starts.
"""
import hashlib
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
# candidate neighbourhood 4k(k+m), neighbour state 24k (16 B record + 8 B of
# neighbour list), commit 28
BYTES_PER_RESOLVE = 4 * K * (K + M) + 24 * K + 28
WORKLOAD = "plain: 512x512 RGBA8 sinusoid+noise example (seed 512) -> 1024x1024, defaults (k=50, m=50, P=5, p=0.5)"


def build_session(ts, ex, profile=0):
    sess = ts.Session.builder().add_example(ex).seed(0).output_size(ts.Dims.square(1024)).build()
    if profile:
        sess.set_profiling(profile)
    return sess


def host_cpu():
    """CPU model and the cores this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, avail


def cpu_baseline(ex, threads, runs=3):
    """The oracle on the host's cores, with the reference's own threading
    (ms.rs:815-1036: one worker while fewer than 1000 items are to be redone,
    then `threads` workers with immediate updates and the TreeGrid kNN) on
    the full 1024^2 workload, median of `runs`; beside it the same workload
    under the device's Jacobi schedule on `threads` threads (one run), and
    the reference's own CPU case, 128^2 -> 256^2 at one thread (one pixel
    per batch: the reference's max_thread_count(1) semantics, seed 211)."""
    import tsgpu as ts
    from oracle import refcpu as R
    from tsgpu import synth
    ts_all = []
    for _ in range(runs):
        job = R.Job(R.default_params(seed=0), [ex], 1024, 1024)
        t0 = time.perf_counter()
        job.run_reference_threads(threads)
        ts_all.append(time.perf_counter() - t0)
    dt = statistics.median(ts_all)
    job = R.Job(R.default_params(seed=0), [ex], 1024, 1024)
    t0 = time.perf_counter()
    job.run(threads=threads, **ts.default_schedule())
    t_jac = time.perf_counter() - t0
    c1 = synth.example(211, 128, 128)
    t1 = []
    for _ in range(runs):
        job = R.Job(R.default_params(seed=211), [c1], 256, 256)
        t0 = time.perf_counter()
        job.run(threads=1, px_per_batch=1, redo_shift=0, new_shift=0, max_batch=0, first_new_shift=0)
        t1.append(time.perf_counter() - t0)
    d1 = statistics.median(t1)
    model, avail = host_cpu()
    return dict(value=round(1024 * 1024 / dt / 1e6, 4), unit="Mpix/s", cores=threads, kind="port",
                cpu_model=model, cores_available=avail,
                sample=f"full 1024^2 syntheses of the same workload, the oracle with the reference's threading "
                       f"(serial while fewer than 1000 items are to be redone, then {threads} workers, immediate "
                       f"updates, TreeGrid kNN), median of {runs}: " + ", ".join(f"{t:.2f}" for t in ts_all) + " s",
                jacobi_schedule={"value": round(1024 * 1024 / t_jac / 1e6, 4), "unit": "Mpix/s", "cores": threads,
                                 "seconds": round(t_jac, 2),
                                 "note": "the same workload under the device's Jacobi batch schedule"},
                one_thread_c1={"value": round(256 * 256 / d1 / 1e6, 4), "unit": "Mpix/s", "cores": 1,
                               "workload": "128^2 sinusoid+noise example (seed 211) -> 256^2, one pixel per batch "
                                           "(max_thread_count(1) semantics), seed 211",
                               "seconds": [round(t, 2) for t in t1]})


def digest(g):
    h = hashlib.blake2b(digest_size=16)
    for a in (g.image, g.coord, g.ids, g.resolved_flat, g.resolved_score):
        h.update(a.tobytes())
    return h.hexdigest()


def batched_specs(n, seed=4000):
    import numpy as np
    rng = np.random.default_rng(seed)
    ex = rng.choice([64, 128, 256, 512], n)
    out = rng.choice([128, 256, 512, 1024], n)
    k = rng.choice([20, 50], n)
    return [(int(ex[j]), int(out[j]), int(k[j]), seed + j) for j in range(n)]


def batched_leg(ts, synth, n, sequential=True):
    """n syntheses one after another (Session.run) and all at once (Batch);
    sequential=False skips the one-after-another runs (profiling the engine)."""
    specs = batched_specs(n)
    exs = [synth.example(sd, e, e) for e, _, _, sd in specs]

    def builders():
        return [ts.Session.builder().add_example(x).seed(sd).nearest_neighbors(k).random_sample_locations(M)
                .output_size(ts.Dims.square(o)) for x, (e, o, k, sd) in zip(exs, specs)]

    px = sum(o * o for _, o, _, _ in specs)
    # timed, as the plain line: the syntheses with their outputs left in
    # device memory (the downloads are timed apart); the digests after
    t_seq, seq = None, None
    if sequential:
        sess = [b.build() for b in builders()]
        t0 = time.perf_counter()
        for s in sess:
            s.run_resident()
        t_seq = time.perf_counter() - t0
        seq = []
        for s in sess:
            seq.append(digest(s.download()))
            s.close()
        del sess
    sess = [b.build() for b in builders()]
    batch = ts.Batch(ts.Context.default(), max_active=n)
    t0 = time.perf_counter()
    tickets = [batch.submit(s) for s in sess]
    batch.run_all()
    t_bat = time.perf_counter() - t0
    done = dict(batch.poll())
    assert sorted(done) == sorted(tickets) and not any(done.values())
    got = [digest(batch.take(t)) for t in tickets]
    t_pcie = time.perf_counter() - t0
    st = batch.stats()
    prof = batch.engine_profile()
    batch.close()
    # algorithmic bytes of the engine launches: every job's resolves (its
    # stage sizes, (0.5^s N) for s = 0..5) at 4k(k+m) + 24k + 28 bytes each
    alg = sum(sum(int((0.5 ** s_) * o * o) for s_ in range(6)) * (4 * k * (k + M) + 24 * k + 28)
              for _, o, k, _ in specs)
    t_kern = prof["launch_to_done_ms"] / 1e3
    ec = counters_from_profiles("engine_c4_counters.md", "k_engine")
    return {
        "workload": f"{n} syntheses: examples uniform in {{64,128,256,512}}^2, outputs uniform in "
                    f"{{128,256,512,1024}}^2, k in {{20,50}}, m=50, seeds 4000+j",
        "output_mpix": round(px / 1e6, 3),
        "value": round(px / t_bat / 1e6, 3), "unit": "Mpix/s",
        "sequential_value": round(px / t_seq / 1e6, 3) if t_seq else None,
        "speedup_vs_sequential": round(t_seq / t_bat, 2) if t_seq else None,
        "seconds": round(t_bat, 3), "sequential_seconds": round(t_seq, 3) if t_seq else None,
        "pcie_inclusive": {"value": round(px / t_pcie / 1e6, 3), "unit": "Mpix/s", "seconds": round(t_pcie, 3),
                           "note": "the batched syntheses with every result downloaded to host memory"},
        "identical_to_alone": (seq == got) if seq is not None else None,
        "launches": st["launches"], "engine_rounds": st["rounds"],
        "engine_workgroup_ms": {k: (round(v, 1) if isinstance(v, float) else v) for k, v in prof.items()},
        "roofline": {"kernel": "k_engine<false>", "bound": "hbm", "algorithmic_bytes": alg,
                     "achieved": round(alg / t_kern / 1e9, 1) if t_kern > 0 else None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(alg / t_kern / 1e9 / HBM_PEAK_GBS, 4) if t_kern > 0 else None,
                     "over": "launch to completion of the engine launches (host clock)",
                     "traffic": (ec or {}).get("hbm_bytes"), "l2_hit": (ec or {}).get("l2_hit"),
                     "valu_busy": (ec or {}).get("valu_busy"), "waves_per_cu": (ec or {}).get("waves_per_cu"),
                     "limiter": limiter(ec), "counters_from": "profiles/engine_c4_counters.md"},
    }


def plain_2048_leg(ts, synth, steps):
    """The plain path one size up: 512^2 example (seed 512) -> 2048^2, default
    parameters, seed 0, outputs resident; and the job's device bytes after
    its run, the stage-ahead sets included."""
    ex = synth.example(512, 512, 512)

    def build():
        return ts.Session.builder().add_example(ex).seed(0).output_size(ts.Dims.square(2048)).build()

    s = build()
    s.run_resident()
    fp = s.footprint()
    s.close()
    lat = []
    for _ in range(steps):
        s = build()
        t0 = time.perf_counter()
        s.run_resident()
        lat.append(time.perf_counter() - t0)
        s.close()
    return {"workload": "plain: 512x512 sinusoid+noise example (seed 512) -> 2048x2048, defaults, seed 0",
            "value": round(2048 * 2048 * steps / sum(lat) / 1e6, 3), "unit": "Mpix/s",
            "ms_per_step": round(1e3 * sum(lat) / steps, 3), "steps": steps,
            "device_bytes_after_run": fp}


def guided_leg(ts, synth, steps):
    """Multi-example guided synthesis: 4 examples of 300^2 with blob guides,
    a 1024^2 blob target guide, guide_alpha 0.8, random_init 10, k=50, seed 211."""
    exs = [synth.example(1000 + i, 300, 300) for i in range(4)]
    gds = [synth.blobs(2000 + i, 300, 300) for i in range(4)]
    tg = synth.blobs(3000, 1024, 1024)

    def build():
        b = ts.Session.builder().seed(211).guide_alpha(0.8).random_init(10).nearest_neighbors(50)
        for e, g in zip(exs, gds):
            b.add_example(ts.Example(e).with_guide(g))
        return b.load_target_guide(tg).output_size(ts.Dims.square(1024)).build()

    build().run()
    sess = [build() for _ in range(steps)]
    lat = []
    for s in sess:
        t0 = time.perf_counter()
        s.run()
        lat.append(time.perf_counter() - t0)
    return {
        "workload": "guided: 4 x 300^2 examples with 300^2 blob guides (seeds 1000+i, 2000+i), 1024^2 blob "
                    "target guide (seed 3000) -> 1024^2, guide_alpha 0.8, random_init 10, k=50, m=50, seed 211",
        "value": round(1024 * 1024 * steps / sum(lat) / 1e6, 3), "unit": "Mpix/s",
        "ms_per_step": round(1e3 * sum(lat) / steps, 3), "steps": steps,
    }


def serving_request(ts, synth, sd, side, inpaint):
    """A request on a seeded 256^2 example; an inpainting one is resized to
    the output size by the builder (load_image's CatmullRom, utils.rs:55-80)."""
    ex = synth.example(sd, 256, 256)
    b = ts.Session.builder().seed(sd)
    if inpaint:
        m = synth.ellipse_mask(sd, side, side)
        b.inpaint_example(m, ts.Example(ex).set_sample_method(m), ts.Dims.square(side))
    else:
        b.add_example(ex).output_size(ts.Dims.square(side))
    return b


def serving_specs(n, seed=5000):
    import numpy as np
    rng = np.random.default_rng(seed)
    sides = np.exp(rng.uniform(np.log(128), np.log(2048), n)).astype(int)
    inpaint = rng.random(n) < 0.3
    return [(seed + i, int(sides[i]), bool(inpaint[i])) for i in range(n)]


def serving_leg(ts, synth, n, qps, budget_gib, quantum=None, rounds_per_pump=1):
    from tsgpu import serve
    import numpy as np
    reqs = [(sp, sp[1] * sp[1], sp[2]) for sp in serving_specs(n)]
    # the requests' images are generated before the clock starts (a client
    # sends finished images; the generator is not part of the server), on
    # host threads (numpy releases the interpreter lock in its array work)
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max(1, min(16, host_cpu()[1]))) as pool:
        built = list(pool.map(lambda r: serving_request(ts, synth, *r[0]), reqs))
    gaps = np.random.default_rng(5999).exponential(1.0 / qps, n)
    arrivals = np.cumsum(gaps) - gaps[0]
    srv = serve.Server(ts.Context.default(), hbm_budget=int(budget_gib * (1 << 30)), max_active=256,
                       rounds_per_pump=rounds_per_pump, quantum=serve.QUANTUM if quantum is None else quantum)
    fin = []
    busy = 0.0  # wall time with at least one request admitted or waiting
    t0 = time.perf_counter()
    i = 0
    while i < n or not srv.idle():
        now = time.perf_counter() - t0
        while i < n and arrivals[i] <= now:
            srv.submit(built[i], arrival=t0 + arrivals[i])
            built[i] = None
            i += 1
        if srv.idle():
            time.sleep(max(0.0, min(0.01, arrivals[i] - now)))
            continue
        ta = time.perf_counter()
        fin.extend(srv.pump())
        busy += time.perf_counter() - ta
    t_end = time.perf_counter()
    bst = srv.batch.stats()
    srv.close()
    errors = [f for f in fin if f.error is not None]
    lat = [f.latency_s * 1e3 for f in fin if f.error is None]
    px = sum(f.out_px for f in fin if f.error is None)
    return {
        "workload": f"{n} requests, Poisson {qps}/s, outputs log-uniform 128^2..2048^2 from seeded 256^2 "
                    f"examples, {sum(r[2] for r in reqs)} inpainting (1-3 ellipses, 20-40% holes)",
        "value": round(px / (t_end - t0) / 1e6, 3), "unit": "Mpix/s",
        "latency_ms": {"p50": round(serve.percentile(lat, 50), 1), "p99": round(serve.percentile(lat, 99), 1),
                       "max": round(max(lat), 1) if lat else None},
        "offered_mpix_per_s": round(sum(r[1] for r in reqs) / 1e6 * qps / n, 3),
        "seconds": round(t_end - t0, 3), "errors": len(errors),
        "busy_fraction": round(busy / (t_end - t0), 3),
        "hbm_budget_gib": budget_gib, "peak_admitted_gib": round(srv.peak_in_use / (1 << 30), 3),
        "memory_holds": srv.memory_holds, "rejected": srv.rejected,
        "engine": {"launches": bst["launches"], "requests_run_alone": bst["alone"],
                   "requests_on_engine": len(fin) - bst["alone"] - len(errors), "alone_below": srv.alone_below,
                   "quantum_items": srv.quantum, "rounds_per_pump": rounds_per_pump},
    }


def traffic_from_profiles():
    """HBM bytes per batch launch (k_resolve_persist and k_resolve) of this
    workload, from the committed counter summary (FETCH_SIZE + WRITE_SIZE per
    dispatch, copy-calibrated)."""
    import re
    path = os.path.join(ROOT, "profiles", "resolve_c2_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        m = re.search(r"^hbm_bytes_per_launch:\s*([0-9.]+)", f.read(), re.M)
    return float(m.group(1)) if m else None


def counters_from_profiles(name, kernel):
    """Counter figures of `kernel` from a committed summary under profiles/
    (lines "<figure>_<kernel>: value"), or None."""
    import re
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    out = {}
    with open(path) as f:
        for m in re.finditer(r"^([a-z0-9_]+)_" + re.escape(kernel) + r":\s*([0-9.]+)", f.read(), re.M):
            out[m.group(1)] = float(m.group(2))
    return out or None


def note(msg):
    """A progress line on stderr as each leg ends (stdout carries only the
    result line)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def limiter(c):
    """The limiter the counters point at."""
    if not c:
        return None
    if c.get("valu_busy", 0) >= 0.5:
        return "vector ALU (VALU busy %.0f%%), then memory latency" % (100 * c["valu_busy"])
    return "memory latency: waves wait %.0f%% of their cycles, VALU busy %.0f%%, %.1f resident waves per CU" % (
        100 * c.get("wait_fraction", 0), 100 * c.get("valu_busy", 0), c.get("waves_per_cu", 0))


def copy_calibration(mb):
    """A plain device-to-device copy of `mb` MiB (runtime copy kernel), the
    reference point that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mb << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n, "copies": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads (0: the cores this process may use, at most 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="time the library's default launch structure")
    ap.add_argument("--batched", type=int, default=256, help="syntheses in the batched leg (0 = skip)")
    ap.add_argument("--no-sequential", dest="sequential", action="store_false",
                    help="batched leg without the one-after-another runs (for profiling the engine)")
    ap.add_argument("--guided", type=int, default=3, help="runs of the guided leg (0 = skip)")
    ap.add_argument("--plain2048", type=int, default=2, help="runs of the 2048^2 plain leg (0 = skip)")
    ap.add_argument("--serve-requests", type=int, default=2000, help="requests in the serving leg (0 = skip)")
    ap.add_argument("--serve-tight-requests", type=int, default=300,
                    help="requests of a second serving run under a tight memory budget (0 = skip)")
    ap.add_argument("--serve-tight-gib", type=float, default=0.5)
    ap.add_argument("--serve-loaded-requests", type=int, default=400,
                    help="requests of a third serving run at --serve-loaded-qps, where requests pile up "
                         "and share the work-queue engine (0 = skip)")
    ap.add_argument("--serve-loaded-qps", type=float, default=40.0)
    ap.add_argument("--serve-qps", type=float, default=18.0,
                    help="Poisson arrival rate; 18/s keeps the engine busy about 70%% of the time")
    ap.add_argument("--serve-budget-gib", type=float, default=4.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import tsgpu as ts
    from tsgpu import synth

    ex = synth.example(512, 512, 512)
    ctx = ts.Context.default()
    for _ in range(args.warmup):
        build_session(ts, ex).run()
    # launch structure for this workload, picked on the device before the
    # timed steps (results do not depend on it); later sessions of the shape get it
    tuned = None
    if args.autotune:
        from tsgpu.autotune import autotune
        b = ts.Session.builder().add_example(ex).seed(0).output_size(ts.Dims.square(1024))
        tuned = autotune(b, reps=2, ctx=ctx)
    # the timed steps record the resolve kernels' device-clock spans (no events)
    sessions = [build_session(ts, ex, profile=-1) for _ in range(args.steps)]
    prof = build_session(ts, ex, profile=1)
    # a step: the synthesis with its output left in device memory, then the
    # job's release; the download to host memory (PCIe) is timed apart
    lat, lat_pcie = [], []
    res = {"ms": 0.0, "items": 0, "launches": 0}  # batch launches of the timed steps
    for s in sessions:
        t0 = time.perf_counter()
        s.run_resident()
        t1 = time.perf_counter()
        step_res = s.download().kernel_stats["resolve"]
        t2 = time.perf_counter()
        for key in res:
            res[key] += step_res[key]
        s.close()
        t3 = time.perf_counter()
        lat.append(((t1 - t0) + (t3 - t2)) * 1e3)
        lat_pcie.append((t3 - t0) * 1e3)
    t_all = sum(lat) / 1e3
    t_pcie = sum(lat_pcie) / 1e3
    # profiled run (separate, after the timed region): every kernel's device time
    g = prof.run()
    ks = g.kernel_stats
    # the dominant kernel: one launch per batch of more than a few items
    # (k_resolve_persist, or k_resolve for teams; the one-workgroup loop over
    # the smallest batches is reported apart), over the timed steps; durations
    # are the launches' device-clock spans
    resolves = (res["items"] + ks["resolve_runs"]["items"] * args.steps) // args.steps
    kern_s = res["ms"] / 1e3
    achieved = res["items"] * BYTES_PER_RESOLVE / kern_s / 1e9 if kern_s > 0 else 0.0
    per_launch = res["items"] * BYTES_PER_RESOLVE / max(1, res["launches"])
    ms_per_step = t_all * 1e3 / args.steps
    value = 1024 * 1024 * args.steps / t_all / 1e6
    rc_ctr = counters_from_profiles("resolve_c2_counters.md", "k_resolve")
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
                   "schedule": "jacobi " + json.dumps(ts.default_schedule()), "batches_per_step": g.n_batches,
                   "launch_structure": (dict(tuned["tuning"], autotuned=True, default_ms=round(tuned["default_ms"], 3),
                                             tuned_ms=round(tuned["ms"], 3), candidates=len(tuned["tried"]))
                                        if tuned else dict(ts.default_tuning(), autotuned=False))},
        "latency_ms": {"p50": round(statistics.median(lat), 3), "p99": round(p99, 3)},
        "pcie_inclusive": {"value": round(1024 * 1024 * args.steps / t_pcie / 1e6, 3), "unit": "Mpix/s",
                           "ms_per_step": round(t_pcie * 1e3 / args.steps, 3),
                           "note": "the same steps with the output image, coordinate and id maps and the "
                                   "resolved list downloaded to host memory"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(),
                     "l2_hit": (rc_ctr or {}).get("l2_hit"), "lds_bank_conflict": (rc_ctr or {}).get("lds_bank_conflict"),
                     "valu_busy": (rc_ctr or {}).get("valu_busy"), "waves_per_cu": (rc_ctr or {}).get("waves_per_cu"),
                     "limiter": limiter(rc_ctr), "counters_from": "profiles/resolve_c2_counters.md",
                     "kernel": "k_resolve_persist + k_resolve", "bytes_per_resolve": BYTES_PER_RESOLVE,
                     "algorithmic_bytes_per_launch": round(per_launch),
                     "avg_launch_ms": round(res["ms"] / max(1, res["launches"]), 4),
                     "launches": res["launches"], "over": "the timed steps"},
        "kernels_ms": {k: round(v["ms"], 3) for k, v in ks.items()},
        "device_ms_per_step": round(g.run_ms, 3),
    }
    note("plain 1024^2: %.3f Mpix/s" % value)
    if args.guided:
        out["guided"] = guided_leg(ts, synth, args.guided)
        note("guided: %.3f Mpix/s" % out["guided"]["value"])
    if args.plain2048:
        out["plain_2048"] = plain_2048_leg(ts, synth, args.plain2048)
        note("plain 2048^2: %.3f Mpix/s" % out["plain_2048"]["value"])
    if args.batched:
        out["batched"] = batched_leg(ts, synth, args.batched, args.sequential)
        out["batched"]["vs_plain"] = round(out["batched"]["value"] / value, 2)
        note("batched: %.3f Mpix/s" % out["batched"]["value"])
    if args.serve_requests:
        out["serving"] = serving_leg(ts, synth, args.serve_requests, args.serve_qps, args.serve_budget_gib)
        note("serving: %d requests" % args.serve_requests)
    if args.serve_tight_requests:
        out["serving_tight_budget"] = serving_leg(ts, synth, args.serve_tight_requests, 30.0, args.serve_tight_gib)
        note("serving, tight budget: %d requests" % args.serve_tight_requests)
    if args.serve_loaded_requests:
        out["serving_loaded"] = serving_leg(ts, synth, args.serve_loaded_requests, args.serve_loaded_qps,
                                            args.serve_budget_gib)
        note("serving, loaded: %d requests" % args.serve_loaded_requests)
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, host_cpu()[1])
        out["cpu_baseline"] = cpu_baseline(ex, threads)
        out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
        note("cpu baseline: %.3f Mpix/s" % out["cpu_baseline"]["value"])
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
