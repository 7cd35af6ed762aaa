"""bench.py's inputs that do not need a GPU: the committed profile summaries
it reads the roofline's traffic and counters from, and the workload mixes
of its batched and serving legs."""
import math

import numpy as np

import bench


def test_profiles_give_the_roofline_fields():
    t = bench.traffic_from_profiles()
    assert t is not None and 1e6 < t < 1e9  # HBM bytes per batch launch
    rc = bench.counters_from_profiles("resolve_c2_counters.md", "k_resolve")
    ec = bench.counters_from_profiles("engine_c4_counters.md", "k_engine")
    for c in (rc, ec):
        assert c is not None
        for key in ("l2_hit", "lds_bank_conflict", "valu_busy", "waves_per_cu", "wait_fraction"):
            assert key in c
        assert 0.0 < c["l2_hit"] <= 1.0 and 0.0 < c["valu_busy"] <= 1.0 and 0.0 < c["waves_per_cu"] <= 32.0
    assert ec["hbm_bytes"] > 0
    assert bench.limiter(rc).startswith("memory latency")
    assert bench.limiter(ec).startswith("vector ALU")
    assert bench.limiter(None) is None


def test_batched_mix_is_the_stated_one():
    specs = bench.batched_specs(256)
    assert specs == bench.batched_specs(256)  # seeded
    ex, out, k, seeds = (np.array(v) for v in zip(*specs))
    assert set(ex) <= {64, 128, 256, 512} and set(out) <= {128, 256, 512, 1024} and set(k) <= {20, 50}
    assert len(set(ex)) == 4 and len(set(out)) == 4 and len(set(k)) == 2
    assert list(seeds) == list(range(4000, 4256))


def test_serving_mix_is_log_uniform_with_30_percent_inpainting():
    specs = bench.serving_specs(2000)
    sides = np.array([s for _, s, _ in specs])
    inpaint = np.array([p for _, _, p in specs])
    assert sides.min() >= 128 and sides.max() <= 2048
    # log-uniform: the median side is near sqrt(128 * 2048) = 512
    assert abs(math.log(np.median(sides)) - math.log(512.0)) < 0.1
    assert abs(inpaint.mean() - 0.3) < 0.03
