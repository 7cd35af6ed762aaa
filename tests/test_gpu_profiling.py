"""Profiling levels through the C ABI: the recorded times are there, and
recording them changes no result."""
import numpy as np
import pytest

import tsgpu as ts
from tsgpu import synth

pytestmark = pytest.mark.gpu


def build(level):
    s = ts.Session.builder().add_example(synth.example(311, 96, 96)).seed(4).output_size(ts.Dims.square(256)).build()
    if level:
        s.set_profiling(level)
    return s


def test_profiling_levels_keep_results_and_report_times(ctx):
    plain = build(0).run()
    spans = build(-1).run()
    full = build(1).run()
    for g in (spans, full):
        assert np.array_equal(g.coord, plain.coord) and np.array_equal(g.image, plain.image)
    # spans only: the batch launches and dependency-ordered runs, nothing else
    ks = spans.kernel_stats
    timed = ks["resolve"]["launches"] + ks["resolve_runs"]["launches"]
    assert timed > 0
    for name in ("resolve", "resolve_runs"):
        if ks[name]["launches"]:
            assert ks[name]["ms"] > 0 and ks[name]["items"] > 0
    assert ks["commit"]["launches"] == 0 and ks["rand"]["launches"] == 0
    # level 1 times every launch
    kf = full.kernel_stats
    assert kf["commit"]["launches"] > 0 and kf["commit"]["ms"] > 0
    assert kf["resolve"]["launches"] == ks["resolve"]["launches"]
    assert kf["resolve"]["items"] == ks["resolve"]["items"]


def test_phase_clocks_need_level_two(ctx):
    s = build(1)
    s.run_resident()
    with pytest.raises(ts.TsgError):
        import ctypes as C
        out = np.zeros((256 * 256, 16), np.uint64)
        ts._check(ts.lib().tsg_job_phase_clocks(s._job, out.ctypes.data_as(C.c_void_p), 256 * 256))
    s.close()
