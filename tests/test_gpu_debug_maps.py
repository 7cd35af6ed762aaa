"""Debug maps and coordinate transforms on the device.

get_id_maps / get_uncertainty_map (ms.rs:603-653) of a finished job equal the
oracle's for the same job (the states are bit-identical). CoordinateTransform
::apply (lib.rs:178-210) gathers each output pixel from its source map: its
result equals a plain gather, and applying a job's own transform to its
example reproduces the synthesized image (the final stage reads the example
at full resolution), which is what the reference's `repeat` relies on.
"""
import numpy as np
import pytest

import tsgpu as ts
from oracle import refcpu as R
from tsgpu import synth

pytestmark = pytest.mark.gpu

DEFAULT = ts.default_schedule()


def oracle_job(ex, out, seed, **kw):
    job = R.Job(R.default_params(seed=seed), [ex], out, out, **kw)
    job.run(threads=8, **DEFAULT)
    return job


def test_debug_maps_match_the_oracle(ctx):
    ex = synth.example(810, 48, 48)
    job = oracle_job(ex, 96, 810)
    s = ts.Session.builder().add_example(ex).seed(810).output_size(ts.Dims.square(96)).build()
    s.set_schedule(**DEFAULT).run_resident()
    maps = s.debug_maps()
    s.close()
    for got, want in zip((maps["patch_ids"], maps["map_ids"], maps["uncertainty"]), job.debug_maps()):
        assert np.array_equal(got, want)


def test_debug_maps_of_an_inpaint_job(ctx):
    n = 64
    m = synth.ellipse_mask(811, n, n)
    col = synth.example(811, n, n)
    job = oracle_job(col, n, 3, sample_kind=[2], sample_masks=[m], inpaint_mask=m, inpaint_color=col)
    s = ts.Session.builder().inpaint_example(m, ts.Example(col).set_sample_method(m), ts.Dims.square(n)).seed(3).build()
    s.set_schedule(**DEFAULT).run_resident()
    maps = s.debug_maps()
    s.close()
    for got, want in zip((maps["patch_ids"], maps["map_ids"], maps["uncertainty"]), job.debug_maps()):
        assert np.array_equal(got, want)


def test_xform_apply_gathers(ctx):
    rng = np.random.default_rng(812)
    srcs = [rng.integers(0, 256, (30, 20, 4), dtype=np.uint8), rng.integers(0, 256, (7, 11, 4), dtype=np.uint8)]
    w, h = 37, 23
    m = rng.integers(0, 2, w * h)
    x = np.where(m == 0, rng.integers(0, 20, w * h), rng.integers(0, 11, w * h))
    y = np.where(m == 0, rng.integers(0, 30, w * h), rng.integers(0, 7, w * h))
    buf = np.stack([x, y, m], 1).reshape(-1).astype(np.uint32)
    ct = ts.CoordinateTransform(buf, ts.Dims(w, h), [ts.Dims(20, 30), ts.Dims(11, 7)])
    got = ct.apply(srcs, ctx)
    want = np.zeros((w * h, 4), np.uint8)
    for i in range(w * h):
        want[i] = srcs[m[i]][y[i], x[i]]
    assert np.array_equal(got.reshape(-1, 4), want)


def test_xform_apply_rejects_a_coordinate_outside_its_map(ctx):
    ct = ts.CoordinateTransform(np.array([5, 0, 0, 0, 0, 0], np.uint32), ts.Dims(2, 1), [ts.Dims(5, 5)])
    with pytest.raises(ts.TsgError):
        ct.apply([np.zeros((5, 5, 4), np.uint8)], ctx)


def test_own_transform_reproduces_the_image(ctx):
    ex = synth.example(813, 64, 64)
    g = ts.Session.builder().add_example(ex).seed(813).output_size(ts.Dims(120, 90)).build().run()
    back = g.get_coordinate_transform().apply([ex], ctx)
    assert np.array_equal(back, g.image)
