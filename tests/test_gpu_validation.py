"""Validation and degenerate inputs through the C ABI.

Parameter and image checks return the reference's Error variants from
build() (session.rs:450-524); cases where the reference hangs or panics
return an error instead; degenerate sizes (1x1 and one-pixel-wide outputs,
ragged sides, tiny examples) stay bit-exact against the oracle.
"""
import ctypes as C

import numpy as np
import pytest

import tsgpu as ts
from tsgpu import synth
from test_gpu_parity import assert_same, run_pair

pytestmark = pytest.mark.gpu


def builder(seed=1, ex=24, out=32):
    return ts.Session.builder().add_example(synth.example(seed, ex, ex)).seed(seed).output_size(ts.Dims.square(out))


@pytest.mark.parametrize("setter,value", [("cauchy_dispersion", 1.5), ("cauchy_dispersion", -0.1),
                                          ("backtrack_percent", 1.01), ("guide_alpha", -0.5)])
def test_out_of_range_parameters(ctx, setter, value):
    with pytest.raises(ts.TsgError) as e:
        getattr(builder(), setter)(value).build()
    assert e.value.kind == "InvalidRange"


def test_no_usable_example(ctx):
    with pytest.raises(ts.TsgError) as e:
        ts.Session.builder().output_size(ts.Dims.square(16)).build()
    assert e.value.kind == "NoExamples"
    b = ts.Session.builder().add_example(
        ts.Example(synth.example(2, 16, 16)).set_sample_method(ts.SampleMethod.Ignore)).output_size(ts.Dims.square(16))
    with pytest.raises(ts.TsgError) as e:
        b.build()
    assert e.value.kind == "NoExamples"


def test_guides_on_some_examples_only(ctx):
    b = ts.Session.builder().add_example(ts.Example(synth.example(3, 16, 16)).with_guide(synth.blobs(3, 16, 16)))
    b.add_example(synth.example(4, 16, 16)).load_target_guide(synth.blobs(5, 32, 32)).output_size(ts.Dims.square(32))
    with pytest.raises(ts.TsgError) as e:
        b.build()
    assert e.value.kind == "ExampleGuideMismatch"


def test_target_guide_of_the_wrong_size(ctx):
    L = ts.lib()
    ex = np.ascontiguousarray(synth.example(6, 16, 16))
    tg = np.ascontiguousarray(synth.blobs(6, 20, 20))
    eg = np.ascontiguousarray(synth.blobs(7, 16, 16))
    exs = (ts.Image * 1)(ts.Image(16, 16, ex.ctypes.data))
    egs = (ts.Image * 1)(ts.Image(16, 16, eg.ctypes.data))
    tgi = ts.Image(20, 20, tg.ctypes.data)
    p = ts.Params()
    L.tsg_params_default(C.byref(p))
    job = C.c_void_p()
    rc = L.tsg_job_create(ctx._h, C.byref(p), 1, exs, None, None, egs, C.byref(tgi), None, 0, 24, 24, C.byref(job))
    assert rc == -4 and not job.value


def test_transform_needs_every_map(ctx):
    ct = ts.CoordinateTransform(np.zeros(6, np.uint32), ts.Dims(2, 1), [ts.Dims(4, 4), ts.Dims(4, 4)])
    with pytest.raises(ts.TsgError) as e:
        ct.apply([np.zeros((4, 4, 4), np.uint8)], ctx)
    assert e.value.kind == "MapsCountMismatch"


def test_sample_mask_without_a_usable_pixel(ctx):
    # the reference loops forever drawing candidates (ms.rs:562-574)
    mask = np.zeros((16, 16, 4), np.uint8)
    b = ts.Session.builder().add_example(ts.Example(synth.example(8, 16, 16)).set_sample_method(mask))
    with pytest.raises(ts.TsgError) as e:
        b.seed(8).output_size(ts.Dims.square(24)).build().run()
    assert e.value.kind in ("Sampling", "InvalidArgument")


@pytest.mark.parametrize("setter,value", [("nearest_neighbors", 0), ("nearest_neighbors", 65),
                                          ("random_sample_locations", 5000)])
def test_outside_the_device_limits(ctx, setter, value):
    with pytest.raises(ts.TsgError):
        getattr(builder(), setter)(value).build()


def test_examples_beyond_the_gather_offsets(ctx):
    # gathers address an example level by a 32-bit byte offset, so an example
    # (or example guide) holds at most 2^30 pixels; the sizes are checked
    # before any pixel is read, so the descriptors point at a small buffer
    L = ts.lib()
    buf = np.zeros(64, np.uint8)
    p = ts.Params()
    L.tsg_params_default(C.byref(p))
    job = C.c_void_p()
    big = (ts.Image * 1)(ts.Image(32768, 32769, buf.ctypes.data))
    rc = L.tsg_job_create(ctx._h, C.byref(p), 1, big, None, None, None, None, None, 0, 16, 16, C.byref(job))
    assert rc == -11 and not job.value
    ex = np.ascontiguousarray(synth.example(9, 16, 16))
    tg = np.ascontiguousarray(synth.blobs(9, 16, 16))
    exs = (ts.Image * 1)(ts.Image(16, 16, ex.ctypes.data))
    tgi = ts.Image(16, 16, tg.ctypes.data)
    rc = L.tsg_job_create(ctx._h, C.byref(p), 1, exs, None, None, big, C.byref(tgi), None, 0, 16, 16, C.byref(job))
    assert rc == -11 and not job.value


def test_random_init_beyond_the_pixels(ctx):
    with pytest.raises(ts.TsgError):
        builder(out=4).random_init(17).build()


@pytest.mark.parametrize("out_w,out_h", [(1, 1), (1, 37), (37, 1), (13, 7), (3, 200)])
def test_degenerate_outputs_bit_exact(ctx, out_w, out_h):
    gen, job = run_pair([synth.example(20 + out_w, 24, 24)], out_w, out_h, seed=out_h)
    assert_same(gen, job)


def test_tiny_ragged_example_bit_exact(ctx):
    gen, job = run_pair([synth.example(30, 17, 16)], 40, 23, seed=4, backtrack_stages=3)
    assert_same(gen, job)


@pytest.mark.parametrize("knob,value", [("flow", 3), ("flow_team", 3), ("ahead", 3), ("team_max", 16),
                                        ("defer", 2), ("persist", 2), ("flow_head_team", 16)])
def test_tuning_out_of_range(ctx, knob, value):
    s = builder().build()
    with pytest.raises(ts.TsgError) as e:
        s.set_tuning(**{knob: value})
    assert e.value.kind == "InvalidArgument"
