"""Tiling mode on the device against the oracle, bit-exact.

In tiling mode a resolved pixel within 5% of an edge also stands one image
width (height) away on the far side when the kNN searches (flush_resolved,
ms.rs:296-331), and the own pattern wraps around the image (ms.rs:1132-1136).
The cases reach every kNN route with ghost points: the resolved list (few
points), the ring search over 8x8 tiles (sparse, more than 2048 points) and
the offset table (dense), with output sides that are not multiples of 8 so
that edge tiles hold both image pixels and ghosts.
"""
import numpy as np
import pytest

import tsgpu as ts
from tsgpu import synth
from test_gpu_parity import DEFAULT, assert_same, run_pair

pytestmark = pytest.mark.gpu

SEQ = dict(px_per_batch=1, redo_shift=0, new_shift=0, max_batch=0, first_new_shift=0)


@pytest.mark.parametrize("ex,out_w,out_h,k", [
    (48, 100, 76, 20),     # list and table routes, partial edge tiles
    (64, 128, 128, 50),
    (96, 518, 390, 50),    # ring route in the first stages
])
def test_tiling_bit_exact(ctx, ex, out_w, out_h, k):
    gen, job = run_pair([synth.example(700 + ex, ex, ex)], out_w, out_h, seed=17, nearest_neighbors=k,
                        tiling_mode=True)
    assert_same(gen, job)


def test_tiling_one_pixel_per_batch(ctx):
    gen, job = run_pair([synth.example(711, 32, 32)], 44, 36, sched=SEQ, threads=1, seed=5, nearest_neighbors=12,
                        tiling_mode=True)
    assert_same(gen, job)


def test_tiling_guided_and_random_init(ctx):
    exs = [synth.example(720 + i, 40, 40) for i in range(2)]
    gds = [synth.blobs(730 + i, 40, 40) for i in range(2)]
    tg = synth.blobs(740, 90, 70)
    gen, job = run_pair(exs, 90, 70, example_guides=gds, target_guide=tg, seed=9, random_init=7,
                        guide_alpha=0.8, tiling_mode=True)
    assert_same(gen, job)


def test_tiling_inpaint(ctx):
    ex = synth.example(750, 80, 80)
    m = synth.ellipse_mask(750, 80, 80)
    gen, job = run_pair([ex], 80, 80, seed=3, inpaint_mask=m, inpaint_index=0, sample_kind=[2], sample_masks=[m],
                        tiling_mode=True)
    assert_same(gen, job)


def test_tiling_changes_the_result(ctx):
    # the ghosts matter: the same job without tiling differs near the edges
    ex = synth.example(760, 48, 48)
    a, _ = run_pair([ex], 96, 96, seed=11, tiling_mode=True)
    b, _ = run_pair([ex], 96, 96, seed=11)
    diff = np.any(a.coord.reshape(96, 96, 3) != b.coord.reshape(96, 96, 3), axis=2)
    assert diff.any()


def test_tiling_in_the_engine(ctx):
    # the multi-job engine gives a tiling job the result it gets alone
    ex = synth.example(770, 48, 48)
    def build():
        s = ts.Session.builder().add_example(ex).seed(21).tiling_mode(True).output_size(ts.Dims(84, 60)).build()
        s.set_schedule(**DEFAULT)
        return s
    alone = build().run()
    batch = ts.Batch(ctx, max_active=4)
    other = batch.submit(ts.Session.builder().add_example(synth.example(771, 40, 40)).seed(2).output_size(
        ts.Dims.square(64)).build())
    t = batch.submit(build())
    batch.run_all()
    got = batch.take(t)
    batch.take(other)
    batch.close()
    assert np.array_equal(got.coord, alone.coord) and np.array_equal(got.image, alone.image)
