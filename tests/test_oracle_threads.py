"""The oracle's timing mode with the reference's own threading
(ms.rs:815-1036; TreeGrid ms.rs:1313-1530), which the CPU baseline times.

Pins: with one thread the grid is one cell over the image and the run must
equal the one-pixel-per-batch run bit for bit (state and resolved order);
with several threads the result depends on timing, as the reference's does,
so the checks are invariants (every pixel resolved once, sources inside the
example) and the stages' mean costs, seed-averaged, within 5% of the
sequential run (the first stages, with fewer than 1000 items to redo, run
serially and must match exactly)."""
import numpy as np
import pytest

from oracle import refcpu as R
from tsgpu import synth

SEQ = dict(px_per_batch=1, redo_shift=0, new_shift=0, max_batch=0, first_new_shift=0)


def job(seed, ex, out, **kw):
    return R.Job(R.default_params(seed=seed, **kw), [ex], out, out)


@pytest.mark.parametrize("seed,k,out", [(211, 30, 96), (5, 50, 72)])
def test_one_thread_equals_one_pixel_per_batch(seed, k, out):
    ex = synth.example(seed, 48, 48)
    a = job(seed, ex, out, nearest_neighbors=k).run(threads=1, **SEQ)
    b = job(seed, ex, out, nearest_neighbors=k).run_reference_threads(1)
    for x, y in zip(a.state(), b.state()):
        assert np.array_equal(x, y)
    fa, sa = a.resolved()
    fb, sb = b.resolved()
    assert np.array_equal(fa, fb) and np.array_equal(sa.view(np.uint32), sb.view(np.uint32))


def test_threads_resolve_every_pixel_once_near_the_sequential_costs():
    seq, thr = [], []
    for seed in range(4):
        ex = synth.example(300 + seed, 64, 64)
        a = job(seed, ex, 128).run(threads=1, **SEQ)
        b = job(seed, ex, 128).run_reference_threads(4)
        flat, _ = b.resolved()
        assert np.array_equal(np.sort(flat), np.arange(128 * 128))
        _, coord, _ = b.state()
        assert coord[..., 0].max() < 64 and coord[..., 1].max() < 64
        seq.append(a.stage_mean_costs())
        thr.append(b.stage_mean_costs())
    seq, thr = np.mean(seq, axis=0), np.mean(thr, axis=0)
    # 128^2: stages 0 and 1 have fewer than 1000 items to redo (serial)
    assert np.array_equal(seq[:2], thr[:2])
    assert np.all(np.abs(thr / seq - 1.0) < 0.05), thr / seq
