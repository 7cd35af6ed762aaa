"""Style transfer (a target guide without example guides) on the device
against the oracle, bit-exact. The job prepares the guides itself, as
SessionBuilder::build does (session.rs:385-393, lib.rs:572-581): the target
guide becomes a blurred grayscale guide map, and each example's guide map is
matched to its histogram."""
import numpy as np
import pytest

from tsgpu import synth
from test_gpu_parity import assert_same, run_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_ex,out_w,out_h", [(1, 64, 64), (2, 90, 70)])
def test_style_transfer_bit_exact(ctx, n_ex, out_w, out_h):
    exs = [synth.example(900 + i, 40 + 8 * i, 40) for i in range(n_ex)]
    tg = synth.blobs(910, out_w, out_h)
    gen, job = run_pair(exs, out_w, out_h, target_guide=tg, seed=7, guide_alpha=0.8, random_init=5)
    assert_same(gen, job)


def test_style_transfer_follows_the_guide(ctx):
    # the guide steers the result: a different target guide changes it
    ex = [synth.example(920, 48, 48)]
    a, _ = run_pair(ex, 64, 64, target_guide=synth.blobs(921, 64, 64), seed=3)
    b, _ = run_pair(ex, 64, 64, target_guide=synth.blobs(922, 64, 64), seed=3)
    assert not np.array_equal(a.coord, b.coord)
