"""tsg_knn_batch_noalloc (kNN with every device buffer in the library's
fixed device scratch array, no allocation per call) gives the same lists and
counts as tsg_knn_batch, on consecutive calls that reuse the array (the
second smaller than the first), on every kNN route, and rejects inputs that
do not fit the array."""
import numpy as np
import pytest

import tsgpu as ts

pytestmark = pytest.mark.gpu


def _case(w, h, density, n_q, k, seed):
    rng = np.random.default_rng(seed)
    mask = (rng.random((h, w)) < density).astype(np.uint8)
    q = rng.integers(0, w * h, size=n_q, dtype=np.uint32)
    return mask, q, k


@pytest.mark.parametrize("first,second", [
    ((512, 512, 0.3, 4000, 50, 1), (64, 48, 0.5, 300, 20, 2)),      # table route, then list route
    ((1024, 768, 0.002, 3000, 50, 3), (256, 256, 0.01, 100, 64, 4)),  # list / ring routes
    ((2048, 2048, 0.05, 20000, 50, 5), (37, 53, 0.7, 7, 7, 6)),       # large mask, then tiny ragged one
])
def test_noalloc_matches_allocating_entry(ctx, first, second):
    for w, h, dens, n_q, k, seed in (first, second):
        print(f"case: knn noalloc {w}x{h} density {dens} queries {n_q} k {k}", flush=True)
        mask, q, k = _case(w, h, dens, n_q, k, seed)
        a_n, a_c = ts.knn(mask, q, k, ctx)
        b_n, b_c = ts.knn_noalloc(mask, q, k, ctx)
        assert np.array_equal(a_c, b_c)
        assert np.array_equal(a_n, b_n)


def test_noalloc_rejects_inputs_beyond_the_scratch(ctx):
    mask = np.ones((4096, 4096), np.uint8)  # 16 MiB of mask + 2 MiB of tiles
    q = np.arange(30000, dtype=np.uint32)   # + 30000 * (12 * 50 + 24) bytes > 32 MiB
    with pytest.raises(ts.TsgError) as e:
        ts.knn_noalloc(mask, q, 50, ctx)
    assert e.value.kind == "Unsupported"
    # the array still serves a call that fits
    m2, q2, k2 = _case(128, 128, 0.4, 500, 30, 9)
    assert np.array_equal(ts.knn_noalloc(m2, q2, k2, ctx)[0], ts.knn(m2, q2, k2, ctx)[0])
