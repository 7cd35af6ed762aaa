"""The oracle's style-transfer guide steps (utils.rs:101-183), checked
against known answers and closed forms.

transform_to_guide_map blurs with sigma 2 (image 0.23's resampler at the same
size, gaussian(x, 2) with support 4), then takes luma 0.2126 R + 0.7152 G +
0.0722 B (f32, cast to u8). match_histograms moves each value v to the
position before the first target CDF entry above the source CDF at v (v
itself when there is none). The blur's bytes against the image crate are
parity unpinned; its invariants are checked here.
"""
import numpy as np

from oracle import refcpu as R


def solid(rgb, w=24, h=20):
    im = np.zeros((h, w, 4), np.uint8)
    im[..., :3] = rgb
    im[..., 3] = 255
    return im


def test_luma_known_answers():
    # a constant image stays constant under the blur, so the guide is the luma
    for rgb, want in [((0, 0, 0), 0), ((255, 0, 0), 54), ((0, 255, 0), 182), ((0, 0, 255), 18),
                      ((100, 100, 100), None)]:
        g = R.guide_map(solid(rgb))
        assert np.all(g[..., 3] == 255)
        assert np.all(g[..., 0] == g[0, 0, 0]) and np.all(g[..., 1] == g[..., 0]) and np.all(g[..., 2] == g[..., 0])
        if want is not None:
            assert g[0, 0, 0] == want, (rgb, g[0, 0, 0])
        else:
            assert g[0, 0, 0] in (99, 100)  # the coefficients sum to 1 up to f32 rounding


def test_blur_spreads_a_point_symmetrically():
    im = solid((0, 0, 0), 31, 31)
    im[15, 15, :3] = 255
    g = R.guide_map(im)[..., 0].astype(int)
    assert g[15, 15] == g.max() > 0
    assert np.array_equal(g, g[::-1, :]) and np.array_equal(g, g[:, ::-1]) and np.array_equal(g, g.T)
    assert g[15, 15 + 5] == 0  # the kernel's support is 2 sigma = 4 pixels


def test_match_histograms_closed_forms():
    # source: half a, half b; target: every value 0..255 equally often
    src = solid((0, 0, 0), 16, 16)
    src[:8, :, :3] = 40
    src[8:, :, :3] = 200
    tgt = np.zeros((256, 1, 4), np.uint8)
    tgt[:, 0, 0] = np.arange(256)
    out = R.match_histograms(src, tgt)
    # a: source CDF 0.5; the first target CDF above it is at 128 ((q + 1) / 256 > 0.5)
    assert np.all(out[:8, :, 0] == 127)
    # b: source CDF 1.0, no target entry above it: b stays
    assert np.all(out[8:, :, 0] == 200)
    assert np.all(out[..., 1] == out[..., 0]) and np.all(out[..., 3] == 255)
    # matched to itself, a value moves to just below the next value present
    self_m = R.match_histograms(src, src)
    assert np.all(self_m[:8, :, 0] == 199) and np.all(self_m[8:, :, 0] == 200)


def test_match_histograms_wraps_below_the_first_target_entry():
    # position 0 gives 0u8 - 1, which the release build wraps to 255
    src = solid((0, 0, 0), 8, 8)
    src[:2, :, :3] = 10  # 25%
    src[2:, :, :3] = 20  # 75%
    tgt = solid((0, 0, 0), 8, 8)
    tgt[4:, :, :3] = 100  # half 0, half 100
    out = R.match_histograms(src, tgt)
    assert np.all(out[:2, :, 0] == 255)  # source CDF 0.25 < target CDF at 0 (0.5)
    assert np.all(out[2:, :, 0] == 20)   # source CDF 1.0: no entry above, unchanged


def test_match_histograms_is_monotone_past_the_wrap():
    rng = np.random.default_rng(5)
    src = rng.integers(0, 256, (20, 30, 4), dtype=np.uint8)
    tgt = rng.integers(0, 80, (25, 25, 4), dtype=np.uint8)
    out = R.match_histograms(src, tgt)
    v, o = src[..., 0].ravel().astype(int), out[..., 0].ravel().astype(int)
    t0 = np.mean(tgt[..., 0] == 0)
    cdf = np.cumsum(np.bincount(v, minlength=256)) / v.size
    keep = cdf[v] >= t0  # values whose position is past 0
    order = np.argsort(v[keep], kind="stable")
    assert np.all(np.diff(o[keep][order]) >= 0)
