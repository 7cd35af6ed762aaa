"""Pins the CPU oracle (oracle/refcpu.cpp) without a GPU.

Each check uses a source independent of the oracle's own code: published
known-answer vectors of the PRNG crates, closed forms and special cases,
brute force on tiny inputs, and invariants of the reference's schedule
(lib/src/ms.rs).
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import refcpu as R
from tsgpu import synth

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def kat():
    with open(os.path.join(GOLD, "prng_kat.json")) as f:
        return json.load(f)


# ---------------------------------------------------------------- PRNG layer
def test_pcg32_demo_vector():
    k = kat()["pcg32_demo"]
    assert list(R.pcg_new_u32(k["state"], k["stream"], 6)) == k["next_u32"]


def test_from_seed_and_seed_from_u64_vectors():
    k = kat()
    assert int(R.pcg_from_seed_u64(bytes(k["from_seed_1_to_16"]["seed"]), 1)[0]) == k["from_seed_1_to_16"]["next_u64"]
    assert int(R.pcg_seed_u64(0, 1)[0]) == k["seed_from_u64_0"]["next_u64"]


def test_gen_range_value_stability_vector():
    k = kat()["value_stability_u32"]
    assert list(R.range_u32(k["state"], k["stream"], k["lo"], k["hi"], 3)) == k["sample_single"]


@pytest.mark.parametrize("bits", [0, 1, 5, 9, 31])
def test_gen_range_u32_power_of_two_closed_form(bits):
    # range 2^b: result = top b bits of the first draw whose bit (31-b) is clear
    # (Lemire zone = 2^31 - 1 for every power of two).
    st, stream = 1234567, 99
    n = 40
    got = R.range_u32(st, stream, 0, 1 << bits, n) if bits < 32 else None
    raw = R.pcg_new_u32(st, stream, 400)
    out, i = [], 0
    while len(out) < n:
        v = int(raw[i])
        i += 1
        if (v >> (31 - bits)) & 1 == 0:
            out.append(v >> (32 - bits) if bits else 0)
    assert list(got) == out


@pytest.mark.parametrize("bits", [0, 3, 20, 63])
def test_gen_range_u64_power_of_two_closed_form(bits):
    st, stream = 42, 7
    n = 30
    got = R.range_u64(st, stream, 0, 1 << bits, n)
    raw = R.pcg_new_u32(st, stream, 800)
    out, i = [], 0
    while len(out) < n:
        v = int(raw[i]) | (int(raw[i + 1]) << 32)  # next_u64: low word first
        i += 2
        if (v >> (63 - bits)) & 1 == 0:
            out.append(v >> (64 - bits) if bits else 0)
    assert [int(x) for x in got] == out


def test_gen_range_bounds_and_offset():
    v = R.range_u32(5, 6, 100, 107, 2000)
    assert v.min() >= 100 and v.max() < 107
    assert set(v.tolist()) == set(range(100, 107))


# ---------------------------------------------------------------- cost tables
def test_cauchy_table_closed_forms():
    my, g = R.luts_full(1.0, 0.0, False)
    assert np.all(np.diag(my) == 0.0)
    assert np.array_equal(my, my.T)
    # depends only on |a-b|
    for d in (0, 1, 17, 128, 255):
        vals = {float(my[a, a + d]) for a in range(0, 256 - d, 13)}
        assert len(vals) == 1
    assert my[0, 255] == np.float32(math.log(2.0))  # log1p(1) = ln 2
    assert np.all(g == 0.0)
    my2, _ = R.luts_full(0.5, 0.0, False)
    assert abs(float(my2[0, 255]) - math.log(5.0)) < 1e-6  # log1p(1 / 0.25)
    my3, g3 = R.luts_full(1.0, 0.25, True)
    assert g3[0, 255] == np.float32(0.25)  # alpha * ((0-255)/255)^2
    assert my3[0, 255] == np.float32(np.float32(0.75) * np.float32(math.log(2.0)))
    # monotone in |a-b|
    row = my[0, :]
    assert np.all(np.diff(row) > 0)


# ---------------------------------------------------------------- pyramid
@pytest.mark.parametrize("w,h,levels", [(64, 64, 5), (300, 200, 5), (37, 29, 3), (16, 16, 1)])
def test_pyramid_invariants(w, h, levels):
    img = synth.example(7, w, h)
    pyr = R.pyramid(img, levels)
    assert pyr.shape == (max(levels, 1), h, w, 4)
    assert np.array_equal(pyr[-1], img)
    const = np.full((h, w, 4), 77, np.uint8)
    assert np.all(R.pyramid(const, levels) == 77)
    # coarser levels are smoother: total variation decreases towards level 0
    tv = [np.abs(np.diff(p[..., :3].astype(int), axis=1)).mean() for p in pyr]
    assert all(tv[i] <= tv[i + 1] + 1e-9 for i in range(len(tv) - 1))


def _resize_f64(img, nw, nh, support, kern):
    """Independent float64 separable resampler (same sampling geometry)."""
    def weights(n_in, n_out):
        ratio = n_in / n_out
        sr = max(1.0, ratio)
        rows = []
        for o in range(n_out):
            c = (o + 0.5) * ratio
            lo = min(max(int(math.floor(c - support * sr)), 0), n_in - 1)
            hi = min(max(int(math.ceil(c + support * sr)), lo + 1), n_in)
            ws = np.array([kern((i - (c - 0.5)) / sr) for i in range(lo, hi)])
            rows.append((lo, ws / ws.sum()))
        return rows
    x = img.astype(np.float64)
    tmp = np.zeros((nh, img.shape[1], 4))
    for o, (lo, ws) in enumerate(weights(img.shape[0], nh)):
        tmp[o] = np.tensordot(ws, x[lo:lo + len(ws)], axes=1)
    out = np.zeros((nh, nw, 4))
    for o, (lo, ws) in enumerate(weights(img.shape[1], nw)):
        out[:, o] = np.tensordot(ws, tmp[:, lo:lo + len(ws)], axes=(0, 1))
    return np.clip(np.round(out), 0, 255)


@pytest.mark.parametrize("filt,support,kern", [
    ("gaussian", 3.0, lambda x: math.exp(-x * x / 0.5)),
    ("triangle", 1.0, lambda x: max(0.0, 1.0 - abs(x))),
    ("catmullrom", 2.0, lambda x: ((1.5 * abs(x) ** 3 - 2.5 * x * x + 1) if abs(x) < 1 else
                                   (-0.5 * abs(x) ** 3 + 2.5 * x * x - 4 * abs(x) + 2) if abs(x) < 2 else 0.0)),
])
@pytest.mark.parametrize("shape", [(64, 48, 32, 24), (50, 40, 100, 80), (33, 21, 7, 5)])
def test_resize_matches_float64_resampler(filt, support, kern, shape):
    w, h, nw, nh = shape
    img = synth.example(3, w, h)
    got = R.resize(img, nw, nh, filt).astype(int)
    ref = _resize_f64(img, nw, nh, support, kern)
    assert np.abs(got - ref).max() <= 1


# ---------------------------------------------------------------- kNN
def _brute_knn(mask, qx, qy, k):
    ys, xs = np.nonzero(mask)
    dx, dy = xs - qx, ys - qy
    order = np.lexsort((dx, dy, dx * dx + dy * dy))
    return np.stack([xs[order][:k], ys[order][:k]], axis=1)


@pytest.mark.parametrize("w,h,density,k", [(16, 16, 0.3, 5), (40, 30, 0.05, 12), (70, 70, 0.6, 50),
                                           (200, 200, 0.02, 50), (150, 90, 0.001, 8)])
def test_knn_matches_brute_force(w, h, density, k):
    rng = np.random.default_rng(w * 1000 + k)
    mask = (rng.random((h, w)) < density).astype(np.uint8)
    q = rng.integers(0, w * h, size=60, dtype=np.uint32)
    nbrs, counts = R.knn(mask, q, k)
    for i, f in enumerate(q):
        ref = _brute_knn(mask, int(f % w), int(f // w), k)
        assert counts[i] == len(ref)
        assert np.array_equal(nbrs[i, :counts[i]], ref)


# ---------------------------------------------------------------- schedule
def test_stage_sizes_and_permutation():
    ex = synth.example(1, 32, 32)
    job = R.Job(R.default_params(seed=3), [ex], 64, 64).run()
    sums, cnts, items = job.stage_stats()
    N = 64 * 64
    # S_s = p^s N for s = P..0; 63/32 N resolves in total (32 | N)
    assert list(items) == [N // 32, N // 16, N // 8, N // 4, N // 2, N]
    assert items.sum() == 63 * N // 32
    flat, score = job.resolved()
    assert len(flat) == N and len(np.unique(flat)) == N
    m = job.mask()
    assert m.all()


def test_jacobi_batch_of_one_is_sequential():
    ex = synth.example(2, 24, 24)
    a = R.Job(R.default_params(seed=9), [ex], 40, 40).run()
    b = R.Job(R.default_params(seed=9), [ex], 40, 40).run(threads=4, px_per_batch=1, redo_shift=0, new_shift=0)
    for x, y in zip(a.state(), b.state()):
        assert np.array_equal(x, y)


def test_jacobi_is_thread_count_independent():
    ex = synth.example(4, 32, 32)
    kw = dict(px_per_batch=0, redo_shift=2, new_shift=3, max_batch=0)
    a = R.Job(R.default_params(seed=5), [ex], 48, 48).run(threads=1, **kw)
    b = R.Job(R.default_params(seed=5), [ex], 48, 48).run(threads=8, **kw)
    for x, y in zip(a.state(), b.state()):
        assert np.array_equal(x, y)


def _u32_stream_of_seed(seed, n):
    """next_u32 draws of Pcg32::seed_from_u64(seed), from its next_u64 draws."""
    out = []
    for v in R.pcg_seed_u64(seed, (n + 1) // 2):
        v = int(v)
        out += [v & 0xFFFFFFFF, v >> 32]
    return out[:n]


def test_first_pixel_resolves_at_random_with_stage_seed():
    # ms.rs:1003-1009: no neighbour -> resolve_at_random(p_stage_seed): map,
    # x and y each from a fresh Pcg32::seed_from_u64(p_stage_seed).
    ex = synth.example(5, 40, 40)
    job = R.Job(R.default_params(seed=11), [ex], 32, 32)
    job.run_until(0, 1)
    flat, _ = job.resolved()
    _, coord, ids = job.state()
    pss = job.stage_info()["stage_seed"]
    # gen_range(0..40) on u32: Lemire zone (40 << 26) - 1
    zone = ((40 << 26) & 0xFFFFFFFF) - 1
    x = None
    for v in _u32_stream_of_seed(pss, 64):
        m = v * 40
        if (m & 0xFFFFFFFF) <= zone:
            x = m >> 32
            break
    first = int(flat[0])
    assert (coord[first][0], coord[first][1], coord[first][2]) == (x, x, 0)
    assert ids[first][0] == first and ids[first][1] == 0


# ---------------------------------------------------------------- frozen-state description
def _describe(seed=13, out=64, stage_t=3, item=40, n=64, k=50, m=50):
    ex = synth.example(seed, 48, 48)
    job = R.Job(R.default_params(seed=seed, nearest_neighbors=k, random_sample_locations=m), [ex], out, out)
    job.run_until(stage_t, item)
    info = job.stage_info()
    pyr = R.pyramid(ex, 5)
    rgba, coord, ids = job.state()
    d = job.describe(n)
    return job, info, pyr[info["level"]], rgba, coord, ids, d


def test_best_match_is_first_minimum_of_full_costs():
    _, _, _, _, _, _, d = _describe()
    for i in range(len(d["flat"])):
        nc = d["cand_count"][i]
        fc = d["full_cost"][i, :nc]
        b = d["best"][i]
        assert fc[b] == fc.min()
        assert b == int(np.argmin(fc))  # first index among equal minima
        assert d["best_score"][i] == fc.min()


def test_weights_normalisation_and_self_weight():
    job, info, _, _, _, _, d = _describe(stage_t=2, item=5)
    for i in range(len(d["flat"])):
        kf = d["nbr_count"][i]
        w = d["weights"][i, :kf].astype(np.float64)
        assert np.all(w > 0) and np.all(w <= 1.0)
        assert np.all(np.diff(w) <= 0)  # neighbours come nearest first
        # redo items are their own nearest neighbour at distance 0 -> w = 1
        x, y = d["flat"][i] % 64, d["flat"][i] // 64
        if d["nbrs"][i, 0, 0] == x and d["nbrs"][i, 0, 1] == y:
            assert w[0] == 1.0
        # mean of d/mean is 1: mean of -ln(w) == 1 (up to f32 rounding of w)
        assert abs(np.mean(-np.log(w)) - 1.0) < 1e-5


def test_candidate_costs_match_float64_recomputation():
    """Recompute every candidate cost in float64 from the reference's rules:
    coherence patterns at cand + (n - p), random patterns mirrored at
    cand - (n - p) (ms.rs:534, 588), outside pixels [0,0,0,255]."""
    job, info, level, rgba, coord, ids, d = _describe(stage_t=3, item=30, n=24)
    my, _ = job.stage_luts()
    lut = my.astype(np.float64)
    H, W = level.shape[:2]
    mirrored_differs = 0
    for i in range(len(d["flat"])):
        p = np.array([d["flat"][i] % 64, d["flat"][i] // 64])
        kf = d["nbr_count"][i]
        nb = d["nbrs"][i, :kf]
        pat = rgba[nb[:, 1], nb[:, 0]].astype(int)
        w = d["weights"][i, :kf].astype(np.float64)
        for c in range(d["cand_count"][i]):
            cx, cy, mp, kind = d["cands"][i, c, :4]
            sign = -1 if kind == 1 else 1

            def cost(sg):
                pc = np.array([cx, cy]) + sg * (nb - p)
                inb = (pc[:, 0] >= 0) & (pc[:, 1] >= 0) & (pc[:, 0] < W) & (pc[:, 1] < H)
                px = np.tile(np.array([0, 0, 0, 255]), (kf, 1))
                px[inb] = level[pc[inb, 1], pc[inb, 0]]
                return float((lut[np.abs(pat - px)].sum(axis=1) * w).sum())
            ref = cost(sign)
            got = float(d["full_cost"][i, c])
            assert abs(got - ref) <= 1e-5 * max(1.0, ref)
            if kind == 1 and abs(cost(-sign) - got) > 1e-4 * max(1.0, got):
                mirrored_differs += 1
    assert mirrored_differs > 0  # the mirrored random pattern is in effect


def test_coherence_candidates_follow_neighbour_sources():
    job, info, level, rgba, coord, ids, d = _describe(stage_t=4, item=100, n=32)
    H, W = level.shape[:2]
    for i in range(len(d["flat"])):
        p = np.array([d["flat"][i] % 64, d["flat"][i] // 64])
        kf = d["nbr_count"][i]
        exp = []
        for j in range(kf):
            n = d["nbrs"][i, j]
            f = n[1] * 64 + n[0]
            c = coord[f, :2].astype(int) + (p - n)
            if 0 <= c[0] < W and 0 <= c[1] < H:
                exp.append((c[0], c[1], ids[f, 1], ids[f, 0]))
        got = [tuple(int(v) for v in d["cands"][i, c, [0, 1, 2, 4]]) for c in range(len(exp))]
        assert got == exp
        assert all(d["cands"][i, c, 3] == 0 for c in range(len(exp)))
        assert all(d["cands"][i, c, 3] == 1 for c in range(len(exp), d["cand_count"][i]))


def test_sample_mask_restricts_random_candidates():
    ex = synth.example(8, 40, 40)
    mask = np.zeros((40, 40, 4), np.uint8)
    mask[:, 20:, 0] = 255
    job = R.Job(R.default_params(seed=4), [ex], 48, 48, sample_kind=[2], sample_masks=[mask])
    job.run_until(2, 10)
    d = job.describe(32)
    for i in range(len(d["flat"])):
        for c in range(d["cand_count"][i]):
            assert d["cands"][i, c, 0] >= 20


def test_inpaint_keeps_locked_pixels():
    ex = synth.example(9, 48, 48)
    m = synth.ellipse_mask(9, 48, 48)
    job = R.Job(R.default_params(seed=2), [ex], 48, 48, inpaint_mask=m, inpaint_color=ex, inpaint_index=0)
    job.run()
    rgba, coord, ids = job.state()
    keep = m[..., 0] == 255
    pyr = R.pyramid(ex, 5)
    assert np.array_equal(rgba[keep], pyr[-1][keep])
    ys, xs = np.nonzero(keep)
    assert np.array_equal(coord.reshape(48, 48, 3)[keep][:, 0], xs)
    assert job.locked() == int(keep.sum())


def test_random_init_locks_pixels_from_finest_level():
    exs = [synth.example(20 + i, 32, 32) for i in range(3)]
    job = R.Job(R.default_params(seed=211, random_init=10), exs, 40, 40)
    job.run_until(0, 0)
    assert job.locked() == 10
    flat, score = job.resolved()
    rgba, coord, _ = job.state()
    for f in flat[:10]:
        x, y, mp = coord[f]
        assert x == y  # identical streams draw the same first value on square maps
        assert np.array_equal(rgba.reshape(-1, 4)[f], exs[mp][y, x])


def test_guided_alpha_schedule():
    # adaptive alpha (ms.rs:846-851): (alpha * (1 - resolved / N))^3 for
    # p_stage > 0, 0 at the last stage.
    ex = synth.example(1, 32, 32)
    g = synth.blobs(2, 32, 32)
    tg = synth.blobs(3, 48, 48)
    job = R.Job(R.default_params(seed=1), [ex], 48, 48, example_guides=[g], target_guide=tg)
    alphas = []
    for t in range(6):
        job.run_until(t, 0)
        alphas.append(job.stage_info()["alpha"])
    b = np.float32(0.8)
    assert np.float32(alphas[0]) == (b * b) * b
    b1 = np.float32(0.8) * (np.float32(1.0) - np.float32(72) / np.float32(48 * 48))
    assert np.float32(alphas[1]) == (b1 * b1) * b1
    assert alphas[-1] == 0.0
    assert all(alphas[i] > alphas[i + 1] for i in range(5))
