"""Parity of the HIP path with the CPU oracle, through the C ABI.

Bit-exact everywhere: pyramids (u8), kNN lists (canonical order), random
candidate streams, per-candidate costs (f32, reference summation order),
argmin (first minimum), and whole images under the same Jacobi schedule.
With one pixel per batch the device reproduces the reference at
max_thread_count(1).
"""
import numpy as np
import pytest

import tsgpu as ts
from oracle import refcpu as R
from tsgpu import synth

pytestmark = pytest.mark.gpu

DEFAULT = ts.default_schedule()


def run_pair(examples, out_w, out_h, sched=DEFAULT, threads=8, sample_kind=None, sample_masks=None,
             example_guides=None, target_guide=None, inpaint_mask=None, inpaint_index=0, tuning=None, **params):
    """Same job on the device (through the Session mirror) and on the oracle."""
    prm = R.default_params(**params)
    job = R.Job(prm, examples, out_w, out_h, sample_kind=sample_kind, sample_masks=sample_masks,
                example_guides=example_guides, target_guide=target_guide, inpaint_mask=inpaint_mask,
                inpaint_color=examples[inpaint_index] if inpaint_mask is not None else None,
                inpaint_index=inpaint_index)
    job.run(threads=threads, **sched)
    b = ts.Session.builder()
    for i, e in enumerate(examples):
        if inpaint_mask is not None and i == inpaint_index:
            continue
        ex = ts.Example(e)
        if example_guides is not None:
            ex.with_guide(example_guides[i])
        kind = (sample_kind or [0] * len(examples))[i]
        if kind == 1:
            ex.set_sample_method(ts.SampleMethod.Ignore)
        elif kind == 2:
            ex.set_sample_method(sample_masks[i])
        if inpaint_mask is not None and i > inpaint_index:
            raise ValueError("keep the inpainted example last in these tests")
        b.add_example(ex)
    if inpaint_mask is not None:
        ex = ts.Example(examples[inpaint_index])
        kind = (sample_kind or [0] * len(examples))[inpaint_index]
        if kind == 2:
            ex.set_sample_method(sample_masks[inpaint_index])
        b.inpaint_example(inpaint_mask, ex, ts.Dims(out_w, out_h))
    if target_guide is not None:
        b.load_target_guide(target_guide)
    b.output_size(ts.Dims(out_w, out_h))
    for k, v in params.items():
        getattr(b, k)(v)
    sess = b.build()
    sess.set_schedule(**sched)
    if tuning:
        sess.set_tuning(**tuning)
    gen = sess.run()
    return gen, job


def assert_same(gen, job):
    rgba, coord, ids = job.state()
    flat, score = job.resolved()
    assert np.array_equal(gen.coord, coord), "coord map"
    assert np.array_equal(gen.ids, ids), "id map"
    assert np.array_equal(gen.image, rgba), "colour map"
    assert np.array_equal(gen.resolved_flat[:len(flat)], flat), "resolved list order"
    assert np.array_equal(gen.resolved_score[:len(flat)].view(np.uint32), score.view(np.uint32)), "scores"
    sums, cnts, items = job.stage_stats()
    gs, gc, gi = gen.stage_stats
    assert np.array_equal(gc, cnts) and np.array_equal(gi, items)
    assert np.allclose(gs, sums, rtol=1e-9, atol=0)


# ------------------------------------------------------------------ pyramid
@pytest.mark.parametrize("w,h,levels", [(64, 64, 5), (300, 300, 5), (512, 512, 5), (37, 29, 3), (100, 60, 1)])
def test_pyramid_bit_exact(ctx, w, h, levels):
    img = synth.example(w + h, w, h)
    assert np.array_equal(ts.pyramid(img, levels, ctx), R.pyramid(img, levels))


# ------------------------------------------------------------------ kNN
def _hole_mask(w, h, r):
    ys, xs = np.mgrid[0:h, 0:w]
    return (((xs - w // 2) ** 2 + (ys - h // 2) ** 2) > r * r).astype(np.uint8)


@pytest.mark.parametrize("w,h,density,k", [
    (64, 64, 0.2, 50),      # list route (resolved <= 2048)
    (64, 64, 0.9, 50),      # table route, dense
    (300, 200, 0.05, 64),   # table route
    (512, 512, 0.012, 50),  # table route with ring fallback
    (1024, 512, 0.002, 20),  # list route, sparse
    (256, 256, 0.5, 1),
    (37, 53, 0.7, 7),
])
def test_knn_bit_exact(ctx, w, h, density, k):
    rng = np.random.default_rng(w * 7 + h + k)
    mask = (rng.random((h, w)) < density).astype(np.uint8)
    q = rng.integers(0, w * h, size=3000, dtype=np.uint32)
    g_n, g_c = ts.knn(mask, q, k, ctx)
    o_n, o_c = R.knn(mask, q, k)
    assert np.array_equal(g_c, o_c)
    for i in range(len(q)):
        assert np.array_equal(g_n[i, :g_c[i]], o_n[i, :o_c[i]]), i


def test_knn_bit_exact_inside_a_hole(ctx):
    # inpainting: queries deep inside an unresolved disc use the ring search
    w = h = 600
    mask = _hole_mask(w, h, 200)
    ys, xs = np.nonzero(mask == 0)
    rng = np.random.default_rng(5)
    sel = rng.choice(len(xs), 500, replace=False)
    q = (ys[sel] * w + xs[sel]).astype(np.uint32)
    g_n, g_c = ts.knn(mask, q, 50, ctx)
    o_n, o_c = R.knn(mask, q, 50)
    assert np.array_equal(g_c, o_c)
    assert np.array_equal(g_n, o_n)


# ------------------------------------------------------------------ frozen state
JOB_KW = ("sample_kind", "sample_masks", "example_guides", "target_guide")


def _frozen(examples, out, stage_t, item, n, threads=8, **kw):
    """Oracle state frozen at (stage, item) under the default Jacobi schedule,
    and its description of the next n work items."""
    params = {k: v for k, v in kw.items() if k not in JOB_KW}
    job = R.Job(R.default_params(**params), examples, out, out, **{k: v for k, v in kw.items() if k in JOB_KW})
    job.run_until(stage_t, item, threads=threads, **DEFAULT)
    info = job.stage_info()
    rgba, coord, ids = job.state()
    my, gl = job.stage_luts()
    d = job.describe(n)
    return job, info, rgba, coord, ids, my, gl, d


def _score_on_device(ctx, info, rgba, coord, ids, my, gl, d, examples, k, sample_kind=None, example_guides=None,
                     target_guide=None, team=1):
    lv = info["level"]
    kinds = sample_kind or [0] * len(examples)
    maps = [R.pyramid(e, 5)[lv] for e, kd in zip(examples, kinds) if kd != 1]
    guides = tg = None
    if target_guide is not None:
        guides = [R.pyramid(g, 5)[lv] for g in example_guides][:len(maps)]
        tg = R.pyramid(target_guide, 5)[lv]
    return ts.score_batch(rgba, coord, ids, maps, my, k, d["flat"], d["nbr_count"], d["nbrs"], d["cand_count"],
                          d["cands"], guides=guides, target_guide=tg, lut_guide=gl if guides else None, ctx=ctx,
                          team=team)


def _check_scores(d, full, best, bs, bp, bsp):
    ok = d["best"] != 0xFFFFFFFF
    for i in np.nonzero(ok)[0]:
        nc = d["cand_count"][i]
        assert np.array_equal(full[i, :nc].view(np.uint32), d["full_cost"][i, :nc].view(np.uint32)), i
    # unpruned lane-per-candidate evaluation
    assert np.array_equal(best[ok], d["best"][ok])
    assert np.array_equal(bs[ok].view(np.uint32), d["best_score"][ok].view(np.uint32))
    # production evaluation: coherence dedup + exact pruning gives the same winner
    assert np.array_equal(bp[ok], d["best"][ok])
    assert np.array_equal(bsp[ok].view(np.uint32), d["best_score"][ok].view(np.uint32))


@pytest.mark.parametrize("stage_t,item,team", [(4, 0, 1), (4, 16384, 1), (2, 3000, 1), (4, 0, 8), (2, 3000, 2),
                                                (4, 16384, 4)])
def test_frozen_state_scoring_bit_exact_256(ctx, stage_t, item, team):
    # team > 1: the candidate list split between waves, winner merged
    exs = [synth.example(211, 128, 128)]
    job, info, rgba, coord, ids, my, gl, d = _frozen(exs, 256, stage_t, item, 4096, seed=211)
    assert len(d["flat"]) > 100
    full, best, bs, bp, bsp = _score_on_device(ctx, info, rgba, coord, ids, my, gl, d, exs, 50, team=team)
    _check_scores(d, full, best, bs, bp, bsp)


def test_frozen_state_scoring_guided_multi_example(ctx):
    exs = [synth.example(1000 + i, 60, 60) for i in range(3)]
    gds = [synth.blobs(2000 + i, 60, 60) for i in range(3)]
    tg = synth.blobs(3000, 96, 96)
    job, info, rgba, coord, ids, my, gl, d = _frozen(exs, 96, 3, 500, 2048, seed=211, example_guides=gds,
                                                     target_guide=tg)
    assert info["alpha"] > 0
    for team in (1, 8):
        full, best, bs, bp, bsp = _score_on_device(ctx, info, rgba, coord, ids, my, gl, d, exs, 50,
                                                   example_guides=gds, target_guide=tg, team=team)
        _check_scores(d, full, best, bs, bp, bsp)


def test_random_candidates_bit_exact(ctx):
    exs = [synth.example(31, 50, 40), synth.example(32, 70, 64)]
    mk = np.zeros((40, 50, 4), np.uint8)
    mk[5:30, 10:45, 0] = 255
    job, info, *_, d = _frozen(exs, 64, 3, 200, 512, seed=7, sample_kind=[2, 0], sample_masks=[mk, None])
    lv = info["level"]
    maps = [R.pyramid(e, 5)[lv] for e in exs]
    got = ts.random_candidates(maps, [mk, None], info["stage_seed"], info["cursor"], len(d["flat"]), 50, ctx)
    for i in range(len(d["flat"])):
        nc = d["cand_count"][i]
        if nc == 0:
            continue
        rnd = d["cands"][i, nc - 50:nc, :3].astype(np.uint32)
        assert np.array_equal(got[i], rnd), i


def test_frozen_state_scoring_bit_exact_1024(ctx):
    # the C2 shape: 512^2 example -> 1024^2, largest stage, 4096 items per batch
    exs = [synth.example(512, 512, 512)]
    job, info, rgba, coord, ids, my, gl, d = _frozen(exs, 1024, 5, 0, 4096, threads=16, seed=0)
    full, best, bs, bp, bsp = _score_on_device(ctx, info, rgba, coord, ids, my, gl, d, exs, 50)
    _check_scores(d, full, best, bs, bp, bsp)
    job2, info2, rgba2, coord2, ids2, my2, gl2, d2 = _frozen(exs, 1024, 5, 600000, 4096, threads=16, seed=0)
    full, best, bs, bp, bsp = _score_on_device(ctx, info2, rgba2, coord2, ids2, my2, gl2, d2, exs, 50)
    _check_scores(d2, full, best, bs, bp, bsp)


# ------------------------------------------------------------------ whole images
def test_one_pixel_per_batch_equals_reference_single_thread_64(ctx):
    ex = synth.example(211, 48, 48)
    seq = dict(px_per_batch=1, redo_shift=0, new_shift=0, max_batch=0)
    gen, job = run_pair([ex], 64, 64, sched=seq, threads=1, seed=211)
    assert_same(gen, job)


@pytest.mark.parametrize("seed", [0, 211])
def test_default_schedule_bit_exact_c1(ctx, seed):
    ex = synth.example(211, 128, 128)
    gen, job = run_pair([ex], 256, 256, seed=seed)
    assert_same(gen, job)


@pytest.mark.parametrize("tuning", [
    dict(small_max=0, fuse_max=0, team_max=1, prefetch_max=0, ahead=0, coop_wgs=0, coop_max=0),
    dict(small_max=4, fuse_max=64, team_max=2, prefetch_max=64, ahead=0, coop_wgs=0, coop_max=0),
    dict(small_max=16, fuse_max=100000, team_max=8, team_waves=1 << 20, prefetch_max=100000, ahead=0,
         coop_wgs=0, coop_max=0),
    dict(small_max=16, fuse_max=2048, team_max=4, ahead=1, coop_wgs=0, coop_max=0),
    dict(coop_wgs=8, coop_max=128, ahead=0),
    dict(coop_wgs=256, coop_max=4096, team_max=8, flow=0),
    dict(flow=0),
    dict(flow=0, ahead=2),
    dict(ahead=1),
    dict(flow=1, flow_team=1, flow_wgs=3),
    dict(flow=1, flow_team=2, flow_wgs=1),
    dict(flow=1, flow_team=4, prefetch_max=64),
    dict(flow=2),
    dict(flow=2, flow_team=1, flow_wgs=3),
    dict(flow=2, flow_team=2, flow_wgs=1, prefetch_max=64),
    dict(flow=2, ahead=2),
    dict(defer=0),
    dict(defer=1, flow=0, ahead=0, prefetch_max=0, small_max=0, team_max=1),
    dict(defer=1, flow=0, ahead=0, prefetch_max=64, small_max=16, fuse_max=64),
    dict(defer=1, coop_wgs=8, coop_max=128, ahead=0),
    dict(persist=1),
    dict(persist=1, team_max=1, ahead=0, flow=0, prefetch_max=0, small_max=0),
    dict(persist=1, team_waves=64),
    dict(flow_head=0),
    dict(flow_head=64, flow_head_team=2, flow_wgs=5),
    dict(flow=2, flow_head=100, flow_head_team=8),
])
def test_launch_tuning_does_not_change_results(ctx, tuning):
    # every launch structure resolves the same batches against the same states
    ex = synth.example(211, 128, 128)
    prm = R.default_params(seed=211)
    job = R.Job(prm, [ex], 256, 256)
    job.run(threads=8, **DEFAULT)
    sess = ts.Session.builder().add_example(ex).seed(211).output_size(ts.Dims.square(256)).build()
    sess.set_schedule(**DEFAULT).set_tuning(**tuning)
    assert_same(sess.run(), job)


@pytest.mark.parametrize("flow", [1, 2])
def test_flow_runs_guided_random_init_bit_exact(ctx, flow):
    # redo runs read locked (random-init) pixels, earlier-batch redos and
    # not-yet-redone pixels from three different sources
    exs = [synth.example(1300 + i, 56, 56) for i in range(2)]
    gds = [synth.blobs(1310 + i, 56, 56) for i in range(2)]
    tg = synth.blobs(1320, 200, 168)
    gen, job = run_pair(exs, 200, 168, seed=5, example_guides=gds, target_guide=tg, random_init=40,
                        guide_alpha=0.7, tuning=dict(flow=flow))
    assert_same(gen, job)


def test_redo_runs_at_512_bit_exact(ctx):
    # at 512x512 the third stage's redo batches (about 600 items) still fit
    # one dependency-ordered run
    ex = synth.example(1330, 128, 128)
    gen, job = run_pair([ex], 512, 512, seed=13, tuning=dict(flow=2, prefetch_max=1024))
    assert_same(gen, job)


@pytest.mark.parametrize("defer", [0, 1])
def test_progress_callback_bit_exact(ctx, defer):
    # the callback sees committed images (pending commits are flushed first),
    # and reading them changes nothing
    ex = synth.example(212, 96, 96)
    prm = R.default_params(seed=3)
    job = R.Job(prm, [ex], 192, 160)
    job.run(threads=8, **DEFAULT)
    seen = []
    sess = ts.Session.builder().add_example(ex).seed(3).output_size(ts.Dims(192, 160)).build()
    sess.set_schedule(**DEFAULT).set_tuning(defer=defer)
    gen = sess.run(progress=lambda p: seen.append((p["total"][0], int((p["image"][..., 3] > 0).sum()))))
    assert_same(gen, job)
    assert len(seen) > 10
    assert all(a[0] <= b[0] and a[1] <= b[1] for a, b in zip(seen, seen[1:]))
    assert seen[-1][1] == 192 * 160


def test_multi_example_random_init_bit_exact(ctx):
    exs = [synth.example(40 + i, 50 + 10 * i, 50) for i in range(4)]
    gen, job = run_pair(exs, 96, 80, seed=211, random_init=10, nearest_neighbors=20)
    assert_same(gen, job)


def test_sample_masks_and_ignore_bit_exact(ctx):
    exs = [synth.example(51, 64, 64), synth.example(52, 40, 40)]
    mk = np.full((64, 64, 4), 255, np.uint8)
    mk[:, :32, 0] = 0
    gen, job = run_pair(exs, 72, 72, seed=211, sample_kind=[2, 1], sample_masks=[mk, None])
    assert_same(gen, job)


def test_inpaint_bit_exact(ctx):
    ex = synth.example(61, 96, 96)
    m = synth.ellipse_mask(61, 96, 96)
    gen, job = run_pair([ex], 96, 96, seed=3, inpaint_mask=m, inpaint_index=0, sample_kind=[2], sample_masks=[m])
    assert_same(gen, job)


def test_guided_bit_exact(ctx):
    exs = [synth.example(1000 + i, 60, 60) for i in range(2)]
    gds = [synth.blobs(2000 + i, 60, 60) for i in range(2)]
    tg = synth.blobs(3000, 100, 100)
    gen, job = run_pair(exs, 100, 100, seed=211, example_guides=gds, target_guide=tg, random_init=10,
                        guide_alpha=0.8)
    assert_same(gen, job)


def test_other_parameters_bit_exact(ctx):
    ex = synth.example(71, 40, 56)
    gen, job = run_pair([ex], 70, 50, seed=99, nearest_neighbors=64, random_sample_locations=80,
                        cauchy_dispersion=0.3, backtrack_percent=0.7, backtrack_stages=3)
    assert_same(gen, job)


def test_c2_workload_bit_exact(ctx):
    # 512^2 -> 1024^2 with defaults: the benchmark workload, whole image
    ex = synth.example(512, 512, 512)
    gen, job = run_pair([ex], 1024, 1024, threads=16, seed=0)
    assert_same(gen, job)
