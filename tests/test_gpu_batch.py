"""Many syntheses in flight at once (tsg_batch) give each synthesis exactly
the result it gets alone, and alone the device matches the oracle.

The mix covers what a serving queue holds at once: different output and
example sizes, k and m, guided and unguided jobs (they run in separate
launches of the same round), inpainting, sample masks, several examples,
and jobs on the one-pixel-per-batch schedule next to default-schedule jobs.
"""
import numpy as np
import pytest

import tsgpu as ts
from oracle import refcpu as R
from tsgpu import synth

pytestmark = pytest.mark.gpu

DEFAULT = ts.default_schedule()
SEQ = dict(px_per_batch=1, redo_shift=0, new_shift=0, max_batch=0, first_new_shift=0)


def spec_list():
    """Job descriptions: (builder kwargs, schedule)."""
    out = []
    for i, (ex, o, k) in enumerate([(32, 64, 20), (48, 96, 50), (64, 128, 50), (40, 80, 8), (96, 160, 30)]):
        out.append((dict(kind="plain", ex=ex, out=o, seed=4000 + i, k=k), DEFAULT))
    out.append((dict(kind="plain", ex=24, out=40, seed=4100, k=20), SEQ))
    out.append((dict(kind="multi", out=72, seed=4200), DEFAULT))
    out.append((dict(kind="guided", out=80, seed=4300), DEFAULT))
    out.append((dict(kind="inpaint", out=72, seed=4400), DEFAULT))
    out.append((dict(kind="guided", out=56, seed=4500), SEQ))
    return out


def build(spec):
    d, sched = spec
    b = ts.Session.builder().seed(d["seed"])
    if d["kind"] == "plain":
        b.add_example(synth.example(d["seed"], d["ex"], d["ex"]))
        b.nearest_neighbors(d["k"]).output_size(ts.Dims.square(d["out"]))
    elif d["kind"] == "multi":
        mk = np.full((40, 40, 4), 255, np.uint8)
        mk[:, :20, 0] = 0
        b.add_example(synth.example(d["seed"], 48, 40))
        b.add_example(ts.Example(synth.example(d["seed"] + 1, 40, 40)).set_sample_method(mk))
        b.random_init(12).random_sample_locations(30).output_size(ts.Dims.square(d["out"]))
    elif d["kind"] == "guided":
        for i in range(2):
            b.add_example(ts.Example(synth.example(d["seed"] + i, 48, 48)).with_guide(
                synth.blobs(d["seed"] + 10 + i, 48, 48)))
        b.load_target_guide(synth.blobs(d["seed"] + 20, d["out"], d["out"]))
        b.guide_alpha(0.8).random_init(10).output_size(ts.Dims.square(d["out"]))
    elif d["kind"] == "inpaint":
        n = d["out"]
        m = synth.ellipse_mask(d["seed"], n, n)
        b.inpaint_example(m, ts.Example(synth.example(d["seed"], n, n)).set_sample_method(m), ts.Dims.square(n))
    sess = b.build()
    sess.set_schedule(**sched)
    return sess


def assert_same_result(a, b, what):
    assert np.array_equal(a.image, b.image), what
    assert np.array_equal(a.coord, b.coord), what
    assert np.array_equal(a.ids, b.ids), what
    assert np.array_equal(a.resolved_flat, b.resolved_flat), what
    assert np.array_equal(a.resolved_score.view(np.uint32), b.resolved_score.view(np.uint32)), what
    for x, y in zip(a.stage_stats, b.stage_stats):
        assert np.array_equal(x, y), what


@pytest.mark.parametrize("max_active,spread", [(64, False), (3, False), (64, True)])
def test_batch_equals_alone(ctx, max_active, spread):
    # 64: the unguided jobs' launches home jobs on XCDs; 3 (fewer jobs than
    # XCDs) and the forced flag: spread launches, every XCD on every job
    specs = spec_list()
    alone = [build(s).run() for s in specs]
    batch = ts.Batch(ctx, max_active=max_active)
    if spread:
        batch.debug_engine(0, 1)
    tickets = [batch.submit(build(s)) for s in specs]
    batch.run_all()
    finished = dict(batch.poll())
    assert sorted(finished) == sorted(tickets) and not any(finished.values())
    for t, s, a in zip(tickets, specs, alone):
        assert_same_result(batch.take(t), a, f"{s[0]}")
    st = batch.stats()
    assert st["units"] > 0 and st["rounds"] > 0
    batch.close()


@pytest.mark.parametrize("quantum,max_active", [(1, 64), (300, 4), (5000, 64)])
def test_batch_quantum_equals_alone(ctx, quantum, max_active):
    # a planning quantum cuts a job's program at Jacobi batch boundaries, so
    # its stages span several launches (quantum 1: one batch per launch);
    # with at most 8 jobs at a time the launches also plan neighbour lists ahead
    specs = spec_list()
    alone = [build(s).run() for s in specs]
    batch = ts.Batch(ctx, max_active=max_active, quantum=quantum)
    tickets = [batch.submit(build(s)) for s in specs]
    batch.run_all()
    finished = dict(batch.poll())
    assert sorted(finished) == sorted(tickets) and not any(finished.values())
    for t, s, a in zip(tickets, specs, alone):
        assert_same_result(batch.take(t), a, f"{s[0]} quantum {quantum}")
    # the stages really were cut: more launches than there are stages
    assert batch.stats()["launches"] > 12
    batch.close()


@pytest.mark.parametrize("alone_below,max_active,ahead", [(100, 64, 0), (100, 64, 1 << 34), (6, 3, 1 << 34)])
def test_batch_alone_lane_equals_alone(ctx, alone_below, max_active, ahead):
    # jobs that start while few are in the batch run alone on the single-job
    # path (100: all of them; 6 with 3 slots: the first ones share the
    # engine, the last ones run alone), with or without stage-ahead buffers
    specs = spec_list()
    alone = [build(s).run() for s in specs]
    batch = ts.Batch(ctx, max_active=max_active)
    batch.set_alone(alone_below, ahead)
    tickets = [batch.submit(build(s)) for s in specs]
    batch.run_all()
    finished = dict(batch.poll())
    assert sorted(finished) == sorted(tickets) and not any(finished.values())
    for t, s, a in zip(tickets, specs, alone):
        assert_same_result(batch.take(t), a, f"{s[0]} alone_below {alone_below}")
    launches = batch.stats()["launches"]
    assert (launches == 0) if alone_below > len(specs) else (launches > 0)
    batch.close()


def test_batch_stepwise_with_late_submissions(ctx):
    # continuous batching: jobs join while others are half way through
    specs = spec_list()
    alone = {i: build(s).run() for i, s in enumerate(specs)}
    batch = ts.Batch(ctx, max_active=4)
    tick = {}
    for i in range(3):
        tick[batch.submit(build(specs[i]))] = i
    got = {}
    nxt = 3
    for _ in range(100000):
        left = batch.step(7)
        if nxt < len(specs):
            tick[batch.submit(build(specs[nxt]))] = nxt
            nxt += 1
        for t, status in batch.poll():
            assert status == 0
            got[tick[t]] = batch.take(t)
        if nxt == len(specs) and left == 0 and len(got) == len(specs):
            break
    assert len(got) == len(specs)
    for i, g in got.items():
        assert_same_result(g, alone[i], f"{specs[i][0]}")
    batch.close()


def test_batch_matches_oracle(ctx):
    # one batched job against the oracle directly (alone == oracle is
    # covered job by job in test_gpu_parity)
    ex = synth.example(4600, 64, 64)
    prm = R.default_params(seed=4600, nearest_neighbors=30)
    job = R.Job(prm, [ex], 112, 112)
    job.run(threads=8, **DEFAULT)
    batch = ts.Batch(ctx)
    others = [batch.submit(build(s)) for s in spec_list()[:4]]
    sess = ts.Session.builder().add_example(ex).seed(4600).nearest_neighbors(30).output_size(
        ts.Dims.square(112)).build()
    sess.set_schedule(**DEFAULT)
    t = batch.submit(sess)
    batch.run_all()
    g = batch.take(t)
    rgba, coord, ids = job.state()
    flat, score = job.resolved()
    assert np.array_equal(g.coord, coord) and np.array_equal(g.ids, ids) and np.array_equal(g.image, rgba)
    assert np.array_equal(g.resolved_flat[:len(flat)], flat)
    assert np.array_equal(g.resolved_score[:len(flat)].view(np.uint32), score.view(np.uint32))
    for o in others:
        batch.take(o)
    batch.close()


def test_batch_rejects_reuse_and_unknown_ticket(ctx):
    batch = ts.Batch(ctx)
    s = build(spec_list()[0])
    batch.submit(s)
    with pytest.raises(RuntimeError):
        batch.submit(s)
    with pytest.raises(ts.TsgError):
        batch.take(12345)
    batch.close()
