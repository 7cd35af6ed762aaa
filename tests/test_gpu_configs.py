"""Parity at the full sizes of the measured workloads, through the C ABI.

* guided: 4 x 300^2 examples (seeds 1000+i) with 300^2 blob guides (seeds
  2000+i), a 1024^2 blob target guide (seed 3000), guide_alpha 0.8,
  random_init 10, k = m = 50, seed 211 -> 1024^2; whole image bit-exact
  against the oracle under the same Jacobi schedule.
* batched: the 256 syntheses of the batched benchmark leg (examples uniform
  in {64..512}^2, outputs uniform in {128..1024}^2, k in {20, 50}, m = 50,
  seeds 4000+j) through one tsg_batch; every result equals the job run
  alone (blake2b digest of every output array), and two 1024^2 jobs (one
  with k = 20, one 512^2 -> 1024^2 with k = 50) equal the oracle.
* serving: requests on seeded 256^2 examples with outputs up to 2048^2,
  inpainting ones resized to the output size with CatmullRom as load_image
  does (utils.rs:55-80); each result equals the request run alone, and the
  2048^2 inpainting request equals the oracle.
"""
import hashlib
import os

import numpy as np
import pytest

import tsgpu as ts
from oracle import refcpu as R
from tsgpu import serve, synth

pytestmark = pytest.mark.gpu

DEFAULT = ts.default_schedule()
THREADS = min(16, os.cpu_count() or 1)


def digest(g):
    """(digest of the output maps and the resolved list, digest of the stage statistics)"""
    h = hashlib.blake2b(digest_size=16)
    for a in (g.image, g.coord, g.ids, g.resolved_flat, g.resolved_score):
        h.update(np.ascontiguousarray(a).tobytes())
    hs = hashlib.blake2b(digest_size=16)
    for a in g.stage_stats:
        hs.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest(), hs.hexdigest()


def same_digest(got, want, what):
    assert got[0] == want[0], f"{what}: output maps differ from the run alone"
    assert got[1] == want[1], f"{what}: stage statistics differ from the run alone"


def same_as_oracle(g, job):
    rgba, coord, ids = job.state()
    flat, score = job.resolved()
    assert np.array_equal(g.coord, coord), "coord map"
    assert np.array_equal(g.ids, ids), "id map"
    assert np.array_equal(g.image, rgba), "colour map"
    assert np.array_equal(g.resolved_flat[:len(flat)], flat), "resolved list order"
    assert np.array_equal(g.resolved_score[:len(flat)].view(np.uint32), score.view(np.uint32)), "scores"


# ------------------------------------------------------------------ guided
def test_guided_4x300_to_1024_bit_exact(ctx):
    exs = [synth.example(1000 + i, 300, 300) for i in range(4)]
    gds = [synth.blobs(2000 + i, 300, 300) for i in range(4)]
    tg = synth.blobs(3000, 1024, 1024)
    print("case: guided 4x300^2 -> 1024^2, seed 211", flush=True)
    b = ts.Session.builder().seed(211).guide_alpha(0.8).random_init(10).nearest_neighbors(50)
    b.random_sample_locations(50)
    for e, g in zip(exs, gds):
        b.add_example(ts.Example(e).with_guide(g))
    sess = b.load_target_guide(tg).output_size(ts.Dims.square(1024)).build()
    gen = sess.run()
    prm = R.default_params(seed=211, guide_alpha=0.8, random_init=10, nearest_neighbors=50,
                           random_sample_locations=50)
    job = R.Job(prm, exs, 1024, 1024, example_guides=gds, target_guide=tg)
    job.run(threads=THREADS, **DEFAULT)
    same_as_oracle(gen, job)
    sums, cnts, items = job.stage_stats()
    gs, gc, gi = gen.stage_stats
    assert np.array_equal(gc, cnts) and np.array_equal(gi, items)
    assert np.allclose(gs, sums, rtol=1e-9, atol=0)


# ------------------------------------------------------------------ batched
def batched_specs(n=256, seed=4000):
    rng = np.random.default_rng(seed)
    ex = rng.choice([64, 128, 256, 512], n)
    out = rng.choice([128, 256, 512, 1024], n)
    k = rng.choice([20, 50], n)
    return [(int(ex[j]), int(out[j]), int(k[j]), seed + j) for j in range(n)]


def test_batched_mix_256_equals_alone_and_oracle(ctx):
    specs = batched_specs()
    exs = {}

    def builder(spec):
        e, o, k, sd = spec
        if sd not in exs:
            exs[sd] = synth.example(sd, e, e)
        return (ts.Session.builder().add_example(exs[sd]).seed(sd).nearest_neighbors(k)
                .random_sample_locations(50).output_size(ts.Dims.square(o)))

    print("case: 256-job batched mix, alone runs", flush=True)
    alone = [digest(builder(s).build().run()) for s in specs]
    print("case: 256-job batched mix, one tsg_batch", flush=True)
    batch = ts.Batch(ctx, max_active=len(specs))
    tickets = [batch.submit(builder(s).build()) for s in specs]
    got = {}
    while len(got) < len(specs):
        batch.step(8)
        for t, status in batch.poll():
            assert status == 0
            got[t] = batch.take(t)
    # oracle spot checks: a 1024^2 job with k = 20 and a 512^2 -> 1024^2 job with k = 50
    k20 = next(j for j, s in enumerate(specs) if s[1] == 1024 and s[2] == 20)
    k50 = next(j for j, s in enumerate(specs) if s[0] == 512 and s[1] == 1024 and s[2] == 50)
    for j in range(len(specs)):
        g = got[tickets[j]]
        same_digest(digest(g), alone[j], f"job {j} {specs[j]}")
        if j in (k20, k50):
            e, o, k, sd = specs[j]
            print(f"case: oracle for batched job {j} {specs[j]}", flush=True)
            job = R.Job(R.default_params(seed=sd, nearest_neighbors=k, random_sample_locations=50), [exs[sd]], o, o)
            job.run(threads=THREADS, **DEFAULT)
            same_as_oracle(g, job)
        del got[tickets[j]]
    batch.close()


# ------------------------------------------------------------------ serving
def test_catmullrom_resize_bit_exact(ctx):
    ex = synth.example(5001, 256, 256)
    for side in (128, 300, 1000, 2048):
        assert np.array_equal(ts.resize(ex, side, side, "catmullrom", ctx), R.resize(ex, side, side, "catmullrom"))
    assert np.array_equal(ts.resize(ex, 200, 90, "triangle", ctx), R.resize(ex, 200, 90, "triangle"))
    assert np.array_equal(ts.resize(ex, 256, 256, "catmullrom", ctx), ex)  # same size: unchanged


def serving_request(seed, side, inpaint):
    ex = synth.example(seed, 256, 256)
    b = ts.Session.builder().seed(seed)
    if inpaint:
        m = synth.ellipse_mask(seed, side, side)
        # the example is resized to the output size by the builder (load_image, CatmullRom)
        b.inpaint_example(m, ts.Example(ex).set_sample_method(m), ts.Dims.square(side))
    else:
        b.add_example(ex).output_size(ts.Dims.square(side))
    return b


def test_serving_mix_to_2048_equals_alone_and_oracle(ctx):
    rng = np.random.default_rng(17)
    sides = np.exp(rng.uniform(np.log(128), np.log(2048), 10)).astype(int)
    reqs = [(5200 + i, int(s), bool(rng.random() < 0.3)) for i, s in enumerate(sides)]
    reqs += [(5300, 2048, False), (5301, 2048, True)]
    print("case: serving mix alone runs", [r[1:] for r in reqs], flush=True)
    alone = [digest(serving_request(*r).build().run()) for r in reqs]
    srv = serve.Server(ctx, hbm_budget=6 << 30, max_active=64, rounds_per_pump=4)
    got = {}
    pending = list(enumerate(reqs))
    print("case: serving mix through the server", flush=True)
    while pending or not srv.idle():
        for _ in range(3):
            if pending:
                i, r = pending.pop(0)
                assert srv.submit(serving_request(*r)) == i
        for f in srv.pump():
            assert f.error is None, f.error
            got[f.rid] = f.result
    assert sorted(got) == list(range(len(reqs)))
    for i, r in enumerate(reqs):
        same_digest(digest(got[i]), alone[i], f"request {i} {r}")
    srv.close()
    seed, side, _ = reqs[-1]
    print("case: oracle for the 2048^2 inpainting request", flush=True)
    ex = R.resize(synth.example(seed, 256, 256), side, side, "catmullrom")
    m = synth.ellipse_mask(seed, side, side)
    job = R.Job(R.default_params(seed=seed), [ex], side, side, sample_kind=[2], sample_masks=[m], inpaint_mask=m,
                inpaint_color=ex, inpaint_index=0)
    job.run(threads=THREADS, **DEFAULT)
    same_as_oracle(got[len(reqs) - 1], job)
