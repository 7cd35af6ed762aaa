"""Serving loop: requests arrive while others run, are admitted by device
memory, and each result is bit-identical to the same request run alone."""
import numpy as np
import pytest

import tsgpu as ts
from tsgpu import serve, synth

pytestmark = pytest.mark.gpu


def request(seed, side, inpaint):
    """A plain or inpainting request on a seeded 64x64 example."""
    ex = synth.example(seed, 64, 64)
    b = ts.Session.builder().seed(seed)
    if inpaint:
        idx = (np.arange(side) * 64) // side
        img = ex[idx][:, idx]
        m = synth.ellipse_mask(seed, side, side)
        b.inpaint_example(m, ts.Example(img).set_sample_method(m), ts.Dims.square(side))
    else:
        b.add_example(ex).output_size(ts.Dims.square(side))
    return b


def mix(n=14, seed=7):
    rng = np.random.default_rng(seed)
    sides = np.exp(rng.uniform(np.log(24), np.log(160), n)).astype(int)
    return [(5000 + i, int(s), bool(rng.random() < 0.3)) for i, s in enumerate(sides)]


def same(a, b):
    return (np.array_equal(a.image, b.image) and np.array_equal(a.coord, b.coord)
            and np.array_equal(a.ids, b.ids) and np.array_equal(a.resolved_flat, b.resolved_flat)
            and np.array_equal(a.resolved_score.view(np.uint32), b.resolved_score.view(np.uint32)))


@pytest.mark.parametrize("budget_mb,max_active,alone_below",
                         [(4096, 64, serve.ALONE_BELOW), (0, 3, serve.ALONE_BELOW), (4096, 64, 0), (0, 3, 0)])
def test_serving_results_equal_alone(ctx, budget_mb, max_active, alone_below):
    # alone_below 0: every request shares the work-queue engine; the default
    # runs these few requests alone on the single-job path
    reqs = mix()
    alone = [request(*r).build().run() for r in reqs]
    # 0: a budget that just fits the largest request, so admissions wait for memory
    budget = budget_mb << 20 if budget_mb else max(request(*r).footprint_estimate() for r in reqs)
    srv = serve.Server(ctx, hbm_budget=budget, max_active=max_active, rounds_per_pump=5, alone_below=alone_below)
    got = {}
    pending = list(enumerate(reqs))
    # arrivals interleaved with engine rounds: two requests join per pump
    while pending or not srv.idle():
        for _ in range(2):
            if pending:
                i, r = pending.pop(0)
                rid = srv.submit(request(*r))
                assert rid == i
        assert srv.in_use <= max(srv.hbm_budget, max(f.footprint for f in srv.running.values()) if srv.running
                                 else 0)
        for f in srv.pump():
            assert f.error is None, f.error
            got[f.rid] = f.result
    assert sorted(got) == list(range(len(reqs)))
    for i, r in enumerate(reqs):
        assert same(got[i], alone[i]), f"request {i} {r}"
    assert srv.in_use == 0 and srv.rejected == 0
    if not budget_mb:
        assert srv.memory_holds > 0
    st = srv.batch.stats()
    assert (st["alone"] == 0 and st["launches"] > 0) if alone_below == 0 else st["alone"] == len(reqs)
    srv.close()


def test_serving_reports_invalid_requests(ctx):
    srv = serve.Server(ctx)
    bad = ts.Session.builder().add_example(synth.example(1, 32, 32)).output_size(ts.Dims.square(32))
    bad.cauchy_dispersion(2.0)  # outside 0..1: InvalidRange, as Session::build
    ok = request(5100, 40, False)
    srv.submit(bad)
    srv.submit(ok)
    fin = {f.rid: f for f in srv.drain()}
    assert isinstance(fin[0].error, ts.TsgError) and fin[0].error.kind == "InvalidRange"
    assert fin[1].error is None and fin[1].result.image.shape == (40, 40, 4)
    srv.close()


def test_footprint_estimate_covers_the_job(ctx):
    # the estimate made before building bounds what the job holds after
    # running under the engine (all scratch grown)
    for seed, side, inp in mix(6, seed=3):
        b = request(seed, side, inp)
        est = b.footprint_estimate()
        batch = ts.Batch(ctx)
        t = batch.submit(b.build())
        batch.run_all()
        g = batch.take(t)
        batch.close()
        assert g.image.shape[0] == side
        assert side * side * 28 <= g.device_bytes <= est, (side, inp, g.device_bytes, est)
