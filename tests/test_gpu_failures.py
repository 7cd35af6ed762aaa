"""Failure handling of the many-job forms: a job that cannot get device
memory, a cancelled job and a request larger than the memory budget each
end with their own status, and every other job still finishes with exactly
the result it gets alone. The failures are forced from the host (the
library's allocation hook, cancellation between launches, a small budget);
no kernel is made to fail."""
import numpy as np
import pytest

import tsgpu as ts
from tsgpu import serve, synth

pytestmark = pytest.mark.gpu


def builder(seed, side, ex=48, k=30):
    return (ts.Session.builder().add_example(synth.example(seed, ex, ex)).seed(seed).nearest_neighbors(k)
            .output_size(ts.Dims.square(side)))


def same(a, b):
    return (np.array_equal(a.image, b.image) and np.array_equal(a.coord, b.coord)
            and np.array_equal(a.ids, b.ids) and np.array_equal(a.resolved_flat, b.resolved_flat)
            and np.array_equal(a.resolved_score.view(np.uint32), b.resolved_score.view(np.uint32)))


SPECS = [(7000 + i, s) for i, s in enumerate([64, 96, 40, 128, 80, 56])]


def test_out_of_memory_job_fails_alone(ctx):
    alone = [builder(*sp).build().run() for sp in SPECS]
    sessions = [builder(*sp).build() for sp in SPECS]  # created before the hook
    batch = ts.Batch(ctx, max_active=64)
    tickets = [batch.submit(s) for s in sessions]
    ts.debug_fail_allocs(1)  # the first admitted job's first allocation fails
    try:
        batch.run_all()
    finally:
        ts.debug_fail_allocs(0)
    status = dict(batch.poll())
    assert sorted(status) == sorted(tickets)
    assert status[tickets[0]] == -21, status  # OutOfMemory
    with pytest.raises(ts.TsgError) as e:
        batch.take(tickets[0])
    assert e.value.kind == "OutOfMemory"
    for t, a in zip(tickets[1:], alone[1:]):
        assert status[t] == 0
        assert same(batch.take(t), a)
    batch.close()


def test_cancelled_jobs_leave_the_others_intact(ctx):
    alone = [builder(*sp).build().run() for sp in SPECS]
    batch = ts.Batch(ctx, max_active=3)
    tickets = [batch.submit(builder(*sp).build()) for sp in SPECS]
    batch.cancel(tickets[5])  # still waiting for a slot
    batch.step(1)             # the first three are admitted and run their first launch
    batch.cancel(tickets[1])  # running
    with pytest.raises(ts.TsgError):
        batch.cancel(987654)
    batch.run_all()
    status = dict(batch.poll())
    assert sorted(status) == sorted(tickets)
    for i, t in enumerate(tickets):
        if i in (1, 5):
            assert status[t] == -23, (i, status[t])
            with pytest.raises(ts.TsgError) as e:
                batch.take(t)
            assert e.value.kind == "Cancelled"
        else:
            assert status[t] == 0
            assert same(batch.take(t), alone[i]), i
    batch.close()


def test_server_rejects_what_can_never_fit_and_cancels(ctx):
    small = builder(7100, 48)
    big = builder(7101, 512)
    budget = small.footprint_estimate() * 3
    assert big.footprint_estimate() > budget
    alone = builder(7102, 64).build().run()
    srv = serve.Server(ctx, hbm_budget=budget, max_active=8, rounds_per_pump=2)
    r_small = srv.submit(small)
    r_big = srv.submit(big)
    r_ok = srv.submit(builder(7102, 64))
    r_cancel = srv.submit(builder(7103, 64))
    assert srv.cancel(r_cancel)
    fin = {f.rid: f for f in srv.drain()}
    assert sorted(fin) == [r_small, r_big, r_ok, r_cancel]
    assert isinstance(fin[r_big].error, ts.TsgError) and fin[r_big].error.kind == "AdmissionRejected"
    assert fin[r_cancel].error.kind == "Cancelled"
    assert fin[r_small].error is None and fin[r_ok].error is None
    assert same(fin[r_ok].result, alone)
    assert srv.rejected == 1 and srv.in_use == 0
    srv.close()
