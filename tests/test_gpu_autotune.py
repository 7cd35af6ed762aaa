"""The autotuner picks among launch structures by device time, stores the
winner for jobs of the same shape, and never changes a result."""
import numpy as np
import pytest

import tsgpu as ts
from tsgpu import synth
from tsgpu.autotune import autotune

pytestmark = pytest.mark.gpu


def builder(out=160, seed=31):
    return ts.Session.builder().add_example(synth.example(seed, 64, 64)).seed(seed).output_size(ts.Dims.square(out))


def test_autotune_stores_the_winner_for_the_shape(ctx):
    ctx.clear_tunings()
    b = builder().context(ctx)
    res = autotune(b, reps=1, space=(("flow_team", (2, 8)), ("small_max", (0,))), ctx=ctx)
    assert res["ms"] <= res["default_ms"]
    assert any(t == res["tuning"] for t, _ in res["tried"]) and len(res["tried"]) == 4
    s = b.build()
    assert s.tuning() == res["tuning"]
    # the stored structure gives the default structure's result
    got = s.run()
    plain = b.build()
    plain.set_tuning(**ts.default_tuning())
    ref = plain.run()
    assert np.array_equal(got.coord, ref.coord) and np.array_equal(got.image, ref.image)
    # another shape keeps the default; clearing forgets the stored one
    assert builder(out=96).context(ctx).build().tuning() == ts.default_tuning()
    ctx.clear_tunings()
    assert b.build().tuning() == ts.default_tuning()


def test_autotune_rejects_unknown_knobs(ctx):
    with pytest.raises(KeyError):
        autotune(builder(out=64).context(ctx), reps=1, space=(("no_such_knob", (1,)),), ctx=ctx, store=False)
