"""Autotuner of the launch structure (tsg_tuning).

Results never depend on the launch structure: every choice resolves the same
Jacobi batches against the same frozen states. What it changes is launch
count, latency and occupancy, so the best choice depends on the job's shape
and the device. ``autotune`` runs the job a SessionBuilder describes under
candidate structures, timing each on the device (the run's event-bracketed
device time, best of ``reps``), walks one knob at a time from the library's
default keeping a change only when it is at least ``min_gain`` faster, and
stores the winner in the context, which gives it to every later job of the
same shape (output size, k, m, sampled maps, guided, tiling, stages).

The reference has no counterpart: its thread count is the only launch knob
(``max_thread_count``, lib.rs:342-359).
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple

from . import Context, SessionBuilder, default_tuning

# knobs in search order, with the values tried for each
SPACE: Sequence[Tuple[str, Sequence[int]]] = (
    ("flow", (1, 2, 0)),
    ("flow_team", (2, 4, 8)),
    ("team_max", (4, 8)),
    ("small_max", (0, 4, 16)),
    ("prefetch_max", (1024, 2048, 4096)),
    ("ahead", (1, 2)),
    ("defer", (1, 0)),
    ("persist", (0, 1)),
    ("flow_head", (1024, 0, 4096)),
)


def _time(builder: SessionBuilder, tuning: Dict[str, int], reps: int) -> float:
    best = float("inf")
    for _ in range(reps):
        sess = builder.build()
        sess.set_tuning(**tuning)
        best = min(best, sess.run().run_ms)
    return best


def autotune(builder: SessionBuilder, reps: int = 2, space: Sequence[Tuple[str, Sequence[int]]] = SPACE,
             min_gain: float = 0.01, ctx: Optional[Context] = None, store: bool = True) -> dict:
    """Picks a launch structure for jobs shaped like `builder`'s.

    Returns {"tuning": best, "ms": its device time, "default_ms": the default's,
    "tried": [(tuning, ms), ...]}; with `store`, the context keeps the best for
    later jobs of the same shape."""
    if reps < 1:
        raise ValueError("reps must be at least 1")
    ctx = ctx or builder._ctx or Context.default()
    builder.context(ctx)
    cur = default_tuning()
    cur_ms = _time(builder, cur, reps)
    default_ms = cur_ms
    tried: List[Tuple[Dict[str, int], float]] = [(dict(cur), cur_ms)]
    for knob, values in space:
        if knob not in cur:
            raise KeyError(f"unknown tuning knob {knob}")
        for v in values:
            if cur[knob] == v:
                continue
            cand = dict(cur, **{knob: int(v)})
            ms = _time(builder, cand, reps)
            tried.append((cand, ms))
            if ms < cur_ms * (1.0 - min_gain):
                cur, cur_ms = cand, ms
    if store:
        like = builder.build()
        try:
            ctx.store_tuning(like, **cur)
        finally:
            like.close()
    return {"tuning": cur, "ms": cur_ms, "default_ms": default_ms, "tried": tried}
