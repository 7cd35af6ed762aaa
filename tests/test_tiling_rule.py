"""Tiling mode's neighbour set, pinned independently of the oracle's search
code: a numpy brute force over the resolved pixels plus the ghost copies
that flush_resolved inserts in tiling mode (lib/src/ms.rs:296-331), against
the oracle's k-neighbour lists on frozen states.

The rule as restated here, straight from ms.rs:
* margins x_l = (W as f32 * 0.05) as i32, x_r = W - x_l, and likewise y_b,
  y_t from H (TILING_BOUNDARY_PERCENTAGE);
* a flushed pixel (x, y) with x < x_l is also present at (x + W, y), one
  with x > x_r at (x - W, y); independently y < y_b adds (x, y + H) and
  y > y_t adds (x, y - H); no diagonal copy;
* pixels resolved before resolve starts (random_init) are flushed with
  ghosts when resolve begins (ms.rs:760-779); a pixel resolved by
  resolve_at_random because it had no resolved neighbour gets none
  (ms.rs:1003-1009 -> resolve_at_random, ms.rs:447-475);
* the k nearest by squared distance, ties in the declared (d^2, dy, dx)
  order.
"""
import numpy as np
import pytest

from oracle import refcpu as R
from tsgpu import synth


def ghost_points(W, H, flats, with_ghosts):
    xl = int(np.float32(W) * np.float32(0.05))
    yb = int(np.float32(H) * np.float32(0.05))
    xr, yt = W - xl, H - yb
    pts = []
    for p, g in zip(flats, with_ghosts):
        x, y = int(p) % W, int(p) // W
        pts.append((x, y))
        if not g:
            continue
        if x < xl:
            pts.append((x + W, y))
        elif x > xr:
            pts.append((x - W, y))
        if y < yb:
            pts.append((x, y + H))
        elif y > yt:
            pts.append((x, y - H))
    return np.array(pts, np.int64)


def brute_knn(pts, x, y, k):
    dx = pts[:, 0] - x
    dy = pts[:, 1] - y
    d2 = dx * dx + dy * dy
    order = np.lexsort((dx, dy, d2))[:k]
    return pts[order]


@pytest.mark.parametrize("W,H,k,random_init,stops", [
    (40, 36, 20, None, [(0, 1), (0, 30), (2, 100), (5, 400)]),
    (64, 48, 50, 7, [(0, 0), (1, 50), (3, 300), (5, 1500)]),
    (33, 70, 8, None, [(0, 2), (4, 200), (5, 2000)]),
])
def test_oracle_knn_follows_the_ghost_rule(W, H, k, random_init, stops):
    ex = synth.example(W * 3 + H, 24, 20)
    prm = R.default_params(seed=W + H, nearest_neighbors=k, tiling_mode=1,
                           random_init=R.U64_NONE if random_init is None else random_init)
    checked = ghosts_seen = 0
    for t, item in stops:
        job = R.Job(prm, [ex], W, H)
        job.run_until(t, item, threads=1, **R.SEQUENTIAL)
        flats, _ = job.resolved()
        # only the very first resolve of a job without random_init has no
        # resolved neighbour (every later query finds the pixels before it)
        with_ghosts = np.ones(len(flats), bool)
        if random_init is None and len(flats):
            with_ghosts[0] = False
        pts = ghost_points(W, H, flats, with_ghosts)
        d = job.describe(24)
        for q in range(len(d["flat"])):
            x, y = int(d["flat"][q]) % W, int(d["flat"][q]) // W
            n = int(d["nbr_count"][q])
            want = brute_knn(pts, x, y, k) if len(pts) else np.zeros((0, 2), np.int64)
            assert n == len(want), (t, item, q)
            assert np.array_equal(d["nbrs"][q, :n].astype(np.int64), want), (t, item, q)
            outside = (want[:, 0] < 0) | (want[:, 0] >= W) | (want[:, 1] < 0) | (want[:, 1] >= H)
            ghosts_seen += int(outside.sum())
            checked += 1
    assert checked > 0 and ghosts_seen > 0  # the lists do reach across the edges


def test_ghost_rule_margins_and_no_diagonals():
    # closed form of the rule on a 100 x 60 image: margins 5 and 3
    pts = ghost_points(100, 60, [0, 99, 60 * 100 - 1, 50 * 100 + 50, 2 * 100 + 50], [True] * 5)
    got = {tuple(p) for p in pts}
    assert got == {(0, 0), (100, 0), (0, 60),          # corner: x and y copies, no diagonal
                   (99, 0), (-1, 0), (99, 60),
                   (99, 59), (-1, 59), (99, -1),
                   (50, 50),                            # interior: no copy
                   (50, 2), (50, 62)}                   # top edge only
