"""The oracle's debug maps (Generator::get_id_maps / get_uncertainty_map,
ms.rs:603-653), checked against what the definitions fix.

Each id-map channel is gen_range(0..255) of a stream seeded from an
expression of the id. For u8 ranges rand 0.8.5 draws u32 values and rejects
only those whose low product word exceeds u32::MAX - (2^32 - 255) % 255 =
2^32 - 2, so the first draw decides: channel = (first u32 * 255) >> 32. The
seeded PCG32 stream is the one pinned by the known-answer vectors. A u32-style
zone ((255 << 24) - 1) would reject about one draw in 256 and is told apart.
"""
import numpy as np

import tsgpu as ts
from oracle import refcpu as R
from tsgpu import synth


def first_channel(seed):
    # next_u64 draws its low word first: that word is the first u32 draw
    v = int(R.pcg_seed_u64(seed & (2 ** 64 - 1), 1)[0]) & 0xFFFFFFFF
    return (v * 255) >> 32


def test_id_map_colours_follow_the_definition():
    ex = synth.example(801, 32, 32)
    job = R.Job(R.default_params(seed=801), [ex], 40, 36)
    job.run(threads=4, **ts.default_schedule())
    patch_map, map_map, _ = job.debug_maps()
    _, _, ids = job.state()
    m32 = 2 ** 32 - 1
    checked = 0
    for i in range(0, 40 * 36, 7):
        patch, mid = int(ids[i, 0]), int(ids[i, 1])
        y, x = divmod(i, 40)
        want_p = [first_channel(patch), first_channel((patch * 5 + 21) & m32), first_channel(patch // 4 + 12), 255]
        want_m = [first_channel(mid * 200), first_channel((mid * 5 + 341) & m32),
                  first_channel((mid * 1200 - 35412) & m32), 255]
        assert list(patch_map[y, x]) == want_p
        assert list(map_map[y, x]) == want_m
        checked += 1
    assert checked > 100


def test_uncertainty_map_follows_the_resolved_list():
    ex = synth.example(802, 32, 32)
    m = synth.ellipse_mask(802, 48, 48)
    job = R.Job(R.default_params(seed=802), [synth.example(803, 48, 48)], 48, 48, sample_kind=[2],
                sample_masks=[m], inpaint_mask=m, inpaint_color=synth.example(803, 48, 48))
    job.run(threads=4, **ts.default_schedule())
    _, _, unc = job.debug_maps()
    flat, score = job.resolved()
    assert len(flat) == 48 * 48
    n = np.clip(np.minimum(score, 1.0) * 255.0, 0, 255).astype(np.uint8)
    u = unc.reshape(-1, 4)[flat]
    assert np.array_equal(u[:, 0], n) and np.array_equal(u[:, 1], 255 - n)
    assert np.all(u[:, 2] == 0) and np.all(u[:, 3] == 255)
    # the kept (locked) pixels enter the list with score 0: pure green
    locked = job.locked()
    assert np.all(u[:locked, :2] == [0, 255])
