"""ctypes binding of oracle/librefcpu.so (test infrastructure only).

The oracle restates lib/src/ms.rs (Generator::resolve and helpers) and
lib/src/img_pyramid.rs of EmbarkStudios/texture-synthesis on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "librefcpu.so")
_lib = None

U64_NONE = (1 << 64) - 1


class Params(C.Structure):
    _fields_ = [
        ("nearest_neighbors", C.c_uint32),
        ("random_sample_locations", C.c_uint64),
        ("cauchy_dispersion", C.c_float),
        ("backtrack_percent", C.c_float),
        ("backtrack_stages", C.c_uint32),
        ("guide_alpha", C.c_float),
        ("seed", C.c_uint64),
        ("tiling_mode", C.c_uint32),
        ("random_init", C.c_uint64),
    ]


class Sched(C.Structure):
    _fields_ = [
        ("px_per_batch", C.c_uint32),
        ("redo_shift", C.c_uint32),
        ("new_shift", C.c_uint32),
        ("max_batch", C.c_uint32),
        ("first_new_shift", C.c_uint32),
    ]


class Image(C.Structure):
    _fields_ = [("w", C.c_uint32), ("h", C.c_uint32), ("rgba", C.c_void_p)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        L.orc_job_new.restype = C.c_void_p
        L.orc_job_new.argtypes = [P(Params), C.c_uint32, P(Image), P(C.c_int), P(Image), P(Image), P(Image),
                                  P(Image), P(Image), C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_job_free.argtypes = [C.c_void_p]
        L.orc_job_error.argtypes = [C.c_void_p]
        L.orc_job_error.restype = C.c_char_p
        L.orc_job_run.argtypes = [C.c_void_p, P(Sched), C.c_uint32]
        L.orc_job_run_until.argtypes = [C.c_void_p, P(Sched), C.c_uint32, C.c_uint32, C.c_uint64]
        L.orc_job_run_reference_threads.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_job_get.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_job_resolved.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
        L.orc_job_resolved.restype = C.c_uint64
        L.orc_job_debug_maps.argtypes = [C.c_void_p] * 4
        L.orc_job_debug_maps.restype = None
        L.orc_job_locked.argtypes = [C.c_void_p]
        L.orc_job_locked.restype = C.c_uint64
        L.orc_job_stage_stats.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_job_stage_stats.restype = C.c_uint32
        L.orc_job_mask.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_job_describe.argtypes = [C.c_void_p, C.c_uint64] + [C.c_void_p] * 9
        L.orc_job_describe.restype = C.c_uint64
        L.orc_job_stage_info.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.orc_job_stage_luts.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_pcg_new_u32.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_void_p]
        L.orc_pcg_seed_u64.argtypes = [C.c_uint64, C.c_uint32, C.c_void_p]
        L.orc_pcg_from_seed_u64.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
        L.orc_range_u32.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
        L.orc_range_u64.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, C.c_void_p]
        L.orc_exp_compare.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint32, C.c_void_p]
        L.orc_exp_compare.restype = C.c_uint64
        L.orc_pick_index.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_pick_index.restype = C.c_uint64
        L.orc_luts_full.argtypes = [C.c_float, C.c_float, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_resize.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_void_p]
        L.orc_pyramid.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
        L.orc_guide_map.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
        L.orc_guide_map.restype = None
        L.orc_match_histograms.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32]
        L.orc_match_histograms.restype = None
        L.orc_knn.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32,
                              C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ---- PRNG ----
def pcg_new_u32(state: int, stream: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().orc_pcg_new_u32(state, stream, n, _ptr(out))
    return out


def pcg_seed_u64(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint64)
    lib().orc_pcg_seed_u64(seed, n, _ptr(out))
    return out


def pcg_from_seed_u64(seed16: bytes, n: int) -> np.ndarray:
    b = np.frombuffer(seed16, np.uint8).copy()
    out = np.zeros(n, np.uint64)
    lib().orc_pcg_from_seed_u64(_ptr(b), n, _ptr(out))
    return out


def range_u32(state: int, stream: int, lo: int, hi: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().orc_range_u32(state, stream, lo, hi, n, _ptr(out))
    return out


def range_u64(state: int, stream: int, lo: int, hi: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint64)
    lib().orc_range_u64(state, stream, lo, hi, n, _ptr(out))
    return out


def exp_compare(i0: int, dev: np.ndarray, threads: int = 8):
    """(mismatches, first differing index) of the device's weight exps against the host's."""
    dev = np.ascontiguousarray(dev, np.float32)
    first = C.c_uint64()
    n = int(lib().orc_exp_compare(i0, len(dev), _ptr(dev), threads, C.byref(first)))
    return n, first.value


def pick_index(seed: int, length: int) -> int:
    return int(lib().orc_pick_index(seed, length))


def luts_full(cauchy: float, alpha_s: float, guided: bool):
    my = np.zeros(65536, np.float32)
    g = np.zeros(65536, np.float32)
    lib().orc_luts_full(cauchy, alpha_s, int(guided), _ptr(my), _ptr(g))
    return my.reshape(256, 256), g.reshape(256, 256)


FILTERS = {"triangle": 0, "catmullrom": 1, "gaussian": 2}


def resize(img: np.ndarray, nw: int, nh: int, filt: str) -> np.ndarray:
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape[:2]
    out = np.zeros((nh, nw, 4), np.uint8)
    rc = lib().orc_resize(_ptr(img), w, h, nw, nh, FILTERS[filt], _ptr(out))
    assert rc == 0
    return out


def guide_map(img: np.ndarray) -> np.ndarray:
    """transform_to_guide_map without a resize: blur sigma 2, grayscale, RGBA."""
    img = np.ascontiguousarray(img, np.uint8)
    out = np.zeros_like(img)
    lib().orc_guide_map(_ptr(img), img.shape[1], img.shape[0], _ptr(out))
    return out


def match_histograms(src: np.ndarray, tgt: np.ndarray) -> np.ndarray:
    out = np.ascontiguousarray(src, np.uint8).copy()
    tgt = np.ascontiguousarray(tgt, np.uint8)
    lib().orc_match_histograms(_ptr(out), out.shape[1], out.shape[0], _ptr(tgt), tgt.shape[1], tgt.shape[0])
    return out


def pyramid(img: np.ndarray, levels: int) -> np.ndarray:
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape[:2]
    n = max(levels, 1)
    out = np.zeros((n, h, w, 4), np.uint8)
    rc = lib().orc_pyramid(_ptr(img), w, h, levels, _ptr(out))
    if rc != 0:
        raise ValueError("image too small for the pyramid")
    return out


def knn(mask: np.ndarray, queries: np.ndarray, k: int):
    mask = np.ascontiguousarray(mask, np.uint8)
    h, w = mask.shape
    q = np.ascontiguousarray(queries, np.uint32)
    nbrs = np.zeros((len(q), k, 2), np.int32)
    counts = np.zeros(len(q), np.uint32)
    lib().orc_knn(_ptr(mask), w, h, len(q), _ptr(q), k, _ptr(nbrs), _ptr(counts))
    return nbrs, counts


def default_params(**kw) -> Params:
    p = Params(50, 50, 1.0, 0.5, 5, 0.8, 0, 0, U64_NONE)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def sched(px_per_batch=0, redo_shift=1, new_shift=3, max_batch=0, first_new_shift=0) -> Sched:
    return Sched(px_per_batch, redo_shift, new_shift, max_batch, first_new_shift)


SEQUENTIAL = dict(px_per_batch=1, redo_shift=0, new_shift=0, max_batch=0)


class Job:
    """One synthesis on the CPU oracle (Session::run restated)."""

    def __init__(self, params: Params, examples, out_w, out_h, sample_kind=None, sample_masks=None,
                 example_guides=None, target_guide=None, inpaint_mask=None, inpaint_color=None,
                 inpaint_index=0):
        self._keep = []
        n = len(examples)

        def imgs(lst):
            arr = (Image * len(lst))()
            for i, im in enumerate(lst):
                if im is None:
                    arr[i] = Image(0, 0, None)
                    continue
                im = np.ascontiguousarray(im, np.uint8)
                self._keep.append(im)
                arr[i] = Image(im.shape[1], im.shape[0], im.ctypes.data)
            return arr

        ex = imgs(examples)
        kinds = (C.c_int * n)(*(sample_kind or [0] * n))
        sm = imgs(sample_masks if sample_masks is not None else [None] * n)
        eg = imgs(example_guides) if example_guides is not None else None
        tg = imgs([target_guide]) if target_guide is not None else None
        im = imgs([inpaint_mask]) if inpaint_mask is not None else None
        ic = imgs([inpaint_color]) if inpaint_color is not None else None
        self.params = params
        self.w, self.h = out_w, out_h
        L = lib()
        self._h = L.orc_job_new(C.byref(params), n, ex, kinds, sm, eg, tg, im, ic, inpaint_index, out_w, out_h)
        err = L.orc_job_error(self._h)
        if err:
            msg = err.decode()
            L.orc_job_free(self._h)
            self._h = None
            raise ValueError(msg)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_job_free(self._h)
            self._h = None

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(lib().orc_job_error(self._h).decode())

    def run(self, threads=1, **sched_kw):
        s = sched(**sched_kw) if sched_kw else sched(**SEQUENTIAL)
        self._check(lib().orc_job_run(self._h, C.byref(s), threads))
        return self

    def run_reference_threads(self, threads=1):
        """Session::run with the reference's own threading (timing mode):
        serial while fewer than 1000 items are to be redone, then `threads`
        workers with immediate updates and the TreeGrid kNN. With threads > 1
        the result depends on timing; with 1 it equals run() one pixel per
        batch."""
        self._check(lib().orc_job_run_reference_threads(self._h, int(threads)))
        return self

    def run_until(self, stage_t, item, threads=1, **sched_kw):
        s = sched(**sched_kw) if sched_kw else sched(**SEQUENTIAL)
        self._check(lib().orc_job_run_until(self._h, C.byref(s), threads, stage_t, item))
        return self

    def state(self):
        n = self.w * self.h
        rgba = np.zeros((self.h, self.w, 4), np.uint8)
        coord = np.zeros((n, 3), np.uint32)
        ids = np.zeros((n, 2), np.uint32)
        lib().orc_job_get(self._h, _ptr(rgba), _ptr(coord), _ptr(ids))
        return rgba, coord, ids

    def resolved(self):
        n = int(lib().orc_job_resolved(self._h, None, None, 0))
        flat = np.zeros(n, np.uint32)
        score = np.zeros(n, np.float32)
        lib().orc_job_resolved(self._h, _ptr(flat), _ptr(score), n)
        return flat, score

    def locked(self):
        return int(lib().orc_job_locked(self._h))

    def mask(self):
        m = np.zeros((self.h, self.w), np.uint8)
        lib().orc_job_mask(self._h, _ptr(m))
        return m

    def debug_maps(self):
        """get_id_maps and get_uncertainty_map: (patch ids, map ids, uncertainty) RGBA8."""
        maps = [np.zeros((self.h, self.w, 4), np.uint8) for _ in range(3)]
        lib().orc_job_debug_maps(self._h, *[_ptr(m) for m in maps])
        return tuple(maps)

    def stage_stats(self):
        P = self.params.backtrack_stages
        sums = np.zeros(P + 1, np.float64)
        cnts = np.zeros(P + 1, np.uint64)
        items = np.zeros(P + 1, np.uint64)
        lib().orc_job_stage_stats(self._h, _ptr(sums), _ptr(cnts), _ptr(items))
        return sums, cnts, items

    def stage_mean_costs(self):
        sums, cnts, _ = self.stage_stats()
        return sums / np.maximum(cnts, 1)

    def stage_info(self):
        lv = C.c_uint32()
        a = C.c_float()
        ss = C.c_uint64()
        rd = C.c_uint64()
        sz = C.c_uint64()
        cur = C.c_uint64()
        lib().orc_job_stage_info(self._h, C.byref(lv), C.byref(a), C.byref(ss), C.byref(rd), C.byref(sz),
                                 C.byref(cur))
        return dict(level=lv.value, alpha=a.value, stage_seed=ss.value, redo=rd.value, stage_size=sz.value,
                    cursor=cur.value)

    def stage_luts(self):
        my = np.zeros(256, np.float32)
        g = np.zeros(256, np.float32)
        lib().orc_job_stage_luts(self._h, _ptr(my), _ptr(g))
        return my, g

    def describe(self, n):
        k = self.params.nearest_neighbors
        cm = k + self.params.random_sample_locations
        flat = np.zeros(n, np.uint32)
        nc = np.zeros(n, np.uint32)
        nbrs = np.zeros((n, k, 2), np.int32)
        w = np.zeros((n, k), np.float32)
        cc = np.zeros(n, np.uint32)
        cands = np.zeros((n, cm, 6), np.int32)
        fc = np.zeros((n, cm), np.float32)
        best = np.zeros(n, np.uint32)
        bs = np.zeros(n, np.float32)
        got = int(lib().orc_job_describe(self._h, n, _ptr(flat), _ptr(nc), _ptr(nbrs), _ptr(w), _ptr(cc),
                                         _ptr(cands), _ptr(fc), _ptr(best), _ptr(bs)))
        return dict(flat=flat[:got], nbr_count=nc[:got], nbrs=nbrs[:got], weights=w[:got], cand_count=cc[:got],
                    cands=cands[:got], full_cost=fc[:got], best=best[:got], best_score=bs[:got])
