"""tsgpu — MI355X patch-search engine for example-based texture synthesis.

Python mirror of texture-synthesis's public API (lib/src/session.rs,
lib/src/lib.rs) over the C ABI in include/tsgpu.h:

    import tsgpu as ts
    sess = (ts.Session.builder()
            .add_example(img)                 # (H, W, 4) uint8 RGBA
            .seed(211)
            .output_size(ts.Dims.square(256))
            .build())
    generated = sess.run()
    generated.image                           # (H, W, 4) uint8

The device library (libtsgpu.so next to this file) must be built; the
module raises ImportError-free but every call fails loudly when it is not.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtsgpu.so")
U64_NONE = (1 << 64) - 1

TSG_SAMPLE_ALL, TSG_SAMPLE_IGNORE, TSG_SAMPLE_IMAGE = 0, 1, 2

STATUS_NAMES = {
    -1: "InvalidRange", -2: "NoExamples", -3: "ExampleGuideMismatch", -4: "SizeMismatch",
    -5: "MapsCountMismatch", -10: "InvalidArgument", -11: "Unsupported", -20: "Device",
    -21: "OutOfMemory", -22: "Sampling", -23: "Cancelled", -24: "AdmissionRejected",
}


class TsgError(Exception):
    """A non-zero status from the C ABI; mirrors texture_synthesis::Error."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code
        self.kind = STATUS_NAMES.get(code, str(code))


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
        ("max_thread_count", C.c_uint32),
    ]


class Schedule(C.Structure):
    _fields_ = [
        ("px_per_batch", C.c_uint32),
        ("redo_shift", C.c_uint32),
        ("new_shift", C.c_uint32),
        ("max_batch", C.c_uint32),
        ("first_new_shift", C.c_uint32),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("small_max", C.c_uint32),
        ("fuse_max", C.c_uint32),
        ("team_max", C.c_uint32),
        ("team_waves", C.c_uint32),
        ("prefetch_max", C.c_uint32),
        ("ahead", C.c_uint32),
        ("coop_wgs", C.c_uint32),
        ("coop_max", C.c_uint32),
        ("flow", C.c_uint32),
        ("flow_wgs", C.c_uint32),
        ("flow_team", C.c_uint32),
        ("defer", C.c_uint32),
        ("persist", C.c_uint32),
        ("flow_head", C.c_uint32),
        ("flow_head_team", C.c_uint32),
    ]


def default_schedule() -> dict:
    """The library's default Jacobi schedule as keyword arguments."""
    s = Schedule()
    lib().tsg_schedule_default(C.byref(s))
    return {n: getattr(s, n) for n, _ in Schedule._fields_}


def default_tuning() -> dict:
    t = Tuning()
    lib().tsg_tuning_default(C.byref(t))
    return {n: getattr(t, n) for n, _ in Tuning._fields_}


class Image(C.Structure):
    _fields_ = [("w", C.c_uint32), ("h", C.c_uint32), ("rgba", C.c_void_p)]


PROGRESS_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_uint8), C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64,
                          C.c_uint64, C.c_uint64)

_lib = None


def lib():
    """Loads libtsgpu.so; raises RuntimeError when the extension is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"tsgpu device library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P, vp, u32, u64 = C.POINTER, C.c_void_p, C.c_uint32, C.c_uint64
    L.tsg_last_error.restype = C.c_char_p
    L.tsg_version.restype = C.c_char_p
    L.tsg_params_default.argtypes = [P(Params)]
    L.tsg_schedule_default.argtypes = [P(Schedule)]
    L.tsg_ctx_create.argtypes = [C.c_int, P(vp)]
    L.tsg_ctx_destroy.argtypes = [vp]
    L.tsg_job_create.argtypes = [vp, P(Params), u32, P(Image), P(C.c_int), P(Image), P(Image), P(Image), P(Image),
                                 u32, u32, u32, P(vp)]
    L.tsg_job_set_schedule.argtypes = [vp, P(Schedule)]
    L.tsg_tuning_default.argtypes = [P(Tuning)]
    L.tsg_job_set_tuning.argtypes = [vp, P(Tuning)]
    L.tsg_job_get_tuning.argtypes = [vp, P(Tuning)]
    L.tsg_ctx_store_tuning.argtypes = [vp, vp, P(Tuning)]
    L.tsg_ctx_clear_tunings.argtypes = [vp]
    L.tsg_job_debug_maps.argtypes = [vp, vp, vp, vp]
    L.tsg_xform_apply.argtypes = [vp, vp, u32, u32, u32, P(Image), vp]
    L.tsg_job_run.argtypes = [vp, PROGRESS_FN, vp]
    L.tsg_job_download.argtypes = [vp, vp, vp, vp, vp, vp]
    L.tsg_job_download_packed.argtypes = [vp, vp, vp, vp]
    L.tsg_job_stage_stats.argtypes = [vp, vp, vp, vp, u32]
    L.tsg_job_timing.argtypes = [vp, vp, vp]
    L.tsg_job_set_profiling.argtypes = [vp, C.c_int]
    L.tsg_job_kernel_stats.argtypes = [vp, vp, vp, vp]
    L.tsg_job_phase_clocks.argtypes = [vp, vp, u64]
    L.tsg_job_destroy.argtypes = [vp]
    L.tsg_pyramid_build.argtypes = [vp, P(Image), u32, vp]
    L.tsg_resize.argtypes = [vp, P(Image), u32, u32, C.c_int, vp]
    L.tsg_knn_batch.argtypes = [vp, vp, u32, u32, u32, vp, u32, vp, vp]
    L.tsg_knn_batch_noalloc.argtypes = [vp, vp, u32, u32, u32, vp, u32, vp, vp]
    L.tsg_random_candidates.argtypes = [vp, u32, P(Image), P(Image), u64, u32, u32, u32, vp]
    L.tsg_score_batch.argtypes = [vp, u32, u32, vp, vp, vp, u32, P(Image), P(Image), P(Image), vp, vp, u32, u32,
                                  u32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, u32]
    L.tsg_job_footprint.argtypes = [vp, P(u64)]
    L.tsg_footprint_estimate.argtypes = [P(Params), u32, vp, vp, vp, u32, u32, P(u64)]
    L.tsg_batch_create.argtypes = [vp, u32, P(vp)]
    L.tsg_batch_submit.argtypes = [vp, vp, P(u64)]
    L.tsg_batch_step.argtypes = [vp, u32, P(u32)]
    L.tsg_batch_run_all.argtypes = [vp]
    L.tsg_batch_poll.argtypes = [vp, vp, vp, u32, P(u32)]
    L.tsg_batch_take.argtypes = [vp, u64, P(vp)]
    L.tsg_batch_stats.argtypes = [vp, P(u64), P(u64), P(u64), P(u64)]
    L.tsg_batch_cancel.argtypes = [vp, u64]
    L.tsg_debug_fail_allocs.argtypes = [C.c_int64]
    L.tsg_batch_debug_engine.argtypes = [vp, u32, u32]
    L.tsg_batch_set_quantum.argtypes = [vp, u64]
    L.tsg_batch_set_alone_below.argtypes = [vp, u32, u64]
    L.tsg_batch_engine_profile.argtypes = [vp, vp]
    L.tsg_debug_exp_sweep.argtypes = [vp, u64, u64, vp]
    L.tsg_debug_exp_sweep_host.argtypes = [u64, u64, vp]
    L.tsg_batch_destroy.argtypes = [vp]
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise TsgError(rc, lib().tsg_last_error().decode())


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def as_rgba8(img) -> np.ndarray:
    """to_rgba8 (utils.rs:62): gray replicated, alpha 255 when absent."""
    a = np.asarray(img)
    if a.dtype != np.uint8:
        raise TypeError("images are uint8")
    if a.ndim == 2:
        a = np.stack([a, a, a, np.full_like(a, 255)], axis=-1)
    elif a.shape[2] == 1:
        a = np.concatenate([a, a, a, np.full_like(a, 255)], axis=-1)
    elif a.shape[2] == 3:
        a = np.concatenate([a, np.full(a.shape[:2] + (1,), 255, np.uint8)], axis=-1)
    return np.ascontiguousarray(a)


class _Images:
    """Keeps numpy buffers alive while a tsg_image array is in use."""

    def __init__(self, imgs: Sequence[Optional[np.ndarray]]):
        self.keep = []
        self.arr = (Image * max(1, len(imgs)))()
        for i, im in enumerate(imgs):
            if im is None:
                self.arr[i] = Image(0, 0, None)
                continue
            im = as_rgba8(im)
            self.keep.append(im)
            self.arr[i] = Image(im.shape[1], im.shape[0], im.ctypes.data)


class Context:
    """One HIP device and stream (tsg_ctx)."""

    _default = None

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().tsg_ctx_create(device, C.byref(h)))
        self._h = h

    @classmethod
    def default(cls) -> "Context":
        if cls._default is None:
            cls._default = Context(0)
        return cls._default

    def store_tuning(self, like: "Session", **tuning) -> None:
        """Launch structure for jobs shaped like `like` created on this context
        from now on (keys of default_tuning(); the rest stay default)."""
        t = default_tuning()
        unknown = set(tuning) - set(t)
        if unknown:
            raise TypeError(f"unknown tuning keys {sorted(unknown)}")
        t.update(tuning)
        _check(lib().tsg_ctx_store_tuning(self._h, like._job, C.byref(Tuning(**t))))

    def clear_tunings(self) -> None:
        _check(lib().tsg_ctx_clear_tunings(self._h))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.tsg_ctx_destroy(self._h)
            self._h = None


# --------------------------------------------------------------------------
# public API mirror (lib.rs / session.rs)
# --------------------------------------------------------------------------
@dataclass(frozen=True)
class Dims:
    width: int
    height: int

    @staticmethod
    def square(size: int) -> "Dims":
        return Dims(size, size)

    @staticmethod
    def new(width: int, height: int) -> "Dims":
        return Dims(width, height)


class SampleMethod:
    """GenericSampleMethod: All, Ignore, or Image(mask)."""

    All = "all"
    Ignore = "ignore"

    @staticmethod
    def image(mask) -> "np.ndarray":
        return as_rgba8(mask)


class Example:
    def __init__(self, img):
        self.img = as_rgba8(img)
        self.guide = None
        self.sample_method = SampleMethod.All

    @staticmethod
    def builder(img) -> "Example":
        return Example(img)

    def with_guide(self, guide) -> "Example":
        self.guide = as_rgba8(guide)
        return self

    def set_sample_method(self, method) -> "Example":
        self.sample_method = method if isinstance(method, str) else as_rgba8(method)
        return self


class GeneratedImage:
    """GeneratedImage (lib.rs:378-455): image plus coord / id maps.

    Holds the packed per-pixel state as downloaded (four u32 per pixel:
    rgba, x | y << 16, coord-map id | id-map id << 16, patch id); `image` is
    a view of it, `coord` (x, y, map) and `ids` (patch, map) are unpacked on
    first use."""

    KERNELS = ("resolve", "commit", "rand", "recolor", "stats", "prefetch", "resolve_runs")

    def __init__(self, state, size, resolved_flat, resolved_score, stage_stats, run_ms, n_batches,
                 original_maps, kernel_stats=None):
        self._state = state
        self._size = size
        self._coord = None
        self._ids = None
        self.resolved_flat = resolved_flat
        self.resolved_score = resolved_score
        self.stage_stats = stage_stats
        self.run_ms = run_ms
        self.n_batches = n_batches
        self.original_maps = original_maps
        self.kernel_stats = kernel_stats

    @property
    def image(self) -> np.ndarray:
        w, h = self._size
        return self._state.view(np.uint8).reshape(h, w, 16)[:, :, :4]

    @property
    def coord(self) -> np.ndarray:
        if self._coord is None:
            st = self._state
            c = np.empty((st.shape[0], 3), np.uint32)
            c[:, 0] = st[:, 1] & 0xFFFF
            c[:, 1] = st[:, 1] >> 16
            c[:, 2] = st[:, 2] & 0xFFFF
            self._coord = c
        return self._coord

    @property
    def ids(self) -> np.ndarray:
        if self._ids is None:
            st = self._state
            d = np.empty((st.shape[0], 2), np.uint32)
            d[:, 0] = st[:, 3]
            d[:, 1] = st[:, 2] >> 16
            self._ids = d
        return self._ids

    def into_image(self) -> np.ndarray:
        return np.ascontiguousarray(self.image)

    def stage_mean_costs(self) -> np.ndarray:
        sums, cnts, _ = self.stage_stats
        return sums / np.maximum(cnts, 1)

    def get_coordinate_transform(self) -> "CoordinateTransform":
        h, w = self.image.shape[:2]
        return CoordinateTransform(self.coord.reshape(-1).copy(), Dims(w, h), list(self.original_maps))


TRANSFORM_MAGIC = 0x1234_0001


class CoordinateTransform:
    """The .xform format of lib.rs:162-325."""

    def __init__(self, buffer: np.ndarray, output_size: Dims, original_maps: List[Dims]):
        self.buffer = np.asarray(buffer, np.uint32)
        self.output_size = output_size
        self.original_maps = original_maps

    def write(self) -> bytes:
        if self.buffer.size != self.output_size.width * self.output_size.height * 3:
            raise ValueError("buffer length doesn't match dimensions")
        hdr = np.array([TRANSFORM_MAGIC, self.output_size.width, self.output_size.height, len(self.original_maps)],
                       "<u4")
        dims = np.array([[d.width, d.height] for d in self.original_maps], "<u4").reshape(-1)
        return hdr.tobytes() + dims.tobytes() + self.buffer.astype("<u4").tobytes()

    @staticmethod
    def read(data: bytes) -> "CoordinateTransform":
        magic = int(np.frombuffer(data[:4], "<u4")[0])
        if magic >> 16 != 0x1234:
            raise ValueError("invalid magic")
        if magic & 0xFFFF != 0x1:
            raise ValueError("invalid version")
        w, h, n = (int(v) for v in np.frombuffer(data[4:16], "<u4"))
        dims = np.frombuffer(data[16:16 + 8 * n], "<u4").reshape(n, 2)
        off = 16 + 8 * n
        buf = np.frombuffer(data[off:off + 12 * w * h], "<u4").copy()
        return CoordinateTransform(buf, Dims(w, h), [Dims(int(a), int(b)) for a, b in dims])

    def apply(self, sources: Sequence[np.ndarray], ctx: Optional["Context"] = None) -> np.ndarray:
        imgs = [as_rgba8(s) for s in sources]
        if len(imgs) != len(self.original_maps):
            raise TsgError(-5, f"{len(imgs)} map(s) were provided, but {len(self.original_maps)} is/are required")
        for im, d in zip(imgs, self.original_maps):
            if im.shape[1] != d.width or im.shape[0] != d.height:
                raise ValueError("source must match the original map size (resize on the caller's side)")
        # the gather runs on the device (tsg_xform_apply)
        w, h = self.output_size.width, self.output_size.height
        out = np.zeros((h, w, 4), np.uint8)
        srcs = _Images(imgs)
        coord = np.ascontiguousarray(self.buffer, np.uint32)
        _check(lib().tsg_xform_apply((ctx or Context.default())._h, _ptr(coord), w, h, len(imgs), srcs.arr, _ptr(out)))
        return out


class Session:
    """Session (session.rs:22-66). run() consumes it, as the reference does."""

    def __init__(self, ctx: Context, job: C.c_void_p, out: Dims, original_maps):
        self._ctx = ctx
        self._job = job
        self._out = out
        self._original_maps = original_maps

    @staticmethod
    def builder() -> "SessionBuilder":
        return SessionBuilder()

    def set_schedule(self, px_per_batch=0, redo_shift=5, new_shift=6, max_batch=0, first_new_shift=7) -> "Session":
        """Jacobi batch schedule (defaults: the library's, see default_schedule())."""
        s = Schedule(px_per_batch, redo_shift, new_shift, max_batch, first_new_shift)
        _check(lib().tsg_job_set_schedule(self._job, C.byref(s)))
        return self

    def tuning(self) -> dict:
        """The launch structure this job will run with."""
        t = Tuning()
        _check(lib().tsg_job_get_tuning(self._job, C.byref(t)))
        return {n: getattr(t, n) for n, _ in Tuning._fields_}

    def set_tuning(self, **kw) -> "Session":
        """Launch structure (results do not depend on it); keys of default_tuning()."""
        t = default_tuning()
        unknown = set(kw) - set(t)
        if unknown:
            raise TypeError(f"unknown tuning keys {sorted(unknown)}")
        t.update(kw)
        _check(lib().tsg_job_set_tuning(self._job, C.byref(Tuning(**t))))
        return self

    def set_profiling(self, level: int = 1) -> "Session":
        _check(lib().tsg_job_set_profiling(self._job, int(level)))
        self._profiled = level
        return self

    def run(self, progress=None) -> GeneratedImage:
        if self._job is None:
            raise RuntimeError("a session runs once")
        L = lib()
        if progress is not None:
            def _cb(user, rgba, w, h, tc, tt, sc, st):
                img = np.ctypeslib.as_array(rgba, shape=(h * w * 4,)).reshape(h, w, 4).copy()
                progress(dict(image=img, total=(tc, tt), stage=(sc, st)))
            cb = PROGRESS_FN(_cb)
        else:
            cb = PROGRESS_FN()
        try:
            _check(L.tsg_job_run(self._job, cb, None))
            return self._collect()
        finally:
            L.tsg_job_destroy(self._job)
            self._job = None

    def run_resident(self) -> "Session":
        """Runs the job and keeps its maps in device memory; download() then
        copies them out and close() frees the job."""
        if self._job is None or getattr(self, "_ran", False):
            raise RuntimeError("a session runs once")
        try:
            _check(lib().tsg_job_run(self._job, PROGRESS_FN(), None))
        except Exception:
            self.close()
            raise
        self._ran = True
        return self

    def debug_maps(self) -> dict:
        """Generator::get_id_maps / get_uncertainty_map (ms.rs:603-653) of a job
        finished by run_resident(), computed on the device: RGBA8 maps
        "patch_ids", "map_ids" and "uncertainty"."""
        if self._job is None or not getattr(self, "_ran", False):
            raise RuntimeError("debug_maps() needs a job finished by run_resident()")
        w, h = self._out.width, self._out.height
        maps = {k: np.zeros((h, w, 4), np.uint8) for k in ("patch_ids", "map_ids", "uncertainty")}
        _check(lib().tsg_job_debug_maps(self._job, _ptr(maps["patch_ids"]), _ptr(maps["map_ids"]),
                                        _ptr(maps["uncertainty"])))
        return maps

    def download(self) -> GeneratedImage:
        if self._job is None or not getattr(self, "_ran", False):
            raise RuntimeError("download() needs a job finished by run_resident()")
        return self._collect()

    def close(self) -> None:
        if self._job is not None:
            lib().tsg_job_destroy(self._job)
            self._job = None

    def footprint(self) -> int:
        """Device bytes the job holds now (inputs, state and scratch)."""
        v = C.c_uint64()
        _check(lib().tsg_job_footprint(self._job, C.byref(v)))
        return v.value

    def _collect(self) -> GeneratedImage:
        """Downloads the finished job's maps, statistics and timings."""
        L = lib()
        w, h = self._out.width, self._out.height
        state = np.empty((h * w, 4), np.uint32)  # written whole by the download
        rf = np.zeros(h * w, np.uint32)
        rs = np.zeros(h * w, np.float32)
        _check(L.tsg_job_download_packed(self._job, _ptr(state), _ptr(rf), _ptr(rs)))
        n = 16
        sums = np.zeros(n, np.float64)
        cnts = np.zeros(n, np.uint64)
        items = np.zeros(n, np.uint64)
        got = L.tsg_job_stage_stats(self._job, _ptr(sums), _ptr(cnts), _ptr(items), n)
        if got < 0:
            _check(got)
        ms = C.c_double()
        nb = C.c_uint64()
        _check(L.tsg_job_timing(self._job, C.byref(ms), C.byref(nb)))
        kst = None
        if getattr(self, "_profiled", False):
            kms = np.zeros(7, np.float64)
            kit = np.zeros(7, np.uint64)
            kla = np.zeros(7, np.uint64)
            _check(L.tsg_job_kernel_stats(self._job, _ptr(kms), _ptr(kit), _ptr(kla)))
            kst = {n: dict(ms=float(kms[i]), items=int(kit[i]), launches=int(kla[i]))
                   for i, n in enumerate(GeneratedImage.KERNELS)}
            if getattr(self, "_profiled", 0) >= 2:
                clk = np.zeros((h * w, 16), np.uint64)
                _check(L.tsg_job_phase_clocks(self._job, _ptr(clk), h * w))
                kst["phase_clocks"] = clk
        gen = GeneratedImage(state, (w, h), rf, rs, (sums[:got], cnts[:got], items[:got]), ms.value,
                             nb.value, self._original_maps, kst)
        gen.device_bytes = self.footprint()  # every buffer the run grew
        return gen

    def __del__(self):
        if getattr(self, "_job", None) and _lib is not None:
            _lib.tsg_job_destroy(self._job)
            self._job = None


class Batch:
    """Many sessions in flight on one device.

    Each round plans every active session as far as its pick orders allow and
    runs everything planned in one launch of the work-queue engine; a
    session's result is identical to Session.run on it alone. No counterpart
    in the reference, whose Session::run (session.rs:37-66) synthesises one
    image at a time."""

    def __init__(self, ctx: Optional[Context] = None, max_active: int = 256, quantum: int = 0,
                 alone_below: int = 0):
        """quantum: resolve items each session may plan per round (0: no
        bound); a bound keeps rounds short for sessions submitted meanwhile.
        alone_below: a session that starts while fewer than this many are in
        the batch runs alone, with the whole device (0: never)."""
        self._ctx = ctx or Context.default()
        self._sessions = {}
        self._status = {}
        h = C.c_void_p()
        _check(lib().tsg_batch_create(self._ctx._h, int(max_active), C.byref(h)))
        self._h = h
        if quantum:
            _check(lib().tsg_batch_set_quantum(h, int(quantum)))
        self._alone_below = int(alone_below)
        if alone_below:
            _check(lib().tsg_batch_set_alone_below(h, int(alone_below), 0))

    def set_alone(self, below: int, ahead_bytes: int = 0):
        """Sessions starting while fewer than `below` are in the batch run
        alone; their stage-ahead buffers are used when they fit ahead_bytes."""
        self._alone_below = int(below)
        _check(lib().tsg_batch_set_alone_below(self._h, int(below), max(0, int(ahead_bytes))))

    def submit(self, session: Session) -> int:
        """Hands the session over; returns its ticket."""
        if session._job is None:
            raise RuntimeError("a session runs once")
        t = C.c_uint64()
        _check(lib().tsg_batch_submit(self._h, session._job, C.byref(t)))
        # the batch owns the job from here on
        session._job = None
        self._sessions[t.value] = session
        return t.value

    def step(self, max_rounds: int = 1) -> int:
        """Runs up to max_rounds rounds; returns the sessions not yet finished."""
        n = C.c_uint32()
        _check(lib().tsg_batch_step(self._h, int(max_rounds), C.byref(n)))
        return n.value

    def run_all(self):
        _check(lib().tsg_batch_run_all(self._h))

    def engine_profile(self) -> dict:
        """Workgroup milliseconds of the engine launches so far, by activity, the
        unit counts, and host milliseconds by step of the rounds."""
        v = np.zeros(11, np.float64)
        _check(lib().tsg_batch_engine_profile(self._h, _ptr(v)))
        return dict(find_ms=v[0], resolve_ms=v[1], other_ms=v[2], close_ms=v[3], resolve_units=int(v[4]),
                    other_units=int(v[5]), host_admit_ms=v[6], host_plan_ms=v[7], host_pack_ms=v[8],
                    launch_to_done_ms=v[9], host_finish_ms=v[10])

    def debug_engine(self, wgs_cap: int = 0, flags: int = 0):
        """Test hook: cap the engine's workgroups; flags bit 0 spreads every launch over all XCDs."""
        _check(lib().tsg_batch_debug_engine(self._h, int(wgs_cap), int(flags)))

    def cancel(self, ticket: int):
        """Stops a session that has not finished: poll then reports it with
        status Cancelled, and take raises it."""
        _check(lib().tsg_batch_cancel(self._h, int(ticket)))

    def poll(self):
        """[(ticket, status)] of finished sessions not yet taken."""
        cap = max(1, len(self._sessions))
        tk = np.zeros(cap, np.uint64)
        st = np.zeros(cap, np.int32)
        n = C.c_uint32()
        _check(lib().tsg_batch_poll(self._h, _ptr(tk), _ptr(st), cap, C.byref(n)))
        out = [(int(tk[i]), int(st[i])) for i in range(n.value)]
        self._status.update(out)
        return out

    def take(self, ticket: int) -> GeneratedImage:
        """Result of a finished session; raises the session's run error, if any."""
        if ticket not in self._status:
            self.poll()
        status = self._status.pop(ticket, None)
        job = C.c_void_p()
        _check(lib().tsg_batch_take(self._h, ticket, C.byref(job)))
        s = self._sessions.pop(ticket)
        s._job = job
        try:
            if status:
                raise TsgError(status, "session failed in the batch")
            return s._collect()
        finally:
            lib().tsg_job_destroy(job)
            s._job = None

    def stats(self):
        r, la, u, al = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(lib().tsg_batch_stats(self._h, C.byref(r), C.byref(la), C.byref(u), C.byref(al)))
        return dict(rounds=r.value, launches=la.value, units=u.value, alone=al.value)

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.tsg_batch_destroy(self._h)
            self._h = None
            self._sessions.clear()

    def __del__(self):
        self.close()


class SessionBuilder:
    """SessionBuilder (session.rs:70-448) over already-decoded images."""

    def __init__(self):
        self.examples: List[Example] = []
        self.target_guide = None
        self.inpaint = None  # (mask, example_index, dims)
        self.params = Params()
        lib().tsg_params_default(C.byref(self.params))
        self._resize_input: Optional[Dims] = None
        self._output_size = Dims.square(500)
        self._ctx: Optional[Context] = None

    def add_example(self, example) -> "SessionBuilder":
        self.examples.append(example if isinstance(example, Example) else Example(example))
        return self

    def add_examples(self, examples) -> "SessionBuilder":
        for e in examples:
            self.add_example(e)
        return self

    def inpaint_example(self, inpaint_mask, example, size: Dims) -> "SessionBuilder":
        self.inpaint = (as_rgba8(inpaint_mask), len(self.examples), size)
        self.add_example(example)
        return self

    def load_target_guide(self, guide) -> "SessionBuilder":
        self.target_guide = as_rgba8(guide)
        return self

    def resize_input(self, dims: Dims) -> "SessionBuilder":
        self._resize_input = dims
        return self

    def seed(self, value: int) -> "SessionBuilder":
        self.params.seed = value
        return self

    def tiling_mode(self, is_tiling: bool) -> "SessionBuilder":
        self.params.tiling_mode = int(bool(is_tiling))
        return self

    def nearest_neighbors(self, count: int) -> "SessionBuilder":
        self.params.nearest_neighbors = count
        return self

    def random_sample_locations(self, count: int) -> "SessionBuilder":
        self.params.random_sample_locations = count
        return self

    def random_init(self, count: int) -> "SessionBuilder":
        self.params.random_init = count
        return self

    def cauchy_dispersion(self, value: float) -> "SessionBuilder":
        self.params.cauchy_dispersion = value
        return self

    def guide_alpha(self, value: float) -> "SessionBuilder":
        self.params.guide_alpha = value
        return self

    def backtrack_percent(self, value: float) -> "SessionBuilder":
        self.params.backtrack_percent = value
        return self

    def backtrack_stages(self, stages: int) -> "SessionBuilder":
        self.params.backtrack_stages = stages
        return self

    def output_size(self, dims: Dims) -> "SessionBuilder":
        self._output_size = dims
        return self

    def max_thread_count(self, count: int) -> "SessionBuilder":
        self.params.max_thread_count = count
        return self

    def context(self, ctx: Context) -> "SessionBuilder":
        self._ctx = ctx
        return self

    def _sized(self, img: np.ndarray, dims: Optional[Dims]) -> np.ndarray:
        """load_image(src, Some(dims)) (utils.rs:55-80): CatmullRom resize on
        the device when the size differs, else the image unchanged."""
        if dims is None or (img.shape[1] == dims.width and img.shape[0] == dims.height):
            return img
        return resize(img, dims.width, dims.height, "catmullrom", self._ctx)

    def _out_dims(self) -> Dims:
        return self.inpaint[2] if self.inpaint is not None else self._output_size

    def _in_dims(self) -> Optional[Dims]:
        return self.inpaint[2] if self.inpaint is not None else self._resize_input

    def footprint_estimate(self) -> int:
        """Device bytes the built session holds while it runs under a Batch."""
        out = self._out_dims()
        ind = self._in_dims()

        def wh_of(img):
            return [ind.width, ind.height] if ind is not None else [img.shape[1], img.shape[0]]

        wh = np.array([wh_of(e.img) for e in self.examples], np.uint32).reshape(-1)
        kinds = np.array([TSG_SAMPLE_IMAGE if not isinstance(e.sample_method, str) else
                          TSG_SAMPLE_IGNORE if e.sample_method == SampleMethod.Ignore else TSG_SAMPLE_ALL
                          for e in self.examples], np.int32)
        gwh = None
        if self.target_guide is not None and self.examples and all(e.guide is not None for e in self.examples):
            gwh = np.array([wh_of(e.guide) for e in self.examples], np.uint32).reshape(-1)
        v = C.c_uint64()
        _check(lib().tsg_footprint_estimate(C.byref(self.params), len(self.examples), _ptr(wh), _ptr(kinds),
                                            _ptr(gwh), out.width, out.height, C.byref(v)))
        return v.value

    def build(self) -> Session:
        if self.inpaint is not None:
            mask, idx, dims = self.inpaint
            out, in_size = dims, dims
        else:
            out, in_size = self._output_size, self._resize_input
        exs = [self._sized(e.img, in_size) for e in self.examples]
        kinds, masks, guides = [], [], []
        for e in self.examples:
            if isinstance(e.sample_method, str):
                kinds.append(TSG_SAMPLE_IGNORE if e.sample_method == SampleMethod.Ignore else TSG_SAMPLE_ALL)
                masks.append(None)
            else:
                kinds.append(TSG_SAMPLE_IMAGE)
                masks.append(self._sized(e.sample_method, in_size))
            guides.append(None if e.guide is None else self._sized(e.guide, in_size))
        n_guides = sum(g is not None for g in guides)
        if n_guides and n_guides != len(self.examples):
            raise TsgError(-3, f"{len(self.examples)} examples were provided, but only {n_guides} guides were")
        ctx = self._ctx or Context.default()
        E = _Images(exs)
        M = _Images(masks)
        G = _Images(guides) if n_guides else None
        TG = _Images([self._sized(self.target_guide, out)]) if self.target_guide is not None else None
        IM = _Images([self._sized(self.inpaint[0], out)]) if self.inpaint is not None else None
        kind_arr = (C.c_int * max(1, len(kinds)))(*kinds)
        h = C.c_void_p()
        _check(lib().tsg_job_create(ctx._h, C.byref(self.params), len(exs), E.arr, kind_arr, M.arr,
                                    G.arr if G else None, TG.arr if TG else None, IM.arr if IM else None,
                                    self.inpaint[1] if self.inpaint else 0, out.width, out.height, C.byref(h)))
        # input_dimensions: every example, ignored ones included (ms.rs:722-731)
        original = [Dims(e.shape[1], e.shape[0]) for e in exs]
        return Session(ctx, h, out, original)


# --------------------------------------------------------------------------
# single-step entry points (parity tests, tools)
# --------------------------------------------------------------------------
def debug_fail_allocs(count: int):
    """Test hook: the library's next `count` device allocations fail (out of memory)."""
    _check(lib().tsg_debug_fail_allocs(int(count)))


def debug_exp_sweep(i0: int, n: int, ctx: Optional[Context] = None) -> np.ndarray:
    """Test hook: the device's (float)exp(-x_i) for the sweep arguments (tsg_debug_exp_sweep)."""
    ctx = ctx or Context.default()
    out = np.empty(n, np.float32)
    _check(lib().tsg_debug_exp_sweep(ctx._h, int(i0), int(n), _ptr(out)))
    return out


def debug_exp_sweep_host(i0: int, n: int) -> np.ndarray:
    """Test hook: the library's weight exp on the host for the sweep arguments."""
    out = np.empty(n, np.float32)
    _check(lib().tsg_debug_exp_sweep_host(int(i0), int(n), _ptr(out)))
    return out


def pyramid(img: np.ndarray, levels: int, ctx: Optional[Context] = None) -> np.ndarray:
    ctx = ctx or Context.default()
    I = _Images([img])
    h, w = I.keep[0].shape[:2]
    out = np.zeros((max(levels, 1), h, w, 4), np.uint8)
    _check(lib().tsg_pyramid_build(ctx._h, I.arr, levels, _ptr(out)))
    return out


FILTERS = {"triangle": 0, "catmullrom": 1, "gaussian": 2}


def resize(img: np.ndarray, width: int, height: int, filt: str = "catmullrom",
           ctx: Optional[Context] = None) -> np.ndarray:
    """image::imageops::resize as load_image applies it (utils.rs:55-80), on
    the device: RGBA8 in, (height, width, 4) RGBA8 out."""
    ctx = ctx or Context.default()
    I = _Images([img])
    out = np.zeros((height, width, 4), np.uint8)
    _check(lib().tsg_resize(ctx._h, I.arr, width, height, FILTERS[filt], _ptr(out)))
    return out


def knn(mask: np.ndarray, queries: np.ndarray, k: int, ctx: Optional[Context] = None):
    ctx = ctx or Context.default()
    mask = np.ascontiguousarray(mask, np.uint8)
    h, w = mask.shape
    q = np.ascontiguousarray(queries, np.uint32)
    nbrs = np.zeros((len(q), k, 2), np.int32)
    counts = np.zeros(len(q), np.uint32)
    _check(lib().tsg_knn_batch(ctx._h, _ptr(mask), w, h, len(q), _ptr(q), k, _ptr(nbrs), _ptr(counts)))
    return nbrs, counts


def knn_noalloc(mask: np.ndarray, queries: np.ndarray, k: int, ctx: Optional[Context] = None):
    """knn through tsg_knn_batch_noalloc (no device allocation per call)."""
    ctx = ctx or Context.default()
    mask = np.ascontiguousarray(mask, np.uint8)
    h, w = mask.shape
    q = np.ascontiguousarray(queries, np.uint32)
    nbrs = np.zeros((len(q), k, 2), np.int32)
    counts = np.zeros(len(q), np.uint32)
    _check(lib().tsg_knn_batch_noalloc(ctx._h, _ptr(mask), w, h, len(q), _ptr(q), k, _ptr(nbrs), _ptr(counts)))
    return nbrs, counts


def random_candidates(maps, sample_masks, stage_seed: int, item0: int, n: int, m: int,
                      ctx: Optional[Context] = None) -> np.ndarray:
    ctx = ctx or Context.default()
    M = _Images(maps)
    S = _Images(sample_masks) if sample_masks is not None else None
    out = np.zeros((n, m, 3), np.uint32)
    _check(lib().tsg_random_candidates(ctx._h, len(maps), M.arr, S.arr if S else None, stage_seed, item0, n, m,
                                       _ptr(out)))
    return out


def score_batch(color, coord, ids, maps, lut_color, k, flat, nbr_count, nbrs, cand_count, cands, guides=None,
                target_guide=None, lut_guide=None, ctx: Optional[Context] = None, team: int = 1):
    ctx = ctx or Context.default()
    h, w = color.shape[:2]
    color = np.ascontiguousarray(color, np.uint8)
    coord = np.ascontiguousarray(coord, np.uint32)
    ids = np.ascontiguousarray(ids, np.uint32)
    M = _Images(maps)
    G = _Images(guides) if guides is not None else None
    T = _Images([target_guide]) if target_guide is not None else None
    lc = np.ascontiguousarray(lut_color, np.float32)
    lg = np.ascontiguousarray(lut_guide, np.float32) if lut_guide is not None else None
    flat = np.ascontiguousarray(flat, np.uint32)
    nbr_count = np.ascontiguousarray(nbr_count, np.uint32)
    nbrs = np.ascontiguousarray(nbrs, np.int32)
    cand_count = np.ascontiguousarray(cand_count, np.uint32)
    cands = np.ascontiguousarray(cands, np.int32)
    n, cmax = cands.shape[0], cands.shape[1]
    full = np.zeros((n, cmax), np.float32)
    best = np.zeros(n, np.uint32)
    bs = np.zeros(n, np.float32)
    bp = np.zeros(n, np.uint32)
    bsp = np.zeros(n, np.float32)
    _check(lib().tsg_score_batch(ctx._h, w, h, _ptr(color), _ptr(coord), _ptr(ids), len(maps), M.arr,
                                 G.arr if G else None, T.arr if T else None, _ptr(lc), _ptr(lg), k, cmax, n,
                                 _ptr(flat), _ptr(nbr_count), _ptr(nbrs), _ptr(cand_count), _ptr(cands),
                                 _ptr(full), _ptr(best), _ptr(bs), _ptr(bp), _ptr(bsp), int(team)))
    return full, best, bs, bp, bsp
