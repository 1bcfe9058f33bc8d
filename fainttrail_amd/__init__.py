"""MI355X (gfx950) FaintTrail search backend.

Python view of the C-ABI in include/ftk_hip.h (libftk_hip.so, built in-tree by
fainttrail_amd/build.py). The names mirror the reference's backend interface
(include/fainttrail/backend.hpp:56-166): ShiftTable, StackPass, StackResult,
linear_shift_table, SearchBackend.run, make_backend, backend_names.

Only the "hip" backend lives here. There is no CPU fallback: when the library is
missing or no device is usable, the calls raise.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from pathlib import Path
from typing import Callable, Optional, Sequence

import numpy as np

__all__ = [
    "lib", "FtkError", "Cancelled", "ShiftTable", "StackPass", "StackResult", "linear_shift_table",
    "HipBackend", "make_backend", "backend_names", "parse_options", "device_count", "emulate_pass", "Plan",
    "velocity_grid",
]

_LIB_PATH = Path(__file__).resolve().parent / "libftk_hip.so"
_lib = None

FTK_OK, FTK_ERROR, FTK_CANCELLED = 0, 1, 2
PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.c_double, C.c_char_p, C.c_void_p)


class FtkError(RuntimeError):
    """An FTK_ERROR status (the reference throws std::runtime_error)."""


class Cancelled(FtkError):
    """FTK_CANCELLED: the progress callback asked to stop (types.hpp:70-72)."""


class _Options(C.Structure):
    _fields_ = [("device", C.c_int), ("max_mb", C.c_double), ("launch_adds", C.c_double), ("pad", C.c_int),
                ("tile_rows", C.c_int), ("queue", C.c_int)]


class _Pass(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("w", C.c_int), ("h", C.c_int),
        ("frames", C.POINTER(C.POINTER(C.c_float))),
        ("nv", C.c_int),
        ("dx", C.POINTER(C.c_int32)), ("dy", C.POINTER(C.c_int32)),
        ("min_count", C.c_int),
    ]


class _Timings(C.Structure):
    _fields_ = [("upload_s", C.c_double), ("kernel_s", C.c_double), ("download_s", C.c_double),
                ("setup_s", C.c_double)]


def lib() -> C.CDLL:
    """Load libftk_hip.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise FtkError(f"{_LIB_PATH} is not built; run python -m fainttrail_amd.build")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7. Loading torch
    # first makes the dynamic loader bind this library to that same copy (equal
    # SONAME); loading /opt/rocm's copy first leaves torch without a usable device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(_LIB_PATH))
    L.ftk_parse_options.argtypes = [C.c_char_p, C.POINTER(_Options), C.c_char_p, C.c_size_t]
    L.ftk_device_count.argtypes = [C.c_char_p, C.c_size_t]
    L.ftk_device_description.argtypes = [C.c_int, C.c_char_p, C.c_size_t]
    L.ftk_stack_pass.argtypes = [C.POINTER(_Pass), C.POINTER(_Options), C.POINTER(C.c_float), C.POINTER(C.c_int32),
                                 C.POINTER(_Timings), PROGRESS_FN, C.c_void_p, C.c_char_p, C.c_size_t]
    L.ftk_plan_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                  C.c_int, C.POINTER(_Options), C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]
    L.ftk_plan_run.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p,
                               C.c_size_t]
    L.ftk_plan_units.argtypes = [C.c_void_p]
    L.ftk_plan_units.restype = C.c_double
    L.ftk_plan_blocks.argtypes = [C.c_void_p]
    L.ftk_plan_destroy.argtypes = [C.c_void_p]
    L.ftk_plan_store.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_char_p, C.c_size_t]
    L.ftk_plan_run_stored.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_size_t]
    L.ftk_plan_tune.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_char_p, C.c_size_t]
    L.ftk_plan_scales.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
    L.ftk_stack_fp64.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32), C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_char_p, C.c_size_t]
    L.ftk_copy_f32.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
    L.ftk_probe_lds.argtypes = [C.c_int, C.POINTER(C.c_double)]
    L.ftk_emulate_pass.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32), C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int32),
                                   C.c_char_p, C.c_size_t]
    L.ftk_plan_shape.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int,
                                 C.POINTER(C.c_int), C.c_char_p, C.c_size_t]
    _lib = L
    return L


def _err() -> C.Array:
    return C.create_string_buffer(1024)


def _raise(status: int, err) -> None:
    if status == FTK_OK:
        return
    msg = err.value.decode(errors="replace")
    if status == FTK_CANCELLED:
        raise Cancelled(msg or "cancelled")
    raise FtkError(msg)


def parse_options(spec: str = "") -> _Options:
    o, e = _Options(), _err()
    _raise(lib().ftk_parse_options(spec.encode(), C.byref(o), e, len(e)), e)
    return o


def probe_lds(width_bytes: int = 4) -> float:
    """Measured LDS read bandwidth of the current device, bytes/s (ftk_probe_lds)."""
    v = C.c_double()
    if lib().ftk_probe_lds(int(width_bytes), C.byref(v)) != FTK_OK:
        raise FtkError("LDS probe failed")
    return v.value


def device_count() -> int:
    e = _err()
    return int(lib().ftk_device_count(e, len(e)))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(C.POINTER(t))


# ------------------------------------------------------------------ reference-shaped types

@dataclass
class ShiftTable:
    """Integer shifts, motion j and frame i at j * frames + i (backend.hpp:56-69)."""
    motions: int
    frames: int
    dx: np.ndarray
    dy: np.ndarray

    @staticmethod
    def make(dx, dy) -> "ShiftTable":
        dx, dy = _i32(dx), _i32(dy)
        if dx.ndim != 2 or dx.shape != dy.shape:
            raise FtkError("stack pass: malformed shift table")
        return ShiftTable(dx.shape[0], dx.shape[1], dx, dy)


def linear_shift_table(t_h: Sequence[float], velocities: Sequence[Sequence[float]]) -> ShiftTable:
    """Straight-line shifts lround(v t) in double, halves away from zero
    (backend.hpp:87-97, backend.cpp:114-123)."""
    t = np.asarray(t_h, dtype=np.float64)
    v = np.asarray(velocities, dtype=np.float64).reshape(-1, 2)

    def lround(x):
        r = np.trunc(x)
        return (r + np.where(np.abs(x - r) >= 0.5, np.sign(x), 0.0)).astype(np.int64)

    dx = lround(v[:, :1] * t[None, :])
    dy = lround(v[:, 1:] * t[None, :])
    return ShiftTable.make(dx, dy)


@dataclass
class StackPass:
    """Inputs of one pass (backend.hpp:71-78). frames: (n, h, w) float32 or a list of (h, w)."""
    frames: object
    t_h: Optional[Sequence[float]] = None
    velocities: Optional[Sequence[Sequence[float]]] = None
    shifts: Optional[ShiftTable] = None
    min_count: int = 3
    threads: int = 0

    def frame_list(self) -> list[np.ndarray]:
        if isinstance(self.frames, np.ndarray) and self.frames.ndim == 3:
            return [_f32(f) for f in self.frames]
        return [_f32(f) for f in self.frames]

    def table(self) -> ShiftTable:
        """validate_pass (backend.cpp:89-108) and the shifts of the pass."""
        fr = self.frame_list()
        if not fr:
            raise FtkError("stack pass: no frames")
        h, w = fr[0].shape
        if any(f.shape != (h, w) for f in fr):
            raise FtkError("stack pass: frames differ in size")
        if self.shifts is not None:
            s = self.shifts
            if s.frames != len(fr):
                raise FtkError("stack pass: shift table and frame counts differ")
            if s.motions < 1 or s.dx.shape != (s.motions, s.frames) or s.dy.shape != s.dx.shape:
                raise FtkError("stack pass: malformed shift table")
            if self.velocities is not None and len(self.velocities) != s.motions:
                raise FtkError("stack pass: shift table and velocity counts differ")
            tab = s
        else:
            if self.t_h is None or self.velocities is None:
                raise FtkError("stack pass: missing times or velocities")
            if len(self.t_h) != len(fr):
                raise FtkError("stack pass: frame and time counts differ")
            tab = linear_shift_table(self.t_h, self.velocities)
        if self.min_count < 1:
            raise FtkError("stack pass: min_count must be at least 1")
        return tab


@dataclass
class StackResult:
    """max_snr (h, w) float32, best_v (h, w) int32, timings [(name, seconds)] (backend.hpp:99-105)."""
    max_snr: np.ndarray
    best_v: np.ndarray
    timings: list = field(default_factory=list)


ProgressFn = Callable[[float, str], bool]


class HipBackend:
    """SearchBackend "hip" (backend.hpp:107-117) over ftk_stack_pass."""

    def __init__(self, options: str = ""):
        self.options_spec = options
        self.opt = parse_options(options)
        why = _err()
        if lib().ftk_device_count(why, len(why)) <= 0:
            raise FtkError("HIP backend unavailable: " + why.value.decode(errors="replace"))

    def name(self) -> str:
        return "hip"

    def description(self) -> str:
        b = _err()
        lib().ftk_device_description(self.opt.device, b, len(b))
        return b.value.decode(errors="replace")

    def run(self, p: StackPass, progress: Optional[ProgressFn] = None) -> StackResult:
        tab = p.table()
        fr = p.frame_list()
        n, (h, w) = len(fr), fr[0].shape
        ptrs = (C.POINTER(C.c_float) * n)(*[_ptr(f, C.c_float) for f in fr])
        ps = _Pass(n, w, h, ptrs, tab.motions, _ptr(tab.dx, C.c_int32), _ptr(tab.dy, C.c_int32), int(p.min_count))
        out_m = np.empty((h, w), np.float32)
        out_b = np.empty((h, w), np.int32)
        tim = _Timings()
        cb = PROGRESS_FN(0)
        if progress is not None:
            cb = PROGRESS_FN(lambda f, s, u: 1 if progress(f, s.decode()) else 0)
        e = _err()
        st = lib().ftk_stack_pass(C.byref(ps), C.byref(self.opt), _ptr(out_m, C.c_float), _ptr(out_b, C.c_int32),
                                  C.byref(tim), cb, None, e, len(e))
        _raise(st, e)
        return StackResult(out_m, out_b, [("upload", tim.upload_s), ("kernel", tim.kernel_s),
                                          ("download", tim.download_s), ("setup", tim.setup_s)])


def backend_names() -> list[str]:
    """Backends this package provides: "hip" when a device is usable (backend.cpp:46-52)."""
    try:
        return ["hip"] if device_count() > 0 else []
    except FtkError:
        return []


def make_backend(name: str = "auto") -> HipBackend:
    """make_backend (backend.cpp:54-87) for the names this package serves:
    "hip", "hip:key=value,..." and "auto" (= "hip"). Unknown names raise."""
    if name in ("", "auto", "hip"):
        return HipBackend("")
    if name.startswith("hip:"):
        return HipBackend(name[4:])
    raise FtkError(f"unknown or unavailable search backend '{name}' (available: {', '.join(backend_names() + ['auto'])})")


def plan_shape(n: int, w: int, h: int, dx, dy, min_count: int) -> dict:
    """Host-only planning summary of a pass (ftk_plan_shape): velocity blocks, lattice
    blocks (stacked by the lattice kernel) and the motions in them, the largest vertical
    spread of a block, segments per tile, blocks per segment, and the motions not
    stacked because a lower-index motion has the same shifts in every frame."""
    dx, dy = _i32(dx), _i32(dy)
    v = (C.c_int * 7)()
    e = _err()
    _raise(lib().ftk_plan_shape(n, w, h, dx.shape[0], _ptr(dx, C.c_int32), _ptr(dy, C.c_int32), int(min_count), v, e,
                                len(e)), e)
    return dict(zip(["blocks", "lattice_blocks", "lattice_motions", "max_sy", "segments", "blocks_per_segment",
                     "shadowed_motions"], list(v)))


def emulate_pass(frames: np.ndarray, dx, dy, min_count: int) -> tuple[np.ndarray, np.ndarray]:
    """Host emulation of the HIP algorithm (ftk_emulate_pass)."""
    fr = _f32(frames)
    n, h, w = fr.shape
    dx, dy = _i32(dx), _i32(dy)
    out_m = np.empty((h, w), np.float32)
    out_b = np.empty((h, w), np.int32)
    e = _err()
    st = lib().ftk_emulate_pass(n, w, h, _ptr(fr, C.c_float), dx.shape[0], _ptr(dx, C.c_int32), _ptr(dy, C.c_int32),
                                int(min_count), _ptr(out_m, C.c_float), _ptr(out_b, C.c_int32), e, len(e))
    _raise(st, e)
    return out_m, out_b


class Plan:
    """Device-resident pass (ftk_plan_create / ftk_plan_run): frames already in HBM."""

    def __init__(self, n: int, w: int, h: int, dx, dy, min_count: int, options: str = ""):
        self._dx, self._dy = _i32(dx), _i32(dy)
        self.n, self.w, self.h = n, w, h
        o = parse_options(options)
        hnd, e = C.c_void_p(), _err()
        _raise(lib().ftk_plan_create(n, w, h, self._dx.shape[0], _ptr(self._dx, C.c_int32), _ptr(self._dy, C.c_int32),
                                     int(min_count), C.byref(o), C.byref(hnd), e, len(e)), e)
        self._h = hnd

    @property
    def units(self) -> float:
        return float(lib().ftk_plan_units(self._h))

    @property
    def blocks(self) -> int:
        return int(lib().ftk_plan_blocks(self._h))

    def run_ptr(self, d_frames: int, frame_stride: int, d_max: int, d_best: int, stream: int = 0) -> None:
        e = _err()
        _raise(lib().ftk_plan_run(self._h, C.c_void_p(d_frames), frame_stride, C.c_void_p(d_max), C.c_void_p(d_best),
                                  C.c_void_p(stream), e, len(e)), e)

    def run(self, frames, max_snr=None, best_v=None, stream=None):
        """frames: torch tensor (n, h, w) float32 on the device. Returns (max_snr, best_v) tensors."""
        import torch
        if frames.dtype != torch.float32 or not frames.is_cuda or frames.dim() != 3:
            raise FtkError("Plan.run needs a (n, h, w) float32 device tensor")
        if tuple(frames.shape[1:]) != (self.h, self.w) or frames.shape[0] != self.n:
            raise FtkError("Plan.run: frame shape differs from the plan")
        if frames.stride(2) != 1 or frames.stride(1) != self.w:
            raise FtkError("Plan.run: frames must be row-contiguous")
        if max_snr is None:
            max_snr = torch.empty((self.h, self.w), dtype=torch.float32, device=frames.device)
        if best_v is None:
            best_v = torch.empty((self.h, self.w), dtype=torch.int32, device=frames.device)
        s = stream if stream is not None else torch.cuda.current_stream(frames.device)
        self.run_ptr(frames.data_ptr(), frames.stride(0), max_snr.data_ptr(), best_v.data_ptr(), s.cuda_stream)
        return max_snr, best_v

    def store(self, frames, storage: str = "f16", stream=None) -> None:
        """Encodes (n, h, w) fp32 device frames into the plan's 16-bit storage ("f16" or
        "bf16"): q = RNE(g / s_i) with s_i = 2^(e - 14), max|g_i| = m 2^e."""
        import torch
        self._check_frames(frames)
        code = STORAGES[storage]
        if code == 0:
            raise FtkError("Plan.store: storage must be 'f16' or 'bf16'")
        s = stream if stream is not None else torch.cuda.current_stream(frames.device)
        e = _err()
        _raise(lib().ftk_plan_store(self._h, C.c_void_p(frames.data_ptr()), frames.stride(0), code, C.c_void_p(s.cuda_stream),
                                    e, len(e)), e)
        self._stored_device = frames.device

    def run_stored(self, max_snr=None, best_v=None, stream=None):
        """A pass over the stored 16-bit frames. Returns (max_snr, best_v) tensors."""
        import torch
        dev = getattr(self, "_stored_device", None)
        if dev is None:
            raise FtkError("Plan.run_stored: no frames stored")
        if max_snr is None:
            max_snr = torch.empty((self.h, self.w), dtype=torch.float32, device=dev)
        if best_v is None:
            best_v = torch.empty((self.h, self.w), dtype=torch.int32, device=dev)
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        e = _err()
        _raise(lib().ftk_plan_run_stored(self._h, C.c_void_p(max_snr.data_ptr()), C.c_void_p(best_v.data_ptr()),
                                         C.c_void_p(s.cuda_stream), e, len(e)), e)
        return max_snr, best_v

    def tune(self, frames=None, reps: int = 2, stream=None):
        """Times the pass at each launch form and tile height (ftk_plan_tune) on these
        fp32 frames, or on the stored 16-bit frames when frames is None, and keeps the
        fastest for that storage's later runs. Returns (tile_rows, queue, seconds)."""
        import torch
        if frames is not None:
            self._check_frames(frames)
            dev = frames.device
        else:
            dev = getattr(self, "_stored_device", None)
            if dev is None:
                raise FtkError("Plan.tune: give fp32 frames or store 16-bit frames first")
        m = torch.empty((self.h, self.w), dtype=torch.float32, device=dev)
        b = torch.empty((self.h, self.w), dtype=torch.int32, device=dev)
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        tr, q, t, e = C.c_int(), C.c_int(), C.c_double(), _err()
        _raise(lib().ftk_plan_tune(self._h, C.c_void_p(frames.data_ptr() if frames is not None else None),
                                   frames.stride(0) if frames is not None else 0, C.c_void_p(m.data_ptr()),
                                   C.c_void_p(b.data_ptr()), reps, C.c_void_p(s.cuda_stream), C.byref(tr), C.byref(q),
                                   C.byref(t), e, len(e)), e)
        return tr.value, q.value, t.value

    def scales(self) -> np.ndarray:
        """The per-frame scales of the stored frames."""
        out = np.empty(self.n, np.float32)
        if lib().ftk_plan_scales(self._h, out.ctypes.data_as(C.POINTER(C.c_float))) != 0:
            raise FtkError("Plan.scales: no frames stored")
        return out

    def _check_frames(self, frames):
        import torch
        if frames.dtype != torch.float32 or not frames.is_cuda or frames.dim() != 3:
            raise FtkError("Plan needs a (n, h, w) float32 device tensor")
        if tuple(frames.shape) != (self.n, self.h, self.w) or frames.stride(2) != 1 or frames.stride(1) != self.w:
            raise FtkError("Plan: frames must be row-contiguous (n, h, w) of the plan's shape")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.ftk_plan_destroy(h)
            self._h = None


STORAGES = {"f32": 0, "f16": 1, "bf16": 2}


def stack_fp64(frames, dx, dy, min_count: int, y0: int = 0, y1=None):
    """The stacking contract in FP64 on the GPU (ftk_stack_fp64) for output rows
    [y0, y1) of (n, h, w) fp32 device frames: returns (max_snr float64, best_v int32)
    device tensors of (y1 - y0, w)."""
    import torch
    n, h, w = frames.shape
    y1 = h if y1 is None else y1
    dx, dy = _i32(dx), _i32(dy)
    if frames.dtype != torch.float32 or not frames.is_cuda or frames.stride(2) != 1 or frames.stride(1) != w:
        raise FtkError("stack_fp64 needs row-contiguous (n, h, w) float32 device frames")
    out = torch.empty((max(0, y1 - y0), w), dtype=torch.float64, device=frames.device)
    best = torch.empty((max(0, y1 - y0), w), dtype=torch.int32, device=frames.device)
    s = torch.cuda.current_stream(frames.device)
    e = _err()
    _raise(lib().ftk_stack_fp64(n, w, h, C.c_void_p(frames.data_ptr()), frames.stride(0), dx.shape[0],
                                _ptr(dx, C.c_int32), _ptr(dy, C.c_int32), int(min_count), int(y0), int(y1),
                                C.c_void_p(out.data_ptr()), C.c_void_p(best.data_ptr()), C.c_void_p(s.cuda_stream),
                                e, len(e)), e)
    return out, best


def velocity_grid(span_h: float, vmin_pxh: float, vmax_pxh: float, grid_tol_px: float):
    """Trial motions of the search (search.cpp:39-58): dv = 4 max(0.05, tol) / span,
    K = ceil(vmax / dv), vy-major then vx over -K..K, kept when
    max(0, vmin - dv/2) <= hypot(vx, vy) <= vmax + dv/2. Returns ((m, 2) array, dv).
    The row order defines the motion indices and therefore the tie-breaking."""
    if not span_h > 0.0:
        raise FtkError("invalid time span")
    if vmax_pxh <= 0.0 or vmax_pxh < vmin_pxh:
        raise FtkError("invalid velocity range")
    dv = 4.0 * max(0.05, grid_tol_px) / span_h
    K = int(np.ceil(vmax_pxh / dv))
    lo, hi = max(0.0, vmin_pxh - 0.5 * dv), vmax_pxh + 0.5 * dv
    idx = np.arange(-K, K + 1, dtype=np.float64)
    vy, vx = np.meshgrid(idx * dv, idx * dv, indexing="ij")
    sp = np.hypot(vx, vy)
    keep = (sp >= lo) & (sp <= hi)
    grid = np.stack([vx[keep], vy[keep]], axis=1)
    if grid.shape[0] > 5_000_000:
        raise FtkError("velocity grid too large; narrow the range")
    return grid, dv
