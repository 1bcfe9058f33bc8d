"""Device-resident search server (ftk_server_*): continuous batching of search jobs
of different frame sizes, frame counts and velocity grids on one GPU.

Jobs queue in arrival order; a free worker slot admits the first queued job whose HBM
reservation fits the budget beside the running jobs (a job may overtake the head of
the queue only while the head has been overtaken fewer than 16 times); each slot runs
its job's batched search on its own streams, so jobs overlap on the device. A job's
results depend on its inputs only, the same alone and in any mix.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import FtkError, _err, _raise, lib
from .search import METHODS, Peaks, SearchParams, _Candidate, _Params, _Summary, read_candidates

_bound = False


def _bind():
    global _bound
    L = lib()
    if _bound:
        return L
    vp, sz, ll = C.c_void_p, C.c_size_t, C.c_longlong
    L.ftk_server_create.argtypes = [C.c_int, C.c_double, C.c_int, C.POINTER(vp), C.c_char_p, sz]
    L.ftk_server_destroy.argtypes = [vp]
    L.ftk_server_submit.argtypes = [vp, vp, sz, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(_Params),
                                    C.POINTER(ll), C.c_char_p, sz]
    L.ftk_server_wait.argtypes = [vp, ll, C.c_double, C.POINTER(_Summary), C.POINTER(C.c_double), C.c_char_p, sz]
    L.ftk_server_peaks.argtypes = [vp, ll, C.c_int, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_float), C.POINTER(C.c_int32)]
    L.ftk_server_candidates.argtypes = [vp, ll, C.c_int, C.POINTER(_Candidate)]
    L.ftk_server_release.argtypes = [vp, ll]
    L.ftk_server_stats.argtypes = [vp, C.POINTER(C.c_double)]
    L.ftk_server_fusion.argtypes = [vp, C.c_int, C.c_double]
    _bound = True
    return L


@dataclass
class JobResult:
    job: int
    threshold_snr: float
    threshold_method: str
    n_velocities: int
    peaks: Peaks
    selected: Peaks
    null_peaks: np.ndarray
    t_submit: float
    t_admit: float
    t_end: float
    reserved_bytes: float
    passes_per_launch: int
    held_bytes: float = 0.0  # device bytes the job's searcher held when it ended
    candidates: list = field(default_factory=list)  # Candidate records, the reference's order
    extra: dict = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.t_end - self.t_submit


class Server:
    """A device-resident search server on one GPU: `slots` jobs at once within
    budget_mb of HBM (0: 85 % of the free memory)."""

    def __init__(self, device: int = 0, budget_mb: float = 0.0, slots: int = 4):
        L = _bind()
        hnd, e = C.c_void_p(), _err()
        _raise(L.ftk_server_create(device, float(budget_mb), int(slots), C.byref(hnd), e, len(e)), e)
        self._h = hnd
        self._frames = {}

    def submit(self, frames, mjd, params: SearchParams) -> int:
        """frames: row-contiguous (n, h, w) float32 device tensor sorted by time (kept
        alive by the server until the job is waited for); mjd: n mid-exposure MJDs."""
        import torch
        if (not isinstance(frames, torch.Tensor) or not frames.is_cuda or frames.dtype != torch.float32
                or frames.dim() != 3 or frames.stride(2) != 1 or frames.stride(1) != frames.shape[2]):
            raise FtkError("Server.submit needs a row-contiguous (n, h, w) float32 device tensor")
        n, h, w = frames.shape
        m = np.ascontiguousarray(mjd, dtype=np.float64)
        if m.size != n:
            raise FtkError("frame and time counts differ")
        torch.cuda.current_stream(frames.device).synchronize()  # the frames are complete before the job reads them
        prm, job, e = params.c(), C.c_longlong(), _err()
        _raise(lib().ftk_server_submit(self._h, frames.data_ptr(), frames.stride(0), n, w, h,
                                       m.ctypes.data_as(C.POINTER(C.c_double)), C.byref(prm), C.byref(job), e,
                                       len(e)), e)
        self._frames[job.value] = frames
        return job.value

    def _peaks(self, job: int, which: int, count: int) -> Peaks:
        x, y, v = (np.empty(count, np.int32) for _ in range(3))
        s = np.empty(count, np.float32)
        I32 = C.POINTER(C.c_int32)
        lib().ftk_server_peaks(self._h, job, which, count, x.ctypes.data_as(I32), y.ctypes.data_as(I32),
                               s.ctypes.data_as(C.POINTER(C.c_float)), v.ctypes.data_as(I32))
        return Peaks(x, y, s, v)

    def wait(self, job: int, timeout: float = -1.0):
        """The job's result, or None if it has not ended within timeout seconds."""
        summ, times, e = _Summary(), (C.c_double * 5)(), _err()
        st = lib().ftk_server_wait(self._h, job, float(timeout), C.byref(summ), times, e, len(e))
        if st == 3:
            return None
        self._frames.pop(job, None)
        try:
            _raise(st, e)
            return JobResult(job, summ.threshold_snr, METHODS[summ.method], summ.n_velocities,
                             self._peaks(job, 0, summ.n_peaks), self._peaks(job, 1, summ.n_selected),
                             self._peaks(job, 2, summ.n_pooled).snr, times[0], times[1], times[2], times[3],
                             summ.passes_per_launch, times[4],
                             read_candidates(lambda cap, recs: lib().ftk_server_candidates(self._h, job, cap, recs),
                                             summ.n_candidates))
        finally:
            # ended jobs are forgotten by the server, failed ones included
            lib().ftk_server_release(self._h, job)

    def stats(self) -> dict:
        v = (C.c_double * 13)()
        lib().ftk_server_stats(self._h, v)
        keys = ["reserved_bytes", "peak_reserved_bytes", "budget_bytes", "idle_searcher_bytes", "running", "queued",
                "completed", "failed", "evicted", "max_running", "stack_launches", "fused_launches", "fused_passes"]
        return {k: (int(x) if i >= 4 else float(x)) for i, (k, x) in enumerate(zip(keys, v))}

    def fusion(self, on: bool = True, window_ms: float = 3.0) -> None:
        """Cross-job fusion of persistent stacking launches (ftk_server_fusion)."""
        if lib().ftk_server_fusion(self._h, int(bool(on)), float(window_ms)) != 0:
            raise FtkError("ftk_server_fusion: window_ms must be >= 0")

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().ftk_server_destroy(h)
            self._h = None
        self._frames = {}

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: module globals may already be cleared
            pass


@dataclass
class JobSpec:
    arrival_s: float
    side: int
    frames: int
    velocities: int
    nulls: int
    movers: int
    seed: int


def poisson_jobs(count: int, rate_hz: float, seed: int, side=(256, 4096), frames=(10, 120),
                 velocities=(500, 50000), nulls=(4, 64), movers=(0, 15)):
    """A seeded Poisson stream of search jobs: exponential inter-arrival times of mean
    1/rate; side, frame count, velocity count and null runs log-uniform in the given
    ranges; movers uniform."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []

    def logu(lo, hi):
        return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))

    for k in range(count):
        t += float(rng.exponential(1.0 / rate_hz))
        out.append(JobSpec(t, logu(*side), logu(*frames), logu(*velocities), logu(*nulls),
                           int(rng.integers(movers[0], movers[1] + 1)), int(rng.integers(0, 2**31 - 1))))
    return out


def job_inputs(spec: JobSpec, device: str = "cuda:0", vmin_pxh: float = 12.0, vmax_pxh: float = 66.0):
    """Synthetic frames (60 s cadence, 40 stars, the spec's movers at per-frame S/N 1-3)
    and search parameters whose velocity grid has about spec.velocities motions: the
    grid of an annulus vmin..vmax spaced dv has about pi (vmax^2 - vmin^2) / dv^2
    motions, and dv = 4 grid_tol / span."""
    from . import synth
    frames, mjd, _ = synth.make_sequence(spec.frames, spec.side, spec.side, nstars=40, nmovers=spec.movers,
                                         mover_snr=(1.0, 3.0), seed=spec.seed, device=device)
    span_h = (mjd[-1] - mjd[0]) * 24.0
    dv = math.sqrt(math.pi * (vmax_pxh ** 2 - vmin_pxh ** 2) / spec.velocities)
    tol = max(0.05, dv * span_h / 4.0)
    prm = SearchParams(vmin_pxh=vmin_pxh, vmax_pxh=vmax_pxh, grid_tol_px=tol, fwhm_px=1.5 * 2.354820045030949,
                       null_runs=spec.nulls, seed=spec.seed)
    return frames, mjd, prm
