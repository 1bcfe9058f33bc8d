"""Device-resident blind search (ftk_searcher_*): prepare, the real pass, the
time-shuffled null runs, the null-tail threshold and peak selection on the GPU.

Mirrors the reference's search() (src/search.cpp:716-809) and the C API's frame
handling (stable sort by MJD, src/c_api.cpp:178-180).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace
from typing import Optional, Sequence

import numpy as np

from . import FtkError, _err, _raise, lib

_bound = False


class _Params(C.Structure):
    _fields_ = [("vmin_pxh", C.c_double), ("vmax_pxh", C.c_double), ("grid_tol_px", C.c_double),
                ("fwhm_px", C.c_double), ("clip_sigma", C.c_double), ("cr_sigma", C.c_double),
                ("static_exclude", C.c_int), ("mask_scatter", C.c_double), ("weight_scatter", C.c_double),
                ("min_coverage", C.c_double), ("threshold_snr", C.c_double), ("min_snr_floor", C.c_double),
                ("null_runs", C.c_int), ("target_false_per_run", C.c_double), ("max_candidates", C.c_int),
                ("edge_px", C.c_int), ("seed", C.c_ulonglong), ("fit_on_device", C.c_int),
                ("passes_per_launch", C.c_int), ("segments", C.c_int), ("storage", C.c_int),
                ("null_tested", C.c_int), ("max_frame_fraction", C.c_double),
                ("reg_dx", C.POINTER(C.c_double)), ("reg_dy", C.POINTER(C.c_double)),
                ("candidates", C.c_int), ("half_tolerance", C.c_double), ("tile_rows", C.c_int),
                ("queue", C.c_int), ("build_lanes", C.c_int), ("neo_min_dist_au", C.c_double),
                ("neo_max_trials", C.c_int), ("par_x", C.POINTER(C.c_double)), ("par_y", C.POINTER(C.c_double))]


class _Summary(C.Structure):
    _fields_ = [("threshold_snr", C.c_double), ("method", C.c_int), ("null_a", C.c_double), ("null_k", C.c_double),
                ("null_inv_s2", C.c_double), ("null_b", C.c_double), ("dv_pxh", C.c_double),
                ("t_ref_mjd", C.c_double), ("span_h", C.c_double), ("n_velocities", C.c_int), ("n_peaks", C.c_int),
                ("n_selected", C.c_int), ("n_pooled", C.c_int), ("t_prepare", C.c_double),
                ("t_stack_real", C.c_double), ("t_null_rebuild", C.c_double), ("t_stack_null", C.c_double),
                ("t_peaks", C.c_double), ("t_threshold", C.c_double), ("t_plan", C.c_double),
                ("t_passes", C.c_double), ("passes_per_launch", C.c_int), ("segments", C.c_int),
                ("n_screened", C.c_int), ("n_artefacts", C.c_int), ("n_candidates", C.c_int),
                ("t_refine", C.c_double), ("tile_rows", C.c_int), ("queue", C.c_int), ("n_distances", C.c_int),
                ("dq_inv_au", C.c_double)]


# Candidate's fields (include/fainttrail/search.hpp:131-154) that the record carries, in
# ftk_candidate's order after id and ok
CANDIDATE_FIELDS = ["x", "y", "vx_pxh", "vy_pxh", "pos_err_px", "vel_err_pxh", "snr", "snr_coarse", "snr_half1",
                    "snr_half2", "max_frame_fraction", "false_alarms_per_run", "flux", "mag_inst", "ra_deg",
                    "dec_deg", "rate_arcsec_min", "pa_deg", "inv_dist_au", "inv_dist_err_au"]


class _Candidate(C.Structure):
    _fields_ = [("id", C.c_int), ("ok", C.c_int)] + [(f, C.c_double) for f in CANDIDATE_FIELDS] + \
        [("flags", C.c_char * 64)]


METHODS = ["null-fit gauss", "null-fit linear", "null-quantile", "fixed", "gaussian"]


def _bind():
    global _bound
    L = lib()
    if _bound:
        return L
    vp, sz = C.c_void_p, C.c_size_t
    L.ftk_search_params_default.argtypes = [C.POINTER(_Params)]
    L.ftk_searcher_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(vp), C.c_char_p, sz]
    L.ftk_searcher_destroy.argtypes = [vp]
    L.ftk_searcher_run.argtypes = [vp, vp, sz, C.POINTER(C.c_double), C.POINTER(_Params), C.POINTER(_Summary), vp,
                                   vp, vp, C.c_char_p, sz]
    L.ftk_searcher_tune.argtypes = [vp, vp, sz, C.POINTER(C.c_double), C.POINTER(_Params), C.c_int,
                                    C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_double), vp, C.c_char_p, sz]
    L.ftk_searcher_peaks.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_float), C.POINTER(C.c_int32)]
    L.ftk_searcher_candidates.argtypes = [vp, C.c_int, C.POINTER(_Candidate)]
    L.ftk_searcher_refit.argtypes = [vp, C.POINTER(_Params), C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int), vp,
                                     C.c_char_p, sz]
    I32 = C.POINTER(C.c_int32)
    L.ftk_find_peaks.argtypes = [vp, vp, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, I32, I32,
                                 C.POINTER(C.c_float), I32, C.POINTER(C.c_int), vp, C.c_char_p, sz]
    _bound = True
    return L


@dataclass
class SearchParams:
    """The SearchParams fields on the path (include/fainttrail/search.hpp:18-90), same
    names and defaults; fwhm_px must be given."""
    vmin_pxh: float = 4.0
    vmax_pxh: float = 40.0
    grid_tol_px: float = 0.5
    fwhm_px: float = 3.0
    clip_sigma: float = 8.0
    cr_sigma: float = 5.0
    static_exclude: int = -1
    mask_scatter: float = 2.5
    weight_scatter: float = 1.5
    min_coverage: float = 0.7
    threshold_snr: float = 0.0
    min_snr_floor: float = 5.0
    null_runs: int = 3
    target_false_per_run: float = 0.02
    max_candidates: int = 300
    edge_px: int = 4
    seed: int = 12345
    # where the FP64 null-tail fit sums run: True on the GPU (threshold within ~m eps
    # relative of the host fit, same method), False on the host (the reference's bits),
    # None on the GPU for 16-bit storage, else on the host
    fit_on_device: Optional[bool] = None
    # batched form (no result depends on these): passes per stacking launch (1 = one
    # pass at a time with a plan per pass; 0 = tuned, else all passes) and
    # velocity-block segments per tile (0 = planner / tuned)
    passes_per_launch: int = 0
    segments: int = 0
    # frame storage of the stacking passes: "f32" (exact), "f16" or "bf16" (a
    # power-of-two scale per frame, fp32 accumulation)
    storage: str = "f32"
    # null-peak screen: the strongest null_tested peaks of each null run go through the
    # artefact tests (0 = none); max_frame_fraction is the single-frame test's limit
    null_tested: int = 3000
    max_frame_fraction: float = 0.25
    # registration offsets of the frames (detector = registered + offset), one per
    # frame, or None: with them the detector-fixed test runs
    reg_dx: Optional[Sequence[float]] = None
    reg_dy: Optional[Sequence[float]] = None
    # candidate stage: refine the selected peaks into candidates (refine and
    # finish_candidates); half_tolerance is the halves test's allowed shortfall (sigma)
    candidates: bool = True
    half_tolerance: float = 3.0
    # stacking work shape (no result depends on these): output rows per tile (8, 16
    # or 32; 0 = tuned, else automatic) and the launch form (1: one persistent launch
    # per group of passes whose workgroups take work items from device counters; 0:
    # one workgroup per work item, 16-row tiles; -1 = tuned, else automatic)
    tile_rows: int = 0
    queue: int = -1
    # stage builds of a group in parallel on this many streams (1-4; 0 = by frame size)
    build_lanes: int = 0
    # near-Earth search (neo_min_dist_au > 0): curved tracks over the inverse distances
    # of inverse_distance_grid times the velocity grid; par_x / par_y: each frame's
    # parallax displacement at 1 AU in pixels, less its least-squares line in t
    # (Prepared::par_x / par_y), one per frame
    neo_min_dist_au: float = 0.0
    neo_max_trials: int = 64
    par_x: Optional[Sequence[float]] = None
    par_y: Optional[Sequence[float]] = None

    def c(self) -> _Params:
        from . import STORAGES
        if self.storage not in STORAGES:
            raise FtkError(f"unknown frame storage '{self.storage}' (f32, f16, bf16)")
        if (self.reg_dx is None) != (self.reg_dy is None):
            raise FtkError("give both registration offsets (reg_dx, reg_dy) or none")
        p = _Params(self.vmin_pxh, self.vmax_pxh, self.grid_tol_px, self.fwhm_px, self.clip_sigma, self.cr_sigma,
                    self.static_exclude, self.mask_scatter, self.weight_scatter, self.min_coverage,
                    self.threshold_snr, self.min_snr_floor, self.null_runs, self.target_false_per_run,
                    self.max_candidates, self.edge_px, self.seed,
                    -1 if self.fit_on_device is None else int(bool(self.fit_on_device)),
                    self.passes_per_launch, self.segments, STORAGES[self.storage], self.null_tested,
                    self.max_frame_fraction, None, None, int(bool(self.candidates)), self.half_tolerance,
                    self.tile_rows, int(self.queue), int(self.build_lanes), float(self.neo_min_dist_au),
                    int(self.neo_max_trials), None, None)
        keep = []  # the arrays live as long as the struct
        for a, b, what in ((self.reg_dx, self.reg_dy, "registration offsets (reg_dx, reg_dy)"),
                           (self.par_x, self.par_y, "parallax displacements (par_x, par_y)")):
            if (a is None) != (b is None):
                raise FtkError(f"give both {what} or none")
            if a is None:
                continue
            ax = np.ascontiguousarray(a, dtype=np.float64)
            ay = np.ascontiguousarray(b, dtype=np.float64)
            if ax.shape != ay.shape or ax.ndim != 1:
                raise FtkError(f"{what}: one value per frame in each")
            keep += [ax, ay]
        it = iter(keep)
        if self.reg_dx is not None:
            p.reg_dx, p.reg_dy = (next(it).ctypes.data_as(C.POINTER(C.c_double)) for _ in range(2))
        if self.par_x is not None:
            p.par_x, p.par_y = (next(it).ctypes.data_as(C.POINTER(C.c_double)) for _ in range(2))
        p._keep = keep
        return p


@dataclass
class Peaks:
    x: np.ndarray
    y: np.ndarray
    snr: np.ndarray
    vidx: np.ndarray


@dataclass
class Candidate:
    """A candidate (Candidate, include/fainttrail/search.hpp:131-154): 0-based position
    at the reference epoch, motion, S/N figures, photometry and flags."""
    id: int
    ok: bool
    x: float
    y: float
    vx_pxh: float
    vy_pxh: float
    pos_err_px: float
    vel_err_pxh: float
    snr: float
    snr_coarse: float
    snr_half1: float
    snr_half2: float
    max_frame_fraction: float
    false_alarms_per_run: float
    flux: float
    mag_inst: float
    ra_deg: float
    dec_deg: float
    rate_arcsec_min: float
    pa_deg: float
    inv_dist_au: float
    inv_dist_err_au: float
    flags: str


def read_candidates(fetch, count: int) -> list:
    """Candidates through fetch(cap, records) -> count (ftk_searcher_candidates or
    ftk_server_candidates)."""
    if count <= 0:
        return []
    recs = (_Candidate * count)()
    got = fetch(count, recs)
    if got != count:
        raise FtkError(f"candidate count changed ({got} != {count})")
    return [Candidate(r.id, bool(r.ok), *(getattr(r, f) for f in CANDIDATE_FIELDS), r.flags.decode())
            for r in recs]


@dataclass
class SearchResult:
    threshold_snr: float
    threshold_method: str
    null_a: float
    null_k: float
    null_inv_s2: float
    null_b: float
    n_velocities: int
    dv_pxh: float
    t_ref_mjd: float
    peaks: Peaks
    selected: Peaks
    null_peaks: np.ndarray
    timings: dict = field(default_factory=dict)
    max_snr: object = None
    best_v: object = None
    passes_per_launch: int = 1
    segments: int = 0
    null_screened: int = 0   # null peaks through the artefact tests
    null_artefacts: int = 0  # of those, left out of null_peaks
    candidates: list = field(default_factory=list)  # Candidate records, the reference's order
    tile_rows: int = 16      # the stacking work shape the run used
    queue: int = 0
    n_distances: int = 1     # inverse distances of the motions (near-Earth search; 1: straight tracks)
    dq_inv_au: float = 0.0   # their step


class Searcher:
    """Device scratch for searches over n frames of w x h."""

    def __init__(self, n: int, w: int, h: int, device: int = 0):
        L = _bind()
        hnd, e = C.c_void_p(), _err()
        _raise(L.ftk_searcher_create(n, w, h, device, C.byref(hnd), e, len(e)), e)
        self._h, self.n, self.w, self.h = hnd, n, w, h

    def _peaks(self, which: int, count: int) -> Peaks:
        L = lib()
        x, y, v = (np.empty(count, np.int32) for _ in range(3))
        s = np.empty(count, np.float32)
        I32 = C.POINTER(C.c_int32)
        L.ftk_searcher_peaks(self._h, which, count, x.ctypes.data_as(I32), y.ctypes.data_as(I32),
                             s.ctypes.data_as(C.POINTER(C.c_float)), v.ctypes.data_as(I32))
        return Peaks(x, y, s, v)

    def _check(self, frames, mjd):
        import torch
        if (not isinstance(frames, torch.Tensor) or not frames.is_cuda or frames.dtype != torch.float32
                or tuple(frames.shape) != (self.n, self.h, self.w) or frames.stride(2) != 1
                or frames.stride(1) != self.w):
            raise FtkError("Searcher.run needs a row-contiguous (n, h, w) float32 device tensor")
        m = np.ascontiguousarray(mjd, dtype=np.float64)
        if m.size != self.n:
            raise FtkError("frame and time counts differ")
        return m

    def tune(self, frames, mjd, params: SearchParams, reps: int = 2, stream=None):
        """Times the batched form over group sizes and segment counts, then launch forms
        and tile heights, on these frames and keeps the fastest for runs with
        passes_per_launch = 0 / segments = 0 / tile_rows = 0 / queue = -1. Returns
        (passes_per_launch, segments, tile_rows, queue, seconds per search); tile_rows 0
        / queue -1 when the automatic shape was fastest."""
        import torch
        m = self._check(frames, mjd)
        s = stream if stream is not None else torch.cuda.current_stream(frames.device)
        prm, e = params.c(), _err()
        g, sg, tr, q, t = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_double()
        _raise(lib().ftk_searcher_tune(self._h, frames.data_ptr(), frames.stride(0),
                                       m.ctypes.data_as(C.POINTER(C.c_double)), C.byref(prm), reps, C.byref(g),
                                       C.byref(sg), C.byref(tr), C.byref(q), C.byref(t), s.cuda_stream, e, len(e)), e)
        return g.value, sg.value, tr.value, q.value, t.value

    def run(self, frames, mjd, params: SearchParams, keep_maps: bool = True, stream=None) -> SearchResult:
        """frames: (n, h, w) float32 device tensor sorted by time; mjd: n mid-exposure MJDs."""
        import torch
        m = self._check(frames, mjd)
        mx = bx = None
        if keep_maps:
            mx = torch.empty((self.h, self.w), dtype=torch.float32, device=frames.device)
            bx = torch.empty((self.h, self.w), dtype=torch.int32, device=frames.device)
        s = stream if stream is not None else torch.cuda.current_stream(frames.device)
        summ, prm, e = _Summary(), params.c(), _err()
        _raise(lib().ftk_searcher_run(self._h, frames.data_ptr(), frames.stride(0),
                                      m.ctypes.data_as(C.POINTER(C.c_double)), C.byref(prm), C.byref(summ),
                                      mx.data_ptr() if mx is not None else None,
                                      bx.data_ptr() if bx is not None else None, s.cuda_stream, e, len(e)), e)
        pooled = self._peaks(2, summ.n_pooled).snr
        return SearchResult(
            summ.threshold_snr, METHODS[summ.method], summ.null_a, summ.null_k, summ.null_inv_s2, summ.null_b,
            summ.n_velocities, summ.dv_pxh, summ.t_ref_mjd, self._peaks(0, summ.n_peaks),
            self._peaks(1, summ.n_selected), pooled,
            {"prepare": summ.t_prepare, "stack (real)": summ.t_stack_real, "null rebuild": summ.t_null_rebuild,
             "stack (null)": summ.t_stack_null, "peaks": summ.t_peaks, "threshold fit": summ.t_threshold,
             "plan": summ.t_plan, "passes (batched)": summ.t_passes, "refine": summ.t_refine},
            mx, bx, summ.passes_per_launch, summ.segments, summ.n_screened, summ.n_artefacts,
            read_candidates(lambda cap, recs: lib().ftk_searcher_candidates(self._h, cap, recs), summ.n_candidates),
            summ.tile_rows, summ.queue, summ.n_distances, summ.dq_inv_au)

    def refit(self, params: SearchParams, on_device: bool, stream=None):
        """The last run's null-tail threshold fitted again with the FP64 fit sums on the
        GPU (on_device) or the host: (threshold, method)."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        prm, e, t, m = params.c(), _err(), C.c_double(), C.c_int()
        _raise(lib().ftk_searcher_refit(self._h, C.byref(prm), int(bool(on_device)), C.byref(t), C.byref(m),
                                        s.cuda_stream, e, len(e)), e)
        return t.value, METHODS[m.value]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().ftk_searcher_destroy(h)
            except Exception:  # interpreter shutdown: module globals may already be cleared
                pass
            self._h = None


def find_peaks(max_snr, best_v, floor: float, radius: int, edge: int, max_peaks: int, stream=None) -> Peaks:
    """find_peaks (src/search.cpp:105-135) on device maps: (h, w) float32 max_snr and
    int32 best_v tensors on the GPU. Returns the peaks in the reference's order."""
    import torch
    L = _bind()
    if (not isinstance(max_snr, torch.Tensor) or not max_snr.is_cuda or max_snr.dtype != torch.float32
            or not max_snr.is_contiguous() or max_snr.dim() != 2):
        raise FtkError("find_peaks needs a contiguous (h, w) float32 device tensor")
    if (not isinstance(best_v, torch.Tensor) or best_v.dtype != torch.int32 or best_v.shape != max_snr.shape
            or not best_v.is_contiguous() or best_v.device != max_snr.device):
        raise FtkError("find_peaks needs an int32 best_v tensor of the map's shape on the same device")
    h, w = max_snr.shape
    s = stream if stream is not None else torch.cuda.current_stream(max_snr.device)
    cap = max(1, min(max_peaks, w * h))
    x, y, v = (np.empty(cap, np.int32) for _ in range(3))
    q = np.empty(cap, np.float32)
    cnt, e = C.c_int(), _err()
    I32 = C.POINTER(C.c_int32)
    _raise(L.ftk_find_peaks(max_snr.data_ptr(), best_v.data_ptr(), w, h, floor, radius, edge, max_peaks, cap,
                            x.ctypes.data_as(I32), y.ctypes.data_as(I32), q.ctypes.data_as(C.POINTER(C.c_float)),
                            v.ctypes.data_as(I32), C.byref(cnt), s.cuda_stream, e, len(e)), e)
    c = cnt.value
    return Peaks(x[:c], y[:c], q[:c], v[:c])


def search(frames, mjd, params: SearchParams, device: int = 0) -> SearchResult:
    """Search a sequence given as a host (n, h, w) array or device tensor: frames are
    stable-sorted by MJD first, as ft_search does (src/c_api.cpp:178-180)."""
    import torch
    m = np.asarray(mjd, dtype=np.float64)
    order = np.argsort(m, kind="stable")
    if isinstance(frames, torch.Tensor):
        dev = frames[torch.as_tensor(order, device=frames.device)].contiguous()
    else:
        dev = torch.from_numpy(np.ascontiguousarray(np.asarray(frames, np.float32)[order])).to(f"cuda:{device}")
    n, h, w = dev.shape
    if params.reg_dx is not None and params.reg_dy is not None:
        # the offsets belong to the frames: same order
        params = replace(params, reg_dx=np.asarray(params.reg_dx, np.float64)[order],
                         reg_dy=np.asarray(params.reg_dy, np.float64)[order])
    if params.par_x is not None and params.par_y is not None:
        params = replace(params, par_x=np.asarray(params.par_x, np.float64)[order],
                         par_y=np.asarray(params.par_y, np.float64)[order])
    return Searcher(n, w, h, device).run(dev, m[order], params)
