"""Device preprocessing (ftk_stage_*): background and build_stage on frames in HBM.

The Python view of the reference's prepare() / with_times() steps that produce the
frames of each stacking pass (src/preprocess.cpp:286-486, 528-585), running on the
GPU and bit-identical to the reference.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import FtkError, _err, _raise, lib

_bound = False


class _StageParams(C.Structure):
    _fields_ = [("fwhm_px", C.c_double), ("clip_sigma", C.c_double), ("cr_sigma", C.c_double),
                ("static_exclude", C.c_int), ("mask_scatter", C.c_double), ("weight_scatter", C.c_double)]


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    vp, sz = C.c_void_p, C.c_size_t
    L.ftk_stage_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(vp), C.c_char_p, sz]
    L.ftk_stage_destroy.argtypes = [vp]
    L.ftk_stage_background.argtypes = [vp, vp, sz, vp, sz, C.POINTER(C.c_double), vp, C.c_char_p, sz]
    L.ftk_stage_build.argtypes = [vp, vp, sz, C.POINTER(C.c_double), C.POINTER(_StageParams), vp, sz, vp,
                                  C.POINTER(C.c_double), vp, C.POINTER(C.c_int32), vp, C.c_char_p, sz]
    _bound = True
    return L


@dataclass
class StageParams:
    """SearchParams fields of build_stage (include/fainttrail/search.hpp:18-90)."""
    fwhm_px: float = 3.0
    clip_sigma: float = 8.0
    cr_sigma: float = 5.0
    static_exclude: int = -1
    mask_scatter: float = 2.5
    weight_scatter: float = 1.5

    def c(self) -> _StageParams:
        return _StageParams(self.fwhm_px, self.clip_sigma, self.cr_sigma, self.static_exclude, self.mask_scatter,
                            self.weight_scatter)


def _dev_ptr(t, name):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32:
        raise FtkError(f"{name} must be a float32 device tensor")
    return t.data_ptr()


class Stage:
    """Scratch for one n x h x w frame stack on the device."""

    def __init__(self, n: int, w: int, h: int, device: int = 0):
        L = _bind()
        hnd, e = C.c_void_p(), _err()
        _raise(L.ftk_stage_create(n, w, h, device, C.byref(hnd), e, len(e)), e)
        self._h, self.n, self.w, self.h = hnd, n, w, h

    def _check(self, t, name):
        if tuple(t.shape) != (self.n, self.h, self.w) or t.stride(2) != 1 or t.stride(1) != self.w:
            raise FtkError(f"{name}: expected a row-contiguous ({self.n}, {self.h}, {self.w}) tensor")
        return _dev_ptr(t, name), t.stride(0)

    def background(self, raw, sub=None, stream=None):
        """(raw device frames) -> (sub device frames, background per frame)."""
        import torch
        if sub is None:
            sub = torch.empty_like(raw)
        pr, sr = self._check(raw, "raw")
        ps, ss = self._check(sub, "sub")
        bg = np.empty(self.n, np.float64)
        s = stream if stream is not None else torch.cuda.current_stream(raw.device)
        e = _err()
        _raise(lib().ftk_stage_background(self._h, pr, sr, ps, ss, bg.ctypes.data_as(C.POINTER(C.c_double)),
                                          s.cuda_stream, e, len(e)), e)
        return sub, bg

    def build(self, sub, t_h, params: StageParams, filtered=None, want_details: bool = False, stream=None):
        """build_stage for frame times t_h: returns the filtered device frames and, with
        want_details, (static_sky, mask) device maps and (sigma, cr_pixels) host arrays."""
        import torch
        if filtered is None:
            filtered = torch.empty_like(sub)
        ps, ss = self._check(sub, "sub")
        pf, sf = self._check(filtered, "filtered")
        t = np.ascontiguousarray(t_h, dtype=np.float64)
        if t.size != self.n:
            raise FtkError("build_stage: frame and time counts differ")
        prm = params.c()
        sky = mask = None
        sig = cr = None
        if want_details:
            sky = torch.empty((self.h, self.w), dtype=torch.float32, device=sub.device)
            mask = torch.empty((self.h, self.w), dtype=torch.float32, device=sub.device)
            sig = np.empty(self.n, np.float64)
            cr = np.empty(self.n, np.int32)
        s = stream if stream is not None else torch.cuda.current_stream(sub.device)
        e = _err()
        _raise(lib().ftk_stage_build(
            self._h, ps, ss, t.ctypes.data_as(C.POINTER(C.c_double)), C.byref(prm), pf, sf,
            sky.data_ptr() if sky is not None else None,
            sig.ctypes.data_as(C.POINTER(C.c_double)) if sig is not None else None,
            mask.data_ptr() if mask is not None else None,
            cr.ctypes.data_as(C.POINTER(C.c_int32)) if cr is not None else None,
            s.cuda_stream, e, len(e)), e)
        if want_details:
            return filtered, {"static_sky": sky, "mask": mask, "sigma": sig, "cr_pixels": cr}
        return filtered

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().ftk_stage_destroy(h)
            except Exception:  # interpreter shutdown: module globals may already be cleared
                pass
            self._h = None
