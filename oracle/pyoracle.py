"""ctypes view of the CPU oracle (oracle/liboracle.so) and of the compiled reference
(oracle/_ref/libftref.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_ORA = HERE / "liboracle.so"
_REF = HERE / "_ref" / "libftref.so"
_o = None
_r = None

F32P = C.POINTER(C.c_float)
I32P = C.POINTER(C.c_int32)
F64P = C.POINTER(C.c_double)


class StageParams(C.Structure):
    _fields_ = [("fwhm_px", C.c_double), ("clip_sigma", C.c_double), ("cr_sigma", C.c_double),
                ("static_exclude", C.c_int), ("mask_scatter", C.c_double), ("weight_scatter", C.c_double),
                ("threads", C.c_int)]


def stage_params(fwhm_px: float, static_exclude: int = -1) -> StageParams:
    """SearchParams defaults (include/fainttrail/search.hpp:18-90) for build_stage."""
    return StageParams(fwhm_px, 8.0, 5.0, static_exclude, 2.5, 1.5, 0)


def _p(a, t):
    return a.ctypes.data_as(t)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def ora() -> C.CDLL:
    global _o
    if _o is None:
        if not _ORA.exists():
            raise RuntimeError(f"{_ORA} is not built (make -C oracle)")
        _o = C.CDLL(str(_ORA))
        _o.ora_test_normals.argtypes = [C.c_uint64, C.c_size_t, F32P]
        _o.ora_null_time_orders.argtypes = [F64P, C.c_int, C.c_int, C.c_uint64, F64P]
        _o.ora_velocity_grid.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, F64P, C.c_int, F64P]
        _o.ora_null_threshold.argtypes = [F32P, C.c_int, C.c_int, C.c_double, C.c_double, F64P]
        _o.ora_find_peaks_rowmajor.argtypes = [F32P, I32P, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_int,
                                               I32P, I32P, F32P, I32P]
        _o.ora_sort_peaks.argtypes = [C.c_int, C.c_int, I32P, I32P, F32P, I32P]
        _o.ora_build_stage.argtypes = [C.c_int, C.c_int, C.c_int, F32P, F64P, C.POINTER(StageParams), F32P, F32P,
                                       F64P, F32P, I32P]
        _o.ora_background.argtypes = [C.c_int, C.c_int, C.c_int, F32P, F32P, F64P, C.c_int]
        _o.ora_check_backend.argtypes = [C.c_int, C.c_int, C.c_int, F32P, C.c_int, I32P, I32P, C.c_int, F32P, I32P,
                                         F32P, I32P, F64P]
    return _o


def ref_available() -> bool:
    return _REF.exists()


def ref() -> C.CDLL:
    global _r
    if _r is None:
        if not _REF.exists():
            raise RuntimeError(f"{_REF} is not built (make -C oracle ref, needs the reference sources)")
        _r = C.CDLL(str(_REF))
        _r.ref_velocity_grid.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, F64P, C.c_int, F64P]
        _r.ref_null_time_orders.argtypes = [F64P, C.c_int, C.c_int, C.c_uint64, F64P]
        _r.ref_null_threshold.argtypes = [F32P, C.c_int, C.c_int, C.c_double, C.c_double]
        _r.ref_null_threshold.restype = C.c_double
        _r.ref_find_peaks.argtypes = [F32P, I32P, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_int, I32P, I32P,
                                      F32P, I32P]
        _r.ref_build_stage.argtypes = [C.c_int, C.c_int, C.c_int, F32P, F64P, C.c_double, C.c_int, F32P, F32P, F64P,
                                       F32P, I32P]
        _r.ref_prepare.argtypes = [C.c_int, C.c_int, C.c_int, F32P, F64P, C.c_double, F32P, F64P, F32P, F64P]
        _r.ref_emulate_cuda.argtypes = [C.c_int, C.c_int, C.c_int, F32P, C.c_int, I32P, I32P, C.c_int, C.c_double,
                                        C.c_double, F32P, I32P]
        _r.ref_check_backend_inputs.argtypes = [C.c_int, I32P, F32P, F64P, F64P, C.c_int]
        _r.ref_linear_shifts.argtypes = [C.c_int, F64P, C.c_int, F64P, I32P, I32P]
    return _r


# ------------------------------------------------------------------ stacking

def _stack(fn, frames, dx, dy, min_count, *extra):
    fr = f32(frames)
    n, h, w = fr.shape
    dx, dy = i32(dx), i32(dy)
    m = np.empty((h, w), np.float32)
    b = np.empty((h, w), np.int32)
    st = fn(n, w, h, _p(fr, F32P), dx.shape[0], _p(dx, I32P), _p(dy, I32P), int(min_count), *extra,
            _p(m, F32P), _p(b, I32P))
    if st != 0:
        raise RuntimeError("oracle stacking failed")
    return m, b


def stack_naive(frames, dx, dy, min_count):
    return _stack(ora().ora_stack_naive, frames, dx, dy, min_count)


def stack_blocked(frames, dx, dy, min_count, threads=0, band_rows=0, vel_block=0):
    return _stack(ora().ora_stack_blocked, frames, dx, dy, min_count, C.c_int(threads), C.c_int(band_rows),
                  C.c_int(vel_block))


def stack_blocked_rows(window, h, fr0, dx, dy, min_count, y0, y1, threads=0):
    """Blocked pass for output rows [y0, y1) of h-row frames; window = frame rows
    [fr0, fr0 + window.shape[1]) of every frame."""
    fr = f32(window)
    n, rows, w = fr.shape
    dx, dy = i32(dx), i32(dy)
    m = np.empty((y1 - y0, w), np.float32)
    b = np.empty((y1 - y0, w), np.int32)
    st = ora().ora_stack_blocked_rows(n, w, h, _p(fr, F32P), int(fr0), int(fr0 + rows), dx.shape[0], _p(dx, I32P),
                                      _p(dy, I32P), int(min_count), int(threads), int(y0), int(y1), _p(m, F32P),
                                      _p(b, I32P))
    if st != 0:
        raise RuntimeError(f"oracle stacking failed ({st})")
    return m, b


def linear_shifts(t_h, vxy):
    t = f64(t_h)
    v = f64(vxy).reshape(-1, 2)
    dx = np.empty((v.shape[0], t.size), np.int32)
    dy = np.empty_like(dx)
    ora().ora_linear_shifts(t.size, _p(t, F64P), v.shape[0], _p(v, F64P), _p(dx, I32P), _p(dy, I32P))
    return dx, dy


def check_backend(frames, dx, dy, min_count, ref, cand):
    fr = f32(frames)
    n, h, w = fr.shape
    dx, dy = i32(dx), i32(dy)
    out = np.zeros(5, np.float64)
    ora().ora_check_backend(n, w, h, _p(fr, F32P), dx.shape[0], _p(dx, I32P), _p(dy, I32P), int(min_count),
                            _p(f32(ref[0]), F32P), _p(i32(ref[1]), I32P), _p(f32(cand[0]), F32P),
                            _p(i32(cand[1]), I32P), _p(out, F64P))
    return {"ok": bool(out[0]), "max_abs_diff": out[1], "bad_values": int(out[2]), "best_v_mismatch": int(out[3]),
            "best_v_unexplained": int(out[4])}


# ------------------------------------------------------------------ data + search steps

def test_normals(seed: int, shape) -> np.ndarray:
    out = np.empty(int(np.prod(shape)), np.float32)
    ora().ora_test_normals(C.c_uint64(seed), out.size, _p(out, F32P))
    return out.reshape(shape)


def velocity_grid(span_h, vmin, vmax, tol, lib=None):
    L = lib or ora()
    fn = L.ora_velocity_grid if lib is None else L.ref_velocity_grid
    dv = C.c_double()
    cnt = fn(span_h, vmin, vmax, tol, None, 0, C.byref(dv))
    if cnt < 0:
        raise RuntimeError("invalid velocity grid")
    out = np.empty((cnt, 2), np.float64)
    fn(span_h, vmin, vmax, tol, _p(out, F64P), cnt, C.byref(dv))
    return out, dv.value


def null_time_orders(t_h, runs, seed, lib=None):
    t = f64(t_h)
    out = np.empty((runs, t.size), np.float64)
    if lib is None:
        ora().ora_null_time_orders(_p(t, F64P), t.size, runs, C.c_uint64(seed), _p(out, F64P))
    else:
        lib.ref_null_time_orders(_p(t, F64P), t.size, runs, C.c_uint64(seed), _p(out, F64P))
    return out


def find_peaks(max_snr, best_v, floor, radius, edge, max_peaks):
    m, b = f32(max_snr), i32(best_v)
    h, w = m.shape
    cap = h * w
    px, py, pv = (np.empty(cap, np.int32) for _ in range(3))
    ps = np.empty(cap, np.float32)
    cnt = ora().ora_find_peaks_rowmajor(_p(m, F32P), _p(b, I32P), w, h, floor, radius, edge, cap, _p(px, I32P),
                                        _p(py, I32P), _p(ps, F32P), _p(pv, I32P))
    keep = ora().ora_sort_peaks(cnt, max_peaks, _p(px, I32P), _p(py, I32P), _p(ps, F32P), _p(pv, I32P))
    return px[:keep].copy(), py[:keep].copy(), ps[:keep].copy(), pv[:keep].copy()


def null_threshold(pooled, runs, target=0.02, floor=5.0):
    p = f32(pooled)
    out = np.zeros(6, np.float64)
    ora().ora_null_threshold(_p(p, F32P), p.size, runs, target, floor, _p(out, F64P))
    return out


def build_stage(sub, t_h, fwhm_px, static_exclude=-1):
    s = f32(sub)
    n, h, w = s.shape
    t = f64(t_h)
    prm = stage_params(fwhm_px, static_exclude)
    filt = np.empty_like(s)
    sky = np.empty((h, w), np.float32)
    sig = np.empty(n, np.float64)
    mask = np.empty((h, w), np.float32)
    cr = np.empty(n, np.int32)
    st = ora().ora_build_stage(n, w, h, _p(s, F32P), _p(t, F64P), C.byref(prm), _p(filt, F32P), _p(sky, F32P),
                               _p(sig, F64P), _p(mask, F32P), _p(cr, I32P))
    if st != 0:
        raise RuntimeError("oracle build_stage failed")
    return {"filtered": filt, "static_sky": sky, "sigma": sig, "mask": mask, "cr_pixels": cr}


def background(frames):
    fr = f32(frames)
    n, h, w = fr.shape
    sub = np.empty_like(fr)
    bg = np.empty(n, np.float64)
    ora().ora_background(n, w, h, _p(fr, F32P), _p(sub, F32P), _p(bg, F64P), 0)
    return sub, bg


# ------------------------------------------------------------------ the compiled reference

def ref_stack(kind, frames, dx, dy, min_count, threads=0, band_rows=0, vel_block=0, budget=1e9, launch_adds=4e9):
    L = ref()
    if kind == "naive":
        return _stack(L.ref_stack_naive, frames, dx, dy, min_count)
    if kind == "simple":
        return _stack(L.ref_stack_simple, frames, dx, dy, min_count, C.c_int(threads))
    if kind == "cpu":
        return _stack(L.ref_stack_cpu, frames, dx, dy, min_count, C.c_int(threads), C.c_int(band_rows),
                      C.c_int(vel_block))
    if kind == "cuda-emulation":
        return _stack(L.ref_emulate_cuda, frames, dx, dy, min_count, C.c_double(budget), C.c_double(launch_adds))
    raise ValueError(kind)


def ref_build_stage(sub, t_h, fwhm_px, static_exclude=-1):
    s = f32(sub)
    n, h, w = s.shape
    t = f64(t_h)
    filt = np.empty_like(s)
    sky = np.empty((h, w), np.float32)
    sig = np.empty(n, np.float64)
    mask = np.empty((h, w), np.float32)
    cr = np.empty(n, np.int32)
    ref().ref_build_stage(n, w, h, _p(s, F32P), _p(t, F64P), fwhm_px, static_exclude, _p(filt, F32P), _p(sky, F32P),
                          _p(sig, F64P), _p(mask, F32P), _p(cr, I32P))
    return {"filtered": filt, "static_sky": sky, "sigma": sig, "mask": mask, "cr_pixels": cr}


def ref_prepare(frames, mjd, fwhm_px):
    fr = f32(frames)
    n, h, w = fr.shape
    m = f64(mjd)
    sub = np.empty_like(fr)
    filt = np.empty_like(fr)
    bg = np.empty(n, np.float64)
    t = np.empty(n, np.float64)
    ref().ref_prepare(n, w, h, _p(fr, F32P), _p(m, F64P), fwhm_px, _p(sub, F32P), _p(bg, F64P), _p(filt, F32P),
                      _p(t, F64P))
    return {"sub": sub, "background": bg, "filtered": filt, "t_h": t}


def ref_find_peaks(max_snr, best_v, floor, radius, edge, max_peaks):
    m, b = f32(max_snr), i32(best_v)
    h, w = m.shape
    cap = max(1, min(h * w, max_peaks))
    px, py, pv = (np.empty(cap, np.int32) for _ in range(3))
    ps = np.empty(cap, np.float32)
    cnt = ref().ref_find_peaks(_p(m, F32P), _p(b, I32P), w, h, floor, radius, edge, max_peaks, _p(px, I32P),
                               _p(py, I32P), _p(ps, F32P), _p(pv, I32P))
    return px[:cnt].copy(), py[:cnt].copy(), ps[:cnt].copy(), pv[:cnt].copy()


def ref_null_threshold(pooled, runs, target=0.02, floor=5.0):
    p = f32(pooled)
    return ref().ref_null_threshold(_p(p, F32P), p.size, runs, target, floor)


def ref_check_backend_inputs(round_: int):
    """The frames, times and velocity grid of `fainttrail check-backend` (cli/main.cpp:697-742)."""
    L = ref()
    dims = np.zeros(4, np.int32)
    nv = L.ref_check_backend_inputs(round_, _p(dims, I32P), None, None, None, 0)
    w, h, n, mc = (int(v) for v in dims)
    fr = np.empty((n, h, w), np.float32)
    t = np.empty(n, np.float64)
    v = np.empty((nv, 2), np.float64)
    L.ref_check_backend_inputs(round_, _p(dims, I32P), _p(fr, F32P), _p(t, F64P), _p(v, F64P), nv)
    return fr, t, v, mc


CAND_FIELDS = ["id", "x", "y", "vx_pxh", "vy_pxh", "pos_err_px", "vel_err_pxh", "snr", "snr_coarse", "snr_half1",
               "snr_half2", "max_frame_fraction", "false_alarms_per_run", "flux", "mag_inst", "ok", "inv_dist_au",
               "inv_dist_err_au"]
METHODS = ["null-fit gauss", "null-fit linear", "null-quantile", "fixed", "gaussian"]


def _ref_search_run(frames, mjd, vmin, vmax, tol, fwhm, null_runs, seed, threshold, floor, target, backend,
                    null_tested, reg, tile, max_candidates, extras=True, neo=None):
    L = ref()
    if not getattr(L, "_neo_bound", False):
        L.ref_set_neo.argtypes = [C.c_int, F64P, F64P, C.c_double, C.c_int]
        L.ref_last_neo.argtypes = [F64P]
        L._neo_bound = True
    if not getattr(L, "_search_bound", False):
        L.ref_search_run.argtypes = [C.c_int, C.c_int, C.c_int, F32P, F64P, F64P, C.c_int, C.c_uint64, C.c_char_p,
                                     C.c_int, F64P, C.c_int, C.c_int, C.c_int]
        L.ref_last_error.restype = C.c_char_p
        L.ref_last_counts.argtypes = [I32P]
        L.ref_last_summary.argtypes = [F64P]
        L.ref_last_maps.argtypes = [F32P, I32P]
        L.ref_last_pooled.argtypes = [F32P]
        L.ref_last_peaks.argtypes = [C.c_int, I32P, I32P, F32P, I32P]
        L.ref_last_candidates.argtypes = [F64P, C.c_char_p]
        L._search_bound = True
    if backend == "hip":
        import fainttrail_amd  # one HIP runtime: torch's, loaded before the adapter
        fainttrail_amd.lib()
    fr = f32(frames)
    n, h, w = fr.shape
    m = f64(mjd)
    prm = f64([vmin, vmax, tol, fwhm, threshold, floor, target])
    rg = None
    if reg is not None:
        rg = f64(np.concatenate([np.asarray(reg[0], np.float64), np.asarray(reg[1], np.float64)]))
        assert rg.size == 2 * n
    if neo is not None:
        px, py = f64(neo["par_x"]), f64(neo["par_y"])
        assert px.size == n and py.size == n
        L.ref_set_neo(n, _p(px, F64P), _p(py, F64P), float(neo["min_dist_au"]), int(neo.get("max_trials", 64)))
    try:
        rc = L.ref_search_run(n, w, h, _p(fr, F32P), _p(m, F64P), _p(prm, F64P), null_runs, C.c_uint64(seed),
                              backend.encode(), null_tested, _p(rg, F64P) if rg is not None else None, tile,
                              max_candidates, int(extras))
    finally:
        if neo is not None:
            L.ref_set_neo(0, None, None, 0.0, 64)
    if rc != 0:
        raise RuntimeError(f"reference search failed: {L.ref_last_error().decode()}")
    cnt = np.zeros(5, np.int32)
    L.ref_last_counts(_p(cnt, I32P))
    summ = np.zeros(9, np.float64)
    L.ref_last_summary(_p(summ, F64P))
    mx = np.empty((h, w), np.float32)
    bv = np.empty((h, w), np.int32)
    L.ref_last_maps(_p(mx, F32P), None if tile > 0 else _p(bv, I32P))
    pooled = np.empty(max(int(cnt[0]), 1), np.float32)
    L.ref_last_pooled(_p(pooled, F32P))

    def peaks(which, c):
        x, y, v = (np.empty(max(c, 1), np.int32) for _ in range(3))
        s = np.empty(max(c, 1), np.float32)
        L.ref_last_peaks(which, _p(x, I32P), _p(y, I32P), _p(s, F32P), _p(v, I32P))
        return {"x": x[:c], "y": y[:c], "snr": s[:c], "vidx": v[:c]}

    nc = int(cnt[3])
    vals = np.zeros((max(nc, 1), len(CAND_FIELDS)), np.float64)
    flags = C.create_string_buffer(int(cnt[4]))
    L.ref_last_candidates(_p(vals, F64P), flags)
    fl = flags.value.decode().split(";")[:nc]
    cands = [dict(zip(CAND_FIELDS, vals[k]), flags=fl[k]) for k in range(nc)]
    for c in cands:
        c["id"], c["ok"] = int(c["id"]), bool(c["ok"])
    out = {"threshold": summ[0], "null_a": summ[1], "null_k": summ[2], "null_inv_s2": summ[3], "null_b": summ[4],
           "n_velocities": int(summ[5]), "dv_pxh": summ[6], "t_ref_mjd": summ[7],
           "method": METHODS[int(summ[8])] if summ[8] >= 0 else "", "pooled": pooled[:int(cnt[0])], "max_snr": mx,
           "candidates": cands}
    nq = np.zeros(2, np.float64)
    L.ref_last_neo(_p(nq, F64P))
    out.update(n_distances=int(nq[0]), dq_inv_au=float(nq[1]))
    if tile <= 0 and extras:
        out.update(best_v=bv, peaks=peaks(0, int(cnt[1])), selected=peaks(1, int(cnt[2])))
    return out


def ref_search(frames, mjd, vmin, vmax, tol, fwhm, null_runs, seed=12345, threshold=0.0, floor=5.0, target=0.02,
               backend="cpu", null_tested=3000, reg=None, max_candidates=0, extras=True, neo=None):
    """The reference's search() on prepare(align off), with the given stacking backend
    ("cpu", or "hip" through the linked adapter), null_tested (the SearchParams
    default 3000 screens the strongest null peaks; 0 screens none) and optional
    registration offsets reg = (reg_dx, reg_dy), one per frame each. Returns the
    threshold and tail fit, the real pass's max_snr and best_v maps, the pooled null
    peaks in the search's order, the real peak list (find_peaks order), the selected
    peaks and the refined candidates (dicts of Candidate's fields). extras=False skips
    the second real pass that exposes best_v and the peak lists (to time search()
    alone). neo = {"par_x", "par_y", "min_dist_au", "max_trials"}: the near-Earth
    search with these parallax displacements set on the prepared sequence."""
    return _ref_search_run(frames, mjd, vmin, vmax, tol, fwhm, null_runs, seed, threshold, floor, target, backend,
                           null_tested, reg, 0, max_candidates, extras, neo)


def ref_search_tiled(frames, mjd, vmin, vmax, tol, fwhm, null_runs, tile, seed=12345, threshold=0.0, floor=5.0,
                     target=0.02, null_tested=3000):
    """The reference's tiled search (search_tiled) with core tiles of `tile` px: the
    same outputs as ref_search except best_v and the peak lists, which it keeps
    internal."""
    return _ref_search_run(frames, mjd, vmin, vmax, tol, fwhm, null_runs, seed, threshold, floor, target, "cpu",
                           null_tested, None, tile, 0)
