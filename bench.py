"""Stacking throughput of the gfx950 search backend on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--cpu-seconds S]

A step is one full stacking pass (every trial motion over every frame and pixel) of
the workload's synthetic frames, already resident in HBM. The timed region is K
steps bracketed by device synchronisation; the kernel's own duration is taken with
HIP events on the stream the passes run on. Prints one JSON line.

Workloads (sizes of the search configurations, 60 s cadence):
  tiled-6144x4096  60 frames of 6144 x 4096, rates 0.1-2 px/min, grid_tol 0.372 px
                   -> 19992 trial motions (the bit-exact GPU configuration; default)
  plain-512        20 frames of 512 x 512, rates 0.2-1.1 px/min, grid_tol 0.125 px
                   -> 5468 trial motions
  nulls-2048       40 frames of 2048 x 2048, rates 0.2-1.1 px/min, grid_tol 0.212 px
                   -> 7996 trial motions

The CPU baseline is the oracle's cache-blocked restatement of the reference CPU
backend, timed on this host on a band of output rows of the same workload, all
motions, and reported in the same unit (scaled per unit).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "stacked px·frames·velocities/s (% HBM roofline); null-calibrated searches/s; p99"
UNIT = "stacked px·frames·velocities/s"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
BYTES_PER_UNIT = 4  # one fp32 frame sample per stacked unit

WORKLOADS = {
    # name: (n, w, h, vmin_pxh, vmax_pxh, grid_tol_px)
    "tiled-6144x4096": (60, 6144, 4096, 6.0, 120.0, 0.372),
    "plain-512": (20, 512, 512, 12.0, 66.0, 0.125),
    "nulls-2048": (40, 2048, 2048, 12.0, 66.0, 0.212),
}


def workload_shifts(name):
    import fainttrail_amd as ft
    n, w, h, vmin, vmax, tol = WORKLOADS[name]
    t = (np.arange(n) - (n - 1) / 2.0) / 60.0  # hours about the mid epoch, 60 s cadence
    grid, dv = ft.velocity_grid(float(t[-1] - t[0]), vmin, vmax, tol)
    tab = ft.linear_shift_table(t, grid)
    min_count = max(3, int(np.ceil(0.7 * n)))
    return t, grid, dv, tab, min_count


def traffic_from_profiles(workload: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), if a summary for this workload exists."""
    p = ROOT / "profiles" / f"stack_kernel_{workload}_traffic.md"
    if not p.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def lds_bytes_from_profiles(workload: str):
    """LDS bytes the pass reads (SQ_INSTS_LDS x 64 lanes x 8 B over its kernels, from
    rocprofv3 --pmc), if a counter summary exists; the headline workload's is in
    stack_kernel_sq_counters.md."""
    if workload != "tiled-6144x4096":
        return None
    p = ROOT / "profiles" / "stack_kernel_sq_counters.md"
    if not p.exists():
        return None
    m = re.search(r"lds_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(frames_dev, tab, min_count, target_s: float):
    """Time the oracle's blocked CPU pass on output rows in the middle of the frame."""
    from oracle import pyoracle as po
    n, h, w = frames_dev.shape
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    lo_dy, hi_dy = int(tab.dy.min()), int(tab.dy.max())

    def window(y0, y1):
        r0, r1 = max(0, y0 + lo_dy), min(h, y1 + hi_dy)
        return r0, frames_dev[:, r0:r1].cpu().numpy()

    # the oracle parallelises over bands of rows: give every thread rows from the start
    # (one row per thread to calibrate, then whole multiples of the thread count)
    y0 = max(0, h // 2 - threads)
    r0, win = window(y0, min(h, y0 + threads))
    t0 = time.perf_counter()
    po.stack_blocked_rows(win, h, r0, tab.dx, tab.dy, min_count, y0, min(h, y0 + threads), threads)
    one = time.perf_counter() - t0
    reps = int(max(1, round(target_s / max(one, 1e-3))))
    rows = int(min(h - y0, threads * reps))
    r0, win = window(y0, y0 + rows)
    t0 = time.perf_counter()
    po.stack_blocked_rows(win, h, r0, tab.dx, tab.dy, min_count, y0, y0 + rows, threads)
    secs = time.perf_counter() - t0
    units = float(tab.motions) * n * w * rows
    return {"value": units / secs, "unit": UNIT, "cores": threads, "kind": "port",
            "sample": f"output rows {y0}-{y0 + rows - 1} of {h} (all {w} columns, all {tab.motions} motions, "
                      f"{n} frames): {units:.3e} units in {secs:.1f} s"}


def host_pass_measure(frames_dev, tab, min_count, dev_max, dev_best):
    """The same pass through the reference-shaped drop-in, ftk_stack_pass (CudaBackend::run,
    src/backend_cuda.cu:99-195): frames in host memory, maps returned to host memory, so
    the PCIe upload and download are inside the timed call. Its maps must equal the
    device-resident pass's."""
    import fainttrail_amd as ft
    host = frames_dev.cpu().numpy()
    n, h, w = host.shape
    be = ft.make_backend("hip")
    t0 = time.perf_counter()
    r = be.run(ft.StackPass(frames=host, shifts=tab, min_count=int(min_count)))
    wall = time.perf_counter() - t0
    same = bool(np.array_equal(r.max_snr.view(np.uint32), dev_max.cpu().numpy().view(np.uint32)) and
                np.array_equal(r.best_v, dev_best.cpu().numpy()))
    units = float(tab.motions) * n * w * h
    tim = dict(r.timings)
    return {"units_per_s": units / wall, "wall_s": wall, "upload_s": tim["upload"], "kernel_s": tim["kernel"],
            "download_s": tim["download"], "setup_s": tim["setup"],
            "upload_gb_per_s": host.nbytes / max(tim["upload"], 1e-9) / 1e9, "maps_equal_device_pass": same,
            "note": "PCIe-inclusive drop-in rate (host frames in, host maps out); never the headline value"}


def search_measure(reps: int, with_plain: bool, nmovers: int = 3, mover_snr=(1.5, 1.5), tune: bool = True):
    """Null-calibrated searches/s on the 512^2 configuration: 20 frames of 512 x 512 at
    60 s cadence, sky 1000 ADU + N(0, 32^2), 150 stars (sigma_psf 1.5 px, flux
    log-uniform 1e3-1e6 ADU), nmovers movers (3 of per-frame S/N 1.5 as stated; the
    bright form 8 of S/N 2-3, all detected) at 0.2-1.1 px/min, grid spaced 0.5 px of
    total drift (grid_tol 0.125 px -> 5468 motions), 8 null runs, seed 7, the
    reference's defaults otherwise (the strongest 3000 peaks of each null run screened
    for artefacts, the selected peaks refined into candidates). The device search runs
    everything on the GPU; the plain form is the reference's own search() with the "hip"
    stacking backend (host preprocessing, screen, refinement and per-pass uploads, as
    the reference drives its CUDA backend), and its candidates are compared with the
    device search's field by field."""
    import torch

    from fainttrail_amd import synth
    from fainttrail_amd.search import SearchParams, Searcher
    n, w, h = 20, 512, 512
    frames, mjd, _ = synth.make_sequence(n, w, h, nstars=150, nmovers=nmovers, mover_snr=mover_snr,
                                         rate_px_min=(0.2, 1.1), seed=7, device="cuda:0")
    prm = SearchParams(vmin_pxh=12.0, vmax_pxh=66.0, grid_tol_px=0.125, fwhm_px=1.5 * 2.354820045030949,
                       null_runs=8, seed=7)
    s = Searcher(n, w, h)

    def timed(p):
        r = s.run(frames, mjd, p)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = s.run(frames, mjd, p)
        torch.cuda.synchronize()
        return r, (time.perf_counter() - t0) / reps

    seq, dt_seq = timed(SearchParams(**{**prm.__dict__, "passes_per_launch": 1}))
    g, seg, tr, q = None, None, None, None
    if tune:
        g, seg, tr, q, _ = s.tune(frames, mjd, prm, reps=1)
    res, dt = timed(prm)
    same = bool(seq.threshold_snr == res.threshold_snr and np.array_equal(seq.null_peaks, res.null_peaks)
                and np.array_equal(seq.peaks.snr, res.peaks.snr) and seq.candidates == res.candidates
                and torch.equal(seq.max_snr, res.max_snr) and torch.equal(seq.best_v, res.best_v))
    res16, dt16 = timed(SearchParams(**{**prm.__dict__, "storage": "f16"}))
    out = {"config": f"512x512, 20 frames, 5468 motions, 8 null runs, {nmovers} movers of per-frame S/N "
                     f"{mover_snr[0]}-{mover_snr[1]}, seed 7",
           "device_searches_per_s": 1.0 / dt, "device_s_per_search": dt, "threshold_snr": res.threshold_snr,
           "method": res.threshold_method, "selected": int(res.selected.x.size),
           "candidates": len(res.candidates), "candidates_ok": sum(c.ok for c in res.candidates),
           "null_peaks_screened": res.null_screened, "null_artefacts_left_out": res.null_artefacts,
           "batched": {"passes_per_launch": g if tune else res.passes_per_launch,
                       "segments": seg if tune else res.segments, "tile_rows": res.tile_rows,
                       "persistent_launch": bool(res.queue), "equal_to_one_pass_form": same},
           "stage_seconds": {k: round(v, 4) for k, v in res.timings.items() if v > 0},
           "one_pass_form_searches_per_s": 1.0 / dt_seq,
           "one_pass_form_stage_seconds": {k: round(v, 4) for k, v in seq.timings.items() if v > 0},
           "f16_storage": {"searches_per_s": 1.0 / dt16, **candidate_agreement(res, res16)}}
    if with_plain:
        from oracle import pyoracle as po
        if po.ref_available():
            host = frames.cpu().numpy()
            dtp = float("inf")
            for _ in range(2):  # the better of two runs (the first also pays one-time set-up)
                t0 = time.perf_counter()
                ref = po.ref_search(host, mjd, 12.0, 66.0, 0.125, prm.fwhm_px, 8, seed=7, backend="hip",
                                    extras=False)
                dtp = min(dtp, time.perf_counter() - t0)
            out["plain_searches_per_s"] = 1.0 / dtp
            out["plain_s_per_search"] = dtp
            out["device_over_plain"] = dtp / dt
            out["plain_threshold_equal"] = bool(ref["threshold"] == res.threshold_snr)
            out["plain_candidates"] = len(ref["candidates"])
            out["candidates_equal_to_plain"] = same_candidate_lists(res.candidates, ref["candidates"])
            out["f16_over_plain"] = dtp / dt16
    return out


def same_candidate_lists(mine, theirs, skip=()) -> bool:
    """Candidate lists equal in order, ids, flags and every field's bits (fields in
    skip left out)."""
    from oracle import pyoracle as po
    if len(mine) != len(theirs):
        return False
    for c, r in zip(mine, theirs):
        if c.id != r["id"] or c.ok != r["ok"] or c.flags != r["flags"]:
            return False
        for f in po.CAND_FIELDS:
            if f in ("id", "ok") or f in skip:
                continue
            if np.float64(getattr(c, f)).tobytes() != np.float64(r[f]).tobytes():
                return False
    return True


def candidate_agreement(exact, other):
    """Selected peaks (position, motion) and S/N of a search against the exact one, and
    its refined candidates: each exact candidate matched to the nearest other one (within
    1 px at the reference epoch and 1 px/h), the largest S/N and position differences."""
    a = {(int(x), int(y), int(v)): float(q) for x, y, v, q in
         zip(exact.selected.x, exact.selected.y, exact.selected.vidx, exact.selected.snr)}
    b = {(int(x), int(y), int(v)): float(q) for x, y, v, q in
         zip(other.selected.x, other.selected.y, other.selected.vidx, other.selected.snr)}
    err = (other.max_snr.double() - exact.max_snr.double()).abs() / exact.max_snr.double().abs().clamp(min=1.0)
    common = set(a) & set(b)
    matched, snr_err, pos_err = 0, 0.0, 0.0
    for c in exact.candidates:
        best = min(other.candidates, key=lambda o: (o.x - c.x) ** 2 + (o.y - c.y) ** 2, default=None)
        if best is None or np.hypot(best.x - c.x, best.y - c.y) > 1.0 or \
                np.hypot(best.vx_pxh - c.vx_pxh, best.vy_pxh - c.vy_pxh) > 1.0:
            continue
        matched += 1
        snr_err = max(snr_err, abs(best.snr - c.snr) / max(1.0, abs(c.snr)))
        pos_err = max(pos_err, float(np.hypot(best.x - c.x, best.y - c.y)))
    return {"selected": len(b), "selected_exact": len(a), "same_selected": a.keys() == b.keys(),
            "threshold": other.threshold_snr, "threshold_exact": exact.threshold_snr,
            "max_rel_snr_error_vs_exact_map": float(err.max()),
            "max_rel_snr_error_selected": max((abs(a[k] - b[k]) / max(1.0, abs(a[k])) for k in common),
                                              default=0.0),
            "candidates": len(other.candidates), "candidates_exact": len(exact.candidates),
            "candidates_matched": matched, "max_rel_snr_error_candidates": snr_err,
            "max_position_difference_px": pos_err}


def local_maxima(m, floor: float):
    """Pixels at or above floor that are the maximum of their 5 x 5 window (torch)."""
    import torch
    md = m.double()
    mx = torch.nn.functional.max_pool2d(md[None, None], 5, stride=1, padding=2)[0, 0]
    ys, xs = torch.nonzero((md >= floor) & (md == mx), as_tuple=True)
    return set(zip(ys.tolist(), xs.tolist()))


def precision_measure(fp64_rows: int):
    """Precision form on the 100-frame 6144 x 4096 configuration: sky 2000 ADU +
    N(0, 45^2), 3000 stars (flux log-uniform 1e3-1.6e6 ADU, so the brightest saturate;
    clipped at 65535 ADU), one frame with 50 cosmic-ray hits, 30 movers of per-frame
    S/N 0.4-1.5 at 0.1-2 px/min; the frames are prepared on the device (background,
    build_stage) and stacked over 29 996 motions (grid_tol 0.509). One pass each with
    fp32, fp16 and bf16 frame storage; the FP64 evaluation on fp64_rows output rows
    (half at the top edge, half in the middle) is the reference for the S/N error
    |dS/N| / max(1, |S/N_fp64|); candidates are 5 x 5 local maxima at S/N >= 6."""
    import torch

    import fainttrail_amd as ft
    from fainttrail_amd import synth
    from fainttrail_amd.stage import Stage, StageParams
    n, w, h = 100, 6144, 4096
    fwhm = 1.5 * 2.354820045030949
    t0 = time.perf_counter()
    raw, mjd, _ = synth.make_sequence(n, w, h, sky=2000.0, noise=45.0, nstars=3000, star_flux=(1e3, 1.6e6),
                                      nmovers=30, mover_snr=(0.4, 1.5), rate_px_min=(0.1, 2.0), seed=41,
                                      device="cuda:0", saturate=65535.0, cr_frames=1, cr_hits=50)
    t_h = (mjd - 0.5 * (mjd[0] + mjd[-1])) * 24.0
    stage = Stage(n, w, h)
    sub, _ = stage.background(raw)
    del raw
    filt = stage.build(sub, t_h, StageParams(fwhm_px=fwhm))
    if isinstance(filt, tuple):
        filt = filt[0]
    del sub, stage
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    t_gen = time.perf_counter() - t0
    grid, _ = ft.velocity_grid(float(t_h[-1] - t_h[0]), 6.0, 120.0, 0.509)
    tab = ft.linear_shift_table(t_h, grid)
    mc = max(3, int(np.ceil(0.7 * n)))
    plan = ft.Plan(n, w, h, tab.dx, tab.dy, mc)
    stream = torch.cuda.current_stream()
    print(f"bench: precision: {n} x {w}x{h}, {tab.motions} motions, {plan.units:.3e} units/pass "
          f"(frames prepared in {t_gen:.1f} s)", file=sys.stderr, flush=True)

    def timed(fn):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        r = fn()
        b.record(stream)
        torch.cuda.synchronize()
        return r, a.elapsed_time(b) / 1e3

    maps, secs = {}, {}
    (m, _), secs["f32"] = timed(lambda: plan.run(filt))
    maps["f32"] = m
    store_s = {}
    for st in ("f16", "bf16"):
        _, store_s[st] = timed(lambda: plan.store(filt, st))
        (m, _), secs[st] = timed(lambda: plan.run_stored())
        maps[st] = m
    half = fp64_rows // 2
    bands = [(0, half), (h // 2 - half // 2, h // 2 - half // 2 + (fp64_rows - half))]
    t0 = time.perf_counter()
    ref = {b: ft.stack_fp64(filt, tab.dx, tab.dy, mc, b[0], b[1])[0] for b in bands}
    t_fp64 = time.perf_counter() - t0
    out = {"config": f"{n} frames of {w}x{h}, sky 2000 + N(0, 45^2), 3000 stars (saturating), 1 frame with 50 "
                     f"cosmic rays, 30 movers S/N 0.4-1.5, {tab.motions} motions, prepared on the device",
           "units_per_pass": plan.units, "fp64_rows": [list(b) for b in bands], "fp64_seconds": t_fp64}
    for st in ("f32", "f16", "bf16"):
        errs = []
        for (y0, y1), r in ref.items():
            d = (maps[st][y0:y1].double() - r).abs() / r.abs().clamp(min=1.0)
            errs.append(float(d.max()))
        cand_ref = set()
        cand = set()
        for (y0, y1), r in ref.items():
            cand_ref |= {(y + y0, x) for y, x in local_maxima(r, 6.0) if 2 <= y < y1 - y0 - 2}
            cand |= {(y + y0, x) for y, x in local_maxima(maps[st][y0:y1], 6.0) if 2 <= y < y1 - y0 - 2}
        out[st] = {"units_per_s": plan.units / secs[st], "pass_seconds": secs[st],
                   "max_rel_snr_error_vs_fp64": max(errs),
                   "candidates_vs_fp64_rows": {"fp64": len(cand_ref), "this": len(cand), "common": len(cand & cand_ref)}}
        if st != "f32":
            out[st]["store_seconds"] = store_s[st]
            full_ref, full = local_maxima(maps["f32"], 6.0), local_maxima(maps[st], 6.0)
            out[st]["candidates_vs_fp32_full_frame"] = {"fp32": len(full_ref), "this": len(full),
                                                        "common": len(full & full_ref)}
    return out


def serving_measure(count: int, rate_hz: float, seed: int, slots: int, sustain_count: int):
    """Serving form: a seeded Poisson stream of search jobs (mean rate_hz arrivals/s;
    per job, log-uniform: side 256-4096 px, 10-120 frames, 500-50 000 motions, 4-64
    null runs; 0-15 movers) fed to the device-resident server with `slots` slots and
    85 % of the free HBM as its budget. Inputs are generated on the device before the
    clock starts; a job's latency runs from its submission (at its arrival time) to the
    end of its search. The stream at the stated rate overloads one GPU (offered load
    rho = arrival rate x mean service time >> 1), so its latency is mostly queueing;
    the first sustain_count jobs are then replayed at 70 % of the measured capacity."""
    import torch

    import fainttrail_amd as ft
    from fainttrail_amd.serve import Server, job_inputs, poisson_jobs

    def run(specs):
        inputs = [job_inputs(sp) for sp in specs]
        units = []
        for sp, (fr, mjd, prm) in zip(specs, inputs):
            span_h = (mjd[-1] - mjd[0]) * 24.0
            nv = ft.velocity_grid(span_h, prm.vmin_pxh, prm.vmax_pxh, prm.grid_tol_px)[0].shape[0]
            units.append(float(nv) * sp.frames * sp.side * sp.side * (1 + sp.nulls))
        torch.cuda.synchronize()
        srv = Server(slots=slots)
        t0 = time.perf_counter()
        ids = []
        for sp, (fr, mjd, prm) in zip(specs, inputs):
            delay = t0 + sp.arrival_s - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(fr, mjd, prm))
        res = [srv.wait(j) for j in ids]
        st = srv.stats()
        srv.close()
        del inputs
        torch.cuda.empty_cache()
        lat = np.array([r.latency for r in res])
        wait = np.array([r.t_admit - r.t_submit for r in res])
        span = max(r.t_end for r in res) - min(r.t_submit for r in res)
        busy = sum(r.t_end - r.t_admit for r in res)
        return {"jobs": len(res), "p50_latency_s": float(np.percentile(lat, 50)),
                "p99_latency_s": float(np.percentile(lat, 99)), "mean_latency_s": float(lat.mean()),
                "max_latency_s": float(lat.max()), "p50_queue_wait_s": float(np.percentile(wait, 50)),
                "jobs_per_s": len(res) / span, "units_per_s": sum(units) / span, "units": sum(units),
                "wall_s": span, "mean_job_s_in_slot": busy / len(res), "max_running": st["max_running"],
                "peak_reserved_gb": st["peak_reserved_bytes"] / 1e9, "budget_gb": st["budget_bytes"] / 1e9,
                "evicted_searchers": st["evicted"], "stack_launches": st["stack_launches"],
                "fused_launches": st["fused_launches"], "fused_passes": st["fused_passes"]}

    specs = poisson_jobs(count, rate_hz, seed)
    over = run(specs)
    service = over["wall_s"] / over["jobs"]  # the GPU's mean time per job, all slots together
    over["offered_load_rho"] = rate_hz * service
    over["arrival_rate_hz"] = rate_hz
    sustain_rate = 0.7 * over["jobs_per_s"]
    sus = run(poisson_jobs(sustain_count, sustain_rate, seed))
    sus["arrival_rate_hz"] = sustain_rate
    sus["offered_load_rho"] = sustain_rate * service
    return {"config": f"Poisson stream, seed {seed}: side 256-4096, 10-120 frames, 500-50000 motions, 4-64 nulls, "
                      f"0-15 movers; {slots} slots, budget 85 % of free HBM",
            "at_stated_rate": over, "at_sustainable_rate": sus}


def serving_stream_measure(count: int, rate_hz: float, seed: int, slots: int, budget_frac: float = 0.6):
    """A long stream of the serving distribution for latency percentiles: count jobs
    (the same per-job distribution as serving_measure) arriving as a seeded Poisson
    stream at rate_hz, below the server's capacity. Each job's frames are generated on
    the device by a producer thread at most two jobs ahead of its arrival and handed to
    the server at the arrival time; a job generated after its arrival time is submitted
    late (counted, and its latency runs from its actual submission). Ended jobs are
    collected as the stream runs, so their frames are released. The server's budget is
    budget_frac of the free HBM (the rest holds the jobs' frames)."""
    import queue
    import threading

    import torch

    import fainttrail_amd as ft
    from fainttrail_amd.serve import Server, job_inputs, poisson_jobs
    specs = poisson_jobs(count, rate_hz, seed)
    free_b, _ = torch.cuda.mem_get_info()
    srv = Server(slots=slots, budget_mb=budget_frac * free_b / 1048576.0)
    ready = queue.Queue(maxsize=2)
    failure = []

    gen_stream = torch.cuda.Stream()  # the producer's work waits for nothing of the server's

    def produce():
        try:
            for sp in specs:
                with torch.cuda.stream(gen_stream):
                    inp = job_inputs(sp)
                gen_stream.synchronize()
                ready.put(inp)
        except Exception as e:  # surfaced by the consumer
            failure.append(e)
            ready.put(None)

    prod = threading.Thread(target=produce, daemon=True)
    t0 = time.perf_counter()
    prod.start()
    pending, results, late, units = {}, {}, [], {}

    def collect(block: bool):
        for j in list(pending):
            r = srv.wait(j, -1.0 if block else 0.0)
            if r is not None:
                results[pending.pop(j)] = r

    for k, sp in enumerate(specs):
        inp = ready.get()
        if inp is None:
            raise failure[0]
        fr, mjd, prm = inp
        span_h = (mjd[-1] - mjd[0]) * 24.0
        nv = ft.velocity_grid(span_h, prm.vmin_pxh, prm.vmax_pxh, prm.grid_tol_px)[0].shape[0]
        units[k] = float(nv) * sp.frames * sp.side * sp.side * (1 + sp.nulls)
        now = time.perf_counter() - t0
        if now < sp.arrival_s:
            time.sleep(sp.arrival_s - now)
        else:
            late.append(now - sp.arrival_s)
        pending[srv.submit(fr, mjd, prm)] = k
        del inp, fr
        collect(False)
        if (k + 1) % 20 == 0:
            print(f"bench: serving stream: {k + 1} of {count} submitted, {len(results)} ended, "
                  f"{time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
    collect(True)
    prod.join()
    st = srv.stats()
    srv.close()
    res = [results[k] for k in range(count)]
    lat = np.array([r.latency for r in res])
    wait = np.array([r.t_admit - r.t_submit for r in res])
    span = max(r.t_end for r in res) - min(r.t_submit for r in res)
    busy = sum(r.t_end - r.t_admit for r in res)
    pct = {f"p{q}_latency_s": float(np.percentile(lat, q)) for q in (50, 90, 95, 99)}
    return {"config": f"Poisson stream, seed {seed}, {count} jobs at {rate_hz} jobs/s: side 256-4096, 10-120 frames, "
                      f"500-50000 motions, 4-64 nulls, 0-15 movers; {slots} slots, budget {budget_frac:.0%} of free HBM",
            "jobs": len(res), **pct, "mean_latency_s": float(lat.mean()), "max_latency_s": float(lat.max()),
            "p50_queue_wait_s": float(np.percentile(wait, 50)), "p99_queue_wait_s": float(np.percentile(wait, 99)),
            "jobs_per_s": len(res) / span, "units_per_s": sum(units.values()) / span, "wall_s": span,
            "mean_job_s_in_slot": busy / len(res), "slot_occupancy": busy / (span * slots),
            "late_submissions": len(late), "max_late_s": max(late, default=0.0),
            "candidates": int(sum(len(r.candidates) for r in res)), "max_running": st["max_running"],
            "peak_reserved_gb": st["peak_reserved_bytes"] / 1e9, "budget_gb": st["budget_bytes"] / 1e9,
            "evicted_searchers": st["evicted"], "failed": st["failed"], "stack_launches": st["stack_launches"],
            "fused_launches": st["fused_launches"], "fused_passes": st["fused_passes"]}


def batched_nulls_measure(with_plain: bool, plain_nulls: int = 2):
    """Batched null calibration on the 2048^2 configuration: 40 frames of 2048 x 2048
    (60 s cadence), 300 stars, 10 movers of per-frame S/N 0.7 at 0.2-1.1 px/min, grid
    tol 0.212 px (7996 motions), 256 shuffled-time null runs plus the real pass, false
    alarms 0.02 per search, seed 7, the batched group size and segment count autotuned
    on a 16-null search first. The plain form (the reference's own search() with the
    "hip" stacking backend) is timed with 0 and plain_nulls null runs and its per-null
    cost extrapolated to 256."""
    import torch

    from fainttrail_amd import synth
    from fainttrail_amd.search import SearchParams, Searcher
    n, w, h = 40, 2048, 2048
    frames, mjd, _ = synth.make_sequence(n, w, h, nstars=300, nmovers=10, mover_snr=(0.7, 0.7),
                                         rate_px_min=(0.2, 1.1), seed=7, device="cuda:0")
    fwhm = 1.5 * 2.354820045030949
    prm = SearchParams(vmin_pxh=12.0, vmax_pxh=66.0, grid_tol_px=0.212, fwhm_px=fwhm, null_runs=256, seed=7)
    s = Searcher(n, w, h)
    g, seg, tr, q, t_tune = s.tune(frames, mjd, SearchParams(**{**prm.__dict__, "null_runs": 16}), reps=1)
    print(f"bench: batched nulls: tuned {g} passes per launch, {seg} segments, tile rows {tr}, launch form {q}",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    res = s.run(frames, mjd, prm)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    units = float(res.n_velocities) * n * w * h * 257
    # the same pooled null peaks through the FP64 tail fit with its sums on the GPU
    t0 = time.perf_counter()
    thr_gpu, m_gpu = s.refit(prm, on_device=True)
    t_fit_gpu = time.perf_counter() - t0
    out = {"config": "2048x2048, 40 frames, 7996 motions, 256 null runs + real, seed 7",
           "n_velocities": res.n_velocities, "device_s_per_search": dt, "device_searches_per_s": 1.0 / dt,
           "stacked_units_per_s": units / dt, "threshold_snr": res.threshold_snr, "method": res.threshold_method,
           "selected": int(res.selected.x.size),
           "tuned_on_16_nulls": {"passes_per_launch": g, "segments": seg, "tile_rows": tr, "queue": q},
           "passes_per_launch": res.passes_per_launch, "segments": res.segments, "tile_rows": res.tile_rows,
           "persistent_launch": bool(res.queue),
           "stage_seconds": {k: round(v, 4) for k, v in res.timings.items() if v > 0},
           "fit_on_gpu": {"threshold_snr": thr_gpu, "method_equal": m_gpu == res.threshold_method,
                          "rel_diff_vs_host_fit": abs(thr_gpu - res.threshold_snr) / abs(res.threshold_snr),
                          "bound_50_m_eps": 50.0 * min(res.null_peaks.size, 3000 * 256) * float(np.finfo(np.float64).eps),
                          "seconds": t_fit_gpu, "pooled_peaks": int(res.null_peaks.size)}}
    if with_plain:
        from oracle import pyoracle as po
        if po.ref_available():
            host = frames.cpu().numpy()
            t0 = time.perf_counter()
            po.ref_search(host, mjd, 12.0, 66.0, 0.212, fwhm, 0, seed=7, backend="hip", extras=False)
            t_real = time.perf_counter() - t0
            t0 = time.perf_counter()
            po.ref_search(host, mjd, 12.0, 66.0, 0.212, fwhm, plain_nulls, seed=7, backend="hip", extras=False)
            t_k = time.perf_counter() - t0
            per_null = max(0.0, (t_k - t_real) / plain_nulls)
            est = t_real + 256 * per_null
            out["plain"] = {"seconds_without_nulls": t_real, f"seconds_with_{plain_nulls}_nulls": t_k,
                            "seconds_per_null_run": per_null, "extrapolated_s_per_search_256_nulls": est,
                            "note": "extrapolated from the measured runs: real pass + 256 x the per-null cost"}
            out["device_over_plain_extrapolated"] = est / dt
    # candidates at this scale: the same sky with 10 movers of per-frame S/N 1.5 (all
    # detectable), searched at the threshold the 256-null calibration set (fixed, so
    # the plain form needs no null runs), device against plain, field by field
    fb, mjdb, _ = synth.make_sequence(n, w, h, nstars=300, nmovers=10, mover_snr=(1.5, 1.5),
                                      rate_px_min=(0.2, 1.1), seed=7, device="cuda:0")
    pb = SearchParams(**{**prm.__dict__, "threshold_snr": res.threshold_snr, "null_runs": 0})
    s.run(fb, mjdb, pb)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rb = s.run(fb, mjdb, pb)
    torch.cuda.synchronize()
    dtb = time.perf_counter() - t0
    br = {"config": "the same frames with 10 movers of per-frame S/N 1.5, threshold fixed at the calibrated one",
          "threshold_snr": res.threshold_snr, "device_s_per_search": dtb, "selected": int(rb.selected.x.size),
          "candidates": len(rb.candidates), "candidates_ok": sum(c.ok for c in rb.candidates),
          "stage_seconds": {k: round(v, 4) for k, v in rb.timings.items() if v > 0}}
    if with_plain:
        from oracle import pyoracle as po
        if po.ref_available():
            host = fb.cpu().numpy()
            t0 = time.perf_counter()
            ref = po.ref_search(host, mjdb, 12.0, 66.0, 0.212, fwhm, 0, seed=7, threshold=res.threshold_snr,
                                backend="hip", extras=False)
            dtp = time.perf_counter() - t0
            br.update({"plain_s_per_search": dtp, "device_over_plain": dtp / dtb,
                       "plain_candidates": len(ref["candidates"]),
                       "candidates_equal_to_plain": same_candidate_lists(rb.candidates, ref["candidates"])})
    out["bright_movers_fixed_threshold"] = br
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="tiled-6144x4096", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-search", action="store_true", help="skip the null-calibrated search measurement")
    ap.add_argument("--search-reps", type=int, default=3)
    ap.add_argument("--options", default="", help="backend options of the timed pass (hip:...), e.g. tile_rows=32")
    ap.add_argument("--no-tune", action="store_true",
                    help="time the plan at its options' work shape instead of the shape ftk_plan_tune picks")
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16"],
                    help="frame storage of the timed stacking passes (16-bit: a power-of-two scale per frame)")
    ap.add_argument("--no-precision", action="store_true", help="skip the fp16/bf16/FP64 precision measurement")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-null batched calibration measurement")
    ap.add_argument("--fp64-rows", type=int, default=16)
    ap.add_argument("--no-serving", action="store_true", help="skip the serving (Poisson stream) measurement")
    ap.add_argument("--no-host-pass", action="store_true",
                    help="skip the host-buffer pass (the drop-in with PCIe upload and download)")
    ap.add_argument("--serve-jobs", type=int, default=24)
    ap.add_argument("--serve-seed", type=int, default=3)
    ap.add_argument("--serving-stream", type=int, default=0, metavar="N",
                    help="measure only a stream of N serving jobs at --serve-rate (latency percentiles); prints its "
                         "JSON object")
    ap.add_argument("--serve-rate", type=float, default=0.3, help="arrival rate of --serving-stream (jobs/s)")
    ap.add_argument("--search-only", action="store_true",
                    help="measure only the null-calibrated search (for profiling); prints its JSON object")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the timed region, copy this many MiB with the library's plain copy kernel "
                         "(the known-size kernel that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this backend has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2

    import torch

    import fainttrail_amd as ft
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    if args.serving_stream > 0:
        print(json.dumps({"serving_stream": serving_stream_measure(args.serving_stream, args.serve_rate,
                                                                   args.serve_seed, 4)}))
        return 0
    if args.search_only:
        print(json.dumps({"search": search_measure(args.search_reps, with_plain=not args.no_cpu)}))
        return 0
    n, w, h = WORKLOADS[args.workload][:3]
    t, grid, dv, tab, min_count = workload_shifts(args.workload)
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(7)
    frames = torch.randn((n, h, w), generator=gen, device=dev, dtype=torch.float32)
    plan = ft.Plan(n, w, h, tab.dx, tab.dy, min_count, args.options)
    out_m = torch.empty((h, w), dtype=torch.float32, device=dev)
    out_b = torch.empty((h, w), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    print(f"bench: {args.workload}: {n} x {w}x{h}, {tab.motions} motions, {plan.blocks} blocks, "
          f"{plan.units:.3e} units/pass", file=sys.stderr, flush=True)
    shape = None
    if args.storage == "f32":
        if not args.no_tune:
            shape = plan.tune(frames, reps=1, stream=stream)

        def one_pass():
            plan.run(frames, out_m, out_b, stream)
    else:
        plan.store(frames, args.storage, stream)
        if not args.no_tune:
            shape = plan.tune(None, reps=1, stream=stream)

        def one_pass():
            plan.run_stored(out_m, out_b, stream)
    if shape is not None:
        print(f"bench: tuned work shape: {shape[0]}-row tiles, launch form {shape[1]} ({shape[2] * 1e3:.1f} ms)",
              file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        one_pass()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        one_pass()
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_s = sum(s.elapsed_time(e) for s, e in zip(starts, ends)) / 1e3 / args.steps
    units = plan.units
    value = units * args.steps / wall
    bytes_per_unit = BYTES_PER_UNIT if args.storage == "f32" else 2  # one stored sample per unit
    achieved = bytes_per_unit * units / kernel_s
    print(f"bench: {wall / args.steps * 1e3:.1f} ms/pass, {value:.3e} units/s", file=sys.stderr, flush=True)
    # the inner loop reads two rows per ds_read2st64_b32 (8 B per lane per instruction):
    # the probe of that form is the kernel's LDS ceiling; 16-byte reads go faster
    lds_peak = ft.probe_lds(8)
    lds_peak16 = ft.probe_lds(16)
    traffic = traffic_from_profiles(args.workload)
    lds_bytes = lds_bytes_from_profiles(args.workload) if args.storage == "f32" else None
    compulsory = float(n) * w * h * (4 if args.storage == "f32" else 2) + 24.0 * w * h  # frames; keys, maps
    add_peak = 256 * 64 * 2.4e9  # unpacked fp32 adds/s: 256 CUs x 4 SIMDs x 16 lanes at 2.4 GHz
    res = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.storage == "f32" else f"{args.storage} storage, f32 accumulation",
        "data": "synthetic (unit-variance fp32 frames, seed 7)",
        "config": {"workload": args.workload, "options": args.options or "defaults",
                   "work_shape": {"tile_rows": shape[0], "persistent_launch": bool(shape[1]), "tuned": True}
                   if shape else "options",
                   "frames": n, "width": w, "height": h, "motions": int(tab.motions),
                   "grid_tol_px": WORKLOADS[args.workload][5], "dv_pxh": dv, "min_count": min_count,
                   "units_per_step": units},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "traffic_over_compulsory": traffic / compulsory if traffic else None,
                     "traffic_rate_gb_s": traffic / kernel_s / 1e9 if traffic else None,
                     "kernel_ms": kernel_s * 1e3,
                     "lds_bytes_per_unit": lds_bytes / units if lds_bytes else None,
                     "lds_peak_measured": lds_peak / 1e9,
                     "lds_frac": lds_bytes / kernel_s / lds_peak if lds_bytes else None,
                     "lds_peak_b128_measured": lds_peak16 / 1e9,
                     "lds_frac_best_form": lds_bytes / kernel_s / lds_peak16 if lds_bytes else None,
                     "add_frac": units / kernel_s / add_peak,
                     "note": "algorithmic bytes = 4 B (one fp32 sample) per stacked unit, so frac > 1 is the reuse of "
                             "LDS-staged frame boxes, not an HBM fraction. The binding figures: lds_frac is the LDS "
                             "bytes the pass reads (rocprofv3 SQ_INSTS_LDS, profiles/) per second against the LDS read "
                             "rate measured for the loop's ds_read2st64_b32 form, lds_frac_best_form the same against "
                             "ds_read_b128; add_frac is one fp32 add per unit against 256 CUs x 64 lanes x 2.4 GHz; "
                             "traffic (FETCH_SIZE + WRITE_SIZE, copy-calibrated) over the compulsory frame and map "
                             "bytes measures re-reads of staged boxes"},
    }
    if args.copy_probe_mib > 0:
        cnt = args.copy_probe_mib * 262144
        src = torch.ones(cnt, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        ft.lib().ftk_copy_f32(src.data_ptr(), dst.data_ptr(), cnt, stream.cuda_stream)
        torch.cuda.synchronize()
        res["copy_probe_bytes"] = cnt * 4
        del src, dst
    if not args.no_cpu:
        print("bench: CPU baseline (oracle, blocked)", file=sys.stderr, flush=True)
        res["cpu_baseline"] = cpu_baseline(frames, tab, min_count, args.cpu_seconds)
    else:
        res["cpu_baseline"] = None
    if not args.no_host_pass and args.storage == "f32":
        print("bench: host-buffer pass (ftk_stack_pass)", file=sys.stderr, flush=True)
        res["host_buffers"] = host_pass_measure(frames, tab, min_count, out_m, out_b)
        print("bench: host_buffers: " + json.dumps(res["host_buffers"]), file=sys.stderr, flush=True)
    del frames, plan
    torch.cuda.empty_cache()
    if not args.no_search:
        print("bench: null-calibrated search (512x512 configuration)", file=sys.stderr, flush=True)
        res["search"] = search_measure(args.search_reps, with_plain=not args.no_cpu)
        print("bench: search (bright movers)", file=sys.stderr, flush=True)
        res["search"]["bright_movers"] = search_measure(args.search_reps, with_plain=not args.no_cpu, nmovers=8,
                                                        mover_snr=(2.0, 3.0), tune=False)
        print("bench: search: " + json.dumps(res["search"]), file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    if not args.no_batched:
        print("bench: batched null calibration (2048x2048, 256 nulls)", file=sys.stderr, flush=True)
        res["batched_nulls"] = batched_nulls_measure(with_plain=not args.no_cpu)
        print("bench: batched_nulls: " + json.dumps(res["batched_nulls"]), file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    if not args.no_precision:
        print("bench: precision form (100 x 6144x4096)", file=sys.stderr, flush=True)
        res["precision"] = precision_measure(args.fp64_rows)
        print("bench: precision: " + json.dumps(res["precision"]), file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    if not args.no_serving:
        print("bench: serving (Poisson stream of search jobs)", file=sys.stderr, flush=True)
        res["serving"] = serving_measure(args.serve_jobs, 20.0, args.serve_seed, 4, 12)
        print("bench: serving: " + json.dumps(res["serving"]), file=sys.stderr, flush=True)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
