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


def search_measure(reps: int, with_plain: bool):
    """Null-calibrated searches/s on the 512^2 configuration: 20 frames of 512 x 512 at
    60 s cadence, sky 1000 ADU + N(0, 32^2), 150 stars (sigma_psf 1.5 px, flux
    log-uniform 1e3-1e6 ADU), 3 movers of per-frame S/N 1.5 at 0.2-1.1 px/min, grid
    spaced 0.5 px of total drift (grid_tol 0.125 px -> 5468 motions), 8 null runs,
    seed 7. The device search runs everything on the GPU; the plain form is the
    reference's own search() with the "hip" stacking backend (host preprocessing and
    per-pass uploads, as the reference drives its CUDA backend)."""
    import torch

    from fainttrail_amd import synth
    from fainttrail_amd.search import SearchParams, Searcher
    n, w, h = 20, 512, 512
    frames, mjd, _ = synth.make_sequence(n, w, h, nstars=150, nmovers=3, mover_snr=(1.5, 1.5),
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
    g, seg, _ = s.tune(frames, mjd, prm, reps=1)
    res, dt = timed(prm)
    same = bool(seq.threshold_snr == res.threshold_snr and np.array_equal(seq.null_peaks, res.null_peaks)
                and np.array_equal(seq.peaks.snr, res.peaks.snr)
                and torch.equal(seq.max_snr, res.max_snr) and torch.equal(seq.best_v, res.best_v))
    out = {"config": "512x512, 20 frames, 5468 motions, 8 null runs, seed 7",
           "device_searches_per_s": 1.0 / dt, "device_s_per_search": dt, "threshold_snr": res.threshold_snr,
           "method": res.threshold_method, "selected": int(res.selected.x.size),
           "batched": {"passes_per_launch": g, "segments": seg, "equal_to_one_pass_form": same},
           "stage_seconds": {k: round(v, 4) for k, v in res.timings.items() if v > 0},
           "one_pass_form_searches_per_s": 1.0 / dt_seq,
           "one_pass_form_stage_seconds": {k: round(v, 4) for k, v in seq.timings.items() if v > 0}}
    if with_plain:
        from oracle import pyoracle as po
        if po.ref_available():
            host = frames.cpu().numpy()
            t0 = time.perf_counter()
            ref = po.ref_search(host, mjd, 12.0, 66.0, 0.125, prm.fwhm_px, 8, seed=7, backend="hip")
            dtp = time.perf_counter() - t0
            out["plain_searches_per_s"] = 1.0 / dtp
            out["plain_s_per_search"] = dtp
            out["device_over_plain"] = dtp / dt
            out["plain_threshold_equal"] = bool(ref["threshold"] == res.threshold_snr)
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
    if args.search_only:
        print(json.dumps({"search": search_measure(args.search_reps, with_plain=not args.no_cpu)}))
        return 0
    n, w, h = WORKLOADS[args.workload][:3]
    t, grid, dv, tab, min_count = workload_shifts(args.workload)
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(7)
    frames = torch.randn((n, h, w), generator=gen, device=dev, dtype=torch.float32)
    plan = ft.Plan(n, w, h, tab.dx, tab.dy, min_count)
    out_m = torch.empty((h, w), dtype=torch.float32, device=dev)
    out_b = torch.empty((h, w), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    print(f"bench: {args.workload}: {n} x {w}x{h}, {tab.motions} motions, {plan.blocks} blocks, "
          f"{plan.units:.3e} units/pass", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        plan.run(frames, out_m, out_b, stream)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        plan.run(frames, out_m, out_b, stream)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_s = sum(s.elapsed_time(e) for s, e in zip(starts, ends)) / 1e3 / args.steps
    units = plan.units
    value = units * args.steps / wall
    achieved = BYTES_PER_UNIT * units / kernel_s
    print(f"bench: {wall / args.steps * 1e3:.1f} ms/pass, {value:.3e} units/s", file=sys.stderr, flush=True)
    # the inner loop reads two rows per ds_read2st64_b32 (8 B per lane per instruction)
    lds_peak = ft.probe_lds(8)
    res = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (unit-variance fp32 frames, seed 7)",
        "config": {"workload": args.workload, "frames": n, "width": w, "height": h, "motions": int(tab.motions),
                   "grid_tol_px": WORKLOADS[args.workload][5], "dv_pxh": dv, "min_count": min_count,
                   "units_per_step": units},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic_from_profiles(args.workload),
                     "kernel_ms": kernel_s * 1e3,
                     "lds_peak_measured": lds_peak / 1e9, "lds_frac": achieved / lds_peak,
                     "note": "algorithmic bytes = 4 B (one fp32 sample) per stacked unit; frac > 1 is reuse "
                             "from LDS-staged frame boxes; the inner loop reads each unit's sample from LDS (two rows "
                             "per ds_read2st64_b32), so lds_frac, against the LDS read bandwidth measured with "
                             "8-byte-per-lane reads, is the binding figure"},
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
    if not args.no_search:
        del frames, plan
        torch.cuda.empty_cache()
        print("bench: null-calibrated search (512x512 configuration)", file=sys.stderr, flush=True)
        res["search"] = search_measure(args.search_reps, with_plain=not args.no_cpu)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
