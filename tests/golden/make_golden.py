"""Write the golden vectors of tests/golden/ from the compiled reference.

Runs in a container that has the reference sources: `make -C oracle ref` compiles
them where they lie into oracle/_ref/libftref.so, and this script calls the
reference's own functions through oracle/ref_shim.cpp on the inputs of the
reference's tests (tests/refcases.py). Only inputs and outputs are stored.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from oracle import pyoracle as po  # noqa: E402
import refcases  # noqa: E402

OUT = Path(__file__).resolve().parent


def main() -> None:
    if not po.ref_available():
        raise SystemExit("oracle/_ref/libftref.so is missing: run make -C oracle ref first")
    data = {}
    # stacking contract cases: the reference CPU backend's maps (bit-equal to its naive form)
    for name, fr, dx, dy, mc in refcases.all_stack_cases():
        m, b = po.ref_stack("cpu", fr, dx, dy, mc, threads=2)
        mn, bn = po.ref_stack("naive", fr, dx, dy, mc)
        data[f"{name}/frames"] = fr
        data[f"{name}/dx"] = dx
        data[f"{name}/dy"] = dy
        data[f"{name}/min_count"] = np.int32(mc)
        data[f"{name}/max_snr"] = m
        data[f"{name}/best_v"] = b
        data[f"{name}/naive_max_snr"] = mn
        data[f"{name}/naive_best_v"] = bn
    # check-backend round 0 (cli/main.cpp:697-720): simulate(72x52, 12 frames, seed 77) -> prepare -> grid
    fr, t, v, mc = po.ref_check_backend_inputs(0)
    dx = np.empty((v.shape[0], t.size), np.int32)
    dy = np.empty_like(dx)
    po.ref().ref_linear_shifts(t.size, po._p(po.f64(t), po.F64P), v.shape[0], po._p(po.f64(v), po.F64P),
                               po._p(dx, po.I32P), po._p(dy, po.I32P))
    m, b = po.ref_stack("naive", fr, dx, dy, mc)
    data["check0/frames"], data["check0/t_h"], data["check0/vel"] = fr, t, v
    data["check0/dx"], data["check0/dy"], data["check0/min_count"] = dx, dy, np.int32(mc)
    data["check0/max_snr"], data["check0/best_v"] = m, b
    # the same frames with the curved table of check-backend (cli/main.cpp:703-711)
    cdx, cdy = dx.copy(), dy.copy()
    for j in range(cdx.shape[0]):
        for i in range(cdx.shape[1]):
            ti = t[i]
            cdx[j, i] += refcases.lround(((j % 9) - 4) * 2.0 * ti * ti)
            cdy[j, i] += refcases.lround(((j % 7) - 3) * 2.0 * ti * ti)
    m, b = po.ref_stack("naive", fr, cdx, cdy, mc)
    data["check0c/dx"], data["check0c/dy"] = cdx, cdy
    data["check0c/max_snr"], data["check0c/best_v"] = m, b
    # velocity grids (search.cpp:39-58) for the benchmark shapes
    for key, (span, vmin, vmax, tol) in {
        "grid_cfg1": (19 / 60.0, 12.0, 66.0, 0.125),
        "grid_cfg2": (59 / 60.0, 6.0, 120.0, 0.372),
        "grid_default": (1.0, 4.0, 40.0, 0.5),
    }.items():
        g, dv = po.velocity_grid(span, vmin, vmax, tol, lib=po.ref())
        data[f"{key}/vel"], data[f"{key}/dv"] = g, np.float64(dv)
        data[f"{key}/args"] = np.array([span, vmin, vmax, tol])
    # null time orders (search.cpp:559-571), seed 12345 and 7
    t = np.linspace(-0.5, 0.5, 20)
    for seed in (12345, 7):
        data[f"orders_{seed}/t"] = t
        data[f"orders_{seed}/out"] = po.null_time_orders(t, 8, seed, lib=po.ref())
    # threshold fit on a pooled sample (search.cpp:573-612)
    rng = np.random.default_rng(5)
    pooled = (3.0 + rng.exponential(0.6, 20000)).astype(np.float32)
    data["threshold/pooled"] = pooled
    data["threshold/runs"] = np.int32(8)
    data["threshold/value"] = np.float64(po.ref_null_threshold(pooled, 8))
    # build_stage on a small synthetic sequence (preprocess.cpp:286-486)
    n, h, w = 12, 40, 48
    rng = np.random.default_rng(11)
    raw = (1000.0 + 30.0 * rng.standard_normal((n, h, w))).astype(np.float32)
    raw[:, 10:13, 20:23] += 5000.0  # a star: masked by the scatter map
    raw[4, 30, 8] += 4000.0          # a cosmic ray in one frame
    raw[:, :, :2] = np.nan           # a blank border that grows by 3 px
    mjd = 61000.0 + np.arange(n) * (60.0 / 86400.0)
    prep = po.ref_prepare(raw, mjd, 3.0)
    data["prepare/raw"], data["prepare/mjd"] = raw, mjd
    data["prepare/sub"], data["prepare/background"] = prep["sub"], prep["background"]
    data["prepare/filtered"], data["prepare/t_h"] = prep["filtered"], prep["t_h"]
    st = po.ref_build_stage(prep["sub"], prep["t_h"][::-1].copy(), 3.0)
    data["stage_rev/filtered"], data["stage_rev/static_sky"] = st["filtered"], st["static_sky"]
    data["stage_rev/sigma"], data["stage_rev/mask"], data["stage_rev/cr"] = st["sigma"], st["mask"], st["cr_pixels"]
    # a blind search with null calibration (search.cpp:716-809, null_tested = 0)
    from fainttrail_amd import synth
    fr, mjd, _ = synth.make_sequence(12, 96, 64, nstars=12, nmovers=2, mover_snr=(3.0, 3.0), seed=21)
    fr = fr.numpy()
    args = np.array([12.0, 66.0, 0.5, 3.53])
    r = po.ref_search(fr, mjd, *args, null_runs=4, seed=77, null_tested=0)
    data["search/frames"], data["search/mjd"], data["search/args"] = fr, mjd, args
    data["search/runs"] = np.int32(4)
    data["search/threshold"] = np.float64(r["threshold"])
    data["search/pooled"], data["search/max_snr"] = r["pooled"], r["max_snr"]
    data["search/n_velocities"] = np.int32(r["n_velocities"])
    np.savez_compressed(OUT / "golden.npz", **data)
    print("wrote", OUT / "golden.npz", sum(v.nbytes for v in data.values()) // 1024, "KiB raw")


if __name__ == "__main__":
    main()
