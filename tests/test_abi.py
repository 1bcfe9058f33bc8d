"""The C-ABI library loads and exports every symbol include/*.h declares; options
parse like the reference's; the host emulation of the HIP algorithm (same plan,
tiles, -0.0 padding, edge counting, key merge) equals the reference on its test
inputs. No GPU needed."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import fainttrail_amd as ft
from refcases import same_bits

ROOT = Path(__file__).resolve().parents[1]
CASES = ["contract-random", "contract-constant", "shift-linear", "shift-curved", "cuda-linear", "cuda-table"]


def declared_symbols():
    syms = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b(ftk_[a-z0-9_]+)\s*\(", text):
            syms.add(m.group(1))
    return sorted(syms)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(ROOT / "fainttrail_amd" / "libftk_hip.so"))
    syms = declared_symbols()
    assert len(syms) >= 10
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_options_parse_like_the_reference():
    """parse_cuda_options (cuda_plan.cpp:14-33) rules for "hip:..."."""
    o = ft.parse_options("device=1,max_mb=64")
    assert o.device == 1 and o.max_mb == 64
    o = ft.parse_options("")
    assert o.device == -1 and o.max_mb == 0 and o.launch_adds > 0 and o.pad == 1
    assert o.tile_rows == 0 and o.queue == -1  # automatic work shape
    assert ft.parse_options("pad=0").pad == 0
    o = ft.parse_options("tile_rows=32,queue=1")
    assert o.tile_rows == 32 and o.queue == 1
    assert ft.parse_options("queue=0").queue == 0 and ft.parse_options("queue=0,tile_rows=16").tile_rows == 16
    for bad in ("speed=11", "device=", "max_mb=abc", "max_mb=-1", "launch_adds=0", "pad=2", "tile_rows=12",
                "tile_rows=8.5", "queue=2", "tile_rows=8,queue=0"):
        with pytest.raises(ft.FtkError):
            ft.parse_options(bad)


def test_unknown_backend_raises():
    with pytest.raises(ft.FtkError):
        ft.make_backend("no-such-backend")


@pytest.mark.parametrize("name", CASES)
def test_emulation_matches_reference(golden, name):
    g = {k.split("/", 1)[1]: v for k, v in golden.items() if k.startswith(name + "/")}
    m, b = ft.emulate_pass(g["frames"], g["dx"], g["dy"], int(g["min_count"]))
    assert same_bits(m, g["max_snr"]) and np.array_equal(b, g["best_v"])


LATTICE_CASES = [  # (frames, w, h, vmin, vmax, grid_tol, cadence h, least lattice share of the blocks)
    (16, 200, 150, 4.0, 60.0, 0.3, 0.05, 0.6),
    (20, 180, 140, 12.0, 66.0, 0.125, 1 / 60, 0.6),  # many motions with equal shift tables
    (9, 150, 90, 2.0, 30.0, 0.6, 0.3, 0.2),  # a coarse grid: many groups spread too far, general blocks
]


@pytest.mark.parametrize("case", LATTICE_CASES)
def test_emulation_with_lattice_blocks_equals_the_oracle(case):
    """Straight-line grids plan mostly into lattice blocks (4 x 4 motions sharing each
    column's dx and each row's dy in every frame). The host emulation stacks their
    interior tiles the lattice kernel's way (four column reads of 4 + spread rows per
    lane, rows c_b .. c_b + 3 per row class, unread rows as NaN) and motions whose
    shifts equal a lower-index motion's are not stacked: maps and indices still equal
    the blocked CPU oracle bit for bit (NaN sample included)."""
    from oracle import pyoracle as po
    n, w, h, vmin, vmax, tol, dt, share = case
    t = (np.arange(n) - (n - 1) / 2.0) * dt
    frames = po.test_normals(99 + n, (n, h, w))
    frames[3, 40, 50] = np.nan
    grid, _ = ft.velocity_grid(float(t[-1] - t[0]), vmin, vmax, tol)
    tab = ft.linear_shift_table(t, grid)
    mc = max(3, int(np.ceil(0.7 * n)))
    shape = ft.plan_shape(n, w, h, tab.dx, tab.dy, mc)
    assert shape["lattice_blocks"] > share * shape["blocks"]
    m, b = ft.emulate_pass(frames, tab.dx, tab.dy, mc)
    wm, wb = po.stack_blocked(frames, tab.dx, tab.dy, mc)
    assert same_bits(m, wm) and np.array_equal(b, wb)


def test_plan_shapes_of_the_benchmark_grids():
    """Lattice blocks hold nearly all stacked motions of the benchmark grids; equal shift
    tables (neighbouring velocities that round alike in every frame) are stacked once."""
    cases = [(60, 6144, 4096, 6.0, 120.0, 0.372, 0), (20, 512, 512, 12.0, 66.0, 0.125, 2528),
             (40, 2048, 2048, 12.0, 66.0, 0.212, 820)]
    for n, w, h, vmin, vmax, tol, shadowed in cases:
        t = (np.arange(n) - (n - 1) / 2.0) / 60.0
        grid, _ = ft.velocity_grid(float(t[-1] - t[0]), vmin, vmax, tol)
        tab = ft.linear_shift_table(t, grid)
        s = ft.plan_shape(n, w, h, tab.dx, tab.dy, max(3, int(np.ceil(0.7 * n))))
        stacked = len(grid) - s["shadowed_motions"]
        assert s["shadowed_motions"] == shadowed
        assert s["lattice_motions"] >= 0.94 * stacked
        assert s["blocks"] * 16 <= 1.06 * stacked  # little padding
        assert s["max_sy"] <= 8


def test_emulation_check_backend_inputs(golden):
    fr, mc = golden["check0/frames"], int(golden["check0/min_count"])
    for key in ("check0", "check0c"):
        m, b = ft.emulate_pass(fr, golden[f"{key}/dx"], golden[f"{key}/dy"], mc)
        assert same_bits(m, golden[f"{key}/max_snr"]) and np.array_equal(b, golden[f"{key}/best_v"])


def test_emulation_min_count_above_n_gives_undefined(golden):
    g = {k.split("/", 1)[1]: v for k, v in golden.items() if k.startswith("contract-random/")}
    m, b = ft.emulate_pass(g["frames"], g["dx"], g["dy"], 8)  # 7 frames
    assert np.all(b == -1) and np.all(m == 0)


def test_malformed_passes_are_refused():
    """validate_pass messages (backend.cpp:89-108) through the Python mirror and the C-ABI."""
    fr = np.zeros((3, 8, 8), np.float32)
    with pytest.raises(ft.FtkError, match="min_count"):
        ft.emulate_pass(fr, np.zeros((2, 3), np.int32), np.zeros((2, 3), np.int32), 0)
    tab = ft.ShiftTable.make(np.zeros((4, 2), np.int32), np.zeros((4, 2), np.int32))
    with pytest.raises(ft.FtkError, match="shift table and frame counts differ"):
        ft.StackPass(frames=fr, shifts=tab).table()
    tab = ft.ShiftTable.make(np.zeros((4, 3), np.int32), np.zeros((4, 3), np.int32))
    with pytest.raises(ft.FtkError, match="velocity counts differ"):
        ft.StackPass(frames=fr, shifts=tab, velocities=[(0, 0)] * 3).table()
    with pytest.raises(ft.FtkError, match="frames differ in size"):
        ft.StackPass(frames=[np.zeros((8, 8)), np.zeros((8, 9))], t_h=[0, 1], velocities=[(0, 0)]).table()


def test_python_linear_table_equals_oracle(golden):
    t, v = golden["check0/t_h"], golden["check0/vel"]
    tab = ft.linear_shift_table(t, v)
    assert np.array_equal(tab.dx, golden["check0/dx"]) and np.array_equal(tab.dy, golden["check0/dy"])


@pytest.mark.parametrize("key", ["grid_cfg1", "grid_cfg2", "grid_default"])
def test_product_velocity_grid_matches_reference(golden, key):
    span, vmin, vmax, tol = golden[f"{key}/args"]
    g, dv = ft.velocity_grid(span, vmin, vmax, tol)
    assert np.array_equal(g, golden[f"{key}/vel"]) and dv == golden[f"{key}/dv"]


def test_search_params_defaults_are_the_reference_defaults():
    """ftk_search_params_default and the Python SearchParams carry SearchParams'
    defaults (include/fainttrail/search.hpp:18-90): among them null_tested = 3000,
    max_frame_fraction = 0.25 and half_tolerance = 3 (:63-66), no registration offsets
    (fwhm_px = 3: the PSF estimate runs outside this path), and the candidate stage on."""
    from fainttrail_amd.search import SearchParams, _Params
    lib = ctypes.CDLL(str(ROOT / "fainttrail_amd" / "libftk_hip.so"))
    lib.ftk_search_params_default.argtypes = [ctypes.POINTER(_Params)]
    c = _Params()
    lib.ftk_search_params_default(ctypes.byref(c))
    want = {"vmin_pxh": 4.0, "vmax_pxh": 40.0, "grid_tol_px": 0.5, "clip_sigma": 8.0, "cr_sigma": 5.0,
            "static_exclude": -1, "mask_scatter": 2.5, "weight_scatter": 1.5, "min_coverage": 0.7,
            "threshold_snr": 0.0, "min_snr_floor": 5.0, "null_runs": 3, "target_false_per_run": 0.02,
            "max_candidates": 300, "edge_px": 4, "seed": 12345, "null_tested": 3000, "max_frame_fraction": 0.25,
            "half_tolerance": 3.0, "candidates": 1, "fit_on_device": -1}
    py = SearchParams().c()
    for k, v in want.items():
        assert getattr(c, k) == v, k
        assert getattr(py, k) == v, k
    assert not c.reg_dx and not c.reg_dy and not py.reg_dx and not py.reg_dy


def test_search_params_registration_offsets():
    from fainttrail_amd.search import SearchParams
    p = SearchParams(reg_dx=[0.5, 1.5, 2.5], reg_dy=[-1.0, 0.0, 1.0]).c()
    assert [p.reg_dx[i] for i in range(3)] == [0.5, 1.5, 2.5]
    assert [p.reg_dy[i] for i in range(3)] == [-1.0, 0.0, 1.0]
    with pytest.raises(ft.FtkError):
        SearchParams(reg_dx=[0.0, 1.0]).c()
    with pytest.raises(ft.FtkError):
        SearchParams(reg_dx=[0.0, 1.0], reg_dy=[0.0]).c()


def test_ctypes_records_match_the_header_layout(tmp_path):
    """The Python records of the search C-ABI (ftk_search_params, ftk_search_summary,
    ftk_candidate) have the header's size and field offsets, as gcc lays them out."""
    import shutil
    import subprocess
    from fainttrail_amd.search import _Candidate, _Params, _Summary
    if not shutil.which("gcc"):
        pytest.skip("gcc not present")
    recs = {"ftk_search_params": _Params, "ftk_search_summary": _Summary, "ftk_candidate": _Candidate}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ftk_hip.h"', "int main(void) {"]
    for cname, py in recs.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'    printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {(a, b): int(c) for a, b, c in (ln.split() for ln in out if ln)}
    for cname, py in recs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
