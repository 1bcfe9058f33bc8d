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
    assert ft.parse_options("pad=0").pad == 0
    for bad in ("speed=11", "device=", "max_mb=abc", "max_mb=-1", "launch_adds=0", "pad=2"):
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
