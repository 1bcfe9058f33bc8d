"""The CPU oracle against the reference: golden vectors written by the compiled
reference (tests/golden/make_golden.py), the reference tests' known answers and
checker sensitivity, and, where oracle/_ref is built, the live reference."""
import numpy as np
import pytest

import refcases
from oracle import pyoracle as po
from refcases import same_bits

CASES = ["contract-random", "contract-constant", "shift-linear", "shift-curved", "cuda-linear", "cuda-table"]


@pytest.mark.parametrize("name", CASES)
def test_naive_matches_reference_golden(golden, name):
    g = {k.split("/", 1)[1]: v for k, v in golden.items() if k.startswith(name + "/")}
    m, b = po.stack_naive(g["frames"], g["dx"], g["dy"], int(g["min_count"]))
    assert same_bits(m, g["naive_max_snr"]) and np.array_equal(b, g["naive_best_v"])


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("threads,band,vblock", [(1, 0, 0), (2, 1, 1), (3, 3, 5), (2, 64, 64), (2, 7, 2)])
def test_blocked_matches_reference_cpu_backend(golden, name, threads, band, vblock):
    """CpuBackend is bit-identical for any thread count and blocking (test_main.cpp:415-435)."""
    g = {k.split("/", 1)[1]: v for k, v in golden.items() if k.startswith(name + "/")}
    m, b = po.stack_blocked(g["frames"], g["dx"], g["dy"], int(g["min_count"]), threads, band, vblock)
    assert same_bits(m, g["max_snr"]) and np.array_equal(b, g["best_v"])


def test_known_answer_constant_frames():
    """All-equal frames: every fully covered motion ties; the lowest index wins with
    sqrt(7) (test_main.cpp:439-444)."""
    c = refcases.contract_case(True)
    m, b = po.stack_naive(c["frames"], c["dx"], c["dy"], c["min_count"])
    assert b[11, 18] == 0
    assert abs(m[11, 18] - np.sqrt(7.0)) < 1e-5


def test_checker_catches_wrong_results():
    """check_backend flags +0.01 on one pixel and a wrong best_v (test_main.cpp:447-465)."""
    c = refcases.contract_case(False, w=20, h=15, n=6, min_count=4)
    fr, dx, dy = c["frames"], c["dx"], c["dy"]
    ref = po.stack_naive(fr, dx, dy, 4)
    assert po.check_backend(fr, dx, dy, 4, ref, ref)["ok"]
    bad = (ref[0].copy(), ref[1])
    bad[0].reshape(-1)[37] += 0.01
    assert not po.check_backend(fr, dx, dy, 4, ref, bad)["ok"]
    b2 = ref[1].copy().reshape(-1)
    k = int(np.nonzero(b2 >= 0)[0][0])
    b2[k] = (b2[k] + 1) % dx.shape[0]
    assert not po.check_backend(fr, dx, dy, 4, ref, (ref[0], b2.reshape(ref[1].shape)))["ok"]


def test_check_backend_inputs_golden(golden):
    """check-backend round 0 inputs (simulate seed 77 -> prepare -> grid): naive and
    the curved table match the reference."""
    fr, mc = golden["check0/frames"], int(golden["check0/min_count"])
    m, b = po.stack_naive(fr, golden["check0/dx"], golden["check0/dy"], mc)
    assert same_bits(m, golden["check0/max_snr"]) and np.array_equal(b, golden["check0/best_v"])
    m, b = po.stack_blocked(fr, golden["check0c/dx"], golden["check0c/dy"], mc, 2)
    assert same_bits(m, golden["check0c/max_snr"]) and np.array_equal(b, golden["check0c/best_v"])


def test_linear_shifts_closed_form(golden):
    """lround(v t) in double: the oracle equals exact rounding of the double product,
    halves away from zero, and the reference's table."""
    t, v = golden["check0/t_h"], golden["check0/vel"]
    dx, dy = po.linear_shifts(t, v)
    assert np.array_equal(dx, golden["check0/dx"]) and np.array_equal(dy, golden["check0/dy"])
    x = np.array([0.5, -0.5, 1.5, -2.5, 0.49999999999999994, 2.0000000000000004])
    dxx, _ = po.linear_shifts(x, np.array([[1.0, 0.0]]))
    assert dxx.tolist() == [[1, -1, 2, -3, 0, 2]]


@pytest.mark.parametrize("key", ["grid_cfg1", "grid_cfg2", "grid_default"])
def test_velocity_grid(golden, key):
    span, vmin, vmax, tol = golden[f"{key}/args"]
    g, dv = po.velocity_grid(span, vmin, vmax, tol)
    assert np.array_equal(g, golden[f"{key}/vel"]) and dv == golden[f"{key}/dv"]
    # closed form: every grid speed lies in the annulus, vy-major order
    sp = np.hypot(g[:, 0], g[:, 1])
    assert sp.min() >= max(0.0, vmin - dv / 2) and sp.max() <= vmax + dv / 2
    assert np.all(np.diff(g[:, 1]) >= 0)


@pytest.mark.parametrize("seed", [12345, 7])
def test_null_time_orders(golden, seed):
    out = po.null_time_orders(golden[f"orders_{seed}/t"], 8, seed)
    assert np.array_equal(out, golden[f"orders_{seed}/out"])
    for row in out:  # each run is a permutation of the times
        assert np.array_equal(np.sort(row), np.sort(golden[f"orders_{seed}/t"]))


def test_threshold(golden):
    thr = po.null_threshold(golden["threshold/pooled"], int(golden["threshold/runs"]))
    assert thr[0] == golden["threshold/value"]


def test_build_stage_matches_reference(golden):
    """build_stage on reversed time order (a null permutation) and the whole prepare()."""
    st = po.build_stage(golden["prepare/sub"], golden["prepare/t_h"][::-1].copy(), 3.0)
    assert same_bits(st["static_sky"], golden["stage_rev/static_sky"])
    assert np.array_equal(st["sigma"], golden["stage_rev/sigma"])
    assert np.array_equal(st["mask"], golden["stage_rev/mask"])
    assert np.array_equal(st["cr_pixels"], golden["stage_rev/cr"])
    assert same_bits(st["filtered"], golden["stage_rev/filtered"])


def test_background_matches_reference(golden):
    sub, bg = po.background(golden["prepare/raw"])
    assert np.array_equal(bg, golden["prepare/background"])
    assert same_bits(sub, golden["prepare/sub"])
    st = po.build_stage(sub, golden["prepare/t_h"], 3.0)
    assert same_bits(st["filtered"], golden["prepare/filtered"])


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
def test_find_peaks_matches_reference(golden):
    m, b = golden["check0/max_snr"], golden["check0/best_v"]
    for floor, radius, edge, cap in [(3.0, 2, 4, 20000), (1.0, 3, 1, 50), (0.0, 2, 2, 100000)]:
        got = po.find_peaks(m, b, floor, radius, edge, cap)
        want = po.ref_find_peaks(m, b, floor, radius, edge, cap)
        for a_, b_ in zip(got, want):
            assert np.array_equal(a_, b_)


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
def test_blocked_matches_live_reference_mid_size():
    """check-backend round 1 (simulate 384x320, 40 frames, seed 78): the oracle's
    blocked pass equals the reference's run_pass_simple and CpuBackend."""
    fr, t, v, mc = po.ref_check_backend_inputs(1)
    dx, dy = po.linear_shifts(t, v)
    m, b = po.stack_blocked(fr, dx, dy, mc)
    rm, rb = po.ref_stack("simple", fr, dx, dy, mc)
    assert same_bits(m, rm) and np.array_equal(b, rb)


def test_blocked_row_window_equals_full(golden):
    """The CPU-baseline sampler (output rows from a window of frame rows) equals the
    same rows of the full pass."""
    g = {k.split("/", 1)[1]: v for k, v in golden.items() if k.startswith("check0/")}
    fr, dx, dy, mc = g["frames"], g["dx"], g["dy"], int(g["min_count"])
    full = po.stack_blocked(fr, dx, dy, mc)
    y0, y1 = 20, 30
    lo, hi = max(0, y0 + int(dy.min())), min(fr.shape[1], y1 + int(dy.max()))
    m, b = po.stack_blocked_rows(fr[:, lo:hi], fr.shape[1], lo, dx, dy, mc, y0, y1)
    assert same_bits(m, full[0][y0:y1]) and np.array_equal(b, full[1][y0:y1])
