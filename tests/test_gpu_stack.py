"""Parity of the gfx950 stacking kernel with the reference CPU backend, through the
C-ABI (ftk_stack_pass / ftk_plan_run). Bit-identical maps and motion indices are the
bar: the reference's check-backend reports "bit-identical in 4 of 4 checks" for its
GPU backend (cli/main.cpp:676-746), and so must this one."""
import numpy as np
import pytest

from oracle import pyoracle as po
from refcases import same_bits

pytestmark = pytest.mark.gpu

CASES = ["contract-random", "contract-constant", "shift-linear", "shift-curved", "cuda-linear", "cuda-table"]


def case(golden, name):
    return {k.split("/", 1)[1]: v for k, v in golden.items() if k.startswith(name + "/")}


def run_hip(ft, frames, dx, dy, min_count, options=""):
    be = ft.make_backend("hip:" + options if options else "hip")
    tab = ft.ShiftTable.make(dx, dy)
    r = be.run(ft.StackPass(frames=frames, shifts=tab, min_count=int(min_count)))
    return r.max_snr, r.best_v


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("options", ["", "pad=0", "max_mb=0.01,launch_adds=1", "max_mb=0.003,launch_adds=3000",
                                     "max_mb=0.01,launch_adds=1,pad=0", "queue=0", "queue=0,pad=0",
                                     "max_mb=0.01,launch_adds=1,queue=0", "queue=1", "queue=1,pad=0",
                                     "max_mb=0.01,launch_adds=1,queue=1", "tile_rows=8", "tile_rows=32,pad=0",
                                     "tile_rows=8,max_mb=0.01,launch_adds=1,pad=0", "tile_rows=32,max_mb=0.003"])
def test_reference_test_inputs_bit_identical(hip, golden, name, options):
    """test_backend_contract / test_shift_tables / test_cuda_algorithm inputs (NaN,
    shifts beyond the frame, 171 motions, curved tables), unbanded and forced into
    row bands with tiny launches (test_main.cpp:602-609 runs "cuda:max_mb=0.01,
    launch_adds=1"); pad=0 keeps plain frame rows (masked edge staging) instead of
    the zero-bordered layout; queue=0 launches one workgroup per work item, queue=1 the
    persistent kernel (the default picks by frame size); tile_rows 8 and 32 change the
    tile height."""
    g = case(golden, name)
    print(f"case {name} options '{options}'", flush=True)
    m, b = run_hip(hip, g["frames"], g["dx"], g["dy"], g["min_count"], options)
    assert same_bits(m, g["max_snr"]), name
    assert np.array_equal(b, g["best_v"]), name


def test_check_backend_round0_naive_and_curved(hip, golden):
    """check-backend checks 1 and 2 (cli/main.cpp:713-720): 72x52, 12 frames vs the
    naive contract, straight and curved tables."""
    fr, mc = golden["check0/frames"], int(golden["check0/min_count"])
    for key in ("check0", "check0c"):
        m, b = run_hip(hip, fr, golden[f"{key}/dx"], golden[f"{key}/dy"], mc)
        assert same_bits(m, golden[f"{key}/max_snr"]) and np.array_equal(b, golden[f"{key}/best_v"]), key


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
def test_check_backend_round1_simple_and_bands(hip):
    """check-backend checks 3 and 4 (cli/main.cpp:721-741): 384x320, 40 frames vs
    run_pass_simple, and the same forced into row bands (cap 2 MB, 2e6 units/launch)."""
    fr, t, v, mc = po.ref_check_backend_inputs(1)
    dx, dy = po.linear_shifts(t, v)
    rm, rb = po.ref_stack("simple", fr, dx, dy, mc)
    for opts in ("", "max_mb=2,launch_adds=2e6"):
        print(f"check-backend round 1 options '{opts}'", flush=True)
        m, b = run_hip(hip, fr, dx, dy, mc, opts)
        assert same_bits(m, rm) and np.array_equal(b, rb), opts
        chk = po.check_backend(fr, dx, dy, mc, (rm, rb), (m, b))
        assert chk["ok"] and chk["max_abs_diff"] == 0.0


def test_plan_tune_keeps_every_bit(hip):
    """ftk_plan_tune times the launch forms and tile heights and keeps the fastest; the
    tuned plan's fp32 and fp16 passes equal the untuned ones bit for bit (and the fp32
    pass equals the oracle's blocked CPU pass)."""
    import torch
    n, h, w = 24, 200, 328
    rng = np.random.default_rng(5)
    frames = rng.standard_normal((n, h, w), dtype=np.float32)
    frames[3, 10:14, 20:30] = np.nan
    t = (np.arange(n) - (n - 1) / 2.0) / 60.0
    grid, _ = hip.velocity_grid(t[-1] - t[0], 10.0, 60.0, 0.3)
    tab = hip.linear_shift_table(t, grid)
    mc = max(3, int(np.ceil(0.7 * n)))
    om, ob = po.stack_blocked(frames, tab.dx, tab.dy, mc)
    d = torch.from_numpy(frames).cuda()
    plan = hip.Plan(n, w, h, tab.dx, tab.dy, mc)
    plan.store(d, "f16")
    m16, b16 = (x.cpu().numpy() for x in plan.run_stored())
    tr, q, secs = plan.tune(d, reps=1)
    print(f"plan tune fp32: tile rows {tr}, launch form {q}, {secs * 1e3:.3f} ms", flush=True)
    assert tr in (8, 16, 32) and q in (0, 1) and secs > 0 and (q == 1 or tr == 16)
    m, b = plan.run(d)
    assert same_bits(m.cpu().numpy(), om) and np.array_equal(b.cpu().numpy(), ob)
    tr16, q16, secs16 = plan.tune(None, reps=1)
    print(f"plan tune fp16: tile rows {tr16}, launch form {q16}, {secs16 * 1e3:.3f} ms", flush=True)
    m2, b2 = plan.run_stored()
    torch.cuda.synchronize()
    assert same_bits(m2.cpu().numpy(), m16) and np.array_equal(b2.cpu().numpy(), b16)


def test_cfg1_shape_bit_identical_to_blocked_oracle(hip):
    """512 x 512, 20 frames, the 5468-motion grid of the 512^2 configuration
    (dv * span = 0.5 px), unit-variance frames: device-resident plan and the host
    pass both equal the oracle's blocked CPU backend bit for bit."""
    import torch
    n, h, w = 20, 512, 512
    rng = np.random.default_rng(7)
    frames = rng.standard_normal((n, h, w), dtype=np.float32)
    t = (np.arange(n) - (n - 1) / 2.0) / 60.0
    grid, dv = hip.velocity_grid(t[-1] - t[0], 12.0, 66.0, 0.125)
    assert grid.shape[0] == 5468
    tab = hip.linear_shift_table(t, grid)
    mc = max(3, int(np.ceil(0.7 * n)))
    print("cfg1 shape: oracle", flush=True)
    om, ob = po.stack_blocked(frames, tab.dx, tab.dy, mc)
    d = torch.from_numpy(frames).cuda()
    for opts in ("", "pad=0", "queue=0", "queue=1", "tile_rows=8", "tile_rows=32", "tile_rows=32,pad=0"):
        print(f"cfg1 shape: device plan '{opts}'", flush=True)
        plan = hip.Plan(n, w, h, tab.dx, tab.dy, mc, opts)
        m, b = plan.run(d)
        torch.cuda.synchronize()
        assert same_bits(m.cpu().numpy(), om) and np.array_equal(b.cpu().numpy(), ob), opts
    print("cfg1 shape: host pass", flush=True)
    m2, b2 = run_hip(hip, frames, tab.dx, tab.dy, mc)
    assert same_bits(m2, om) and np.array_equal(b2, ob)


def test_full_size_bands_bit_identical_to_blocked_oracle(hip):
    """The full tiled-6144x4096 shape: 60 frames of 6144 x 4096, the 19952-motion grid
    (grid_tol 0.372, 6-120 px/h, 60 s cadence), min_count 42. One device pass over the
    whole frame; the oracle's blocked CPU pass on three bands of output rows (the top
    edge, the middle and the bottom edge, every column: edge tiles on all four sides)
    must give the same bits, and the whole map must keep the contract's form (indices
    in range, undefined pixels 0 / -1)."""
    import os

    import torch
    n, h, w = 60, 4096, 6144
    t = (np.arange(n) - (n - 1) / 2.0) / 60.0
    grid, _ = hip.velocity_grid(t[-1] - t[0], 6.0, 120.0, 0.372)
    tab = hip.linear_shift_table(t, grid)
    assert tab.motions == 19952
    mc = max(3, int(np.ceil(0.7 * n)))
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(11)
    d = torch.randn((n, h, w), generator=gen, device="cuda:0", dtype=torch.float32)
    print("full size: device pass", flush=True)
    plan = hip.Plan(n, w, h, tab.dx, tab.dy, mc)
    m, b = plan.run(d)
    torch.cuda.synchronize()
    mh, bh = m.cpu().numpy(), b.cpu().numpy()
    assert bh.min() >= -1 and bh.max() < tab.motions
    assert np.all(mh[bh < 0] == 0.0) and np.all(np.isfinite(mh))
    threads = len(os.sched_getaffinity(0))
    lo_dy, hi_dy = int(tab.dy.min()), int(tab.dy.max())
    for y0, y1 in ((0, 6), (2045, 2051), (h - 6, h)):
        print(f"full size: oracle rows {y0}-{y1 - 1}", flush=True)
        r0, r1 = max(0, y0 + lo_dy), min(h, y1 + hi_dy)
        om, ob = po.stack_blocked_rows(d[:, r0:r1].cpu().numpy(), h, r0, tab.dx, tab.dy, mc, y0, y1, threads)
        assert same_bits(mh[y0:y1], om) and np.array_equal(bh[y0:y1], ob), (y0, y1)
    del d, m, b
    torch.cuda.empty_cache()


def test_plan_reuse_and_stride(hip, golden):
    """A plan runs repeatedly on resident frames; a padded frame stride gives the same maps."""
    import torch
    g = case(golden, "cuda-linear")
    fr = g["frames"]
    n, h, w = fr.shape
    big = torch.zeros((n, h * w + 77), dtype=torch.float32, device="cuda")
    big[:, : h * w] = torch.from_numpy(fr.reshape(n, -1)).cuda()
    m = torch.empty((h, w), dtype=torch.float32, device="cuda")
    b = torch.empty((h, w), dtype=torch.int32, device="cuda")
    plan = hip.Plan(n, w, h, g["dx"], g["dy"], int(g["min_count"]))
    for k in range(4):
        if k == 3:
            plan = hip.Plan(n, w, h, g["dx"], g["dy"], int(g["min_count"]), "pad=0")
        plan.run_ptr(big.data_ptr(), h * w + 77, m.data_ptr(), b.data_ptr(), torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert same_bits(m.cpu().numpy(), g["max_snr"]) and np.array_equal(b.cpu().numpy(), g["best_v"])


def test_cancellation_between_launches(hip, golden):
    """A progress callback returning false cancels (types.hpp:70-76; backend_cuda.cu:172)."""
    g = case(golden, "contract-random")
    be = hip.make_backend("hip:launch_adds=1")
    calls = []

    def progress(f, stage):
        calls.append(f)
        return len(calls) < 3

    with pytest.raises(hip.Cancelled):
        be.run(hip.StackPass(frames=g["frames"], shifts=hip.ShiftTable.make(g["dx"], g["dy"]), min_count=5), progress)
    assert len(calls) == 3


def test_missing_device_is_an_error(hip, golden):
    g = case(golden, "contract-random")
    be = hip.make_backend("hip:device=57")
    with pytest.raises(hip.FtkError, match="not present"):
        be.run(hip.StackPass(frames=g["frames"], shifts=hip.ShiftTable.make(g["dx"], g["dy"]), min_count=5))


@pytest.mark.parametrize("min_count", [3, 8])
def test_corner_coverage_chains_anti_chains_and_others(hip, min_count):
    """Edge tiles that leave the frame in x and y count their covering frames from the
    two coverage tables when a motion's (dx, dy) pairs form a chain or an anti-chain, and
    walk the frames otherwise. A 150 x 40 frame (every tile an edge tile, corners
    included) with shifts up to 22 px: straight tracks in all four quadrants (chains and
    anti-chains), axis-parallel ones (both), and zig-zag and random tables (neither),
    against the naive contract of the oracle, bit for bit."""
    import fainttrail_amd as ft
    rng = np.random.default_rng(11)
    n, w, h = 12, 150, 40
    frames = rng.normal(size=(n, h, w)).astype(np.float32)
    t = np.arange(n) - 5.5
    rows_x, rows_y = [], []
    for a, b in [(3, 2), (-3, 2), (3, -2), (-3, -2), (0, 3.5), (3.5, 0), (1.3, -3.7), (-3.9, 1.1), (2.2, 2.9),
                 (-1.7, -3.3), (3.8, -0.6), (0.4, 3.9)]:
        rows_x.append(np.round(a * t).astype(np.int32))  # numpy rounds half to even; any table will do
        rows_y.append(np.round(b * t).astype(np.int32))
    for q in range(12):  # zig-zags: dx rises while dy alternates (neither kind)
        rows_x.append(np.round((q - 6) * 0.6 * t).astype(np.int32))
        rows_y.append(np.where(np.arange(n) % 2 == 0, 1, -1).astype(np.int32) * (6 + q))
    for q in range(24):  # random tables
        rows_x.append(rng.integers(-22, 23, n).astype(np.int32))
        rows_y.append(rng.integers(-22, 23, n).astype(np.int32))
    dx, dy = np.stack(rows_x), np.stack(rows_y)
    want_m, want_b = po.stack_naive(frames, dx, dy, min_count)
    for options in ("", "pad=0"):
        m, b = run_hip(ft, frames, dx, dy, min_count, options)
        assert same_bits(m, want_m) and np.array_equal(b, want_b), options
