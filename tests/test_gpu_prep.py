"""Device preprocessing (background + build_stage) against the reference: the golden
prepare() / build_stage outputs written by the compiled reference, and the oracle on
larger seeded sequences with stars, cosmic rays, blank borders and permuted times."""
import numpy as np
import pytest

from oracle import pyoracle as po
from refcases import same_bits

pytestmark = pytest.mark.gpu


def _stage(ft, n, w, h):
    from fainttrail_amd.stage import Stage
    return Stage(n, w, h)


def test_prepare_matches_reference_golden(hip, golden):
    """Background, blank growth and build_stage (real time order) equal the reference's
    prepare() (preprocess.cpp:589-650) bit for bit."""
    import torch
    from fainttrail_amd.stage import StageParams
    raw = golden["prepare/raw"]
    n, h, w = raw.shape
    st = _stage(hip, n, w, h)
    d_raw = torch.from_numpy(raw).cuda()
    sub, bg = st.background(d_raw)
    assert np.array_equal(bg, golden["prepare/background"])
    assert same_bits(sub.cpu().numpy(), golden["prepare/sub"])
    filt = st.build(sub, golden["prepare/t_h"], StageParams(fwhm_px=3.0))
    torch.cuda.synchronize()
    assert same_bits(filt.cpu().numpy(), golden["prepare/filtered"])


def test_null_order_stage_matches_reference_golden(hip, golden):
    """build_stage with reversed times (a null permutation; with_times,
    preprocess.cpp:488-521): static sky, sigma, mask, cosmic-ray counts and frames."""
    import torch
    from fainttrail_amd.stage import StageParams
    sub = torch.from_numpy(golden["prepare/sub"]).cuda()
    n, h, w = sub.shape
    st = _stage(hip, n, w, h)
    filt, det = st.build(sub, golden["prepare/t_h"][::-1].copy(), StageParams(fwhm_px=3.0), want_details=True)
    torch.cuda.synchronize()
    assert same_bits(det["static_sky"].cpu().numpy(), golden["stage_rev/static_sky"])
    assert np.array_equal(det["sigma"], golden["stage_rev/sigma"])
    assert np.array_equal(det["mask"].cpu().numpy(), golden["stage_rev/mask"])
    assert np.array_equal(det["cr_pixels"], golden["stage_rev/cr"])
    assert same_bits(filt.cpu().numpy(), golden["stage_rev/filtered"])


def synthetic_sequence(n, h, w, seed):
    """Sky 1000 ADU + N(0, 32^2), Gaussian stars, a few cosmic rays and a blank border."""
    rng = np.random.default_rng(seed)
    raw = (1000.0 + 32.0 * rng.standard_normal((n, h, w))).astype(np.float32)
    yy, xx = np.mgrid[0:h, 0:w]
    for _ in range(max(3, h * w // 4000)):
        x0, y0 = rng.uniform(0, w), rng.uniform(0, h)
        f = 10 ** rng.uniform(3, 6)
        raw += (f / (2 * np.pi * 1.5 ** 2) * np.exp(-((xx - x0) ** 2 + (yy - y0) ** 2) / (2 * 1.5 ** 2))).astype(np.float32)
    for _ in range(3):
        i, y, x = rng.integers(0, n), rng.integers(2, h - 2), rng.integers(2, w - 2)
        raw[i, y, x] += 3000.0
    raw[:, :, : w // 50 + 1] = np.nan
    return raw


@pytest.mark.parametrize("n,h,w,seed,integer", [(20, 96, 128, 1, False), (33, 64, 200, 2, False),
                                                (12, 48, 96, 3, True), (40, 48, 96, 4, True), (60, 40, 80, 5, True),
                                                (70, 32, 72, 6, True)])
def test_stage_matches_oracle_random_orders(hip, n, h, w, seed, integer):
    """Background and build_stage for the real order and two shuffled orders (the null
    runs of the search) equal the oracle's restatement, which is pinned to the
    reference by the golden tests. Frame counts cover every form of the static-sky
    kernel (samples in registers for n <= 16, 32, 48, 64; in LDS above); integer ADU
    frames (as FITS data are) make equal samples common, so the time-rank order of
    equal values is exercised."""
    import torch
    from fainttrail_amd.stage import StageParams
    raw = synthetic_sequence(n, h, w, seed)
    if integer:
        raw = np.round(raw).astype(np.float32)
    st = _stage(hip, n, w, h)
    sub_d, bg = st.background(torch.from_numpy(raw).cuda())
    sub_o, bg_o = po.background(raw)
    assert np.array_equal(bg, bg_o)
    assert same_bits(sub_d.cpu().numpy(), sub_o)
    t = (np.arange(n) - (n - 1) / 2) / 60.0
    orders = [t] + list(po.null_time_orders(t, 2, 12345))
    for k, tk in enumerate(orders):
        print(f"stage n={n} {w}x{h} order {k}", flush=True)
        filt, det = st.build(sub_d, tk, StageParams(fwhm_px=3.53), want_details=True)
        torch.cuda.synchronize()
        want = po.build_stage(sub_o, tk, 3.53)
        assert same_bits(det["static_sky"].cpu().numpy(), want["static_sky"]), k
        assert np.array_equal(det["sigma"], want["sigma"]), k
        assert np.array_equal(det["mask"].cpu().numpy(), want["mask"]), k
        assert np.array_equal(det["cr_pixels"], want["cr_pixels"]), k
        assert same_bits(filt.cpu().numpy(), want["filtered"]), k
