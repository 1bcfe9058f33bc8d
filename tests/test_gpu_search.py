"""The device-resident search against the reference's own search() (null_tested = 0):
the real pass's max-S/N map, every pooled null-peak S/N value and the calibrated
threshold (an FP64 value) must be identical; with the tail-fit sums on the GPU the
threshold agrees to 1e-12 relative (device log and summation order differ)."""
import numpy as np
import pytest

from oracle import pyoracle as po
from refcases import same_bits

pytestmark = pytest.mark.gpu


def test_search_matches_reference_golden(hip, golden):
    import torch
    from fainttrail_amd.search import SearchParams, Searcher
    fr, mjd = golden["search/frames"], golden["search/mjd"]
    vmin, vmax, tol, fwhm = golden["search/args"]
    runs = int(golden["search/runs"])
    n, h, w = fr.shape
    res = Searcher(n, w, h).run(torch.from_numpy(fr).cuda(), mjd,
                                SearchParams(vmin_pxh=vmin, vmax_pxh=vmax, grid_tol_px=tol, fwhm_px=fwhm,
                                             null_runs=runs, seed=77))
    torch.cuda.synchronize()
    assert res.n_velocities == int(golden["search/n_velocities"])
    assert same_bits(res.max_snr.cpu().numpy(), golden["search/max_snr"])
    assert np.array_equal(np.sort(res.null_peaks), np.sort(golden["search/pooled"]))
    assert res.threshold_snr == float(golden["search/threshold"])


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed,runs", [(3, 3), (11, 5)])
def test_search_matches_live_reference(hip, seed, runs):
    import torch
    from fainttrail_amd import synth
    from fainttrail_amd.search import SearchParams, Searcher
    fr, mjd, _ = synth.make_sequence(20, 160, 120, nstars=40, nmovers=3, mover_snr=(2.5, 4.0), seed=seed)
    frames = fr.numpy()
    want = po.ref_search(frames, mjd, 12.0, 66.0, 0.4, 3.53, runs, seed=1000 + seed)
    print(f"search seed {seed}: reference threshold {want['threshold']}", flush=True)
    prm = SearchParams(vmin_pxh=12.0, vmax_pxh=66.0, grid_tol_px=0.4, fwhm_px=3.53, null_runs=runs, seed=1000 + seed)
    s = Searcher(20, 160, 120)
    got = s.run(torch.from_numpy(frames).cuda(), mjd, prm)
    torch.cuda.synchronize()
    assert got.n_velocities == want["n_velocities"]
    assert same_bits(got.max_snr.cpu().numpy(), want["max_snr"])
    assert np.array_equal(np.sort(got.null_peaks), np.sort(want["pooled"]))
    assert got.threshold_snr == want["threshold"]
    # the FP64 tail-fit sums on the GPU: same decision, threshold within 1e-12 relative
    prm.fit_on_device = True
    got2 = s.run(torch.from_numpy(frames).cuda(), mjd, prm)
    assert got2.threshold_method == got.threshold_method
    assert abs(got2.threshold_snr - want["threshold"]) <= 1e-12 * abs(want["threshold"])
    # selected peaks are the real peaks at or above the threshold, strongest first
    sel = got.selected
    assert np.all(sel.snr >= got.threshold_snr) and np.all(np.diff(sel.snr) <= 0)


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("tile", [48, 64])
def test_search_matches_reference_tiled_search(hip, tile):
    """search_tiled (src/tiled.cpp:272-645) with uneven tiles (150 x 110 px cut into
    48- or 64-px cores with halos, the last ones clipped) equals the device search,
    which keeps the whole frame resident: same max-S/N map, null peaks and threshold."""
    import torch
    from fainttrail_amd import synth
    from fainttrail_amd.search import SearchParams, Searcher
    fr, mjd, _ = synth.make_sequence(16, 150, 110, nstars=30, nmovers=2, mover_snr=(3.0, 3.0), seed=5)
    frames = fr.numpy()
    want = po.ref_search_tiled(frames, mjd, 12.0, 66.0, 0.5, 3.53, 3, tile, seed=9)
    got = Searcher(16, 150, 110).run(torch.from_numpy(frames).cuda(), mjd,
                                     SearchParams(vmin_pxh=12.0, vmax_pxh=66.0, grid_tol_px=0.5, fwhm_px=3.53,
                                                  null_runs=3, seed=9))
    torch.cuda.synchronize()
    assert same_bits(got.max_snr.cpu().numpy(), want["max_snr"])
    assert np.array_equal(np.sort(got.null_peaks), np.sort(want["pooled"]))
    assert got.threshold_snr == want["threshold"]
