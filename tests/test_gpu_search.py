"""The device-resident search against the reference's own search() (its defaults: the
strongest 3000 peaks of each null run screened for artefacts): the real pass's max-S/N
and best-velocity maps, its peak list (find_peaks order), the selected peaks, every
pooled null-peak S/N value in the search's own order (run after run) and the
calibrated threshold (an FP64 value) must be identical; with the tail-fit sums on the
GPU the threshold agrees to 50 m eps relative (m fitted peaks; see fit_tol).
The golden fixture holds a search with no screen (null_tested = 0)."""
import numpy as np
import pytest

from oracle import pyoracle as po
from refcases import same_bits

pytestmark = pytest.mark.gpu

# The FP64 null-tail fit with its sums on the GPU against the host fit. The host adds the
# m fitted terms of each of the nine normal-equation sums one after another (the
# reference's order), so its sums carry up to ~m eps relative rounding; the GPU adds
# them in blocks (and takes its own log), which is closer to exact. The 3 x 3 solve and
# the bisection for the level pass the difference on, amplified by the fit's
# conditioning: 1.4e-12 at m = 1 877 and 4.2e-10 at m = 768 000 measured. The bound is
# 50 m eps; the method (the decision) must be the same.
def fit_tol(pooled, runs) -> float:
    m = min(len(pooled), 3000 * runs)
    return 50.0 * m * np.finfo(np.float64).eps


def _f64_bits(v) -> int:
    return int(np.array([v], np.float64).view(np.uint64)[0])


def _same_candidates(got, want):
    """The candidate list of refine + finish_candidates (search.cpp:246-395, 627-675):
    the same candidates in the same order, every field bit for bit, the same flags."""
    assert len(got) == len(want), (len(got), len(want))
    for c, r in zip(got, want):
        assert c.id == r["id"] and c.ok == r["ok"] and c.flags == r["flags"], (c, r)
        for f in po.CAND_FIELDS:
            if f in ("id", "ok"):
                continue
            assert _f64_bits(getattr(c, f)) == _f64_bits(r[f]), (c.id, f, getattr(c, f), r[f])


def _same_as_reference(got, want):
    """Every output the reference's search exposes, bit for bit and in order."""
    assert got.n_velocities == want["n_velocities"]
    assert same_bits(got.max_snr.cpu().numpy(), want["max_snr"])
    assert same_bits(got.null_peaks, want["pooled"])  # run after run, in the search's order
    assert got.threshold_snr == want["threshold"]
    assert got.threshold_method == want["method"]
    if "best_v" in want:
        assert np.array_equal(got.best_v.cpu().numpy(), want["best_v"])
        for mine, theirs in ((got.peaks, want["peaks"]), (got.selected, want["selected"])):
            for f in ("x", "y", "vidx"):
                assert np.array_equal(getattr(mine, f), theirs[f]), f
            assert same_bits(mine.snr, theirs["snr"])
        _same_candidates(got.candidates, want["candidates"])


def test_search_matches_reference_golden(hip, golden):
    import torch
    from fainttrail_amd.search import SearchParams, Searcher
    fr, mjd = golden["search/frames"], golden["search/mjd"]
    vmin, vmax, tol, fwhm = golden["search/args"]
    runs = int(golden["search/runs"])
    n, h, w = fr.shape
    res = Searcher(n, w, h).run(torch.from_numpy(fr).cuda(), mjd,
                                SearchParams(vmin_pxh=vmin, vmax_pxh=vmax, grid_tol_px=tol, fwhm_px=fwhm,
                                             null_runs=runs, seed=77, null_tested=0))
    torch.cuda.synchronize()
    assert res.n_velocities == int(golden["search/n_velocities"])
    assert same_bits(res.max_snr.cpu().numpy(), golden["search/max_snr"])
    assert same_bits(res.null_peaks, golden["search/pooled"])  # in the search's order
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
    _same_as_reference(got, want)
    # the FP64 tail-fit sums on the GPU: same decision, threshold within 50 m eps relative
    prm.fit_on_device = True
    got2 = s.run(torch.from_numpy(frames).cuda(), mjd, prm)
    assert got2.threshold_method == got.threshold_method
    assert abs(got2.threshold_snr - want["threshold"]) <= fit_tol(got.null_peaks, runs) * abs(want["threshold"])
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
    _same_as_reference(got, want)


def _screen_sequence(n=20, w=160, h=120, seed=5):
    """A sequence whose null runs hold artefacts of every kind the screen tests for:
    PSF-wide flashes in single frames (the single-frame share), registration borders
    (NaN) that grow with the drift (the field edge), and registration offsets that make
    the detector-fixed test run."""
    from fainttrail_amd import synth
    fr, mjd, _ = synth.make_sequence(n, w, h, nstars=40, nmovers=2, mover_snr=(3.0, 3.0), seed=seed)
    f = fr.numpy().copy()
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    for _ in range(12):
        i = rng.integers(n)
        x0, y0 = rng.uniform(10, w - 10), rng.uniform(10, h - 10)
        f[i] += 400.0 * np.exp(-((xx - x0) ** 2 + (yy - y0) ** 2) / (2 * 1.5 ** 2))
    for i in range(n):
        f[i, :, : 2 + i // 3] = np.nan
        f[i, : 1 + i // 4, :] = np.nan
    reg = (0.35 * np.arange(n) - 3.0, -0.25 * np.arange(n) + 2.0)
    return f, mjd, reg


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("with_reg", [False, True])
@pytest.mark.parametrize("group", [1, 0, 2])
def test_null_screen_matches_reference(hip, with_reg, group):
    """is_artefact_peak on the device (src/search.cpp:229-243, 773-778): on a sequence
    where the screen leaves out about half of the null peaks (426 -> 228 in the
    reference, 200 with the detector test) and moves the threshold from 15.3 to 5.7,
    the pooled null peaks and the threshold equal the reference's, one pass at a time
    and batched, with and without registration offsets."""
    import torch
    from fainttrail_amd.search import SearchParams, Searcher
    f, mjd, reg = _screen_sequence()
    n, h, w = f.shape
    rg = reg if with_reg else None
    want = po.ref_search(f, mjd, 12.0, 66.0, 0.5, 3.53, 4, seed=9, reg=rg)
    none = po.ref_search(f, mjd, 12.0, 66.0, 0.5, 3.53, 4, seed=9, null_tested=0)
    print(f"null screen group {group} reg {with_reg}: reference pools {len(want['pooled'])} of "
          f"{len(none['pooled'])}, threshold {want['threshold']}", flush=True)
    assert len(want["pooled"]) < 0.7 * len(none["pooled"])
    prm = SearchParams(vmin_pxh=12.0, vmax_pxh=66.0, grid_tol_px=0.5, fwhm_px=3.53, null_runs=4, seed=9,
                       passes_per_launch=group, reg_dx=rg[0] if rg else None, reg_dy=rg[1] if rg else None)
    got = Searcher(n, w, h).run(torch.from_numpy(f).cuda(), mjd, prm)
    torch.cuda.synchronize()
    _same_as_reference(got, want)
    assert got.null_artefacts == len(none["pooled"]) - len(want["pooled"])
    assert got.null_screened == len(none["pooled"])  # fewer than 3000 peaks per run: all tested
    # null_tested = 0 is the unscreened search
    prm.null_tested = 0
    got0 = Searcher(n, w, h).run(torch.from_numpy(f).cuda(), mjd, prm)
    _same_as_reference(got0, none)
    assert got0.null_artefacts == 0


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
def test_null_screen_limit(hip):
    """null_tested below the run's peak count: only the strongest null_tested peaks of
    each run (in find_peaks order) are tested (null_tested_count, search.cpp:695-698)."""
    import torch
    from fainttrail_amd.search import SearchParams, Searcher
    f, mjd, _ = _screen_sequence(seed=6)
    n, h, w = f.shape
    want = po.ref_search(f, mjd, 12.0, 66.0, 0.5, 3.53, 3, seed=2, null_tested=25)
    got = Searcher(n, w, h).run(torch.from_numpy(f).cuda(), mjd,
                                SearchParams(vmin_pxh=12.0, vmax_pxh=66.0, grid_tol_px=0.5, fwhm_px=3.53,
                                             null_runs=3, seed=2, null_tested=25))
    assert got.null_screened == 3 * 25
    _same_as_reference(got, want)


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("group", [1, 0])
def test_search_selects_the_references_peaks(hip, group):
    """A sequence with 8 movers bright enough to pass the calibrated threshold: the
    selected list (select_peaks, search.cpp:614-620) is non-trivial (the reference
    keeps at least 5) and equal to the reference's, one pass at a time and batched."""
    import torch
    from fainttrail_amd import synth
    from fainttrail_amd.search import SearchParams, Searcher
    fr, mjd, _ = synth.make_sequence(24, 256, 192, nstars=60, nmovers=8, mover_snr=(2.0, 4.0), seed=13)
    frames = fr.numpy()
    want = po.ref_search(frames, mjd, 12.0, 66.0, 0.5, 3.53, 5, seed=21)
    print(f"selected: reference keeps {len(want['selected']['x'])} of {len(want['peaks']['x'])} peaks, "
          f"threshold {want['threshold']}", flush=True)
    assert len(want["selected"]["x"]) >= 5
    got = Searcher(24, 256, 192).run(torch.from_numpy(frames).cuda(), mjd,
                                     SearchParams(vmin_pxh=12.0, vmax_pxh=66.0, grid_tol_px=0.5, fwhm_px=3.53,
                                                  null_runs=5, seed=21, passes_per_launch=group))
    torch.cuda.synchronize()
    _same_as_reference(got, want)
    print(f"candidates: {len(got.candidates)} ({sum(c.ok for c in got.candidates)} ok), flags "
          f"{sorted(set(c.flags for c in got.candidates))}", flush=True)
    assert len(got.candidates) >= 5


def _neo_sequence():
    """24 frames of 160 x 128 over 2 h with a ground observer's diurnal parallax
    (detrended: 0.17 px at 1 AU) and 6 movers on curved tracks at inverse distances
    15-45 / AU (2.6-7.7 px of curvature), bright enough to be selected."""
    from fainttrail_amd import synth
    n, w, h, cad = 24, 160, 128, 300.0
    mjd = 61000.0 + np.arange(n) * cad / 86400.0
    par = synth.diurnal_parallax((mjd - 0.5 * (mjd[0] + mjd[-1])) * 24.0, amp_px=9.0)
    fr, _, movers = synth.make_sequence(n, w, h, nstars=40, nmovers=6, mover_snr=(2.5, 4.0), rate_px_min=(0.2, 0.45),
                                        cadence_s=cad, seed=31, mover_box=(55, 105, 45, 83), parallax=par,
                                        mover_inv_dist=(15.0, 45.0))
    return fr.numpy(), mjd, par, movers


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("form", ["one-pass", "batched", "server"])
def test_near_earth_search_matches_reference(hip, form):
    """The near-Earth search (neo_min_dist_au > 0, src/search.cpp:727-733): curved
    tracks over the reference's inverse_distance_grid (8 distances here, the
    neo_max_trials cap) times the velocity grid, each null run's tables with the
    parallax of the frames whose times it takes (with_times, preprocess.cpp:503-515),
    the artefact screen and the refinement along curved tracks (the inverse distance
    searched and fitted, refine :274-285, 308-316), ghosts by curved tracks. Every
    output equals the reference's search() with the same parallax model, bit for bit,
    one pass at a time, batched, and as a server job."""
    import torch
    from fainttrail_amd.search import SearchParams, Searcher
    frames, mjd, par, movers = _neo_sequence()
    neo = {"par_x": par[0], "par_y": par[1], "min_dist_au": 0.02, "max_trials": 8}
    want = po.ref_search(frames, mjd, 12.0, 30.0, 0.5, 3.53, 3, seed=21, neo=neo)
    assert want["n_distances"] == 8 and want["dq_inv_au"] > 0.0
    assert len(want["candidates"]) >= 5 and any(c["inv_dist_au"] > 10.0 for c in want["candidates"])
    prm = SearchParams(vmin_pxh=12.0, vmax_pxh=30.0, grid_tol_px=0.5, fwhm_px=3.53, null_runs=3, seed=21,
                       neo_min_dist_au=0.02, neo_max_trials=8, par_x=par[0], par_y=par[1],
                       passes_per_launch=1 if form == "one-pass" else 0)
    dev = torch.from_numpy(frames).cuda()
    print(f"near-Earth search, {form}", flush=True)
    if form == "server":
        from fainttrail_amd.serve import Server
        srv = Server(slots=2)
        try:
            res = srv.wait(srv.submit(dev, mjd, prm))
        finally:
            srv.close()
        assert res.threshold_snr == want["threshold"] and res.threshold_method == want["method"]
        assert same_bits(res.null_peaks, want["pooled"])
        for f in ("x", "y", "vidx"):
            assert np.array_equal(getattr(res.selected, f), want["selected"][f]), f
        _same_candidates(res.candidates, want["candidates"])
        return
    n, h, w = frames.shape
    got = Searcher(n, w, h).run(dev, mjd, prm)
    torch.cuda.synchronize()
    assert got.n_distances == want["n_distances"] and got.dq_inv_au == want["dq_inv_au"]
    _same_as_reference(got, want)
    # the velocity grid times the distances: motion indices up to nv * nq
    assert int(got.best_v.max()) >= got.n_velocities


def _plateau_map(seed, h, w):
    """A max-S/N map on a coarse value grid (steps of 0.5): many equal neighbours and
    flat plateaus, so the one-peak-per-plateau rule decides most peaks."""
    rng = np.random.default_rng(seed)
    m = np.round(rng.normal(3.0, 2.0, (h, w)) * 2.0) / 2.0
    for _ in range(20):  # flat-topped plateaus of a few pixels
        y, x = rng.integers(0, h - 4), rng.integers(0, w - 4)
        m[y:y + rng.integers(1, 4), x:x + rng.integers(1, 4)] = 12.0
    return m.astype(np.float32), rng.integers(-1, 5000, (h, w)).astype(np.int32)


@pytest.mark.skipif(not po.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("radius,edge,floor,max_peaks", [(2, 4, 3.0, 100000), (3, 1, 5.0, 100000),
                                                         (2, 4, 3.0, 37), (5, 0, -1e30, 20000)])
def test_find_peaks_matches_reference(hip, radius, edge, floor, max_peaks):
    """The device peak finder (peaks_kernel + the host's row-major scan order and
    std::sort, ftk_find_peaks) equals the reference's find_peaks on maps with plateaus
    and ties: same peaks, same order, same truncation."""
    import torch
    from fainttrail_amd.search import find_peaks
    m, b = _plateau_map(radius * 100 + edge, 157, 203)
    want = po.ref_find_peaks(m, b, floor, radius, edge, max_peaks)
    assert len(want[0]) > 20
    got = find_peaks(torch.from_numpy(m).cuda(), torch.from_numpy(b).cuda(), floor, radius, edge, max_peaks)
    for a, e in zip((got.x, got.y, got.snr, got.vidx), want):
        assert np.array_equal(a, e)
    ora = po.find_peaks(m, b, floor, radius, edge, max_peaks)
    for a, e in zip(ora, want):
        assert np.array_equal(a, e)


def _same_result(a, b):
    assert a.n_velocities == b.n_velocities
    assert same_bits(a.max_snr.cpu().numpy(), b.max_snr.cpu().numpy())
    assert np.array_equal(a.best_v.cpu().numpy(), b.best_v.cpu().numpy())
    for f in ("x", "y", "vidx"):
        assert np.array_equal(getattr(a.peaks, f), getattr(b.peaks, f))
        assert np.array_equal(getattr(a.selected, f), getattr(b.selected, f))
    assert same_bits(a.peaks.snr, b.peaks.snr)
    assert same_bits(a.null_peaks, b.null_peaks)  # run after run, in order
    assert a.threshold_snr == b.threshold_snr and a.threshold_method == b.threshold_method


@pytest.mark.parametrize("group,segments,storage,tile_rows,queue",
                         [(0, 0, "f32", 0, 1), (2, 0, "f32", 0, -1), (3, 5, "f32", 0, 0),
                          (4, 1, "f32", 8, 1), (6, 64, "f32", 32, 1), (0, 0, "f32", 16, 0),
                          (0, 0, "f16", 0, -1), (4, 0, "bf16", 0, 0), (3, 0, "f16", 8, 1),
                          (0, 0, "bf16", 32, 1), (2, 0, "bf16", 16, 1)])
def test_batched_search_equals_one_pass_at_a_time(hip, group, segments, storage, tile_rows, queue):
    """The batched form (one plan whose per-frame rows are permuted for each null run,
    passes fused into one persistent launch per group (queue 1), or multi-pass launches
    of one workgroup per work item (queue 0), or the automatic choice (-1), stage
    builds on a second stream) gives the one-pass-at-a-
    time form's results bit for bit, for any group size, segment count and tile height:
    maps, peak lists, the pooled null peaks in run order and the threshold."""
    import torch
    from fainttrail_amd import synth
    from fainttrail_amd.search import SearchParams, Searcher
    fr, mjd, _ = synth.make_sequence(18, 200, 136, nstars=40, nmovers=3, mover_snr=(4.0, 8.0), seed=21)
    frames = fr.cuda()
    print(f"batched search group {group} segments {segments} storage {storage} tile_rows {tile_rows} queue {queue}",
          flush=True)
    s = Searcher(18, 200, 136)
    base = SearchParams(vmin_pxh=10.0, vmax_pxh=60.0, grid_tol_px=0.5, fwhm_px=3.0, null_runs=5, seed=4,
                        storage=storage)
    one = s.run(frames, mjd, SearchParams(**{**base.__dict__, "passes_per_launch": 1, "queue": 0}))
    many = s.run(frames, mjd, SearchParams(**{**base.__dict__, "passes_per_launch": group, "segments": segments,
                                              "tile_rows": tile_rows, "queue": queue}))
    torch.cuda.synchronize()
    assert one.passes_per_launch == 1 and one.queue == 0
    assert many.passes_per_launch == (group if group else 6)
    assert many.tile_rows == (tile_rows or 16) and (queue < 0 or many.queue == queue)
    assert one.null_peaks.size > 0 and one.peaks.x.size > 0
    _same_result(one, many)


def test_batched_search_equals_one_pass_full_size(hip):
    """The 2048^2 null-calibration shape: 40 frames of 2048 x 2048, the 7980-motion grid
    (grid_tol 0.212, 12-66 px/h), 16 null runs. The batched form (real pass first, null
    runs in fused groups) gives the one-pass-at-a-time form's results bit for bit."""
    import torch
    from fainttrail_amd import synth
    from fainttrail_amd.search import SearchParams, Searcher
    n, w, h = 40, 2048, 2048
    fr, mjd, _ = synth.make_sequence(n, w, h, nstars=300, nmovers=10, mover_snr=(0.7, 0.7), seed=7,
                                     device="cuda:0")
    print("full-size batched search", flush=True)
    s = Searcher(n, w, h)
    base = SearchParams(vmin_pxh=12.0, vmax_pxh=66.0, grid_tol_px=0.212, fwhm_px=3.0, null_runs=16, seed=7)
    one = s.run(fr, mjd, SearchParams(**{**base.__dict__, "passes_per_launch": 1}))
    many = s.run(fr, mjd, SearchParams(**{**base.__dict__, "passes_per_launch": 6}))
    torch.cuda.synchronize()
    assert one.passes_per_launch == 1 and many.passes_per_launch == 6
    assert one.null_peaks.size > 0 and one.peaks.x.size > 0
    _same_result(one, many)


def test_batched_search_fixed_threshold_and_tuner(hip):
    """No null runs (a fixed threshold): the batched form is the real pass alone. The
    autotuner's choice, used by later runs, leaves every result bit unchanged."""
    import torch
    from fainttrail_amd import synth
    from fainttrail_amd.search import SearchParams, Searcher
    fr, mjd, _ = synth.make_sequence(16, 176, 128, nstars=30, nmovers=2, mover_snr=(6.0, 9.0), seed=8)
    frames = fr.cuda()
    s = Searcher(16, 176, 128)
    fixed = SearchParams(vmin_pxh=10.0, vmax_pxh=60.0, grid_tol_px=0.5, fwhm_px=3.0, threshold_snr=6.0)
    a = s.run(frames, mjd, SearchParams(**{**fixed.__dict__, "passes_per_launch": 1}))
    b = s.run(frames, mjd, fixed)
    torch.cuda.synchronize()
    assert b.threshold_method == "fixed" and b.null_peaks.size == 0
    _same_result(a, b)
    nulls = SearchParams(vmin_pxh=10.0, vmax_pxh=60.0, grid_tol_px=0.5, fwhm_px=3.0, null_runs=4, seed=2)
    ref = s.run(frames, mjd, SearchParams(**{**nulls.__dict__, "passes_per_launch": 1}))
    g, seg, tr, q, secs = s.tune(frames, mjd, nulls, reps=1)
    print(f"tuned: {g} passes per launch, {seg} segments, tile rows {tr}, launch form {q}, "
          f"{secs * 1e3:.2f} ms per search", flush=True)
    assert 1 <= g <= 5 and seg >= 1 and tr in (0, 8, 16, 32) and q in (-1, 0, 1) and secs > 0
    tuned = s.run(frames, mjd, nulls)
    torch.cuda.synchronize()
    assert tuned.passes_per_launch == g and tuned.segments == seg
    assert (tr == 0 or tuned.tile_rows == tr) and (q < 0 or tuned.queue == q)
    _same_result(ref, tuned)


def test_sixteen_bit_search_agrees_with_exact_search(hip):
    """Stored-frame searches (fp16, bf16) against the exact fp32 search on a sequence
    with bright movers: the same selected peaks (position and motion), S/N of each
    within 1e-3 (fp16) / 1e-2 (bf16) of max(1, S/N), thresholds within 1 %. The
    candidates are refined on the exact fp32 frames, so the refined tracks, S/N and
    photometry equal the exact search's bit for bit; only the grid S/N differs. The
    null-tail fit of the 16-bit forms runs on the GPU: within 50 m eps relative of the
    host fit of the same pooled peaks, same method."""
    import torch
    from fainttrail_amd import synth
    from fainttrail_amd.search import SearchParams, Searcher
    fr, mjd, _ = synth.make_sequence(20, 224, 160, nstars=40, nmovers=4, mover_snr=(6.0, 9.0), seed=31)
    frames = fr.cuda()
    s = Searcher(20, 224, 160)
    prm = SearchParams(vmin_pxh=10.0, vmax_pxh=60.0, grid_tol_px=0.5, fwhm_px=3.0, null_runs=6, seed=8)
    ref = s.run(frames, mjd, prm)
    assert ref.selected.x.size >= 2
    for st, tol in (("f16", 1e-3), ("bf16", 1e-2)):
        got = s.run(frames, mjd, SearchParams(**{**prm.__dict__, "storage": st}))
        torch.cuda.synchronize()
        print(f"{st}: threshold {got.threshold_snr:.6f} vs {ref.threshold_snr:.6f}, "
              f"{got.selected.x.size} vs {ref.selected.x.size} selected", flush=True)
        assert abs(got.threshold_snr - ref.threshold_snr) <= 0.01 * ref.threshold_snr, st
        a = {(int(x), int(y), int(v)): float(q) for x, y, v, q in zip(ref.selected.x, ref.selected.y, ref.selected.vidx,
                                                                        ref.selected.snr)}
        b = {(int(x), int(y), int(v)): float(q) for x, y, v, q in zip(got.selected.x, got.selected.y, got.selected.vidx,
                                                                        got.selected.snr)}
        assert a.keys() == b.keys(), st
        for k in a:
            assert abs(a[k] - b[k]) <= tol * max(1.0, abs(a[k])), (st, k)
        err = (got.max_snr - ref.max_snr).abs() / ref.max_snr.abs().clamp(min=1.0)
        assert float(err.max()) <= 2 * tol, st
        assert len(got.candidates) == len(ref.candidates) >= 2, st
        for c, r in zip(got.candidates, ref.candidates):
            for f in ("x", "y", "vx_pxh", "vy_pxh", "snr", "snr_half1", "snr_half2", "flux", "mag_inst",
                      "max_frame_fraction", "pos_err_px", "vel_err_pxh"):
                assert getattr(c, f) == getattr(r, f), (st, c.id, f)
            assert abs(c.snr_coarse - r.snr_coarse) <= tol * max(1.0, abs(r.snr_coarse)), (st, c.id)
        host_t, host_m = s.refit(SearchParams(**{**prm.__dict__, "storage": st}), on_device=False)
        assert host_m == got.threshold_method, st
        print(f"{st}: GPU fit vs host fit {abs(host_t - got.threshold_snr) / abs(host_t):.2e} relative", flush=True)
        assert abs(host_t - got.threshold_snr) <= fit_tol(got.null_peaks, prm.null_runs) * abs(host_t), st
