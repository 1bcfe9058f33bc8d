"""Serving form (ftk_server_*): jobs of different frame sizes, frame counts and velocity
grids submitted together or while others run overlap on the device, and each job's
result (peaks, selected candidates, pooled null peaks in run order, threshold) is
bit-identical to the same search run alone; a tight memory budget serialises
admission without changing any result; jobs of the same frame size are stacked in
shared launches; a small job overtakes a large one that does not fit yet; a reused searcher is reserved at the bytes it holds; a job larger than the
budget is refused at submission."""
import numpy as np
import pytest

from refcases import same_bits

pytestmark = pytest.mark.gpu

SPECS = [  # (side_x, side_y, frames, nulls, grid_tol, seed)
    (160, 120, 12, 3, 0.6, 1),
    (224, 176, 16, 4, 0.5, 2),
    (96, 96, 10, 2, 0.4, 3),
    (200, 136, 18, 5, 0.5, 4),
    (128, 200, 14, 3, 0.7, 5),
]


def _inputs():
    from fainttrail_amd import synth
    from fainttrail_amd.search import SearchParams
    jobs = []
    for w, h, n, nulls, tol, seed in SPECS:
        fr, mjd, _ = synth.make_sequence(n, w, h, nstars=30, nmovers=2, mover_snr=(4.0, 8.0), seed=seed,
                                         device="cuda:0")
        prm = SearchParams(vmin_pxh=10.0, vmax_pxh=60.0, grid_tol_px=tol, fwhm_px=3.0, null_runs=nulls, seed=seed)
        jobs.append((fr.contiguous(), mjd, prm))
    return jobs


def _same(alone, mixed):
    assert alone.n_velocities == mixed.n_velocities
    assert alone.threshold_snr == mixed.threshold_snr and alone.threshold_method == mixed.threshold_method
    for a, b in ((alone.peaks, mixed.peaks), (alone.selected, mixed.selected)):
        assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y) and np.array_equal(a.vidx, b.vidx)
        assert same_bits(a.snr, b.snr)
    assert same_bits(alone.null_peaks, mixed.null_peaks)
    # the candidate stage: the same candidates, every field and flag
    assert alone.candidates == mixed.candidates


def test_jobs_alone_and_in_the_mix(hip):
    import torch
    from fainttrail_amd.search import Searcher
    from fainttrail_amd.serve import Server
    jobs = _inputs()
    alone = []
    for fr, mjd, prm in jobs:
        n, h, w = fr.shape
        alone.append(Searcher(n, w, h).run(fr, mjd, prm))
    torch.cuda.synchronize()
    srv = Server(slots=3)
    ids = [srv.submit(fr, mjd, prm) for fr, mjd, prm in jobs]
    got = [srv.wait(j) for j in ids]
    st = srv.stats()
    print("server stats:", st, flush=True)
    srv.close()
    for a, b in zip(alone, got):
        _same(a, b)
        assert b.t_submit <= b.t_admit <= b.t_end
    assert sum(len(a.candidates) for a in alone) >= len(jobs)  # bright movers: the lists are not empty
    assert st["completed"] == len(jobs) and st["failed"] == 0
    assert 1 <= st["max_running"] <= 3
    assert st["peak_reserved_bytes"] <= st["budget_bytes"]


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_same_geometry_jobs_fused_in_one_stacking_launch(hip, storage):
    """Jobs of one frame size and velocity grid (different frames, seeds and null-run
    counts) submitted together: their persistent stacking launches meet and are stacked
    in shared launches (the server's fused-launch count), and every job's result is
    still the same as run alone."""
    import torch
    from fainttrail_amd import synth
    from fainttrail_amd.search import SearchParams, Searcher
    from fainttrail_amd.serve import Server
    jobs = []
    for seed, nulls in ((11, 6), (12, 5), (13, 7), (14, 6)):
        fr, mjd, _ = synth.make_sequence(16, 256, 192, nstars=30, nmovers=2, mover_snr=(4.0, 8.0), seed=seed,
                                         device="cuda:0")
        prm = SearchParams(vmin_pxh=10.0, vmax_pxh=60.0, grid_tol_px=0.5, fwhm_px=3.0, null_runs=nulls, seed=seed,
                           passes_per_launch=3, queue=1, storage=storage)
        jobs.append((fr.contiguous(), mjd, prm))
    alone = [Searcher(16, 256, 192).run(fr, mjd, prm) for fr, mjd, prm in jobs]
    torch.cuda.synchronize()
    srv = Server(slots=4)
    srv.fusion(True, window_ms=20.0)
    ids = [srv.submit(fr, mjd, prm) for fr, mjd, prm in jobs]
    got = [srv.wait(j) for j in ids]
    st = srv.stats()
    print(f"fusion {storage}: server stats {st}", flush=True)
    srv.close()
    for a, b in zip(alone, got):
        _same(a, b)
    assert st["completed"] == len(jobs) and st["failed"] == 0
    assert st["fused_launches"] >= 1 and st["fused_passes"] >= 2
    assert st["stack_launches"] >= st["fused_launches"]


def test_tight_budget_serialises_without_changing_results(hip):
    from fainttrail_amd import FtkError
    from fainttrail_amd.search import Searcher
    from fainttrail_amd.serve import Server
    jobs = _inputs()[:3]
    alone = []
    for fr, mjd, prm in jobs:
        n, h, w = fr.shape
        alone.append(Searcher(n, w, h).run(fr, mjd, prm))
    # a budget that holds the largest job but not two of them at once
    probe = Server(slots=3)
    sizes = []
    for fr, mjd, prm in jobs:
        j = probe.submit(fr, mjd, prm)
        sizes.append(probe.wait(j).reserved_bytes)
    probe.close()
    budget_mb = 1.2 * max(sizes) / 1048576.0
    srv = Server(budget_mb=budget_mb, slots=3)
    ids = [srv.submit(fr, mjd, prm) for fr, mjd, prm in jobs]
    got = [srv.wait(j) for j in ids]
    st = srv.stats()
    print("tight budget stats:", st, flush=True)
    for a, b in zip(alone, got):
        _same(a, b)
    assert st["peak_reserved_bytes"] <= st["budget_bytes"]
    assert sum(sizes) > st["budget_bytes"] and st["max_running"] < len(jobs)
    for g in got:  # the reservation covers what the job's searcher holds
        assert 0 < g.held_bytes <= g.reserved_bytes
    # a job larger than the whole budget is refused when submitted
    tiny = Server(budget_mb=1.0, slots=1)
    fr, mjd, prm = jobs[0]
    with pytest.raises(FtkError):
        tiny.submit(fr, mjd, prm)
    tiny.close()
    srv.close()


def test_job_with_registration_offsets(hip):
    """A job's registration offsets (the detector-fixed test of the null screen) are
    copied at submission: the caller's arrays may change or go away while the job
    queues, and its result is still the one of the same search run alone."""
    import torch
    from fainttrail_amd.search import SearchParams, Searcher
    from fainttrail_amd.serve import Server
    from fainttrail_amd import synth
    n, w, h = 14, 160, 120
    fr, mjd, _ = synth.make_sequence(n, w, h, nstars=30, nmovers=2, mover_snr=(4.0, 8.0), seed=9, device="cuda:0")
    fr = fr.contiguous()
    rx = 0.3 * np.arange(n) - 2.0
    ry = -0.2 * np.arange(n) + 1.0
    prm = SearchParams(vmin_pxh=10.0, vmax_pxh=60.0, grid_tol_px=0.5, fwhm_px=3.0, null_runs=3, seed=9,
                       reg_dx=rx.copy(), reg_dy=ry.copy())
    alone = Searcher(n, w, h).run(fr, mjd, prm)
    no_reg = Searcher(n, w, h).run(fr, mjd, SearchParams(**{**prm.__dict__, "reg_dx": None, "reg_dy": None}))
    torch.cuda.synchronize()
    srv = Server(slots=2)
    ids = [srv.submit(fr, mjd, SearchParams(**{**prm.__dict__, "reg_dx": rx, "reg_dy": ry})) for _ in range(2)]
    rx[:] = 99.0  # the jobs hold their own copies
    ry[:] = -99.0
    got = [srv.wait(j) for j in ids]
    srv.close()
    for g in got:
        _same(alone, g)
    print(f"null screen with offsets: {alone.null_artefacts} artefacts, without: {no_reg.null_artefacts}", flush=True)
    assert alone.null_artefacts >= no_reg.null_artefacts  # the detector test only adds rejections


def test_small_job_overtakes_a_blocked_head(hip):
    """Under a budget that holds one large job and one small one but not two large
    ones, a small job queued behind a large job that does not fit yet is admitted
    first (bounded overtaking); every result equals the same search alone."""
    import time
    from fainttrail_amd import synth
    from fainttrail_amd.search import SearchParams, Searcher
    from fainttrail_amd.serve import Server

    def job(w, h, n, nulls, seed):
        fr, mjd, _ = synth.make_sequence(n, w, h, nstars=30, nmovers=2, mover_snr=(4.0, 8.0), seed=seed,
                                         device="cuda:0")
        return fr.contiguous(), mjd, SearchParams(vmin_pxh=10.0, vmax_pxh=60.0, grid_tol_px=0.3, fwhm_px=3.0,
                                                  null_runs=nulls, seed=seed)

    # the large jobs run for a few tenths of a second: the others arrive while the first runs
    big_a, big_b, small = job(1024, 1024, 32, 8, 11), job(1024, 1024, 32, 8, 12), job(96, 96, 10, 2, 13)
    alone = [Searcher(fr.shape[0], fr.shape[2], fr.shape[1]).run(fr, mjd, prm) for fr, mjd, prm in (big_a, big_b, small)]
    probe = Server(slots=1)
    sizes = []
    for fr, mjd, prm in (big_a, small):
        sizes.append(probe.wait(probe.submit(fr, mjd, prm)).reserved_bytes)
    probe.close()
    budget_mb = (sizes[0] + 1.5 * sizes[1]) / 1048576.0
    assert 2 * sizes[0] > budget_mb * 1048576.0
    srv = Server(budget_mb=budget_mb, slots=3)
    ia = srv.submit(*big_a)
    time.sleep(0.02)  # admitted at once
    ib = srv.submit(*big_b)
    ic = srv.submit(*small)
    ga, gb, gc = srv.wait(ia), srv.wait(ib), srv.wait(ic)
    st = srv.stats()
    srv.close()
    print(f"overtaking: a {ga.t_admit:.3f}-{ga.t_end:.3f}, b {gb.t_admit:.3f}, c {gc.t_admit:.3f}-{gc.t_end:.3f}",
          flush=True)
    for a, g in zip(alone, (ga, gb, gc)):
        _same(a, g)
    assert gc.t_submit < ga.t_end  # the test's premise: the others arrived while the first ran
    assert gc.t_admit < gb.t_admit  # the small job went past the blocked one
    assert gb.t_admit >= ga.t_end - 1e-3  # the second large job waited for the first
    assert st["peak_reserved_bytes"] <= st["budget_bytes"]


def test_reused_searcher_is_reserved_at_what_it_holds(hip):
    """A large job and then a small one of the same frame shape (n, w, h) under a tight
    budget with one slot: the small job reuses the large job's searcher, whose buffers
    only grow, so its reservation is the bytes that searcher holds, not its own smaller
    estimate; the budget is never exceeded and the result equals the search alone."""
    from fainttrail_amd import synth
    from fainttrail_amd.search import SearchParams, Searcher
    from fainttrail_amd.serve import Server
    n, w, h = 16, 320, 256
    fr, mjd, _ = synth.make_sequence(n, w, h, nstars=30, nmovers=2, mover_snr=(4.0, 8.0), seed=17, device="cuda:0")
    fr = fr.contiguous()
    large = SearchParams(vmin_pxh=10.0, vmax_pxh=60.0, grid_tol_px=0.2, fwhm_px=3.0, null_runs=8, seed=3)
    small = SearchParams(vmin_pxh=20.0, vmax_pxh=30.0, grid_tol_px=0.8, fwhm_px=3.0, null_runs=1, seed=3)
    want = Searcher(n, w, h).run(fr, mjd, small)
    srv = Server(slots=1)
    a = srv.wait(srv.submit(fr, mjd, large))
    b = srv.wait(srv.submit(fr, mjd, small))
    st = srv.stats()
    srv.close()
    print(f"reuse: large reserved {a.reserved_bytes / 1e6:.1f} MB holds {a.held_bytes / 1e6:.1f} MB; "
          f"small reserved {b.reserved_bytes / 1e6:.1f} MB holds {b.held_bytes / 1e6:.1f} MB", flush=True)
    assert a.held_bytes <= a.reserved_bytes
    assert b.reserved_bytes >= a.held_bytes  # reserved at what the reused searcher holds
    assert b.held_bytes <= b.reserved_bytes
    assert st["idle_searcher_bytes"] == pytest.approx(b.held_bytes)
    assert st["peak_reserved_bytes"] <= st["budget_bytes"]
    _same(want, b)


def test_staggered_stream_from_the_serving_distribution(hip):
    """Jobs drawn from the serving workload's seeded Poisson stream (side 256-4096,
    10-120 frames, 500-50k motions, 4-64 null runs, 0-15 movers; the first six whose
    search is at most 3e12 stacked units, to keep the test short) are submitted one by
    one while earlier ones run; each result equals the same job run alone."""
    import time
    from fainttrail_amd.search import Searcher
    from fainttrail_amd.serve import Server, job_inputs, poisson_jobs
    specs = [j for j in poisson_jobs(400, 20.0, 5)
             if j.side * j.side * j.frames * j.velocities * (1 + j.nulls) <= 3e12][:6]
    assert len(specs) == 6
    inputs = [job_inputs(sp) for sp in specs]
    alone = []
    for fr, mjd, prm in inputs:
        n, h, w = fr.shape
        alone.append(Searcher(n, w, h).run(fr, mjd, prm))
    srv = Server(slots=4)
    ids = []
    for k, (fr, mjd, prm) in enumerate(inputs):
        if k:
            time.sleep(0.1)  # later jobs arrive while earlier ones run
        ids.append(srv.submit(fr, mjd, prm))
    got = [srv.wait(j) for j in ids]
    st = srv.stats()
    srv.close()
    for sp, g in zip(specs, got):
        print(f"job {sp.side}^2 x {sp.frames} x {sp.velocities} x {sp.nulls} nulls: submitted {g.t_submit:.3f}, "
              f"admitted {g.t_admit:.3f}, ended {g.t_end:.3f}", flush=True)
    for a, g in zip(alone, got):
        _same(a, g)
        assert g.held_bytes <= g.reserved_bytes
    # at least one job arrived while an earlier one was on the device
    assert any(any(e.t_admit <= g.t_submit < e.t_end for e in got[:k]) for k, g in enumerate(got))
    assert st["max_running"] >= 2 and st["failed"] == 0
