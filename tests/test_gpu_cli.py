"""The reference's own command line and C API, unchanged, searching with the "hip" backend.

`oracle/_ref/fainttrail` is the reference CLI (cli/main.cpp and src/*.cpp compiled where
they lie by `make -C oracle cli`) linked with the "hip" SearchBackend adapter
(fainttrail_amd/integration/backend_hip.cpp) over libftk_hip.so. These tests simulate a
FITS sequence with the reference's own simulator (src/simulate.cpp), run
`fainttrail search` once with `--backend cpu` (the reference's CpuBackend) and once with
`--backend hip`, and require the candidates CSV and the max-S/N map FITS to be the same
bytes. That is the whole host pipeline of the reference (FITS reading, registration
check, PSF estimate, build_stage, null calibration, refine, candidate formatting,
src/search.cpp:716-809 and io.cpp:82-115) over this backend's stacking passes.
"""
import ctypes as C
import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "oracle" / "_ref" / "fainttrail"

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not CLI.exists(), reason="reference CLI not built (make -C oracle cli)")]


def _run(args, tmp, timeout=300):
    env = dict(os.environ)
    env.pop("FAINTTRAIL_HIP_OPTIONS", None)
    r = subprocess.run([str(CLI)] + [str(a) for a in args], cwd=tmp, env=env, capture_output=True, text=True,
                       timeout=timeout)
    assert r.returncode == 0, f"fainttrail {' '.join(map(str, args))} failed:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
    return r.stdout


@pytest.fixture(scope="module")
def sequence(tmp_path_factory):
    """20 frames of 512 x 512 from the reference simulator, 3 movers, seed 7."""
    tmp = tmp_path_factory.mktemp("cli")
    _run(["simulate", "-o", tmp / "seq", "--width", 512, "--height", 512, "--frames", 20, "--movers", 3,
          "--seed", 7], tmp)
    return tmp


def test_backends_lists_hip(hip, sequence):
    out = _run(["backends"], sequence)
    assert "hip" in out


@pytest.mark.parametrize("name,extra", [
    ("default", []),
    # the 512^2 benchmark search: 0.5 px total drift over the sequence, 8 null runs
    ("fine_grid_8_nulls", ["--grid-tol", "0.125", "--null-runs", "8"]),
    ("fixed_threshold", ["--threshold", "5.5", "--null-runs", "0"]),
    # the reference's tiled search (uneven tiles: 512 = 2 x 192 + 128)
    ("tiled_192", ["--tile", "192"]),
    # near-Earth search: curved tracks from the diurnal parallax at 6 trial distances,
    # explicit shift tables through the backend (parallax.cpp, search.cpp:73-97)
    ("neo_curved", ["--neo", "0.0002", "--site", "50.8,4.4,100"]),
])
def test_search_csv_and_map_identical(hip, sequence, name, extra):
    tmp = sequence
    outs = {}
    for backend in ("cpu", "hip"):
        csv = tmp / f"{name}_{backend}.csv"
        fits = tmp / f"{name}_{backend}_snr.fits"
        _run(["search", tmp / "seq", "--backend", backend, "-o", csv, "--snr-map", fits, "--quiet",
              "--photometry", "off"] + extra, tmp)
        outs[backend] = (csv.read_bytes(), fits.read_bytes() if fits.exists() else None)
    cpu, gpu = outs["cpu"], outs["hip"]
    assert cpu[0].count(b"\n") >= 3, "the search wrote no candidates CSV"
    assert gpu[0] == cpu[0], f"{name}: candidates CSV differs between --backend hip and --backend cpu"
    assert gpu[1] == cpu[1], f"{name}: max-S/N map differs between --backend hip and --backend cpu"


CLI_HIP = ROOT / "oracle" / "_ref" / "fainttrail-hip"


def _run_hip_cli(args, tmp, timeout=300):
    env = dict(os.environ)
    env.pop("FAINTTRAIL_HIP_OPTIONS", None)
    r = subprocess.run([str(CLI_HIP)] + [str(a) for a in args], cwd=tmp, env=env, capture_output=True, text=True,
                       timeout=timeout)
    assert r.returncode == 0, f"fainttrail-hip {' '.join(map(str, args))} failed:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
    return r.stdout


@pytest.mark.skipif(not CLI_HIP.exists(), reason="reference CLI with the built-in hip backend not built (make -C oracle cli-hip)")
def test_builtin_hip_backend_auto_options_and_four_checks(hip, sequence):
    """The reference built with the "hip" backend in, as it builds CUDA in
    (oracle/hip_integration.sh: FAINTTRAIL_HAVE_HIP): `backends` lists hip first, "auto"
    picks it, "hip:key=value" options reach it, and check-backend runs all four checks,
    the row-band one included, bit-identical; a search with --backend auto writes the
    CPU backend's candidates CSV and map."""
    out = _run_hip_cli(["backends"], sequence)
    print(out, flush=True)
    assert out.split()[0] == "hip"
    for name in ("auto", "hip", "hip:max_mb=64,launch_adds=1e9"):
        out = _run_hip_cli(["check-backend", "--backend", name], sequence)
        print(out, flush=True)
        assert "checking backend 'hip'" in out
        assert "bit-identical to the reference in 4 of 4 checks" in out
    csv_cpu, csv_auto = sequence / "auto_cpu.csv", sequence / "auto_auto.csv"
    for backend, csv in (("cpu", csv_cpu), ("auto", csv_auto)):
        _run_hip_cli(["search", sequence / "seq", "--backend", backend, "-o", csv, "--snr-map",
                      sequence / f"auto_{backend}_snr.fits", "--quiet", "--photometry", "off"], sequence)
    assert csv_auto.read_bytes() == csv_cpu.read_bytes()
    assert (sequence / "auto_auto_snr.fits").read_bytes() == (sequence / "auto_cpu_snr.fits").read_bytes()


def test_reference_regression_cases_strict(hip, tmp_path):
    """The reference's real-data regression suite (three registered sequences: Ceres,
    the Leo triplet with (21996) 1999 XP31, (12631) Mariekebaan; fixtures copied by
    tests/golden/make_regression_fixtures.py) through `fainttrail regress --backend hip
    --strict`: every known asteroid found, nothing else passing, thresholds 6.419 /
    6.399 / 7.599 and the candidate lists equal to the reference's baselines
    (tests/regression/README.md: "same cases, another backend")."""
    cases = sorted((ROOT / "tests" / "golden" / "regression").glob("*.case"))
    assert len(cases) == 3
    out = _run(["regress"] + cases + ["--backend", "hip", "--strict"], tmp_path)
    assert "backend hip" in out
    assert "regression: 3 passed, 0 changed, 0 failed, 0 skipped" in out, out[-3000:]


# --------------------------------------------------------------------- the C API (ft_*)
# include/fainttrail/fainttrail.h of the reference: ft_params and ft_candidate as declared
# there; oracle/_ref/libftref.so is the reference library (src/*.cpp, c_api.cpp among
# them) with the "hip" adapter registered.

class _FtParams(C.Structure):
    _fields_ = [(name, C.c_double) for name in (
        "vmin_pxh", "vmax_pxh", "grid_tol_px", "fwhm_px", "clip_sigma", "cr_sigma", "mask_scatter",
        "weight_scatter", "min_coverage", "threshold_snr", "min_snr_floor", "target_false_per_run",
        "max_frame_fraction", "half_tolerance")] + [
        ("static_exclude", C.c_int), ("null_runs", C.c_int), ("max_candidates", C.c_int), ("edge_px", C.c_int),
        ("threads", C.c_int), ("seed", C.c_ulonglong), ("backend", C.c_char * 16)]


class _FtCandidate(C.Structure):
    _fields_ = [("id", C.c_int), ("ok", C.c_int)] + [(name, C.c_double) for name in (
        "x", "y", "vx_pxh", "vy_pxh", "pos_err_px", "vel_err_pxh", "snr", "snr_coarse", "snr_half1", "snr_half2",
        "max_frame_fraction", "false_alarms_per_run", "flux", "mag_inst", "ra_deg", "dec_deg", "rate_arcsec_min",
        "pa_deg")] + [("flags", C.c_char * 64)]


def _c_api_search(lib, fits_files, backend, null_runs):
    ctx = lib.ft_create()
    assert ctx
    try:
        for f in fits_files:
            assert lib.ft_add_fits(ctx, str(f).encode()) == 0, lib.ft_last_error(ctx)
        prm = _FtParams()
        lib.ft_params_default(C.byref(prm))
        prm.backend = backend.encode()
        prm.null_runs = null_runs
        st = lib.ft_search(ctx, C.byref(prm), None, None)
        assert st == 0, lib.ft_last_error(ctx)
        cands = []
        for i in range(lib.ft_candidate_count(ctx)):
            c = _FtCandidate()
            assert lib.ft_get_candidate(ctx, i, C.byref(c)) == 0
            cands.append(bytes(c))
        return lib.ft_threshold(ctx), lib.ft_reference_mjd(ctx), cands
    finally:
        lib.ft_destroy(ctx)


@pytest.mark.skipif(not (ROOT / "oracle" / "_ref" / "libftref.so").exists(), reason="reference library not built")
@pytest.mark.parametrize("null_runs", [3, 8])
def test_c_api_search_identical(hip, sequence, null_runs):
    lib = C.CDLL(str(ROOT / "oracle" / "_ref" / "libftref.so"))
    lib.ft_create.restype = C.c_void_p
    lib.ft_destroy.argtypes = [C.c_void_p]
    lib.ft_add_fits.argtypes = [C.c_void_p, C.c_char_p]
    lib.ft_params_default.argtypes = [C.POINTER(_FtParams)]
    lib.ft_search.argtypes = [C.c_void_p, C.POINTER(_FtParams), C.c_void_p, C.c_void_p]
    lib.ft_candidate_count.argtypes = [C.c_void_p]
    lib.ft_get_candidate.argtypes = [C.c_void_p, C.c_int, C.POINTER(_FtCandidate)]
    lib.ft_threshold.argtypes = [C.c_void_p]
    lib.ft_threshold.restype = C.c_double
    lib.ft_reference_mjd.argtypes = [C.c_void_p]
    lib.ft_reference_mjd.restype = C.c_double
    lib.ft_last_error.argtypes = [C.c_void_p]
    lib.ft_last_error.restype = C.c_char_p
    files = sorted((sequence / "seq").glob("*.fits"))
    assert len(files) == 20
    thr_c, mjd_c, cands_c = _c_api_search(lib, files, "cpu", null_runs)
    thr_h, mjd_h, cands_h = _c_api_search(lib, files, "hip", null_runs)
    assert len(cands_c) >= 1, "the search found no candidate"
    assert thr_h == thr_c and mjd_h == mjd_c
    assert cands_h == cands_c, "ft_search candidates differ between backend hip and backend cpu"
