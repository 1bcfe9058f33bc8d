"""The reference's own command line, unchanged, searching with the "hip" backend.

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
