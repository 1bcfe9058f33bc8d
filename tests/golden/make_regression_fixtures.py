"""Copy the reference's real-data regression fixtures into tests/golden/regression/.

The reference's own regression suite (tests/regression/ of FaintTrail: `.case` files,
16-bit FITS crops of three real sequences in data/, and the expected candidates in
baselines/) is data; the GPU tests run it with the reference's unchanged
`fainttrail regress --backend hip --strict`, which is how the reference validates a
backend (tests/regression/README.md). Runs where the reference sources are:

    python tests/golden/make_regression_fixtures.py [REFERENCE_ROOT]
"""
from __future__ import annotations

import shutil
import sys
from pathlib import Path

CASES = ["ceres-cut", "leo-triplet-cut", "mariekebaan-cut"]


def main() -> int:
    src = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "tests" / "regression"
    dst = Path(__file__).resolve().parent / "regression"
    if not src.is_dir():
        print(f"{src} not found", file=sys.stderr)
        return 1
    (dst / "baselines").mkdir(parents=True, exist_ok=True)
    (dst / "data").mkdir(exist_ok=True)
    for c in CASES:
        shutil.copyfile(src / f"{c}.case", dst / f"{c}.case")
        shutil.copyfile(src / "baselines" / f"{c}.csv", dst / "baselines" / f"{c}.csv")
        if (dst / "data" / c).exists():
            shutil.rmtree(dst / "data" / c)
        shutil.copytree(src / "data" / c, dst / "data" / c)
        for f in (dst / "data" / c).iterdir():
            f.chmod(0o644)
    print(f"copied {len(CASES)} cases into {dst}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
