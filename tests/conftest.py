import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device; runs through the C-ABI")


@pytest.fixture(scope="session")
def golden():
    with np.load(ROOT / "tests" / "golden" / "golden.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def hip():
    import fainttrail_amd as ft
    if ft.device_count() < 1:
        pytest.fail("no HIP device visible to a test marked gpu")
    return ft
