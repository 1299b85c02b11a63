"""Test configuration: the `gpu` marker and import paths.

Tests marked `gpu` need an MI355X and call the product through its C-ABI;
everything else runs on the CPU (oracle checks, host shell, library loading).
"""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "how-mhd-kokkos_amd", ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); runs the HIP library")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device plumbing only)
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    if not gpu_available():
        pytest.fail("GPU test selected but no GPU is visible")
