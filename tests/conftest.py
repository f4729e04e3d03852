import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

QUERY_FIXTURES = ROOT / "tests" / "golden" / "queries"
FIXTURE_NAMES = ["_2_road2", "_3_road3", "_3_triangle", "_4_road4", "_4_roadY", "_4_square", "_5_plus", "_5_road5"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libforkmatch.so on the device)")


def gpu_available() -> bool:
    try:
        import forkmatch_amd as fm
        return fm.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def engine():
    import forkmatch_amd as fm
    if fm.device_count() == 0:
        pytest.fail("gpu test collected without a visible HIP device")
    e = fm.Engine(soln_prealloc=1 << 30, output_prealloc=1 << 30, bank_size=1 << 15, output_page_bytes=1 << 16)
    yield e
    e.close()


@pytest.fixture(scope="session")
def er_graph():
    """G(1000, 4000) undirected, edge labels uniform in 0..7 (the reference's
    CPU-runnable example size)."""
    import forkmatch_amd as fm
    return fm.Graph.erdos_renyi(1000, 4000, 8, seed=0xF0F1)


def fixture_path(name: str) -> Path:
    return QUERY_FIXTURES / f"{name}.g"
