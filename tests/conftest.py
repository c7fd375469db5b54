import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lulesh-accelerate_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP library on the device)")
    config.addinivalue_line(
        "markers", "slow: longer CPU oracle runs")


@pytest.fixture(scope="session")
def oracle():
    """The C oracle, rebuilt by make when its sources are newer."""
    import lulesh_oracle as lo
    if os.path.exists(os.path.join(ORACLE, "lulesh_oracle.c")):
        try:
            lo.build()
        except (OSError, RuntimeError):
            if not os.path.exists(lo.LIB_PATH):
                raise
    lo.lib()
    return lo


@pytest.fixture(scope="session")
def reference_bins(oracle):
    """Binaries built from the reference sources; skipped where the
    reference tree (and so oracle/_ref) is absent."""
    need = ["lulesh_omp_dump", "lulesh2_dump", "lulesh_omp", "lulesh2"]
    if not all(os.path.exists(os.path.join(oracle.REF_DIR, n)) for n in need):
        src = "/root/reference/reference/C/LULESH-OMP.cc"
        if os.path.exists(src):
            oracle.build()
    if not all(os.path.exists(os.path.join(oracle.REF_DIR, n)) for n in need):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return oracle.REF_DIR


@pytest.fixture(scope="session")
def lulesh_gpu():
    """The HIP library; a GPU test fails (not skips) when it is missing."""
    import lulesh_amd
    lulesh_amd.lib()
    return lulesh_amd
