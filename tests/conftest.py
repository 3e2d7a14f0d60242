import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcindergs_hip.so)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    if not os.path.exists(oracle.LIB_PATH):
        from cindergs_amd.build import build_oracle

        build_oracle(always=False)
    oracle.lib()
    oracle.set_threads(min(16, os.cpu_count() or 1))
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """One GpuPipelineState for the GPU tests (the reference serialises its GPU
    tests on one session as well, gpu_parity_tests.rs:17-50)."""
    from cindergs_amd import GpuPipelineState

    st = GpuPipelineState(0)
    yield st
    st.close()
