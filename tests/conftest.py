import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sph-magnetic-particles-oldroyd-b-in-cuda_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (ROOT, PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


def _ensure_built():
    lib = os.path.join(PKG, "libsph_amd.so")
    if not os.path.exists(lib):
        import build  # the package's build script

        build.build()
    orc = os.path.join(ORACLE, "liboracle.so")
    if not os.path.exists(orc):
        import oracle

        oracle.build_oracle()


_ensure_built()


@pytest.fixture
def tmpdir_path(tmp_path):
    return str(tmp_path)
