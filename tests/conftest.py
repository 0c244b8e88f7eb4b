import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "ref: needs the reference's own CPU build (oracle/_ref)")


def _ensure_built():
    """Build the native libraries if nothing is built yet (build() compiles them)."""
    need = [os.path.join(ROOT, "ppc_amd", "lib", "libppc_hip.so"), os.path.join(ROOT, "ppc_amd", "lib", "libxppc.so"),
            os.path.join(ROOT, "ppc_amd", "bin", "ppc"), os.path.join(ROOT, "oracle", "liboracle.so")]
    if all(os.path.exists(p) for p in need):
        return
    import __graft_entry__

    __graft_entry__.build()


_ensure_built()


@pytest.fixture(scope="session")
def c1_dir():
    from ppc_amd import configs

    return configs.ensure("c1")


@pytest.fixture(scope="session")
def c2_dir():
    from ppc_amd import configs

    return configs.ensure("c2")
