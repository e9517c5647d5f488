import os
import sys

import pytest

# torch (plumbing for device buffers in a few tests) ships its own HIP runtime with the
# same soname as ROCm's; load it before the solver library so both use one runtime.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "ins-solver_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (MI355X)")


@pytest.fixture(scope="session")
def ip():
    import ins_poisson
    return ins_poisson


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    return O
