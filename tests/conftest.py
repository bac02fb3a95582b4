"""Test configuration: import paths, the ``gpu`` marker and shared fixtures.

The oracle under ``oracle/`` is the checker; tests are the only product-side
code allowed to import it."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cutile-stencil-dsl_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP runtime")


@pytest.fixture(autouse=True)
def _require_gpu(request):
    """GPU tests fail (not skip) without a device: a missing GPU or runtime
    must never look like a pass."""
    if request.node.get_closest_marker("gpu") is not None:
        import torch

        from cutile.runtime import rt

        assert rt.available(), "libstencil_rt.so is not built"
        assert torch.cuda.is_available(), "no ROCm device visible"
    yield


@pytest.fixture(scope="session")
def oracle():
    import oracle as o

    return o
