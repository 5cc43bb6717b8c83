import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "gpu-quicksort_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


def _has_gpu():
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu_ctx():
    if not _has_gpu():
        pytest.fail("GPU test collected on a host without a GPU")
    import gqsort

    ctx = gqsort.Context(0)
    yield ctx
    ctx.close()


def gpu_sort(ctx, a, **params):
    """Sort host array a on the GPU through gqs_sort_device (device buffers
    from the library's own runtime); returns the downloaded result."""
    import gqsort

    if params:
        base = gqsort.default_params(gqsort.DTYPE_CODES[a.dtype])
        ctx.set_params(base, **params)
    return gqsort.device_sort(ctx, a)
