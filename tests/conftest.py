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


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a host without a GPU")
    import gqsort

    ctx = gqsort.Context(0)
    yield ctx
    ctx.close()


_TORCH_VIEW = {np.dtype(np.uint32): np.int32, np.dtype(np.float32): np.float32, np.dtype(np.float64): np.float64}


def to_device(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a).view(_TORCH_VIEW[a.dtype])).cuda()


def from_device(t, dtype):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy().view(dtype)


def gpu_sort(ctx, a, **params):
    """Sort host array a on the GPU through gqs_sort_device; returns the result."""
    import torch
    import gqsort

    if params:
        base = gqsort.default_params(gqsort.DTYPE_CODES[a.dtype])
        ctx.set_params(base, **params)
    dt = gqsort.DTYPE_CODES[a.dtype]
    d = to_device(a)
    dn = torch.empty_like(d)
    ctx.sort_device(dt, d, dn, a.size, torch.cuda.current_stream())
    return from_device(d, a.dtype)
