"""Autotuner: finds parameters, persists them, and sorts stay bit-exact with them."""
import os

import pytest

import gqsort
import orc
from conftest import gpu_sort

pytestmark = pytest.mark.gpu


def test_autotune_roundtrip(gpu_ctx, tmp_path):
    n = 1 << 20
    best = gpu_ctx.autotune(gqsort.U32, n, "uniform", reps=1)
    assert best.leaf_target <= best.leaf_max
    path = os.path.join(tmp_path, "tuned.json")
    gpu_ctx.tune_save(path)
    ctx2 = gqsort.Context(0)
    try:
        ctx2.tune_load(path)
        a = gqsort.generate(gqsort.U32, "bucket", 3, n + 7)
        got = gqsort.device_sort(ctx2, a)
        assert orc.mismatch(got, orc.std_sort(a))[0] == 0
    finally:
        ctx2.close()
    # explicit parameters take over again
    gpu_ctx.set_params(gqsort.default_params())


@pytest.mark.parametrize("thr,cut", [(4, 64), (16, 256), (64, 512), (4096, 512)])
def test_leaf_cutoffs(gpu_ctx, thr, cut):
    a = gqsort.generate(gqsort.U32, "uniform", 9, 300001)
    got = gpu_sort(gpu_ctx, a, sort_threshold=thr, bitonic_cutoff=cut)
    assert orc.mismatch(got, orc.std_sort(a))[0] == 0
    gpu_ctx.set_params(gqsort.default_params())
