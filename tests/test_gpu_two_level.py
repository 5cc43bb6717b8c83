"""GPU parity for sorts of two tile-local levels.

A sort of one range that plans two partition levels and a spare runs level 0
tile-locally in place in d and level 1 on gathered segments: every bucket of
level 0 is read back as the runs it left in level 0's tiles (no count pass,
no scan of per-tile counts).  The reference partitions every level over
contiguous ranges (CL12/QuicksortKernels.cl:102-240), so the output, not the
schedule, is what must match: every test compares with std::sort of the same
input (CL12/QuicksortMain.cpp:498-516), under the three schedules of
gqs_set_level_scheme: 0 (the default), 1 (a contiguous level before the
tile-local one) and 2 (as 0, reading runs through 64-entry run tables, the
path taken when runs are too short for one table).
"""
import numpy as np
import pytest

import gqsort
import orc
from conftest import gpu_sort

pytestmark = pytest.mark.gpu

SCHEMES = [0, 1, 2]


def check(got, exp, case):
    bad, first = orc.mismatch(got, exp)
    assert bad == 0, f"{case}: {bad} of {exp.size} keys differ from std::sort, first at {first}"


def sort_schemes(ctx, a, case, schemes=SCHEMES, **params):
    exp = orc.std_sort(a)
    try:
        for sc in schemes:
            ctx.set_level_scheme(sc)
            print(f"case {case} scheme={sc}", flush=True)
            check(gpu_sort(ctx, a, **params), exp, f"{case} scheme={sc}")
            s = ctx.stats()
            assert s.overflow == 0 and s.oversize_leaves == 0, f"{case} scheme={sc}"
    finally:
        ctx.set_level_scheme(0)
        if params:
            ctx.set_params(gqsort.default_params())


@pytest.mark.parametrize("gen", ["uniform", "gaussian", "sorted", "reverse", "equal", "bucket", "permutation"])
@pytest.mark.parametrize("n", [(1 << 22) + 12345, (1 << 25) + 777])
def test_u32_generators(gpu_ctx, gen, n):
    a = gqsort.generate(gqsort.U32, gen, 11, n)
    sort_schemes(gpu_ctx, a, f"u32 {gen} n={n}")


@pytest.mark.parametrize("dt,n", [(gqsort.F32, (1 << 23) + 3), (gqsort.F64, (1 << 22) + 5), (gqsort.F64, 1 << 24)])
def test_float_keys(gpu_ctx, dt, n):
    a = gqsort.generate(dt, "normal", 5, n)
    sort_schemes(gpu_ctx, a, f"dt={dt} normal n={n}")


@pytest.mark.parametrize("dt", [gqsort.U32, gqsort.F64])
@pytest.mark.parametrize("distinct", [3, 17, 1000])
def test_duplicate_heavy(gpu_ctx, dt, distinct):
    """Level 0's equality buckets become fill tasks that run after level 1 has
    read its runs out of d; the other buckets' keys pass through level 1."""
    n = (1 << 23) + 99
    rng = np.random.default_rng(100 + distinct)
    w = 1.0 / np.arange(1, distinct + 1) ** 1.2
    idx = rng.choice(distinct, size=n, p=w / w.sum())
    if dt == gqsort.U32:
        vals = rng.choice(np.iinfo(np.uint32).max, size=distinct, replace=False).astype(np.uint32)
    else:
        vals = rng.standard_normal(distinct) * 1000
    a = vals[idx].astype(gqsort.NP_DTYPES[dt])
    sort_schemes(gpu_ctx, a, f"duplicates dt={dt} distinct={distinct}")


def test_mostly_one_value(gpu_ctx):
    """99% of the keys equal one value: level 0 fills them, level 1 gets the
    few keys around them as small gathered segments with short runs."""
    n = 1 << 24
    rng = np.random.default_rng(9)
    a = np.full(n, 0x80000000, dtype=np.uint32)
    m = rng.random(n) < 0.01
    a[m] = rng.integers(0, 1 << 32, size=int(m.sum()), dtype=np.uint64).astype(np.uint32)
    sort_schemes(gpu_ctx, a, "99% one value")


@pytest.mark.parametrize("n", [600_001, 1_000_003])
def test_small_segments(gpu_ctx, n):
    """Parameters that plan two levels at under a million keys: level 1's
    segments are single partial tiles gathered from every level-0 tile."""
    a = gqsort.generate(gqsort.U32, "uniform", 3, n)
    sort_schemes(gpu_ctx, a, f"small segments n={n}", leaf_max=4096, leaf_target=2048, max_fanout=256,
                 max_levels=8, mode=0, host_driven=0)


def test_host_driven(gpu_ctx):
    """The per-level readback loop over the two tile-local levels."""
    a = gqsort.generate(gqsort.U32, "uniform", 4, (1 << 23) + 1)
    sort_schemes(gpu_ctx, a, "host driven", schemes=[0, 2], host_driven=1)


def test_full_size_u32(gpu_ctx):
    """2^28 uniform keys, the headline size, under schedules 0 and 2."""
    a = gqsort.generate(gqsort.U32, "uniform", 1, 1 << 28)
    sort_schemes(gpu_ctx, a, "2^28 uniform", schemes=[0, 2])
