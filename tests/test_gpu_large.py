"""Sorts beyond one run's 32-bit sizes: runs sorted through the partition
levels, then merged pairwise with 64-bit indices (Merge Path).

The reference caps n at 32 bits (CL12/QuicksortMain.cpp:405).  Its own check,
std::sort of the same input (CL12/QuicksortMain.cpp:498-516), is the oracle
wherever the host can afford it: the run length is lowered with
gqs_set_large_run so that every shape of the merge schedule (ragged last run,
a lone run copied across a round, an odd number of rounds ending with the
copy back to d, ties across run boundaries) runs at small n.  At n > 2^32 the
result is checked by size-independent properties: sortedness, and
order-independent checksums of the key multiset before and after.
"""
import numpy as np
import pytest

import gqsort
import orc
from conftest import gpu_sort

pytestmark = pytest.mark.gpu

CASES = [
    # dtype, generator, n, run length
    (gqsort.U32, "uniform", 1_000_003, 1 << 16),      # 16 runs (the last ragged), 4 merge rounds
    (gqsort.U32, "uniform", 300_001, 100_000),        # 4 runs, the last of one key
    (gqsort.U32, "gaussian", 500_000, 200_000),       # 3 runs: the third is copied across in the first merge round
    (gqsort.U32, "equal", 1 << 20, 1 << 16),          # every comparison a tie
    (gqsort.U32, "sorted", 1 << 20, 3 << 15),
    (gqsort.U32, "reverse", 1 << 20, 3 << 15),
    (gqsort.U32, "bucket", (1 << 20) + 3, 1 << 17),
    (gqsort.U32, "permutation", 1001, 2),             # runs of two keys: 9 merge rounds of tiny merges
    (gqsort.F32, "normal", (1 << 20) + 77, 1 << 18),  # 5 runs, 3 merge rounds: copied back to d
    (gqsort.F64, "normal", 700_001, 123_457),         # 6 runs, 3 rounds
    (gqsort.F64, "ties", 262_144, 65_536),            # 4 runs, 2 rounds, 80 distinct values
]


def check(got, inp, case):
    exp = orc.std_sort(inp)
    bad, first = orc.mismatch(got, exp)
    assert bad == 0, f"{case}: {bad} of {inp.size} keys differ from std::sort, first at {first}"


@pytest.mark.parametrize("dt,gen,n,run", CASES)
def test_run_and_merge_schedule(gpu_ctx, dt, gen, n, run):
    if gen == "ties":
        a = np.random.default_rng(41).integers(-40, 40, size=n).astype(gqsort.NP_DTYPES[dt])
    else:
        a = gqsort.generate(dt, gen, 41, n)
    print(f"case run+merge dt={dt} {gen} n={n} run={run}", flush=True)
    gpu_ctx.set_large_run(run)
    try:
        got = gpu_sort(gpu_ctx, a)
    finally:
        gpu_ctx.set_large_run(0)
    check(got, a, f"run+merge dt={dt} {gen} n={n} run={run}")


def test_run_and_merge_float_specials(gpu_ctx):
    """Merges compare floats through the sort's key order: infinities and
    denormals in place, ties split across runs.  (No -0.0: std::sort leaves
    -0.0 and +0.0 in either order, so bitwise parity is undefined there.)"""
    vals = np.array([np.inf, -np.inf, 0.0, 1.0, -1.0, np.finfo(np.float32).max, -np.finfo(np.float32).max,
                     np.finfo(np.float32).tiny, -np.finfo(np.float32).tiny, 1e-45, -1e-45], dtype=np.float32)
    a = np.random.default_rng(9).choice(vals, size=50_000).astype(np.float32)
    gpu_ctx.set_large_run(4096)
    try:
        got = gpu_sort(gpu_ctx, a)
    finally:
        gpu_ctx.set_large_run(0)
    assert np.array_equal(got.view(np.uint32), orc.std_sort(a).view(np.uint32))


def test_run_and_merge_host_pointers(gpu_ctx):
    """GPUQSort with host pointers takes the same path (staged copies in, runs,
    merges, staged copies out)."""
    n = 3 * (1 << 20) + 5
    d = gqsort.generate(gqsort.U32, "uniform", 43, n)
    inp = d.copy()
    gpu_ctx.set_large_run(1 << 20)
    try:
        gqsort.GPUQSort(gpu_ctx, n, d, np.empty_like(d))
    finally:
        gpu_ctx.set_large_run(0)
    check(d, inp, "run+merge host pointers")


def test_large_run_bounds(gpu_ctx):
    with pytest.raises(gqsort.GQSError):
        gpu_ctx.set_large_run(1)
    with pytest.raises(gqsort.GQSError):
        gpu_ctx.set_large_run((1 << 32) - (1 << 16))
    gpu_ctx.set_large_run(0)


def multiset_sums(a, chunk=1 << 27):
    """Order-independent checksums of a uint32 array (mod 2^64): the sum of
    the keys, of their squares and of a 64-bit mix of each key."""
    m = (1 << 64) - 1
    s1 = s2 = s3 = 0
    for i in range(0, a.size, chunk):
        x = a[i:i + chunk].astype(np.uint64)
        s1 = (s1 + int(x.sum(dtype=np.uint64))) & m
        s2 = (s2 + int((x * x).sum(dtype=np.uint64))) & m
        x ^= x >> np.uint64(29)
        x *= np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(32)
        s3 = (s3 + int(x.sum(dtype=np.uint64))) & m
    return s1, s2, s3


def is_sorted(a, chunk=1 << 28):
    for i in range(0, a.size - 1, chunk):
        j = min(i + chunk + 1, a.size)
        if not np.all(a[i + 1:j] >= a[i:j - 1]):
            return False
    return True


def test_n_above_2p32_u32(gpu_ctx):
    """2^32 + 4099 uint32 keys (16 GiB) through the reference's host-pointer
    entry point: three runs of at most 2^31 keys and two merge rounds."""
    n = (1 << 32) + 4099
    print("case 2^32+4099 u32: generating", flush=True)
    a = np.random.default_rng(2024).integers(0, 1 << 32, size=n, dtype=np.uint32)
    before = multiset_sums(a)
    print("case 2^32+4099 u32: sorting", flush=True)
    gqsort.GPUQSort(gpu_ctx, n, a, np.empty_like(a))
    print("case 2^32+4099 u32: checking", flush=True)
    assert is_sorted(a)
    assert multiset_sums(a) == before
