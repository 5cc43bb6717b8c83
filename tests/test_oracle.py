"""Pins the CPU oracle before it is trusted as the checker (CPU only).

Each oracle function is checked against something independent of its own
formula: numpy, committed fixtures, the reference header compiled as-is
(oracle/_ref), closed forms from the reference's constants, brute force.
"""
import itertools
import os

import numpy as np
import pytest

import gqsort
import orc

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "fixtures.npz")


@pytest.fixture(scope="module")
def fixtures():
    with np.load(FIX, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _cases(fx):
    return sorted(k[:-3] for k in fx if k.endswith("_in"))


def test_std_sort_matches_fixtures(fixtures):
    for c in _cases(fixtures):
        a, e = fixtures[c + "_in"], fixtures[c + "_out"]
        got = orc.std_sort(a)
        assert orc.mismatch(got, e)[0] == 0, c


def test_restatement_matches_fixtures(fixtures):
    """The CPU restatement of GPUQSort (CL12 level loop + kernels) sorts every
    fixture exactly like numpy.sort."""
    for c in _cases(fixtures):
        a, e = fixtures[c + "_in"], fixtures[c + "_out"]
        if a.size == 0:
            continue
        got, st = orc.gpuqsort(a)
        assert orc.mismatch(got, e)[0] == 0, c
        assert st.invariant_violations == 0, c


@pytest.mark.parametrize("gen", ["uniform", "gaussian", "sorted", "reverse", "equal", "bucket", "permutation"])
def test_restatement_levels(gen):
    a = gqsort.generate(gqsort.U32, gen, 5, 1 << 17)
    got, st = orc.gpuqsort(a)
    assert np.array_equal(got, np.sort(a))
    assert st.invariant_violations == 0
    assert st.maxseq == 256
    if gen != "equal":
        assert 1 <= st.max_done_len <= 1728
        assert st.levels >= 5
    else:
        assert st.levels == 1 and st.done_records == 0


@pytest.mark.parametrize("tun", [(1024, 128, 256, 512), (1024, 128, 128, 256), (1536, 128, 128, 256),
                                 (1728, 256, 128, 256)])
def test_restatement_other_tunables(tun):
    """Per-device tunable sets of CL12/Quicksort.h:79-97, CL20/Quicksort.h:84-88, FSY/Quicksort.h:130-136."""
    a = gqsort.generate(gqsort.U32, "uniform", 9, 100003)
    got, _ = orc.gpuqsort(a, *tun)
    assert np.array_equal(got, np.sort(a))


def test_restatement_rejects_bad_threshold():
    a = np.arange(10, dtype=np.uint32)
    with pytest.raises(ValueError):
        orc.gpuqsort(a, 1728, 128, 128, 512)  # SORT_THRESHOLD != 2*LQSORT_WG


def test_par_sort_matches_std_sort():
    for dt, gen in [(gqsort.U32, "uniform"), (gqsort.F32, "normal"), (gqsort.F64, "normal")]:
        a = gqsort.generate(dt, gen, 2, 300007)
        for t in (1, 3, 8):
            assert orc.mismatch(orc.par_sort(a, t), orc.std_sort(a))[0] == 0


def test_device_median_brute_force():
    lib = orc.load()
    for x in itertools.product(range(3), repeat=3):
        assert lib.orc_median_u32(*x) == sorted(x)[1], x


def test_host_median_slip():
    """Host median (CL12/Quicksort.h:34-57): when x1 is strictly the largest it
    returns min(x2, x3); otherwise the true median."""
    lib = orc.load()
    for x in itertools.product(range(3), repeat=3):
        x1, x2, x3 = x
        exp = min(x2, x3) if (x1 >= x2 and x1 >= x3) else sorted(x)[1]
        assert lib.orc_host_median_u32(*x) == exp, x


def test_host_median_against_reference_header():
    ref = orc.load_ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    lib = orc.load()
    for x in itertools.product(range(4), repeat=3):
        assert lib.orc_host_median_u32(*x) == ref.ref_host_median_u32(*x), x
    rng = np.random.default_rng(0)
    for _ in range(2000):
        x = [int(v) for v in rng.integers(0, 2**32, 3, dtype=np.uint64)]
        assert lib.orc_host_median_u32(*x) == ref.ref_host_median_u32(*x)


def test_reference_header_constants():
    ref = orc.load_ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    import ctypes

    sz = (ctypes.c_size_t * 3)()
    ref.ref_record_sizes(sz)
    assert list(sz) == [16, 20, 20]
    t = (ctypes.c_int * 5)()
    ref.ref_tunables(t)
    assert list(t) == [1728, 128, 128, 256, 42]


@pytest.mark.parametrize("n,maxseq", [(1 << 16, 256), (1 << 20, 256), (1 << 22, 512), (1 << 26, 8192),
                                      (1 << 27, 16384), (1 << 28, 32768), (1 << 29, 65536)])
def test_optp_table(n, maxseq):
    """MAXSEQ = optp(N, 0.00009516, 203) (CL12/QuicksortMain.cpp:293-295, 315)."""
    assert orc.load().orc_optp(n, 0.00009516, 203) == maxseq
    # closed form: nearest power of two of k*N + m in the log domain
    assert maxseq == 2 ** int(np.floor(np.log2(n * 0.00009516 + 203) + 0.5))


def test_float_key_fixed_points():
    lib = orc.load()
    bits = lambda f: int(np.float32(f).view(np.uint32))  # noqa: E731
    assert lib.orc_f32_key(bits(0.0)) == 0x80000000
    assert lib.orc_f32_key(bits(-0.0)) == 0x7FFFFFFF
    assert lib.orc_f32_key(bits(np.inf)) == 0xFF800000
    assert lib.orc_f32_key(bits(-np.inf)) == 0x007FFFFF
    assert lib.orc_f32_key(bits(1.0)) == 0xBF800000
    for v in (0.0, -0.0, 1.5, -1.5, np.inf, -np.inf, 1e-45, -3e38):
        assert lib.orc_f32_unkey(lib.orc_f32_key(bits(v))) == bits(v)


def test_float_key_order_matches_numpy():
    lib = orc.load()
    rng = np.random.default_rng(1)
    f = np.concatenate([rng.standard_normal(3000), rng.uniform(-1e30, 1e30, 1000), [0.0, np.inf, -np.inf]])
    for dt, key, ity in ((np.float32, lib.orc_f32_key, np.uint32), (np.float64, lib.orc_f64_key, np.uint64)):
        x = f.astype(dt)
        k = np.array([key(int(b)) for b in x.view(ity)], dtype=ity)
        assert np.array_equal(np.argsort(x, kind="stable"), np.argsort(k, kind="stable"))


def test_kway_single_pivot_closed_forms():
    """With one splitter p with an equality bucket the brute-force classifier
    gives #{<p}, #{=p}, #{>p}: the lt/gt counts of gqsort_kernel
    (CL12/QuicksortKernels.cl:135-145) and its pivot fill (:214-216)."""
    a = gqsort.generate(gqsort.U32, "gaussian", 3, 50000)
    a[::97] = a[5]
    p = a[5]
    ids, counts = orc.kway_classify(a, [p], [1])
    assert list(counts) == [int((a < p).sum()), int((a == p).sum()), int((a > p).sum())]
    assert np.array_equal(ids, np.where(a < p, 0, np.where(a == p, 1, 2)))


def test_kway_invariants():
    rng = np.random.default_rng(4)
    a = rng.integers(0, 1000, 20000, dtype=np.uint64).astype(np.uint32)
    u = np.array([10, 200, 201, 500, 999], dtype=np.uint32)
    eq = np.array([1, 0, 1, 1, 0], dtype=np.uint8)
    ids, counts = orc.kway_classify(a, u, eq)
    assert len(counts) == 6 + int(eq.sum()) and counts.sum() == a.size
    # bucket ids are monotone in the key
    order = np.argsort(a, kind="stable")
    assert np.all(np.diff(ids[order].astype(int)) >= 0)
    # equality buckets hold exactly their splitter
    for b, v in ((1, 10), (4, 201), (6, 500)):
        assert np.all(a[ids == b] == v)


def test_kway_bucket_ids_above_255():
    """255 splitters, every one with an equality bucket: 511 buckets, ids
    beyond one byte.  Closed form: id = 2 * #{u < k} + [k in u]."""
    rng = np.random.default_rng(12)
    u = np.unique(rng.integers(0, 1 << 31, 400, dtype=np.uint64).astype(np.uint32))[:255]
    assert u.size == 255
    a = np.concatenate([rng.integers(0, 1 << 31, 30000, dtype=np.uint64).astype(np.uint32), u, u[::3]])
    ids, counts = orc.kway_classify(a, u, np.ones(255, np.uint8))
    lt = np.searchsorted(u, a, side="left")
    hit = (lt < u.size) & (u[np.minimum(lt, u.size - 1)] == a)
    assert np.array_equal(ids.astype(np.int64), 2 * lt + hit)
    assert len(counts) == 511 and int(counts.sum()) == a.size and ids.max() > 255


def test_ordered_keys_and_segmented_sort():
    """The array forms agree with the pinned scalar bijection (and so order like
    the values), and the segmented std::sort equals numpy.sort per array."""
    rng = np.random.default_rng(21)
    for dt, ity, f in ((np.float32, np.uint32, "orc_f32_key"), (np.float64, np.uint64, "orc_f64_key")):
        x = (rng.standard_normal(5000) * 1e3).astype(dt)
        x[:4] = [0.0, np.inf, -np.inf, 1.0]
        k = orc.ordered_keys(x)
        scalar = getattr(orc.load(), f)
        assert k.dtype == ity
        assert all(int(k[i]) == scalar(int(b)) for i, b in enumerate(x.view(ity)[:300]))
        assert np.array_equal(np.argsort(k, kind="stable"), np.argsort(x, kind="stable"))
    u = rng.integers(0, 1 << 32, 1000, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(orc.ordered_keys(u), u)
    for dt in (np.uint32, np.float32, np.float64):
        a = (rng.integers(0, 1 << 31, 3000, dtype=np.uint64) if dt == np.uint32 else rng.standard_normal(3000)).astype(dt)
        offs = np.array([0, 0, 1, 17, 17, 1500, 2999, 3000], np.uint64)
        got = orc.std_sort_segmented(a, offs)
        exp = np.concatenate([np.sort(a[int(offs[i]):int(offs[i + 1])]) for i in range(len(offs) - 1)])
        assert np.array_equal(got, exp)


def test_std_sort_u64_matches_numpy():
    rng = np.random.default_rng(31)
    a = rng.integers(0, np.iinfo(np.uint64).max, 100001, dtype=np.uint64, endpoint=True)
    a[::7] = a[3]
    assert np.array_equal(orc.std_sort(a), np.sort(a))
