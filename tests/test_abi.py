"""C ABI and host-side logic of libgqsort.so (CPU only: no kernel launches)."""
import ctypes
import os
import re

import numpy as np
import pytest

import gqsort

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if name.endswith(".h"):
            text = open(os.path.join(ROOT, "include", name)).read()
            syms |= set(re.findall(r"^\s*(?:gqs_status|const char\*|void|int)\s+(gqs_\w+)\s*\(", text, re.M))
    return syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(gqsort.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert declared_symbols() <= set(gqsort._SIGS) | set(gqsort.OPTIONAL_SIGS)


def test_default_params():
    p = gqsort.default_params(gqsort.U32)
    assert p.mode == gqsort.MODE_MULTIPIVOT and p.max_fanout == 256
    assert 16 <= p.leaf_target <= p.leaf_max <= 16384
    q = gqsort.default_params(gqsort.F64)
    assert q.leaf_max * 8 <= 65536


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-GPU error path")
def test_create_without_gpu_reports_error():
    h = ctypes.c_void_p()
    st = gqsort.load().gqs_create(ctypes.byref(h), 0)
    assert st == 3 and not h.value  # GQS_EDEVICE, no context


def test_null_arguments_are_rejected():
    lib = gqsort.load()
    assert lib.gqs_sort(None, 0, None, None, 10) == 1
    assert lib.gqs_sort_device(None, 0, None, None, 10, None) == 1
    assert lib.gqs_set_large_run(None, 1 << 20) == 1
    assert lib.gqs_default_params(0, None) == 1
    assert lib.gqs_generate(0, 0, 1, None, 10) == 1
    assert lib.gqs_generate(0, 99, 1, None, 0) == 0


# ------------------------------------------------------------- generators
N = 1 << 16


def test_generators_deterministic_and_seeded():
    for gen in range(7):
        a = gqsort.generate(gqsort.U32, gen, 1, N)
        assert np.array_equal(a, gqsort.generate(gqsort.U32, gen, 1, N))
        if gen != 4:
            assert not np.array_equal(a, gqsort.generate(gqsort.U32, gen, 2, N))


def test_uniform_and_gaussian_moments():
    u = gqsort.generate(gqsort.U32, "uniform", 1, 1 << 20).astype(np.float64)
    g = gqsort.generate(gqsort.U32, "gaussian", 1, 1 << 20).astype(np.float64)
    m = 2.0**31
    assert abs(u.mean() - m) / m < 0.01 and abs(g.mean() - m) / m < 0.01
    # mean of 4 uniforms: standard deviation halves
    assert abs(g.std() / u.std() - 0.5) < 0.02


def test_sorted_reverse_equal():
    s = gqsort.generate(gqsort.U32, "sorted", 3, N)
    r = gqsort.generate(gqsort.U32, "reverse", 3, N)
    assert np.all(np.diff(s.astype(np.int64)) > 0)
    assert np.array_equal(r, s[::-1])
    assert np.all(gqsort.generate(gqsort.U32, "equal", 3, N) == 0x9E3779B9)


def test_bucket_runs():
    n = 256 * 1000
    a = gqsort.generate(gqsort.U32, "bucket", 4, n)
    runs = a.reshape(256, -1) >> 24
    assert np.all(runs == runs[:, :1])  # each run inside one 2^24-wide range
    assert sorted(runs[:, 0].tolist()) == list(range(256))  # every range once
    assert not np.array_equal(runs[:, 0], np.arange(256))  # shuffled


def test_permutation_is_reference_data():
    a = gqsort.generate(gqsort.U32, "permutation", 1, N)
    assert np.array_equal(np.sort(a), np.arange(1, N + 1, dtype=np.uint32))


@pytest.mark.parametrize("dt,npdt", [(gqsort.F32, np.float32), (gqsort.F64, np.float64)])
def test_float_generator(dt, npdt):
    n = 1 << 20
    a = gqsort.generate(dt, "normal", 5, n)
    assert a.dtype == npdt
    assert np.all(np.isfinite(a))
    assert not np.any((a == 0) & np.signbit(a))  # no -0
    wide = np.abs(a) > 1e6
    assert 0.008 < wide.mean() < 0.012  # ~1% outliers in [-1e30, 1e30]
    body = a[~wide].astype(np.float64)
    assert abs(body.mean()) < 0.01 and abs(body.std() - 1) < 0.01
    _, counts = np.unique(a, return_counts=True)
    assert (counts > 1).sum() > n * 0.0005  # ~0.1% duplicated values


def test_generator_rejects_bad_combinations():
    lib = gqsort.load()
    buf = np.empty(16, np.uint32)
    assert lib.gqs_generate(gqsort.U32, 7, 1, buf.ctypes.data, 16) == 1
    assert lib.gqs_generate(gqsort.F32, 0, 1, buf.ctypes.data, 16) == 1


def test_cpp_dropin_compiles_and_links(tmp_path):
    """A C++ caller written like the reference's harness builds against
    include/gqsort.hpp and links libgqsort.so; without a GPU it gets a status,
    not an abort."""
    import subprocess

    src = tmp_path / "caller.cpp"
    src.write_text(
        '#include "gqsort.hpp"\n#include <cstdio>\n#include <vector>\n'
        "int main() {\n"
        "  gqs_ctx* ctx = nullptr;\n"
        "  gqs_status s = gqs_create(&ctx, 0);\n"
        "  if (s != GQS_OK) { std::printf(\"status %d\\n\", (int)s); return 0; }\n"
        "  std::vector<uint32_t> d = {3, 1, 2}, dn(3);\n"
        "  s = GPUQSort(ctx, d.size(), d.data(), dn.data());\n"
        "  std::printf(\"status %d %u %u %u\\n\", (int)s, d[0], d[1], d[2]);\n"
        "  gqs_destroy(ctx);\n  return 0;\n}\n")
    exe = tmp_path / "caller"
    libdir = os.path.dirname(gqsort.LIB_PATH)
    r = subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L", libdir, "-lgqsort", f"-Wl,-rpath,{libdir}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    if os.path.exists("/dev/kfd"):
        assert r.stdout.strip() == "status 0 1 2 3"
    else:
        assert r.stdout.strip() == "status 3"


# ------------------------------------------------------- parameter ranges
def test_check_params_rejects_out_of_range_fields():
    """The range checks shared by gqs_set_params and gqs_tune_load: values the
    kernels index or divide by (a fan-out above the 256-entry splitter table,
    leaf_target 0, a wave network above 512 keys, no work-groups) are refused."""
    for dt in (gqsort.U32, gqsort.F32, gqsort.F64):
        assert gqsort.check_params(gqsort.default_params(dt))
    bad = [("max_fanout", 512), ("max_fanout", 3), ("max_fanout", 1), ("leaf_target", 0), ("oversample", 0),
           ("bitonic_cutoff", 1024), ("bitonic_cutoff", 32), ("part_wgs_per_cu", 0), ("leaf_wgs_per_cu", 33),
           ("leaf_max", 8), ("leaf_max", 32768), ("sort_threshold", 2), ("max_levels", 0), ("max_levels", 48),
           ("mode", 2), ("host_driven", 2)]
    for field, v in bad:
        p = gqsort.default_params()
        setattr(p, field, v)
        assert not gqsort.check_params(p), (field, v)
    p = gqsort.default_params()
    p.max_fanout, p.leaf_max, p.max_levels = 2, 16, 47
    assert gqsort.check_params(p)


# ------------------------------------------------------- server config
def test_server_config_caps_batches_below_one_run():
    """A slot of slot_keys 8-byte keys holds twice as many 4-byte keys; one
    segmented sort takes fewer than 2^32 - 2^16 keys, so the derived (or given)
    slot size is capped to keep a batch of 4-byte keys below that."""
    limit = (1 << 32) - (1 << 16)
    c = gqsort.resolve_server_config(hbm_budget_bytes=64 << 30, slots=1)
    assert c.slot_keys == (limit - 1) // 2 and 2 * c.slot_keys < limit
    c = gqsort.resolve_server_config(slot_keys=1 << 40)
    assert 2 * c.slot_keys < limit
    c = gqsort.resolve_server_config(hbm_budget_bytes=1 << 30, slots=4)
    assert c.slot_keys == (1 << 30) // 4 // 24
    d = gqsort.resolve_server_config()
    assert d.slots == 8 and d.hbm_budget_bytes == 64 << 30 and d.slot_keys == (64 << 30) // 8 // 24
    assert d.batch_bytes == 64 << 20 and gqsort.resolve_server_config(batch_bytes=1 << 20).batch_bytes == 1 << 20
    for kw in ({"slots": 0}, {"slots": 17}, {"max_batch_requests": 0}, {"hbm_budget_bytes": 0}):
        with pytest.raises(gqsort.GQSError):
            gqsort.resolve_server_config(**kw)
