"""The command-line driver keeps the reference's argv and output lines
(CL12/QuicksortMain.cpp:112-136, 485-540; FSY/QuicksortMain.cpp:1175-1195)."""
import os
import re
import subprocess

import pytest

from conftest import ROOT, _has_gpu

CLI = os.path.join(ROOT, "gpu-quicksort_amd", "gqsort", "quicksort")


def run(*args, timeout=300):
    assert os.path.exists(CLI), "quicksort driver not built (gpu-quicksort_amd/build.py)"
    return subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=timeout)


def test_usage_on_wrong_argument_count():
    # the reference prints its usage string and exits with -1 unless argc == 7
    for args in ((), ("1", "gpu", "amd", "8", "8"), ("1", "gpu", "amd", "8", "8", "no_show_CL", "uint", "x")):
        r = run(*args)
        assert r.returncode == 255
        assert r.stdout.startswith("Usage: Quicksort [num test iterations] [cpu|gpu]")


def test_bad_values_and_types():
    assert run(0, "gpu", "amd", 8, 8, "no_show_CL").returncode == 255
    assert run(1, "gpu", "amd", 0, 8, "no_show_CL").returncode == 255
    assert run(1, "gpu", "amd", 8, 8, "no_show_CL", "int64").returncode == 255


def test_cpu_device_is_refused():
    r = run(1, "cpu", "intel", 8, 8, "no_show_CL")
    assert r.returncode == 1 and "runs on the GPU" in r.stderr


@pytest.mark.skipif(_has_gpu(), reason="checks the behaviour on a host without a GPU")
def test_without_gpu_fails_loudly():
    r = run(1, "gpu", "amd", 64, 64, "no_show_CL")
    assert r.returncode == 1 and "no GPU" in r.stderr


@pytest.mark.gpu
def test_reference_output_lines_all_types():
    iters = 2
    r = run(iters, "gpu", "amd", 512, 256, "show_CL", "all")
    print(r.stdout[-2000:], r.stderr[-500:], flush=True)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    # the show_CL report: CUs, work-group sizes, clocks and memory (CL12/OpenCLUtils.h:291-387)
    assert "HIP Device info:" in out and re.search(r"^MAX_COMPUTE_UNITS\s+:\s+256$", out, re.M)
    assert re.search(r"^DEVICE_ARCH\s+:gfx950", out, re.M)
    for key in ("MAX_CLOCK_FREQUENCY (MHz)", "MEMORY_CLOCK (MHz)", "MEMORY_BUS_WIDTH (bits)", "MAX_WORK_GROUP_SIZE"):
        m = re.search(r"^" + re.escape(key) + r"\s*:\s*(\d+)$", out, re.M)
        assert m and int(m.group(1)) > 0, key
    m = re.search(r"^GLOBAL_MEM_SIZE\s+:\s*([0-9.]+)$", out, re.M)
    assert m and float(m.group(1)) > 200e9  # 288 GB of HBM3E
    assert out.count("Allocating array size of 131072") == 3
    for t in ("uint", "float", "double"):
        assert f"Sorting with GPUQSort on the gpu with type {t}: " in out
    # one CPU check and one per GPU iteration, for each of the three types
    assert out.count("verifying: true") == 3 * (1 + iters)
    assert "verifying: false" not in out
    assert out.count(f" Number of failures: 0 out of {iters}") == 3
    assert len(re.findall(r"^Average Time: [0-9.e+-]+ ms$", out, re.M)) == 3
    assert len(re.findall(r"^Time to sort: [0-9.e+-]+ ms$", out, re.M)) == 3 * (2 + iters)
    assert out.count("min time: ") == 3 and out.count("Average speedup over CPU std::sort: ") == 3
    assert out.count("-------done--------------------------------------------------------") == 3


@pytest.mark.gpu
def test_reference_default_uint_run():
    # the reference's own invocation shape: uint keys only, no device report
    r = run(1, "gpu", "amd", 1024, 1024, "no_show_CL")
    assert r.returncode == 0, r.stderr
    assert "HIP Device info:" not in r.stdout
    assert r.stdout.count("verifying: true") == 2
    assert " Number of failures: 0 out of 1" in r.stdout
