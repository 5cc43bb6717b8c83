"""Build recipe for libgqsort.so (gfx950) -- run on any host with hipcc.

    python gpu-quicksort_amd/build.py            # build the library in-tree

The library is written to gpu-quicksort_amd/gqsort/libgqsort.so next to the
Python binding that loads it.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "gqsort", "libgqsort.so")
ARCH = os.environ.get("GQS_OFFLOAD_ARCH", "gfx950")

# (source, extra defines): kernels_typed.hip is compiled once per key type.
UNITS = [("kernels.hip", []), ("kernels_typed.hip", ["-DGQS_DT=0"]), ("kernels_typed.hip", ["-DGQS_DT=1"]),
         ("kernels_typed.hip", ["-DGQS_DT=2"]), ("merge.hip", []), ("gqsort.cpp", []), ("devutil.cpp", []),
         ("generate.cpp", []), ("serve.cpp", []), ("autotune.cpp", []), ("noalloc.hip", [])]
OBJ_DIR = os.path.join(HERE, "_build")

# Packed FP32 instructions are disabled for device code; the host pass of
# hipcc warns that it does not know the feature, which is harmless.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def _compile(unit, verbose):
    src, defs = unit
    tag = "".join(d.split("=")[-1] for d in defs)
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + (f"_{tag}" if tag else "") + ".o")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-pthread", "-c",
           "-Wno-unused-command-line-argument", "-Wno-unused-value", *DEVICE_FLAGS, *defs,
           "-I", os.path.join(HERE, "..", "include"), "-o", obj, os.path.join(CSRC, src)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return None, r.stdout + r.stderr
    return obj, ""


def build(verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    units = [u for u in UNITS if os.path.exists(os.path.join(CSRC, u[0]))]
    jobs = max(1, min(len(units), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda u: _compile(u, verbose), units))
    errs = [msg for obj, msg in results if obj is None]
    if errs:
        sys.stderr.write("\n".join(errs))
        raise RuntimeError("hipcc failed building libgqsort.so")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", OUT + ".tmp",
           *[obj for obj, _ in results]]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libgqsort.so")
    os.replace(OUT + ".tmp", OUT)
    build_cli(verbose)
    return OUT


CLI_SRC = os.path.join(HERE, "cli", "quicksort_main.cpp")
CLI_OUT = os.path.join(HERE, "gqsort", "quicksort")


def build_cli(verbose=False):
    """The command-line driver (reference argv and output lines): host C++ only,
    linked against libgqsort.so next to it."""
    cmd = ["g++", "-O2", "-std=c++17", "-pthread", "-I", os.path.join(HERE, "..", "include"), CLI_SRC,
           "-L", os.path.dirname(OUT), "-lgqsort", "-Wl,-rpath,$ORIGIN", "-o", CLI_OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building the quicksort driver")
    os.replace(CLI_OUT + ".tmp", CLI_OUT)
    return CLI_OUT


if __name__ == "__main__":
    print(build(verbose=True))
