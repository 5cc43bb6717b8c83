"""Build recipe for libgqsort.so (gfx950) -- run on any host with hipcc.

    python gpu-quicksort_amd/build.py            # build the library in-tree

The library is written to gpu-quicksort_amd/gqsort/libgqsort.so next to the
Python binding that loads it.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "gqsort", "libgqsort.so")
ARCH = os.environ.get("GQS_OFFLOAD_ARCH", "gfx950")

SOURCES = ["kernels.hip", "merge.hip", "gqsort.cpp", "devutil.cpp", "generate.cpp", "serve.cpp", "autotune.cpp"]

# Packed FP32 instructions are disabled for device code; the host pass of
# hipcc warns that it does not know the feature, which is harmless.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread",
           "-Wno-unused-command-line-argument", *DEVICE_FLAGS,
           "-I", os.path.join(HERE, "..", "include"), "-o", OUT + ".tmp", *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libgqsort.so")
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
