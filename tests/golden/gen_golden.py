"""Writes tests/golden/fixtures.npz: small inputs and their sorted outputs.

Inputs come from the library's host generators (stored, so a later change of
a generator or of libm cannot alter them); expected outputs are numpy.sort,
independent of both the oracle's std::sort and the GPU path.  The reference
ships no golden vectors (its only check is std::sort at run time,
CL12/QuicksortMain.cpp:498-516), so these are the committed known answers.

    python tests/golden/gen_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "gpu-quicksort_amd"))
import gqsort  # noqa: E402

SIZES = [0, 1, 2, 3, 15, 16, 17, 255, 256, 257, 1727, 1728, 1729, 4096, 8193]
CASES = [(gqsort.U32, g) for g in ("uniform", "gaussian", "sorted", "reverse", "equal", "bucket", "permutation")]
CASES += [(gqsort.F32, "normal"), (gqsort.F64, "normal"), (gqsort.F32, "permutation")]
NAMES = {gqsort.U32: "u32", gqsort.F32: "f32", gqsort.F64: "f64"}


def main():
    out = {}
    for dt, gen in CASES:
        for n in SIZES:
            a = gqsort.generate(dt, gen, 1000 + n, n)
            key = f"{NAMES[dt]}_{gen}_{n}"
            out[key + "_in"] = a
            out[key + "_out"] = np.sort(a, kind="stable")
    # The reference's own data at a slightly larger size (1..N shuffled).
    a = gqsort.generate(gqsort.U32, "permutation", 1, 65537)
    out["u32_permutation_65537_in"] = a
    out["u32_permutation_65537_out"] = np.sort(a)
    path = os.path.join(HERE, "fixtures.npz")
    np.savez_compressed(path, **out)
    print(path, len(out) // 2, "cases", os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
