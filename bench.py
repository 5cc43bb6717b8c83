"""Benchmark of the plain GPUQSort hot path on one MI355X.

Workload (one "step"): sort 2^28 uint32 keys, uniform over [0, 2^32) from
seed 1, already resident in HBM, with gqs_sort_device.  Before every step the
unsorted input is restored from a pristine device copy outside the timed
region; each step is bracketed by HIP events on the sort's own stream and the
K step times are summed.

Prints one JSON line with the throughput (Mkeys/s), the fraction of the
one-read-plus-one-write HBM roofline, a per-kernel roofline for the dominant
kernel (HIP events around each launch, in a separate profiled pass), and a
CPU baseline (std::sort on one core over a bounded sample, timed here).

    python bench.py [--steps K] [--warmup W] [--n N] [--gen uniform] [--dtype u32]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-quicksort_amd"))

import gqsort  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X published HBM3E bandwidth, B/s
DTYPES = {"u32": gqsort.U32, "f32": gqsort.F32, "f64": gqsort.F64}
# Algorithmic HBM bytes per key processed, per kernel kind (key bytes kb):
# count reads each key once, scatter and the local sort read and write it
# once, fill writes it once.
ALG_BYTES_PER_KEY = {"count": 1, "scatter": 2, "leaf": 2, "fill": 1}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(kernel, workload):
    """HBM bytes per sort of `kernel` on `workload` from the committed rocprofv3
    table (profiles/traffic.txt), or None when that workload was not profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and not line.startswith("#") and f[0] == f"k_{kernel}" and f[1] == workload:
            return float(f[2])
    return None


def kernel_roofline(ctx, dt, d, dn, pristine, nbytes, n, stream, kb, workload):
    """Per-kernel event timing of one sort; returns the dominant kernel's roofline."""
    gqsort.load().gqs_memcpy_async(ctx.handle, d.ptr, pristine.ptr, nbytes, stream.handle)
    ctx.set_profiling(True)
    ctx.sort_device(dt, d, dn, n, stream)
    times = ctx.kernel_times()
    ctx.set_profiling(False)
    st = ctx.stats()
    level_keys = [int(k) for k in st.level_keys]
    keys_partitioned = sum(level_keys)
    # keys handled per kind over the whole sort
    keys = {"count": keys_partitioned, "scatter": keys_partitioned, "leaf": n, "fill": 0}
    dom = max((k for k in ALG_BYTES_PER_KEY if times[k][0]), key=lambda k: times[k][1])
    launches, ms = times[dom]
    alg_bytes_per_launch = ALG_BYTES_PER_KEY[dom] * kb * keys[dom] / launches
    avg_s = ms / launches / 1e3
    achieved = alg_bytes_per_launch / avg_s / 1e9
    per_sort = measured_traffic(dom, workload)
    traffic = None if per_sort is None else int(per_sort / launches)
    return {
        "kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
        "frac": round(achieved * 1e9 / HBM_PEAK, 4), "traffic": traffic,
        "alg_bytes_per_launch": int(alg_bytes_per_launch), "avg_launch_ms": round(ms / launches, 4),
        "launches": launches,
    }, {k: {"launches": v[0], "ms": round(v[1], 4)} for k, v in times.items() if v[0]}, level_keys


def cpu_baseline(dt, gen, seed, sample_n, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc

    a = gqsort.generate(dt, gen, seed, sample_n)
    t0 = time.perf_counter()
    orc.std_sort(a)
    t1 = time.perf_counter()
    par = None
    if threads > 1:
        t2 = time.perf_counter()
        orc.par_sort(a, threads)
        t3 = time.perf_counter()
        par = {"value": round(sample_n / (t3 - t2) / 1e6, 2), "unit": "Mkeys/s", "cores": threads,
               "kind": "port", "sample": f"{sample_n} keys, chunked std::sort + merge (tbb::parallel_sort stand-in)"}
    return {"value": round(sample_n / (t1 - t0) / 1e6, 2), "unit": "Mkeys/s", "cores": 1, "kind": "port",
            "sample": f"{sample_n} {gen} keys (seed {seed}), std::sort on one core"}, par


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 28)
    ap.add_argument("--gen", default="uniform")
    ap.add_argument("--dtype", default="u32", choices=sorted(DTYPES))
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=1 << 25)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the sort shards by sample-sort exchange, which this build does not "
              "run (see DESIGN.md); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    dt = DTYPES[args.dtype]
    kb = gqsort.KEY_BYTES[dt]
    n = args.n
    nbytes = n * kb
    log(f"generating {n} {args.gen} {args.dtype} keys (seed {args.seed})")
    host = gqsort.generate(dt, args.gen, args.seed, n)

    ctx = gqsort.Context(0)
    info = ctx.device_info()
    stream = gqsort.Stream(ctx)
    pristine = gqsort.DeviceBuffer.from_array(ctx, host, stream)
    d = gqsort.DeviceBuffer(ctx, nbytes)
    dn = gqsort.DeviceBuffer(ctx, nbytes)
    lib = gqsort.load()
    ev0, ev1 = gqsort.Event(ctx), gqsort.Event(ctx)

    def step():
        lib.gqs_memcpy_async(ctx.handle, d.ptr, pristine.ptr, nbytes, stream.handle)
        ev0.record(stream)
        ctx.sort_device(dt, d, dn, n, stream)
        ev1.record(stream)
        return ev0.elapsed_ms(ev1)

    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    stream.sync()
    log(f"timing {args.steps} steps")
    times = [step() for _ in range(args.steps)]
    stream.sync()
    total_ms = float(sum(times))
    ms_per_step = total_ms / args.steps
    value = n / (ms_per_step / 1e3) / 1e6

    # correctness of the last step: sortedness and multiset checksums (size-independent)
    check = None
    if not args.no_check:
        out = d.download(host.dtype, n, stream)
        iv = host.view(np.uint32 if kb == 4 else np.uint64)
        ov = out.view(iv.dtype)
        sorted_ok = bool(np.all(out[1:] >= out[:-1])) if args.dtype == "u32" else bool(
            np.all(np.diff(out.astype(np.float64)) >= 0))
        sums_ok = int(iv.sum(dtype=np.uint64)) == int(ov.sum(dtype=np.uint64)) and \
            int(np.bitwise_xor.reduce(iv)) == int(np.bitwise_xor.reduce(ov))
        check = {"sorted": sorted_ok, "multiset_checksums": sums_ok}
        if not (sorted_ok and sums_ok):
            log(f"CHECK FAILED: {check}")
            sys.exit(1)

    log("per-kernel profile pass")
    roof, kernels, level_keys = kernel_roofline(ctx, dt, d, dn, pristine, nbytes, n, stream, kb,
                                                f"{args.dtype}:{n}:{args.gen}")

    cpu = par = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on a {args.cpu_sample}-key sample")
        cpu, par = cpu_baseline(dt, args.gen, args.seed, args.cpu_sample, threads)

    sort_frac = (2 * nbytes / (ms_per_step / 1e3)) / HBM_PEAK
    line = {
        "metric": f"Mkeys/s {args.dtype} 2^{int(np.log2(n))} (+% HBM roofline)" if n & (n - 1) == 0 else
        f"Mkeys/s {args.dtype} n={n}",
        "value": round(value, 1),
        "unit": "Mkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": f"plain GPUQSort of {n} {args.dtype} keys, generator {args.gen}, seed {args.seed}, "
                               "device-resident", "n": n, "generator": args.gen, "seed": args.seed,
                   "parallelism": "single-gpu"},
        "sort_roofline_frac": round(sort_frac, 4),
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_parallel": par,
        "kernels_ms": kernels,
        "level_keys": level_keys,
        "step_ms": {"min": round(min(times), 4), "median": round(float(np.median(times)), 4),
                    "max": round(max(times), 4)},
        "check": check,
        "device": info,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
