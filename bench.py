"""Benchmark of the plain GPUQSort hot path on one MI355X.

Workload (one "step"): sort 2^28 uint32 keys, uniform over [0, 2^32) from
seed 1, already resident in HBM, with gqs_sort_device.  Before every step the
unsorted input is restored from a pristine device copy outside the timed
region; each step is bracketed by HIP events on the sort's own stream and the
K step times are summed.

Prints one JSON line with the throughput (Mkeys/s), the fraction of the
one-read-plus-one-write HBM roofline, a per-kernel roofline for the dominant
kernel (HIP events around each launch, in a separate profiled pass), a CPU
baseline (std::sort on one core over a bounded sample, timed here), and the
batched (captured segmented sort vs per-array plain calls) and serving
(continuous batching vs plain per-request calls, p50/p99) measurements.

    python bench.py [--steps K] [--warmup W] [--n N] [--gen uniform] [--dtype u32]
"""
import argparse
import ctypes
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
# once (equality buckets are filled inside the local-sort launches).
ALG_BYTES_PER_KEY = {"count": 1, "scatter": 2, "leaf": 2, "tilepart": 2}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# kernels behind each kind of the library's per-kind timing
KIND_KERNELS = {"count": ("k_count",), "scatter": ("k_scatter",), "leaf": ("k_leaf_big", "k_leaf_small"),
                "tilepart": ("k_tilepart",)}


def measured_traffic(kind, workload):
    """HBM bytes per sort of the kernels of `kind` on `workload` from the
    committed rocprofv3 table (profiles/traffic.txt), or None when that
    workload was not profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    total, found = 0.0, False
    for line in open(path):
        f = line.split()
        # kernel names may carry their template arguments: k_tilepart<0,true>
        if len(f) == 3 and not line.startswith("#") and f[0].split("<")[0] in KIND_KERNELS.get(kind, ()) \
                and f[1] == workload:
            total += float(f[2])
            found = True
    return total if found else None


def kernel_roofline(ctx, dt, d, dn, pristine, nbytes, n, stream, kb, workload):
    """Per-kernel event timing of one sort; returns the dominant kernel's roofline."""
    gqsort.load().gqs_memcpy_async(ctx.handle, d.ptr, pristine.ptr, nbytes, stream.handle)
    ctx.set_profiling(True)
    ctx.sort_device(dt, d, dn, n, stream)
    times = ctx.kernel_times()
    trace = ctx.kernel_trace()
    ctx.set_profiling(False)
    st = ctx.stats()
    level_keys = [int(k) for k in st.level_keys]
    # levels that partition tile-locally (k_tilepart) and those that count and
    # scatter, told apart by the launches of each level (one splitter launch
    # opens every level)
    lvl, local_levels = -1, set()
    for kind, _ in trace:
        if kind == "splitters":
            lvl += 1
        elif kind == "tilepart":
            local_levels.add(lvl)
    local_keys = sum(level_keys[i] for i in local_levels if 0 <= i < len(level_keys))
    keys_partitioned = sum(level_keys) - local_keys
    # keys handled per kind over the whole sort
    keys = {"count": keys_partitioned, "scatter": keys_partitioned, "leaf": n, "tilepart": local_keys}
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


def cpu_info():
    """Host provenance: nproc and the CPU model string."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def cpu_baseline(host, gpu_out, threads):
    """std::sort on one core and a parallel sort on `threads` cores of the same
    input the GPU sorted (the reference's CPU legs, CL12/QuicksortMain.cpp:
    418-437), each timed around the sort call alone; the GPU result must equal
    std::sort's bit for bit."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc

    lib = orc.load()
    suffix = orc._SUFFIX[host.dtype]
    n = host.size
    a = host.copy()
    t0 = time.perf_counter()
    getattr(lib, f"orc_std_sort_{suffix}")(a.ctypes.data, n)
    t1 = time.perf_counter()
    exact = gpu_out is not None and orc.mismatch(gpu_out, a)[0] == 0
    b = host.copy()
    t2 = time.perf_counter()
    getattr(lib, f"orc_par_sort_{suffix}")(b.ctypes.data, n, threads)
    t3 = time.perf_counter()
    par_exact = orc.mismatch(b, a)[0] == 0
    sample = f"the timed input itself ({n} keys)"
    return ({"value": round(n / (t1 - t0) / 1e6, 2), "unit": "Mkeys/s", "cores": 1, "kind": "port",
             "sample": sample + ", std::sort on one core", "ms": round((t1 - t0) * 1e3, 1),
             "gpu_equals_std_sort": exact},
            {"value": round(n / (t3 - t2) / 1e6, 2), "unit": "Mkeys/s", "cores": threads, "kind": "port",
             "sample": sample + f", {threads} threads: chunked std::sort + pairwise merges "
                                "(tbb::parallel_sort stand-in: TBB headers are absent)",
             "ms": round((t3 - t2) * 1e3, 1), "equals_std_sort": par_exact})


def pcie_rates(ctx, nbytes, reps=3):
    """Link rates with pinned host memory, GB/s (median of reps): one
    synchronous copy of nbytes each way, and both directions at once on two
    streams ("duplex", per direction)."""
    lib = ctx._lib
    hp = [ctypes.c_void_p(), ctypes.c_void_p()]
    for p in hp:
        gqsort._check(lib.gqs_host_alloc(ctx.handle, nbytes, ctypes.byref(p)), ctx)
    dev = [gqsort.DeviceBuffer(ctx, nbytes), gqsort.DeviceBuffer(ctx, nbytes)]
    streams = [gqsort.Stream(ctx), gqsort.Stream(ctx)]
    try:
        for p in hp:
            ctypes.memset(p.value, 1, nbytes)
        out = {}
        for name, dst, src in (("h2d", dev[0].ptr, hp[0].value), ("d2h", hp[0].value, dev[0].ptr)):
            ts = []
            for _ in range(reps + 1):
                t0 = time.perf_counter()
                gqsort._check(lib.gqs_memcpy(ctx.handle, dst, src, nbytes, None), ctx)
                ts.append(time.perf_counter() - t0)
            out[name] = nbytes / float(np.median(ts[1:])) / 1e9
        ts = []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            gqsort._check(lib.gqs_memcpy_async(ctx.handle, dev[1].ptr, hp[1].value, nbytes, streams[0].handle), ctx)
            gqsort._check(lib.gqs_memcpy_async(ctx.handle, hp[0].value, dev[0].ptr, nbytes, streams[1].handle), ctx)
            streams[0].sync()
            streams[1].sync()
            ts.append(time.perf_counter() - t0)
        out["duplex"] = nbytes / float(np.median(ts[1:])) / 1e9
        return out
    finally:
        for x in dev:
            x.free()
        for st in streams:
            st.close()
        for p in hp:
            lib.gqs_host_free(ctx.handle, p)


def host_pointer_e2e(ctx, dt, host, reps):
    """The reference's timed call: GPUQSort on host arrays, upload + sort +
    download (CL12/QuicksortMain.cpp:491-494), wall clock per call, with the
    copies overlapping the sort (the default from 2^24 keys) and in turn; and
    the bound a host sort cannot beat: every input key must arrive before the
    first output key is known, so upload + download at the pinned link rates."""
    d = host.copy()
    dn = np.empty_like(d)

    def timed():
        ts = []
        for r in range(reps + 1):
            np.copyto(d, host)
            t0 = time.perf_counter()
            gqsort.GPUQSort(ctx, d.size, d, dn)
            t1 = time.perf_counter()
            if r:
                ts.append((t1 - t0) * 1e3)
        ok = bool(np.all(d[1:] >= d[:-1])) if d.dtype == np.uint32 else bool(np.all(np.diff(d) >= 0))
        return float(np.median(ts)), ok

    ms, ok = timed()
    ctx.set_host_pipeline(2**64 - 1)
    try:
        serial_ms, ok_serial = timed()
    finally:
        ctx.set_host_pipeline(0)
    # the same call on arrays in pinned memory (no staging copy on the host)
    lib = ctx._lib
    hp = [ctypes.c_void_p(), ctypes.c_void_p()]
    for p in hp:
        gqsort._check(lib.gqs_host_alloc(ctx.handle, host.nbytes, ctypes.byref(p)), ctx)
    try:
        ct = ctypes.c_uint32 if host.dtype == np.uint32 else (ctypes.c_float if host.dtype == np.float32 else
                                                              ctypes.c_double)
        d = np.ctypeslib.as_array((ct * host.size).from_address(hp[0].value))
        dn = np.ctypeslib.as_array((ct * host.size).from_address(hp[1].value))
        pinned_ms, ok_pinned = timed()
        ok = ok and ok_pinned
    finally:
        del d, dn
        for p in hp:
            lib.gqs_host_free(ctx.handle, p)
    link = pcie_rates(ctx, host.nbytes)
    bound = host.nbytes / link["h2d"] / 1e6 + host.nbytes / link["d2h"] / 1e6
    return {"ms": round(ms, 2), "Mkeys_s": round(host.size / ms / 1e3, 1),
            "GB_s_moved": round(2 * host.nbytes / (ms / 1e3) / 1e9, 1), "reps": reps, "sorted": ok and ok_serial,
            "serial_ms": round(serial_ms, 2), "pinned_ms": round(pinned_ms, 2),
            "pcie_pinned_GB_s": {k: round(v, 1) for k, v in link.items()},
            "copy_bound_ms": round(bound, 2), "frac_of_copy_bound": round(bound / ms, 3),
            "note": "pageable host buffers staged through pinned chunks; includes PCIe both ways; ms: copies "
                    "overlapped with the sort (runs sorted as they arrive, final merge downloaded by ranges); "
                    "serial_ms: upload, sort, download in turn; pinned_ms: overlapped, on arrays in pinned memory "
                    "(gqs_host_alloc); copy_bound_ms: upload + download alone at the pinned link rates"}


def multiset_ok(inp, out):
    """Order-independent checksums of the key multiset (sum and xor of the bit
    patterns) plus sortedness."""
    iv = inp.view(np.uint32 if inp.itemsize == 4 else np.uint64)
    ov = out.view(iv.dtype)
    sorted_ok = bool(np.all(out[1:] >= out[:-1])) if inp.dtype == np.uint32 else bool(
        np.all(np.diff(out.astype(np.float64)) >= 0))
    sums_ok = int(iv.sum(dtype=np.uint64)) == int(ov.sum(dtype=np.uint64)) and \
        int(np.bitwise_xor.reduce(iv)) == int(np.bitwise_xor.reduce(ov))
    return sorted_ok, sums_ok


def bench_batched(ctx, args, stream, ev0, ev1):
    """65,536 uint32 arrays, lengths floor(2^U), U ~ U[4, 20] (seeded), rescaled
    to about 2^28 keys in total (the configuration's stated total); one captured
    segmented sort per batch vs one plain gqs_sort_device per array."""
    lib = gqsort.load()
    nseg = args.batch_arrays
    lens = gqsort.batch_lengths(nseg, 16, 1 << 20, seed=args.seed, total=args.batch_total)
    total = int(lens.sum())
    offs = np.zeros(nseg + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    keys = gqsort.generate(gqsort.U32, "uniform", args.seed + 1, total)
    pristine = gqsort.DeviceBuffer.from_array(ctx, keys, stream)
    d = gqsort.DeviceBuffer(ctx, keys.nbytes)
    dn = gqsort.DeviceBuffer(ctx, keys.nbytes)
    od = gqsort.DeviceBuffer.from_array(ctx, offs, stream)
    nd = gqsort.DeviceBuffer.from_array(ctx, np.array([nseg], np.uint32), stream)

    def timed(fn, reps):
        ts = []
        for _ in range(reps):
            lib.gqs_memcpy_async(ctx.handle, d.ptr, pristine.ptr, keys.nbytes, stream.handle)
            ev0.record(stream)
            fn()
            ev1.record(stream)
            ts.append(ev0.elapsed_ms(ev1))
        return float(np.median(ts))

    graph = gqsort.SegmentedGraph(ctx, gqsort.U32, d, dn, od, nd, nseg, total)
    graph.launch(stream)
    stream.sync()
    t_graph = timed(lambda: graph.launch(stream), args.batch_reps)
    out = d.download(np.uint32, total, stream)
    ok = True
    bounds = offs[1:-1].astype(np.int64)
    inside = np.ones(total - 1, bool)
    inside[bounds[(bounds > 0) & (bounds < total)] - 1] = False
    ok &= bool(np.all(out[1:][inside] >= out[:-1][inside]))
    ok &= int(keys.sum(dtype=np.uint64)) == int(out.sum(dtype=np.uint64))
    graph.close()
    t_eager = timed(lambda: ctx.sort_segmented(gqsort.U32, d, dn, od, nseg, total, stream), args.batch_reps)

    def plain():
        gqsort._check(lib.gqs_sort_each(ctx.handle, gqsort.U32, d.ptr, dn.ptr, offs.ctypes.data, nseg,
                                        stream.handle), ctx)
    t_plain = timed(plain, 1)
    for b in (pristine, d, dn, od, nd):
        b.free()
    return {
        "workload": f"{nseg} uint32 arrays, lengths floor(2^U) U~U[4,20] rescaled to {total} keys, uniform keys",
        "arrays": nseg, "keys": total,
        "graph_ms": round(t_graph, 3), "eager_ms": round(t_eager, 3), "plain_ms": round(t_plain, 3),
        "sorts_per_s": round(nseg / (t_graph / 1e3), 1), "keys_per_s": round(total / (t_graph / 1e3), 1),
        "plain_sorts_per_s": round(nseg / (t_plain / 1e3), 1),
        "speedup_vs_plain": round(t_plain / t_graph, 2), "check_sorted_and_checksum": ok,
    }


def bench_serving(args):
    """A stream of sort requests, sizes floor(2^U) U~U[10,26], dtypes 60% u32 /
    30% f32 / 10% f64, in pinned host memory, through the continuous-batching
    server (64 GiB budget) and through plain per-request gqs_sort calls (the
    reference's caller, one GPUQSort per array, CL12/QuicksortMain.cpp:485-497).

    Saturated: everything offered at t=0.  Then the comparison at equal tail
    latency: the latency bound is plain's p99 under Poisson arrivals at 80% of
    its saturated rate; each system is swept over offered rates (three Poisson
    streams per rate, the median p99 counts), and its sustainable rate is the
    highest one whose p99 stays within that bound.
    Both directions of PCIe carry every key once, so no host-memory server can
    beat bytes / (per-direction rate with both directions busy); that ceiling
    is reported beside the ratio."""
    lib = gqsort.load()
    rng = np.random.default_rng(args.seed + 7)
    nreq = args.serve_requests
    sizes = np.floor(2.0 ** rng.uniform(10, 26, nreq)).astype(np.int64)
    kinds = rng.choice([gqsort.U32, gqsort.F32, gqsort.F64], size=nreq, p=[0.6, 0.3, 0.1])
    kb = np.array([gqsort.KEY_BYTES[int(k)] for k in kinds], np.int64)
    offs = np.zeros(nreq + 1, np.int64)
    offs[1:] = np.cumsum(sizes * kb)
    total_bytes = int(offs[-1])
    ctx = gqsort.Context(0)
    ins = ctypes.c_void_p()
    outs = ctypes.c_void_p()
    gqsort._check(lib.gqs_host_alloc(ctx.handle, total_bytes, ctypes.byref(ins)), ctx)
    gqsort._check(lib.gqs_host_alloc(ctx.handle, total_bytes, ctypes.byref(outs)), ctx)
    for i in range(nreq):
        gen = gqsort.GENERATORS["uniform" if kinds[i] == gqsort.U32 else "normal"]
        gqsort._check(lib.gqs_generate(int(kinds[i]), gen, 5000 + i, ins.value + int(offs[i]), int(sizes[i])))
    link = pcie_rates(ctx, 1 << 30)

    def run_plain(arrivals):
        lat = np.empty(nreq)
        t0 = time.perf_counter()
        for i in range(nreq):
            while time.perf_counter() - t0 < arrivals[i]:
                pass
            gqsort._check(lib.gqs_sort(ctx.handle, int(kinds[i]), ins.value + int(offs[i]), outs.value + int(offs[i]),
                                       int(sizes[i])), ctx)
            lat[i] = time.perf_counter() - t0 - arrivals[i]
        return time.perf_counter() - t0, lat

    srv = gqsort.Server(0, hbm_budget_bytes=args.serve_budget, slots=args.serve_slots or None,
                        batch_bytes=args.serve_batch_bytes or None)
    srv_cfg = srv.config

    def run_server(arrivals):
        st0 = srv.stats()
        tickets = []
        t0 = time.perf_counter()
        for i in range(nreq):
            while time.perf_counter() - t0 < arrivals[i]:
                pass
            tickets.append(srv.submit(int(kinds[i]), ins.value + int(offs[i]), outs.value + int(offs[i]),
                                      int(sizes[i])))
        for t in tickets:
            srv.wait(t)
        wall = time.perf_counter() - t0
        lat = np.array([(lambda s, a, d: d - s)(*srv.times(t)) / 1e6 for t in tickets])
        for t in tickets:
            srv.release(t)
        st = srv.stats()
        return wall, lat, {k: st[k] - st0[k] for k in ("batches", "completed", "failed")}

    def poisson(rate, seed):
        arr = np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, nreq))
        return arr - arr[0]

    p99 = lambda lat: float(np.percentile(lat, 99))  # noqa: E731
    ms = lambda x: round(x * 1e3, 3)  # noqa: E731
    try:
        # warm both paths once on a few requests
        for i in range(min(8, nreq)):
            gqsort._check(lib.gqs_sort(ctx.handle, int(kinds[i]), ins.value + int(offs[i]),
                                       outs.value + int(offs[i]), int(sizes[i])), ctx)
        run_server(np.zeros(nreq))
        zero = np.zeros(nreq)
        wall_p, _ = run_plain(zero)  # saturated: everything offered at t=0
        wall_s, _, st_s = run_server(zero)
        cap_p, cap_s = nreq / wall_p, nreq / wall_s
        # the latency bound: plain's p99 at 80% of its saturated rate
        _, lat_ref = run_plain(poisson(0.8 * cap_p, 11))
        bound = p99(lat_ref)
        # each offered rate: three Poisson streams (seeds differ); the median p99 decides
        sweep = {"plain": [], "server": []}
        for name, runner, factors in (("plain", run_plain, (0.8, 0.9, 1.0)),
                                      ("server", run_server, (0.8, 1.2, 1.6, 2.0, 2.2, 2.4, 2.6, 2.8, 3.0, 3.2))):
            for f in factors:
                rate = f * cap_p
                if name == "server" and rate > 1.05 * cap_s:
                    break
                reps = [runner(poisson(rate, 100 * rep + 12 + int(f * 10))) for rep in range(3)]
                p99s = sorted(p99(r[1]) for r in reps)
                sweep[name].append({"offered_per_s": round(rate, 1),
                                    "achieved_per_s": round(float(np.median([nreq / r[0] for r in reps])), 1),
                                    "p50_ms": ms(float(np.median([np.median(r[1]) for r in reps]))),
                                    "p99_ms": ms(p99s[1]), "p99_ms_each": [ms(x) for x in p99s]})
                if name == "server":
                    sweep[name][-1]["batches"] = int(np.median([r[2]["batches"] for r in reps]))
        sustain = {k: max([r["offered_per_s"] for r in v if r["p99_ms"] <= ms(bound)] or [0.0])
                   for k, v in sweep.items()}
        # every request's output: sorted, and the same multiset as its input (last server run)
        ok = True
        for i in range(nreq):
            ct = {gqsort.U32: ctypes.c_uint32, gqsort.F32: ctypes.c_float, gqsort.F64: ctypes.c_double}[int(kinds[i])]
            n = int(sizes[i])
            a_in = np.ctypeslib.as_array((ct * n).from_address(ins.value + int(offs[i])))
            a_out = np.ctypeslib.as_array((ct * n).from_address(outs.value + int(offs[i])))
            if i % 8 == 0 or n < (1 << 16):
                so, su = multiset_ok(a_in, a_out)
                ok = ok and so and su
        # every key crosses the link once each way, both directions at once
        link_bound = nreq / (total_bytes / link["duplex"] / 1e9)
    finally:
        srv.close()
        lib.gqs_host_free(ctx.handle, ins.value)
        lib.gqs_host_free(ctx.handle, outs.value)
        ctx.close()
    return {
        "workload": f"{nreq} requests, sizes floor(2^U) U~U[10,26], dtypes 60/30/10 u32/f32/f64, pinned host buffers",
        "keys": int(sizes.sum()), "bytes_each_way": total_bytes, "budget_bytes": args.serve_budget,
        "server_config": {"slots": srv_cfg.slots, "batch_bytes": srv_cfg.batch_bytes},
        "saturated": {"server_sorts_per_s": round(cap_s, 1), "plain_sorts_per_s": round(cap_p, 1),
                      "speedup": round(wall_p / wall_s, 2), "server_batches": st_s["batches"]},
        "p99_bound_ms": ms(bound),
        "sweep": sweep,
        "sustainable_per_s": sustain,
        "ratio_at_equal_p99": round(sustain["server"] / sustain["plain"], 2) if sustain["plain"] else None,
        "pcie_pinned_GB_s": {k: round(v, 1) for k, v in link.items()},
        "pcie_ceiling_per_s": round(link_bound, 1),
        "pcie_ceiling_ratio": round(link_bound / sustain["plain"], 2) if sustain["plain"] else None,
        "outputs_checked": ok,
    }


def organ_pipe(n):
    """0, 1, .., n/2-1, n/2-1, .., 0: median-of-3 pivots (first, middle,
    last) of every segment a stable partition leaves are its two smallest
    keys and its largest, so an unbounded single-pivot schedule needs n/2
    levels."""
    h = np.arange(n // 2, dtype=np.uint32)
    return np.concatenate([h, h[::-1], np.zeros(n % 2, np.uint32)])


def bench_inputs(ctx, stream, ev0, ev1, d, dn, pristine, cases, steps):
    """Throughput of plain sorts over other generators and key types (same
    device buffers, each input restored from its pristine copy before every
    sort): Mkeys/s, fraction of the one-read-plus-one-write HBM roofline, and
    sortedness and multiset checksums of the last output.  A case may carry
    parameters (the classic single-pivot schedule, the reference's)."""
    lib = gqsort.load()
    out = []
    for case in cases:
        dname, gen, n = case[:3]
        params = case[3] if len(case) > 3 else None
        dt = DTYPES[dname]
        kb = gqsort.KEY_BYTES[dt]
        log(f"input {dname} {gen} n={n} {params or ''}")
        host = organ_pipe(n) if gen == "organ" else gqsort.generate(dt, gen, 1, n)
        pristine.upload(host, stream)
        if params:
            ctx.set_params(gqsort.default_params(dt), **params)
        ts = []
        try:
            for r in range(steps + 1):
                lib.gqs_memcpy_async(ctx.handle, d.ptr, pristine.ptr, n * kb, stream.handle)
                ev0.record(stream)
                ctx.sort_device(dt, d, dn, n, stream)
                ev1.record(stream)
                if r:
                    ts.append(ev0.elapsed_ms(ev1))
            st = ctx.stats()
        finally:
            if params:
                ctx.set_params(gqsort.default_params())
                ctx.tune_load(gqsort.TUNED_TABLE)
        res = d.download(host.dtype, n, stream)
        sorted_ok, sums_ok = multiset_ok(host, res)
        ms = float(np.median(ts))
        row = {"dtype": dname, "generator": gen, "n": n, "ms": round(ms, 4), "Mkeys_s": round(n / ms / 1e3, 1),
               "roofline_frac": round(2 * n * kb / (ms / 1e3) / HBM_PEAK, 4), "sorted": sorted_ok,
               "multiset_checksums": sums_ok}
        if params:
            row.update({"params": params, "levels_launched": st.levels_launched, "levels_used": st.levels_used,
                        "oversize_leaves": st.oversize_leaves})
        out.append(row)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 28)
    ap.add_argument("--gen", default="uniform")
    ap.add_argument("--dtype", default="u32", choices=sorted(DTYPES))
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-host", action="store_true", help="skip the host-pointer end-to-end timing")
    ap.add_argument("--host-reps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-inputs", action="store_true", help="skip the other generators and float keys")
    ap.add_argument("--input-steps", type=int, default=3)
    ap.add_argument("--batch-arrays", type=int, default=65536)
    ap.add_argument("--batch-total", type=int, default=1 << 28)
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-budget", type=int, default=64 << 30)
    ap.add_argument("--serve-slots", type=int, default=0, help="server batch slots (0: the library default)")
    ap.add_argument("--serve-batch-bytes", type=int, default=0, help="server batch target (0: the library default)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: across GPUs the sort would shard by a sample-sort exchange of "
              "key ranges, which is not implemented; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    dt = DTYPES[args.dtype]
    kb = gqsort.KEY_BYTES[dt]
    n = args.n
    nbytes = n * kb
    log(f"generating {n} {args.gen} {args.dtype} keys (seed {args.seed})")
    host = gqsort.generate(dt, args.gen, args.seed, n)

    ctx = gqsort.Context(0)
    info = ctx.device_info()
    info["report"] = ctx.device_report()
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
    out = None
    if not args.no_check:
        out = d.download(host.dtype, n, stream)
        sorted_ok, sums_ok = multiset_ok(host, out)
        check = {"sorted": sorted_ok, "multiset_checksums": sums_ok}
        if not (sorted_ok and sums_ok):
            log(f"CHECK FAILED: {check}")
            sys.exit(1)

    log("per-kernel profile pass")
    roof, kernels, level_keys = kernel_roofline(ctx, dt, d, dn, pristine, nbytes, n, stream, kb,
                                                f"{args.dtype}:{n}:{args.gen}")

    cpu = par = None
    if not args.no_cpu:
        # the box's CPU share for one GPU is 16 cores, whatever nproc reports
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baselines on the same {n}-key input (1 and {threads} threads)")
        cpu, par = cpu_baseline(host, out, threads)
        if out is not None and not cpu["gpu_equals_std_sort"]:
            log("CHECK FAILED: GPU output differs from std::sort")
            sys.exit(1)
    out = None

    e2e = None
    if not args.no_host:
        log("host-pointer end-to-end (upload + sort + download)")
        e2e = host_pointer_e2e(ctx, dt, host, args.host_reps)

    inputs = None
    if not args.no_inputs and n == 1 << 28 and args.dtype == "u32":
        # the other u32 generators at the same size, then f32 2^27 and f64 2^26
        # (N(0,1) with outliers and duplicates): one buffer size, 1 GiB / 512 MiB
        cases = [("u32", g, n) for g in ("gaussian", "sorted", "reverse", "equal", "bucket")]
        cases += [("f32", "normal", n >> 1), ("f64", "normal", n >> 2)]
        # the reference's schedule (one median-of-3 pivot per level, 1728-key
        # leaves), enqueued up front, on random keys and on an organ pipe,
        # which defeats median-of-3 pivots (its rescue levels finish it)
        classic = {"mode": gqsort.MODE_CLASSIC, "leaf_max": 1728, "max_levels": 40, "host_driven": 0}
        cases += [("u32", "uniform", n, classic), ("u32", "organ", n, classic), ("u32", "organ", n)]
        inputs = bench_inputs(ctx, stream, ev0, ev1, d, dn, pristine, cases, args.input_steps)
        if not all(c["sorted"] and c["multiset_checksums"] for c in inputs):
            log(f"CHECK FAILED: {inputs}")
            sys.exit(1)

    batched = serving = None
    if not args.no_batched:
        for b in (pristine, d, dn):
            b.free()
        log("batched segmented sort")
        batched = bench_batched(ctx, args, stream, ev0, ev1)
    if not args.no_serving:
        log("serving stream")
        serving = bench_serving(args)

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
        "roofline_sort": {"bound": "hbm", "achieved": round(2 * nbytes / (ms_per_step / 1e3) / 1e9, 1),
                          "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(sort_frac, 4),
                          "traffic": None, "what": "whole sort: 2*N*sizeof(key) / time per sort"},
        "cpu_baseline": cpu,
        "cpu_parallel": par,
        "host": cpu_info(),
        "host_pointer_e2e": e2e,
        "kernels_ms": kernels,
        "level_keys": level_keys,
        "step_ms": {"min": round(min(times), 4), "median": round(float(np.median(times)), 4),
                    "max": round(max(times), 4)},
        "check": check,
        "inputs": inputs,
        "batched": batched,
        "serving": serving,
        "device": info,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
