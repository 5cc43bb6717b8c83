"""One partition level on the GPU against the closed forms of the oracle.

After one level of the reference's gqsort_kernel (CL12/QuicksortKernels.cl:
102-240) every record satisfies: keys of [oldstart, sstart) < pivot, the
pivot fill [sstart, send) == pivot, keys of [send, oldend) > pivot, and the
multiset of the record is preserved; sstart - start = #{< pivot} and
oldend - send = #{> pivot} are fixed by the input.  The k-way level has the
same invariants per bucket: the level's own splitters (read back by
gqs_partition_level) fix every bucket's size by brute force
(orc.kway_classify), every stored key must classify into the bucket whose
range holds it, and the stored keys plus the equality buckets' values are the
input's multiset.  Which splitters are drawn depends on sample positions; the
invariants hold for whatever they are.
"""
import numpy as np
import pytest

import gqsort
import orc

pytestmark = pytest.mark.gpu


def duplicate_heavy(n, distinct, seed):
    rng = np.random.default_rng(seed)
    w = 1.0 / np.arange(1, distinct + 1) ** 1.1
    vals = rng.choice(np.iinfo(np.uint32).max, size=distinct, replace=False).astype(np.uint32)
    return vals[rng.choice(distinct, size=n, p=w / w.sum())]


CASES = [
    # dtype, input, mode
    (gqsort.U32, ("uniform", 1 << 22), "multipivot"),
    (gqsort.U32, ("gaussian", (1 << 21) + 777), "multipivot"),
    (gqsort.U32, ("bucket", 1 << 21), "multipivot"),
    (gqsort.U32, ("sorted", 1 << 21), "multipivot"),
    (gqsort.U32, ("equal", 1 << 20), "multipivot"),
    (gqsort.U32, ("dups", 1 << 21), "multipivot"),
    (gqsort.F32, ("normal", (1 << 21) + 5), "multipivot"),
    (gqsort.F64, ("normal", 1 << 20), "multipivot"),
    (gqsort.U32, ("uniform", 1 << 20), "classic"),
    (gqsort.U32, ("dups", 1 << 20), "classic"),
    (gqsort.F64, ("normal", 300001), "classic"),
]


@pytest.mark.parametrize("dt,inp,mode", CASES)
def test_one_level_partition_invariants(gpu_ctx, dt, inp, mode):
    gen, n = inp
    a = duplicate_heavy(n, 40, 5) if gen == "dups" else gqsort.generate(dt, gen, 17, n)
    print(f"case level dt={dt} {gen} n={n} {mode}", flush=True)
    if mode == "classic":
        gpu_ctx.set_params(gqsort.default_params(dt), mode=gqsort.MODE_CLASSIC, leaf_max=1728)
    try:
        d = gqsort.DeviceBuffer.from_array(gpu_ctx, a)
        dn = gqsort.DeviceBuffer(gpu_ctx, a.nbytes)
        spl, eq, start, size, beq = gpu_ctx.partition_level(dt, d, dn, n)
        keys = orc.ordered_keys(a)
        kd = keys.dtype
        out = dn.download(kd, n)
        d.free()
        dn.free()
    finally:
        if mode == "classic":
            gpu_ctx.set_params(gqsort.default_params())
            gpu_ctx.tune_load(gqsort.TUNED_TABLE)
    m = spl.size
    # splitters: strictly increasing; classic mode has one pivot with its fill bucket
    assert 1 <= m <= 255 and np.all(spl[1:] > spl[:-1])
    if mode == "classic":
        assert m == 1 and eq[0] == 1
    nb = m + 1 + int(eq.sum())
    assert start.size == nb
    # bucket layout: contiguous, covering [0, n), equality buckets right after their splitter's range bucket
    assert start[0] == 0 and np.all(start[1:] == start[:-1] + size[:-1]) and int(size.sum()) == n
    exp_eq = []
    for i in range(m):
        exp_eq += [0, 1] if eq[i] else [0]
    exp_eq.append(0)
    assert np.array_equal(beq.astype(int), np.array(exp_eq))
    # sizes: the brute-force counts of the input against the level's own splitters
    ids_in, counts = orc.kway_classify(keys, spl, eq)
    assert counts.size == nb and np.array_equal(counts.astype(np.int64), size.astype(np.int64)), \
        "bucket sizes differ from the closed-form counts"
    # every stored key lies in its bucket's range; equality ranges get their splitter value
    owner = np.repeat(np.arange(nb), size.astype(np.int64))
    eq_spl = {}
    b = 0
    for i in range(m):
        b += 1
        if eq[i]:
            eq_spl[b] = spl[i]
            b += 1
    filled = out.copy()
    for bk, v in eq_spl.items():
        filled[int(start[bk]):int(start[bk]) + int(size[bk])] = v
    ids_out, _ = orc.kway_classify(filled, spl, eq)
    bad = np.flatnonzero(ids_out.astype(np.int64) != owner)
    assert bad.size == 0, f"{bad.size} keys outside their bucket, first at {bad[:5]}"
    # multiset: the level only moved keys
    got = orc.std_sort(filled)
    assert orc.mismatch(got, orc.std_sort(keys))[0] == 0, "the level changed the key multiset"
    if mode == "classic":
        p = spl[0]
        # the reference's closed forms: #{< p}, #{== p}, #{> p}
        assert list(size) == [int((keys < p).sum()), int((keys == p).sum()), int((keys > p).sum())]
