"""Serving loop: every result is the same alone and in the mix, and equal to
std::sort of its request (bitwise)."""
import threading

import numpy as np
import pytest

import gqsort
import orc

pytestmark = pytest.mark.gpu


def make_requests(count, seed, lo=2**10, hi=2**20):
    rng = np.random.default_rng(seed)
    sizes = np.floor(2.0 ** rng.uniform(np.log2(lo), np.log2(hi), count)).astype(np.int64)
    kinds = rng.choice([gqsort.U32, gqsort.F32, gqsort.F64], size=count, p=[0.6, 0.3, 0.1])
    reqs = []
    for i, (n, dt) in enumerate(zip(sizes, kinds)):
        gen = "uniform" if dt == gqsort.U32 else "normal"
        reqs.append(gqsort.generate(int(dt), gen, 1000 + i, int(n)))
    return reqs


def test_alone_vs_mix():
    reqs = make_requests(40, seed=7)
    with gqsort.Server(0, hbm_budget_bytes=1 << 30, slots=3) as srv:
        alone = []
        for a in reqs:
            t, out = srv.submit_array(a)
            srv.wait(t)
            srv.release(t)
            alone.append(out)
        # the mix: all submitted from two threads at once
        mixed = [None] * len(reqs)
        tickets = [None] * len(reqs)

        def submitter(idx):
            for i in idx:
                tickets[i], mixed[i] = srv.submit_array(reqs[i])

        th = [threading.Thread(target=submitter, args=(range(k, len(reqs), 2),)) for k in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for t in tickets:
            srv.wait(t)
        st = srv.stats()
        print("server stats", st, flush=True)
        assert st["completed"] == 2 * len(reqs) and st["failed"] == 0
        assert st["max_batch_requests"] > 1  # the mix was batched
    for i, a in enumerate(reqs):
        exp = orc.std_sort(a)
        assert orc.mismatch(alone[i], exp)[0] == 0, f"request {i} alone"
        assert orc.mismatch(mixed[i], alone[i])[0] == 0, f"request {i} differs alone vs mix"


def test_admission_limits():
    with gqsort.Server(0, hbm_budget_bytes=64 << 20, slots=2) as srv:
        cap = srv.config.slot_keys
        big = np.zeros(cap + 1, np.float64)
        with pytest.raises(gqsort.GQSError) as ei:
            srv.submit_array(big)
        assert ei.value.status == 5  # GQS_EBUSY: never fits a slot
        for n in (0, 1):
            a = np.arange(n, dtype=np.uint32)
            t, out = srv.submit_array(a)
            srv.wait(t)
            assert np.array_equal(out, a)


@pytest.mark.parametrize("slots,batch_bytes", [(1, 1 << 20), (2, 4 << 20), (8, 64 << 20)])
def test_stage_pipeline(slots, batch_bytes):
    """Batches move through upload, sort and download on three streams as the
    scheduler sees each stage finish: with one slot (every stage in turn),
    two, and the default eight; batch targets below most request sizes (each
    larger request a batch of its own) and above; requests of all three key
    types interleaved.  Every result equals std::sort of its request, and
    every request is marked done only after its keys are back."""
    reqs = make_requests(60, seed=21, lo=2**8, hi=2**21)
    with gqsort.Server(0, hbm_budget_bytes=1 << 30, slots=slots, batch_bytes=batch_bytes) as srv:
        print(f"case stage pipeline slots={slots} batch_bytes={batch_bytes}", flush=True)
        pending = [srv.submit_array(a) for a in reqs]
        for t, _ in pending:
            srv.wait(t)
            s, a, d = srv.times(t)
            assert s <= a <= d
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0 and st["pending"] == 0
        if batch_bytes <= 1 << 20:
            assert st["batches"] >= len([a for a in reqs if a.nbytes > batch_bytes])
    for i, (a, (_, out)) in enumerate(zip(reqs, pending)):
        assert orc.mismatch(out, orc.std_sort(a))[0] == 0, f"request {i} (n={a.size}, dtype={a.dtype})"
