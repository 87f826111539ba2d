"""Batch search captured in a HIP graph, and the autotuner.

Every graph replay is checked against the eager batch search of the same
inputs (ids, distance bits, counts), with eager work and synchronizations
between replays; the live query count and the query buffer are changed
between replays without re-capturing.  The autotuner must keep results
identical to the default launch configuration."""
import sys

import numpy as np
import pytest

from rover_amd import data, native
from rover_amd import index as rindex
from rover_amd.search import Calculator

pytestmark = pytest.mark.gpu

NQ, K, L = 300, 10, 48


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    say("case graph index")
    base, queries = data.workload(8000, 128, 32, 8.0, 2 * NQ, base_seed=61, query_seed=62)
    ix = rindex.build_pq_index(base, m=32, R=32, res_bits=8, iters=6, train_max=8000,
                               n_router=64)
    calc = Calculator()
    calc.load_pq_index(ix, with_router=False)
    rows = rindex.route_on_host(queries, ix.router_centroids, ix.router_ids, 4)
    yield calc, ix, queries, rows
    calc.close()


class Bufs:
    def __init__(self, q, rows):
        self.dq = native.DeviceBuffer.from_array(np.ascontiguousarray(q, dtype=np.float32))
        self.de = native.DeviceBuffer.from_array(np.ascontiguousarray(rows, dtype=np.int32))
        self.ids = native.DeviceBuffer(NQ * K * 4)
        self.d = native.DeviceBuffer(NQ * K * 4)
        self.cnt = native.DeviceBuffer(NQ * 4)

    def clear(self):
        self.ids.upload(np.full(NQ * K, -7, dtype=np.int32))
        self.d.upload(np.zeros(NQ * K, dtype=np.float32))
        self.cnt.upload(np.full(NQ, -9, dtype=np.int32))

    def get(self):
        return (self.ids.download(np.int32, NQ * K).reshape(NQ, K),
                self.d.download(np.float32, NQ * K).reshape(NQ, K),
                self.cnt.download(np.int32, NQ))


def eager(calc, q, rows, correction=1, use_rows=True):
    b = Bufs(q, rows)
    calc.search_batch_device(b.dq.ptr, NQ, L, K, b.ids.ptr, b.d.ptr, b.cnt.ptr, K,
                             correction=correction,
                             d_entry_ids=b.de.ptr if use_rows else None,
                             entry_count=4 if use_rows else 0,
                             entry_stride=4 if use_rows else 0)
    native.synchronize()
    return b.get()


def same(a, b, n=NQ):
    ia, da, ca = a
    ib, db, cb = b
    assert np.array_equal(ca[:n], cb[:n])
    assert np.array_equal(ia[:n], ib[:n])
    assert np.array_equal(da[:n].view(np.uint32), db[:n].view(np.uint32))


@pytest.mark.parametrize("correction,use_rows", [(1, True), (0, True), (1, False)])
def test_graph_replays_match_eager(setup, correction, use_rows):
    calc, ix, queries, rows = setup
    say("case graph correction=%d rows=%s" % (correction, use_rows))
    q1, r1 = queries[:NQ], rows[:NQ]
    q2, r2 = queries[NQ:], rows[NQ:]
    ref1 = eager(calc, q1, r1, correction, use_rows)
    ref2 = eager(calc, q2, r2, correction, use_rows)
    assert (ref1[2] == K).all()
    b = Bufs(q1, r1)
    stream = native.Stream()
    live = native.DeviceBuffer.from_array(np.array([NQ], dtype=np.int32))
    g = calc.search_graph_create(b.dq.ptr, NQ, L, K, b.ids.ptr, b.d.ptr, b.cnt.ptr, K,
                                 stream.ptr, d_num_queries=live.ptr, correction=correction,
                                 d_entry_ids=b.de.ptr if use_rows else None,
                                 entry_count=4 if use_rows else 0,
                                 entry_stride=4 if use_rows else 0)
    try:
        for rep in range(3):
            b.clear()
            # the program's own eager work between replays
            other = eager(calc, q2, r2, correction, use_rows)
            same(other, ref2)
            calc.search_graph_launch(g, stream.ptr)
            stream.synchronize()
            same(b.get(), ref1)
        # a smaller live batch: the first 137 queries, the rest untouched
        live.upload(np.array([137], dtype=np.int32))
        b.clear()
        calc.search_graph_launch(g, stream.ptr)
        stream.synchronize()
        got = b.get()
        same(got, ref1, 137)
        assert (got[2][137:] == -9).all()
        # new queries (and rows) in the same buffers, full batch again
        live.upload(np.array([NQ], dtype=np.int32))
        b.dq.upload(np.ascontiguousarray(q2, dtype=np.float32))
        b.de.upload(np.ascontiguousarray(r2, dtype=np.int32))
        b.clear()
        calc.search_graph_launch(g, stream.ptr)
        stream.synchronize()
        same(b.get(), ref2)
    finally:
        calc.search_graph_destroy(g)


def test_graph_exact_visited_sets(setup):
    """A graph captured with exact visited sets (its own HBM sets) replays
    the eager results while eager exact-set launches (the session's sets)
    run between the replays; the allocation-free entry point refuses them."""
    calc, ix, queries, rows = setup
    say("case graph exact visited sets")
    q1, r1 = queries[:NQ], rows[:NQ]
    q2, r2 = queries[NQ:], rows[NQ:]
    ref1 = eager(calc, q1, r1, 1, True)
    ref2 = eager(calc, q2, r2, 1, True)
    b = Bufs(q1, r1)
    b2 = Bufs(q2, r2)
    stream = native.Stream()
    g = calc.search_graph_create(b.dq.ptr, NQ, L, K, b.ids.ptr, b.d.ptr, b.cnt.ptr, K,
                                 stream.ptr, correction=1, d_entry_ids=b.de.ptr,
                                 entry_count=4, entry_stride=4, hash_log2=native.HASH_EXACT)
    try:
        for rep in range(3):
            b.clear()
            b2.clear()
            calc.search_batch_device(b2.dq.ptr, NQ, L, K, b2.ids.ptr, b2.d.ptr, b2.cnt.ptr, K,
                                     correction=1, d_entry_ids=b2.de.ptr, entry_count=4,
                                     entry_stride=4, hash_log2=native.HASH_EXACT,
                                     grid_waves=7 + rep)
            native.synchronize()
            same(b2.get(), ref2)
            calc.search_graph_launch(g, stream.ptr)
            stream.synchronize()
            same(b.get(), ref1)
    finally:
        calc.search_graph_destroy(g)
    with pytest.raises(native.RoverError) as e:
        calc.search_batch_device(b.dq.ptr, NQ, L, K, b.ids.ptr, b.d.ptr, b.cnt.ptr, K,
                                 correction=1, d_entry_ids=b.de.ptr, entry_count=4,
                                 entry_stride=4, hash_log2=native.HASH_EXACT, noalloc=True)
    assert e.value.code == native.ROVER_EBADPARM


def test_graph_uneven_batch_matches_eager(setup):
    """Per-query beam width and k from device arrays, captured once and
    replayed with a new batch (sizes, L, k, queries) in the same buffers."""
    calc, ix, queries, rows = setup
    say("case graph uneven")
    rng = np.random.default_rng(5)
    KMAX, LMAX = 100, 256
    stream = native.Stream()

    def draw(n):
        k = rng.choice(np.array([1, 10, 100], dtype=np.int32), n)
        lo = np.maximum(16, k)
        L_ = (lo + rng.integers(0, LMAX + 1 - lo)).astype(np.int32)
        sel = rng.integers(0, len(queries), n)
        return (np.ascontiguousarray(queries[sel], dtype=np.float32),
                np.ascontiguousarray(rows[sel], dtype=np.int32), L_, k.astype(np.int32))

    dq = native.DeviceBuffer(NQ * 128 * 4)
    de = native.DeviceBuffer(NQ * 4 * 4)
    dL = native.DeviceBuffer(NQ * 4)
    dK = native.DeviceBuffer(NQ * 4)
    ids = native.DeviceBuffer(NQ * KMAX * 4)
    dd = native.DeviceBuffer(NQ * KMAX * 4)
    cnt = native.DeviceBuffer(NQ * 4)
    live = native.DeviceBuffer(4)

    def load(q, r, L_, k):
        n = len(q)
        dq.upload(q)
        de.upload(r)
        dL.upload(L_)
        dK.upload(k)
        live.upload(np.array([n], dtype=np.int32))
        ids.upload(np.full(NQ * KMAX, -1, dtype=np.int32))
        cnt.upload(np.full(NQ, -9, dtype=np.int32))

    def result(n):
        return (ids.download(np.int32, NQ * KMAX).reshape(NQ, KMAX)[:n],
                dd.download(np.float32, NQ * KMAX).reshape(NQ, KMAX)[:n],
                cnt.download(np.int32, NQ)[:n])

    g = calc.search_graph_create(dq.ptr, NQ, LMAX, KMAX, ids.ptr, dd.ptr, cnt.ptr, KMAX,
                                 stream.ptr, d_num_queries=live.ptr, correction=1,
                                 d_pool_sizes=dL.ptr, d_num_outs=dK.ptr, d_entry_ids=de.ptr,
                                 entry_count=4, entry_stride=4)
    try:
        for n in (NQ, 1, 173):
            batch = draw(n)
            load(*batch)
            calc.search_graph_launch(g, stream.ptr)
            stream.synchronize()
            got = result(n)
            load(*batch)
            calc.search_batch_device(dq.ptr, n, LMAX, KMAX, ids.ptr, dd.ptr, cnt.ptr, KMAX,
                                     correction=1, d_pool_sizes=dL.ptr, d_num_outs=dK.ptr,
                                     d_entry_ids=de.ptr, entry_count=4, entry_stride=4)
            native.synchronize()
            ref = result(n)
            assert np.array_equal(ref[2], batch[3])  # every query returns its k
            assert np.array_equal(got[2], ref[2])
            for i in range(n):
                kq = int(ref[2][i])
                assert np.array_equal(got[0][i, :kq], ref[0][i, :kq])
                assert np.array_equal(got[1][i, :kq].view(np.uint32),
                                      ref[1][i, :kq].view(np.uint32))
    finally:
        calc.search_graph_destroy(g)


def test_graph_rejects_default_stream(setup):
    calc, ix, queries, rows = setup
    b = Bufs(queries[:NQ], rows[:NQ])
    with pytest.raises(native.RoverError):
        calc.search_graph_create(b.dq.ptr, NQ, L, K, b.ids.ptr, b.d.ptr, b.cnt.ptr, K, None)


def test_autotune_keeps_results(setup):
    calc, ix, queries, rows = setup
    say("case autotune")
    ref = eager(calc, queries[:NQ], rows[:NQ], 1, True)
    b = Bufs(queries[:NQ], rows[:NQ])
    b.clear()
    t = calc.autotune(b.dq.ptr, NQ, L, K, b.ids.ptr, b.d.ptr, b.cnt.ptr, K, repeats=2,
                      correction=1, d_entry_ids=b.de.ptr, entry_count=4, entry_stride=4)
    native.synchronize()
    say("tuning: hash 2^%d probes %d grid %d: %.3f ms (default %.3f ms), %d tried, %d "
        "rejected" % (t.hash_log2, t.probe_limit, t.grid_waves, t.ms, t.default_ms, t.tried,
                      t.rejected))
    assert t.tried > 10 and t.rejected == 0
    assert 0 < t.ms <= t.default_ms
    same(b.get(), ref)
    # the chosen configuration reproduces the results as a plain batch setting
    b.clear()
    calc.search_batch_device(b.dq.ptr, NQ, L, K, b.ids.ptr, b.d.ptr, b.cnt.ptr, K,
                             correction=1, d_entry_ids=b.de.ptr, entry_count=4,
                             entry_stride=4, hash_log2=t.hash_log2, probe_limit=t.probe_limit,
                             grid_waves=t.grid_waves)
    native.synchronize()
    same(b.get(), ref)
