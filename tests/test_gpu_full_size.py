"""The benchmark configurations at their own size on the GPU, each against
the oracle (ids, distance bits and counters):

  1M x 128 PQ (32 x 256 codes, the reference's exact search-build graph of
  degree 32, the reference's seeds: medoid + 31 LCG ids):
    plain search with 8-bit and 4-bit residuals, ADC ranking and online
    correction, L in {32, 64, 128, 256}, 1,000 queries;
    uneven batches (sizes 1, 37 and 4096; per-query L in [max(16, k), 512];
    k in {1, 10, 100}) through one call;
    the learned stop: a threshold no probability reaches gives the plain
    search bit for bit, and a trained predictor cuts expansions at the same
    L with recall@10 within 0.02.
The counters compared are the distinct nodes scored / expanded (a visited
hash of 2^14 slots and 64 probes drops nothing at these sizes)."""
import ctypes as C
import sys

import numpy as np
import pytest

import pyoracle
from rover_amd import data, native
from rover_amd import index as rindex
from rover_amd.search import Calculator

pytestmark = pytest.mark.gpu
N, DIM, NQ, K = 1_000_000, 128, 2000, 10
EXACT = dict(hash_log2=14, probe_limit=64)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def big():
    import torch
    dev = torch.device("cuda:0")
    say("case full size: 1M x 128 index build")
    base, queries = data.workload(N, DIM, 1024, 8.0, NQ)
    graph = rindex.build_graph_exact(base, 32, 64, 42, 1)
    ix8 = rindex.build_pq_index(base, m=32, R=32, res_bits=8, iters=20, train_max=100000,
                                device=dev, graph=graph)
    _, res4, qp4 = rindex.encode_on_device(base, ix8.codebook, 4)
    ix4 = rindex.PqIndexData(graph, ix8.entry, ix8.codes, ix8.codebook, res4, qp4, 4)
    gt = data.ground_truth(base, queries, k=100, device=dev)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    seeds = native.make_seed_ids(N, 32, ix8.entry)
    calc = Calculator()
    yield calc, {8: ix8, 4: ix4}, base, queries, gt, seeds
    calc.close()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.mark.parametrize("res_bits", [8, 4])
def test_cfg1_plain_search_matches_oracle(big, res_bits):
    calc, ixs, base, queries, gt, seeds = big
    ix = ixs[res_bits]
    calc.load_pq_index(ix, with_router=False)
    oix = pyoracle.Index.from_pq(ix)
    q = queries[:1000]
    for mode in (0, 1):
        for L in (32, 64, 128, 256):
            say("case cfg1 res%d mode %d L=%d" % (res_bits, mode, L))
            ids, d, cnt, st = calc.search_batch(q, L, K, correction=mode, entry_points=32,
                                                stats=True, **EXACT)
            scorer = pyoracle.SCORE_ONLINE if mode else pyoracle.SCORE_ADC
            oids, od, ocnt, ost = pyoracle.search_batch(oix, scorer, q, seeds, 32, L, K,
                                                        threads=16)
            assert np.array_equal(cnt, ocnt)
            assert np.array_equal(ids, oids)
            assert np.array_equal(bits(d), bits(od))
            assert np.array_equal(st.astype(np.int64), ost.astype(np.int64))
            say("  recall@10 %.4f" % data.recall_at_k(ids, cnt, gt[:1000], K))
            if mode == 1 and L in (64, 256):
                # exact visited sets in HBM
                ids, d, cnt, st = calc.search_batch(q, L, K, correction=mode,
                                                    entry_points=32, stats=True,
                                                    hash_log2=native.HASH_EXACT)
                assert np.array_equal(cnt, ocnt)
                assert np.array_equal(ids, oids)
                assert np.array_equal(bits(d), bits(od))
                assert np.array_equal(st.astype(np.int64), ost.astype(np.int64))


def test_cfg2_uneven_batches_match_oracle(big):
    calc, ixs, base, queries, gt, seeds = big
    ix = ixs[8]
    calc.load_pq_index(ix, with_router=False)
    oix = pyoracle.Index.from_pq(ix)
    rng = np.random.default_rng(21)
    pool = np.concatenate([queries, queries, queries])
    for size in (1, 37, 4096):
        q = np.ascontiguousarray(pool[:size])
        ks = rng.choice([1, 10, 100], size=size).astype(np.int32)
        Ls = np.array([rng.integers(max(16, k), 513) for k in ks], dtype=np.int32)
        say("case cfg2 uneven batch of %d" % size)
        ids, d, cnt, st = calc.search_batch(q, Ls, ks, correction=1, entry_points=32,
                                            stats=True, **EXACT)
        oids, od, ocnt, ost = pyoracle.search_batch(oix, pyoracle.SCORE_ONLINE, q, seeds, 32,
                                                    Ls, ks, threads=16)
        for h in (EXACT, dict(hash_log2=native.HASH_EXACT)):
            if h is not EXACT:
                ids, d, cnt, st = calc.search_batch(q, Ls, ks, correction=1, entry_points=32,
                                                    stats=True, **h)
            assert np.array_equal(cnt, ocnt) and np.array_equal(cnt, ks)
            for i in range(size):
                c = int(cnt[i])
                assert np.array_equal(ids[i, :c], oids[i, :c]), i
                assert np.array_equal(bits(d[i, :c]), bits(od[i, :c])), i
            assert np.array_equal(st.astype(np.int64), ost.astype(np.int64))


def test_cfg3_learned_stop_at_full_size(big):
    calc, ixs, base, queries, gt, seeds = big
    ix = ixs[8]
    calc.load_pq_index(ix, with_router=False)
    say("case cfg3 training")
    train_q = np.ascontiguousarray(data.gaussian_mixture(50_000, DIM, 1024, 8.0, 44,
                                                         centers_seed=42))
    t = native.RoverLearnedStopTrain()
    t.queries = train_q.ctypes.data
    t.num_queries = train_q.shape[0]
    t.pool_size = 512
    t.correction = 1
    t.label_k = 10
    t.entry_points = 32
    t.expand_count = 32
    t.stop_every = 12
    t.min_steps = 12
    t.trace_cap = 24
    t.steps = 800
    t.batch = 8192
    t.lr = 3e-3
    t.seed = 44
    native.check(native.lib().rover_learned_stop_train(calc.h, C.byref(t)), "train")
    q = queries[:1000]
    L = 512
    a = calc.search_batch(q, L, K, correction=1, entry_points=32, stats=True)
    b = calc.search_batch(q, L, K, correction=1, entry_points=32, stats=True,
                          learned_stop=1, stop_threshold=1.5)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(bits(a[1]), bits(b[1]))
    assert np.array_equal(a[3], b[3])
    c = calc.search_batch(q, L, K, correction=1, entry_points=32, stats=True,
                          learned_stop=1, stop_threshold=0.9)
    rp = data.recall_at_k(a[0], a[2], gt[:1000], K)
    rl = data.recall_at_k(c[0], c[2], gt[:1000], K)
    ep, el = a[3][:, 2].mean(), c[3][:, 2].mean()
    say("cfg3 at L=%d: plain recall %.4f, %.1f expansions; learned %.4f, %.1f" %
        (L, rp, ep, rl, el))
    assert rl >= rp - 0.02
    assert el <= 0.7 * ep
    # exact visited sets in HBM: the same stops, ids and bits; suspended
    # after 48 expansions and resumed in predicted order: the same again,
    # counters included
    ce = calc.search_batch(q, L, K, correction=1, entry_points=32, stats=True,
                           learned_stop=1, stop_threshold=0.9, hash_log2=native.HASH_EXACT)
    assert np.array_equal(c[0], ce[0])
    assert np.array_equal(bits(c[1]), bits(ce[1]))
    assert np.array_equal(c[3][:, 2], ce[3][:, 2])
    cr = calc.search_batch(q, L, K, correction=1, entry_points=32, stats=True,
                           learned_stop=1, stop_threshold=0.9, hash_log2=native.HASH_EXACT,
                           reorder_step=48)
    assert np.array_equal(ce[0], cr[0])
    assert np.array_equal(bits(ce[1]), bits(cr[1]))
    assert np.array_equal(ce[3], cr[3])
    # the two-turn pipeline: a threshold no probability reaches gives the
    # plain window-2 search bit for bit
    w = calc.search_batch(q, L, K, correction=1, entry_points=32, stats=True,
                          source_window=2)
    wl = calc.search_batch(q, L, K, correction=1, entry_points=32, stats=True,
                           learned_stop=1, stop_threshold=1.5, source_window=2)
    assert np.array_equal(w[0], wl[0])
    assert np.array_equal(bits(w[1]), bits(wl[1]))
    assert np.array_equal(w[3], wl[3])


def test_window2_at_full_size(big):
    """The live pipeline's two source turns in flight at 1M x 128: ids,
    distance bits and counters equal the oracle's window-2 search on 1,000
    queries (exact visited sets), L = 64 and 256."""
    calc, ixs, base, queries, gt, seeds = big
    ix = ixs[8]
    calc.load_pq_index(ix, with_router=False)
    oix = pyoracle.Index.from_pq(ix)
    q = queries[:1000]
    for L in (64, 256):
        say("case window 2 at 1M, L=%d" % L)
        ids, d, cnt, st = calc.search_batch(q, L, K, correction=1, entry_points=32,
                                            stats=True, hash_log2=native.HASH_EXACT,
                                            source_window=2)
        oids, od, ocnt, ost = pyoracle.search_batch(oix, pyoracle.SCORE_ONLINE, q, seeds, 32,
                                                    L, K, threads=16, window=2)
        assert np.array_equal(cnt, ocnt)
        assert np.array_equal(ids, oids)
        assert np.array_equal(bits(d), bits(od))
        assert np.array_equal(st.astype(np.int64), ost.astype(np.int64))
        say("  recall@10 %.4f" % data.recall_at_k(ids, cnt, gt[:1000], K))
