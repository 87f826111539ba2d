"""Entry router on the GPU: the selected centroids equal an exact
float32 restatement (sequential sum over dimensions, ties by centroid
index), and a routed search equals the oracle run with the same per-query
entries."""
import sys

import numpy as np
import pytest

import pyoracle
from rover_amd import data
from rover_amd import index as rindex
from rover_amd.search import Calculator

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def routed():
    say("case router index")
    base, queries = data.workload(12000, 128, 64, 8.0, 200, base_seed=52, query_seed=53)
    ix = rindex.build_pq_index(base, m=32, R=32, res_bits=8, iters=6, train_max=12000,
                               n_router=200)
    calc = Calculator()
    calc.load_pq_index(ix)
    yield calc, ix, base, queries
    calc.close()


def exact_route(centroids, rep_ids, q, S):
    acc = np.zeros(len(centroids), dtype=np.float32)
    for d in range(q.shape[0]):
        diff = (np.float32(q[d]) - centroids[:, d]).astype(np.float32)
        acc = (acc + diff * diff).astype(np.float32)
    order = np.lexsort((np.arange(len(centroids)), acc))[:S]
    return rep_ids[order]


@pytest.mark.parametrize("S", [1, 4, 16])
def test_router_selection_exact(routed, S):
    calc, ix, base, queries = routed
    say("case router S=%d" % S)
    got = calc.route_entries(queries, S)
    expect = np.stack([exact_route(ix.router_centroids, ix.router_ids, q, S)
                       for q in queries])
    assert np.array_equal(got, expect)


@pytest.mark.parametrize("mode", [0, 1])
def test_routed_search_matches_oracle(routed, mode):
    calc, ix, base, queries = routed
    say("case routed search mode=%d" % mode)
    S = 4
    ents = calc.route_entries(queries, S)
    ids, d, cnt, st = calc.search_batch(queries, 48, 10, correction=mode,
                                        route_entries=S, stats=True, hash_log2=14)
    oix = pyoracle.Index.from_pq(ix)
    oids, od, ocnt, ost = pyoracle.search_batch(oix, mode, queries, ents, 32, 48, 10)
    assert np.array_equal(cnt, ocnt)
    assert np.array_equal(ids, oids)
    assert np.array_equal(d.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(st.astype(np.int64), ost.astype(np.int64))
    # the same rows passed explicitly give the same results
    ids2, d2, cnt2 = calc.search_batch(queries, 48, 10, correction=mode, entry_ids=ents)
    assert np.array_equal(ids2, ids) and np.array_equal(d2.view(np.uint32), d.view(np.uint32))


def test_router_refuses_more_centroids_than_tested(routed):
    """Routing with more centroids than the router is tested at is refused
    before anything is launched; the host-routed rows remain the way in."""
    from rover_amd import native
    calc, ix, base, queries = routed
    say("case router limit")
    big = np.ascontiguousarray(np.repeat(ix.router_centroids, 2, axis=0)[:257])
    reps = np.ascontiguousarray(np.resize(ix.router_ids, 257).astype(np.int32))
    calc.pure_pq_router_load(big, reps)
    try:
        with pytest.raises(native.RoverError) as e:
            calc.route_entries(queries[:8], 4)
        assert e.value.code == native.ROVER_EBADPARM
        assert "centroids" in str(e.value)
    finally:
        calc.pure_pq_router_load(ix.router_centroids, ix.router_ids)
