"""Graphs wider than a wavefront: rows of 96 slots load (the reference
checks no degree), a search reads min(expand_count, degree) of them per
expansion as the reference does (retrieval_sqmd_pure_pq.c:562-567,
824-825), and expansions of up to 64 neighbors equal the oracle's search
on the same graph (ids, distance bits, counters) in ADC and online modes;
a wider expansion is refused; an SQ index with a 96-slot graph loads and
searches."""
import sys

import numpy as np
import pytest

import pyoracle
from rover_amd import data, native
from rover_amd import index as rindex
from rover_amd.search import Calculator

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def wide():
    say("case wide graph: 12k x 128 index, 96-slot rows")
    base, queries = data.workload(12000, 128, 64, 8.0, 200)
    ix = rindex.build_pq_index(base, m=32, R=32, res_bits=8, iters=6,
                               train_max=12000, n_router=0)
    n = ix.graph.shape[0]
    rng = np.random.default_rng(47)
    extra = rng.integers(0, n, size=(n, 64)).astype(np.int32)
    graph = np.ascontiguousarray(np.concatenate([ix.graph, extra], axis=1), dtype=np.int32)
    c = Calculator()
    c.load_pq_index(ix)
    c.pure_pq_graph_load(graph, ix.entry)
    yield c, ix, graph, base, queries
    c.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_wide_graph_matches_oracle(wide, mode):
    calc, ix, graph, base, queries = wide
    oix = pyoracle.Index(graph, ix.codes, ix.codebook, ix.res_codes, ix.res_qparams,
                         ix.res_bits)
    seeds = native.make_seed_ids(graph.shape[0], 32, ix.entry)
    for expand in (16, 32, 48, 64):
        say("case wide graph mode %d expand %d" % (mode, expand))
        ids, d, cnt, st = calc.search_batch(queries, 96, 10, correction=mode, entry_ids=seeds,
                                            expand_count=expand, stats=True,
                                            hash_log2=native.HASH_EXACT)
        scorer = pyoracle.SCORE_ONLINE if mode else pyoracle.SCORE_ADC
        ref = pyoracle.search_batch(oix, scorer, queries, seeds, expand, 96, 10, threads=8)
        assert np.array_equal(cnt, ref[2])
        assert np.array_equal(ids, ref[0])
        assert np.array_equal(bits(d), bits(ref[1]))
        assert np.array_equal(st.astype(np.int64), ref[3].astype(np.int64))


def test_expansion_wider_than_a_wavefront_is_refused(wide):
    calc, ix, graph, base, queries = wide
    with pytest.raises(native.RoverError) as e:
        calc.search_batch(queries[:4], 96, 10, correction=1, entry_points=32, expand_count=96)
    assert e.value.code == native.ROVER_EBADPARM


def test_wide_sq_graph_loads_and_searches(wide):
    calc, ix, graph, base, queries = wide
    scale, zero = rindex.sq_params(base, 8)
    codes = rindex.encode_sq(base, 8, scale, zero)
    other = Calculator()
    try:
        other.rover_configure(l1_bits=8)
        other.retrieval_load(graph, codes, base.shape[1], ix.entry, scale, zero)
        ids, d, cnt = other.search_batch(queries, 64, 10, method=native.METHOD_SQ,
                                         entry_points=32, expand_count=32)
        assert np.all(cnt == 10)
    finally:
        other.close()
