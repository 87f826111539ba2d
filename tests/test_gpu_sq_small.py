"""Plain SQ search (best-first, (distance, id) order) at the small SQ
workload's shape: 100,000 x 128 vectors from a 256-component Gaussian
mixture (sigma 8), SQ8 global-scale codes, degree-32 graph (exact 64-NN +
RNG prune, the reference's search-build run on the device), entry = medoid
with the reference's 32 LCG seeds, 1,000 queries, L=64, k=10.  Every
query's ids, distance bits and counters equal the oracle's (the C
restatement pinned to the reference's scalar search)."""
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


def test_sq8_100k_matches_oracle():
    say("case sq8 100k: building")
    base, queries = data.workload(100_000, 128, 256, 8.0, 1000)
    graph = rindex.build_graph_exact(base, 32, 64, 42, 1)
    entry = rindex.medoid(base)
    scale, zero = rindex.sq_params(base, 8)
    codes = rindex.encode_sq(base, 8, scale, zero)
    calc = Calculator()
    try:
        calc.rover_configure(l1_bits=8)
        calc.retrieval_load(graph, codes, 128, entry, scale, zero)
        say("case sq8 100k: searching")
        ids, d, cnt, st = calc.search_batch(queries, 64, 10, method=native.METHOD_SQ,
                                            entry_points=32, expand_count=32, stats=True,
                                            hash_log2=14,
                                            traversal=native.TRAVERSAL_BEST_FIRST)
        oix = pyoracle.Index(graph, sq_codes=codes, sq_bits=8, sq_scale=float(scale),
                             sq_zero=float(zero), dim=128)
        seeds = pyoracle.seed_ids(100_000, 32, entry, 42)
        oids, od, ocnt, ost = pyoracle.search_batch(oix, pyoracle.SCORE_SQ, queries, seeds,
                                                    32, 64, 10)
        assert np.array_equal(cnt, ocnt)
        assert np.array_equal(ids, oids)
        assert np.array_equal(d.view(np.uint32), od.view(np.uint32))
        assert np.array_equal(st.astype(np.int64), ost.astype(np.int64))
        gt = data.ground_truth(base, queries, k=10)
        say("sq8 100k: recall@10 %.4f, mean expanded %.1f"
            % (data.recall_at_k(ids, cnt, gt, 10), st[:, 2].mean()))
        # the IDL op for single queries: the DSP search op's distance-only
        # traversal over the same seeds
        for qi in (0, 500, 999):
            i1, d1 = calc.retrieval_search(queries[qi], 32, 32, 64, 10)
            oi, od1, _ = pyoracle.sqmd_search(oix, queries[qi], seeds, 32, 32, 64, 10)
            assert np.array_equal(i1, oi)
            assert np.array_equal(d1.view(np.uint32), od1.view(np.uint32))
    finally:
        calc.close()
