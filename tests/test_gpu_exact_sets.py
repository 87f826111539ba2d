"""Exact visited sets in HBM (ROVER_HASH_EXACT, the mode the launch tuner
picks for the large workloads) against the LDS hash and the oracle: a lossy
hash may re-score a node but never changes a result, so both must give the
same ids, distance bits, counts and expansion counts, and the exact sets
also the oracle's counters.  Per-query L and k down to pools of one entry
(every expansion evicts), the final refine, and the learned stop with a
random predictor (the same features at the same steps, so the same stops)."""
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
def mid():
    say("case exact sets: 12k x 128 index")
    base, queries = data.workload(12000, 128, 64, 8.0, 300)
    ix = rindex.build_pq_index(base, m=32, R=32, res_bits=8, iters=6,
                               train_max=12000, n_router=0)
    c = Calculator()
    c.load_pq_index(ix)
    yield c, ix, queries
    c.close()


def same_results(a, b, stats_equal):
    assert np.array_equal(a[2], b[2])
    for i in range(len(a[2])):
        c = int(a[2][i])
        assert np.array_equal(a[0][i, :c], b[0][i, :c]), i
        assert np.array_equal(bits(a[1][i, :c]), bits(b[1][i, :c])), i
    if len(a) > 3:
        assert np.array_equal(a[3][:, 2], b[3][:, 2])  # expansions
        if stats_equal:
            assert np.array_equal(a[3], b[3])


@pytest.mark.parametrize("mode", [0, 1])
def test_exact_sets_match_oracle_and_hash(mid, mode):
    """Per-query L (1..300) and k, the whole pool compared for L = k."""
    calc, ix, queries = mid
    oix = pyoracle.Index(ix.graph, ix.codes, ix.codebook, ix.res_codes, ix.res_qparams,
                         ix.res_bits)
    seeds = native.make_seed_ids(ix.graph.shape[0], 32, ix.entry)
    rng = np.random.default_rng(17)
    for case in ("tiny", "mixed", "whole"):
        if case == "tiny":
            Ls = rng.choice([1, 2, 3, 5, 10], size=len(queries)).astype(np.int32)
            ks = Ls.copy()
        elif case == "mixed":
            ks = rng.choice([1, 10, 100], size=len(queries)).astype(np.int32)
            Ls = np.array([rng.integers(k, 301) for k in ks], dtype=np.int32)
        else:
            Ls = np.full(len(queries), 200, np.int32)
            ks = Ls.copy()
        say("case exact sets mode %d %s" % (mode, case))
        got = calc.search_batch(queries, Ls, ks, correction=mode, entry_ids=seeds,
                                expand_count=32, stats=True, hash_log2=native.HASH_EXACT)
        hsh = calc.search_batch(queries, Ls, ks, correction=mode, entry_ids=seeds,
                                expand_count=32, stats=True, hash_log2=12)
        same_results(got, hsh, stats_equal=False)
        scorer = pyoracle.SCORE_ONLINE if mode else pyoracle.SCORE_ADC
        ref = pyoracle.search_batch(oix, scorer, queries, seeds, 32, Ls, ks, threads=8)
        same_results(got, ref, stats_equal=True)


def test_exact_sets_refine_matches_hash(mid):
    calc, ix, queries = mid
    for L in (20, 128):
        say("case exact sets refine L=%d" % L)
        a = calc.search_batch(queries, L, 10, correction=1, entry_points=32,
                              refine_candidates=20, hash_log2=native.HASH_EXACT)
        b = calc.search_batch(queries, L, 10, correction=1, entry_points=32,
                              refine_candidates=20, hash_log2=12)
        same_results(a, b, stats_equal=False)


def test_exact_sets_learned_stop_matches_hash(mid):
    """A random predictor, evaluated every 1, 4 or 12 expansions, stops each
    query at the same step with either visited set."""
    calc, ix, queries = mid
    P = native.lib().rover_mlp_param_count()
    rng = np.random.default_rng(23)
    params = (rng.normal(size=P) * 0.3).astype(np.float32)
    mean = np.zeros(16, np.float32)
    istd = np.ones(16, np.float32)
    for every, min_steps in ((4, 4), (12, 12), (1, 1)):
        native.check(native.lib().rover_learned_stop_set(
            calc.h, params.ctypes.data, P, mean.ctypes.data, istd.ctypes.data, every, 10,
            min_steps), "rover_learned_stop_set")
        for thr in (0.5, 0.9):
            say("case exact sets learned every %d threshold %.1f" % (every, thr))
            a = calc.search_batch(queries, 200, 10, correction=1, entry_points=32,
                                  stats=True, learned_stop=1, stop_threshold=thr,
                                  hash_log2=native.HASH_EXACT)
            b = calc.search_batch(queries, 200, 10, correction=1, entry_points=32,
                                  stats=True, learned_stop=1, stop_threshold=thr,
                                  hash_log2=12)
            same_results(a, b, stats_equal=False)
            plain = calc.search_batch(queries, 200, 10, correction=1, entry_points=32,
                                      stats=True, hash_log2=native.HASH_EXACT)
            say("  expansions %.1f (plain %.1f)" % (a[3][:, 2].mean(), plain[3][:, 2].mean()))
