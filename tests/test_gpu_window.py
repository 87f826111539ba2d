"""The source window on the GPU (RoverSearchBatch.source_window = 2): the
live pipeline with two source turns in flight (pure_pq_source_pipeline_step,
retrieval_sqmd_pure_pq.c:722-762) through the C-ABI, against the compiled
reference's pipeline (tests/golden/*_w2.npz) and the oracle's restatement
(orc_search_window): ids, distance bits and the turn counts, bit for bit, with
the exact visited sets and with LDS hashes (a lossy hash may re-score a node,
never change a result)."""
import ctypes as C
import sys

import numpy as np
import pytest

import pyoracle
from rover_amd import data, native
from rover_amd import index as rindex
from rover_amd.search import Calculator

pytestmark = pytest.mark.gpu
W2_SETS = ["pq_d128_m32", "pq_d128_m32_r8"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def pq_data(g):
    return rindex.PqIndexData(g["graph"], int(g["entry"]), g["codes"], g["codebook"],
                              g["res_codes"], g["res_qparams"], int(g["res_bits"]))


@pytest.fixture(scope="module")
def calc():
    c = Calculator()
    yield c
    c.close()


@pytest.mark.parametrize("hash_log2", [native.HASH_EXACT, 14, 4])
@pytest.mark.parametrize("name", W2_SETS)
def test_window2_matches_reference_pipeline(golden, calc, name, hash_log2):
    g = golden(name)
    w = golden(name + "_w2")
    say("case", name, "window 2, hash", hash_log2)
    calc.load_pq_index(pq_data(g), chunk_vectors=300)
    n = g["graph"].shape[0]
    for mode, ep, expand, L in w["search_cases"]:
        key = "m%d_e%d_x%d_L%d" % (mode, ep, expand, L)
        ids, d, cnt, st = calc.search_batch(g["queries"], int(L), int(L),
                                            correction=int(mode), entry_points=int(ep),
                                            expand_count=int(expand), stats=True,
                                            hash_log2=hash_log2,
                                            probe_limit=1 if hash_log2 == 4 else 0,
                                            source_window=2)
        seeds = len(np.unique(pyoracle.seed_ids(n, int(ep), int(g["entry"]), 42)))
        for qi in range(len(g["queries"])):
            c = int(w[key + "_counts"][qi])
            assert cnt[qi] == c, (key, qi, cnt[qi])
            assert np.array_equal(ids[qi, :c], w[key + "_ids"][qi][:c]), (key, qi)
            assert np.array_equal(bits(d[qi, :c]), bits(w[key + "_dists"][qi][:c])), (key, qi)
            turns, work = (int(x) for x in w[key + "_stats"][qi][:2])
            assert int(st[qi, 2]) == turns, (key, qi)
            if hash_log2 != 4:
                assert int(st[qi, 0]) == seeds + work, (key, qi)
            else:
                assert int(st[qi, 0]) >= seeds + work, (key, qi)


@pytest.fixture(scope="module")
def mid():
    say("case window 2: 12k x 128 index")
    base, queries = data.workload(12000, 128, 64, 8.0, 400)
    ix = rindex.build_pq_index(base, m=32, R=32, res_bits=8, iters=6,
                               train_max=12000, n_router=0)
    c = Calculator()
    c.load_pq_index(ix)
    yield c, ix, queries
    c.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_window2_matches_oracle(mid, mode):
    """12k points, L up to 300: ADC ranking and the online correction (the
    two-lane and grouped scorers), exact visited sets and an LDS hash."""
    calc, ix, queries = mid
    oix = pyoracle.Index(ix.graph, ix.codes, ix.codebook, ix.res_codes, ix.res_qparams,
                         ix.res_bits)
    seeds = native.make_seed_ids(ix.graph.shape[0], 32, ix.entry)
    for L in (32, 128, 300):
        ref = pyoracle.search_batch(oix, mode, queries, seeds, 32, L, 10, window=2)
        for h in (native.HASH_EXACT, 12):
            say("case window 2 vs oracle: mode", mode, "L", L, "hash", h)
            ids, d, cnt, st = calc.search_batch(queries, L, 10, correction=mode,
                                                entry_ids=seeds, expand_count=32,
                                                stats=True, hash_log2=h, source_window=2)
            assert np.array_equal(cnt, ref[2])
            assert np.array_equal(ids, ref[0])
            assert np.array_equal(bits(d), bits(ref[1]))
            assert np.array_equal(st[:, 2], ref[3][:, 2])
            if h == native.HASH_EXACT:
                assert np.array_equal(st, ref[3])


def test_window_is_validated(mid):
    calc, ix, queries = mid
    with pytest.raises(RuntimeError):
        calc.search_batch(queries[:4], 32, 10, correction=1, source_window=3)
