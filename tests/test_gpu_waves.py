"""Several waves per query and the candidate tile (RoverSearchBatch.
waves_per_query / candidate_tile): every (waves, tile) in {1, 2, 4} x
{16, 32, 64} gives the default kernel's ids, distance bits, counts and
counters, against the reference's golden outputs and with the exact visited
sets, an LDS hash, per-query L and k, the final refine and the learned stop;
the autotuner times these axes; unsupported combinations are refused."""
import ctypes as C
import sys

import numpy as np
import pytest

from rover_amd import data, native
from rover_amd import index as rindex
from rover_amd.search import Calculator

pytestmark = pytest.mark.gpu
WAVES = (1, 2, 4)
TILES = (16, 32, 64)


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


@pytest.mark.parametrize("name", ["pq_d128_m32_r8", "pq_d128_m32"])
def test_waves_match_reference(golden, calc, name):
    g = golden(name)
    calc.load_pq_index(pq_data(g), chunk_vectors=300)
    for mode, ep, expand, L in g["search_cases"]:
        if int(mode) != 1 or int(expand) > 32:
            continue
        key = "m%d_e%d_x%d_L%d" % (mode, ep, expand, L)
        for w in WAVES:
            for t in TILES:
                say("case", name, key, "waves", w, "tile", t)
                ids, d, cnt, st = calc.search_batch(
                    g["queries"], int(L), int(L), correction=1, entry_points=int(ep),
                    expand_count=int(expand), stats=True, hash_log2=native.HASH_EXACT,
                    waves_per_query=w, candidate_tile=t)
                for qi in range(len(g["queries"])):
                    c = int(g[key + "_counts"][qi])
                    assert cnt[qi] == c, (key, qi)
                    assert np.array_equal(ids[qi, :c], g[key + "_ids"][qi][:c]), (key, qi)
                    assert np.array_equal(bits(d[qi, :c]), bits(g[key + "_dists"][qi][:c]))
                    assert tuple(int(x) for x in st[qi]) == \
                        tuple(int(x) for x in g[key + "_stats"][qi]), (key, qi)


@pytest.fixture(scope="module")
def mid():
    say("case waves: 12k x 128 index")
    base, queries = data.workload(12000, 128, 64, 8.0, 500)
    ix = rindex.build_pq_index(base, m=32, R=32, res_bits=8, iters=6,
                               train_max=12000, n_router=0)
    c = Calculator()
    c.load_pq_index(ix)
    yield c, ix, queries
    c.close()


def same(a, b):
    assert np.array_equal(a[2], b[2])
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(bits(a[1]), bits(b[1]))
    if len(a) > 3 and a[3] is not None:
        assert np.array_equal(a[3], b[3])


def test_waves_match_default_kernel(mid):
    """Per-query L (16..300) and k (1, 10, 100), 8-bit residuals, the exact
    visited sets and a 2^12-slot LDS hash; the final refine of the top 20."""
    calc, ix, queries = mid
    rng = np.random.default_rng(9)
    ks = rng.choice([1, 10, 100], size=len(queries)).astype(np.int32)
    Ls = np.array([rng.integers(max(16, k), 301) for k in ks], dtype=np.int32)
    for h in (native.HASH_EXACT, 12):
        ref = calc.search_batch(queries, Ls, ks, correction=1, entry_points=32,
                                stats=True, hash_log2=h)
        for w in WAVES:
            for t in TILES:
                say("case waves vs default: hash", h, "waves", w, "tile", t)
                got = calc.search_batch(queries, Ls, ks, correction=1, entry_points=32,
                                        stats=True, hash_log2=h, waves_per_query=w,
                                        candidate_tile=t)
                same(ref if h == native.HASH_EXACT else ref[:3],
                     got if h == native.HASH_EXACT else got[:3])
    ref = calc.search_batch(queries, 128, 10, correction=1, entry_points=32,
                            refine_candidates=20)
    for w in WAVES:
        got = calc.search_batch(queries, 128, 10, correction=1, entry_points=32,
                                refine_candidates=20, waves_per_query=w, candidate_tile=16)
        same(ref, got)


def test_waves_learned_stop_matches_default_kernel(mid):
    """A predictor with random weights stops queries at various steps; the
    multi-wave kernel evaluates the same features and network at the same
    steps, so ids, distances and expansion counts agree."""
    calc, ix, queries = mid
    P = native.lib().rover_mlp_param_count()
    rng = np.random.default_rng(5)
    params = (rng.normal(size=P) * 0.3).astype(np.float32)
    mean = np.zeros(16, np.float32)
    istd = np.ones(16, np.float32)
    native.check(native.lib().rover_learned_stop_set(calc.h, params.ctypes.data, P,
                                                     mean.ctypes.data, istd.ctypes.data,
                                                     4, 10, 4), "rover_learned_stop_set")
    for thr in (0.5, 0.9):
        ref = calc.search_batch(queries, 200, 10, correction=1, entry_points=32, stats=True,
                                learned_stop=1, stop_threshold=thr,
                                hash_log2=native.HASH_EXACT)
        plain = calc.search_batch(queries, 200, 10, correction=1, entry_points=32,
                                  stats=True, hash_log2=native.HASH_EXACT)
        say("case waves learned, threshold", thr, "expansions", ref[3][:, 2].mean(),
            "plain", plain[3][:, 2].mean())
        for w in WAVES:
            got = calc.search_batch(queries, 200, 10, correction=1, entry_points=32,
                                    stats=True, learned_stop=1, stop_threshold=thr,
                                    hash_log2=native.HASH_EXACT, waves_per_query=w,
                                    candidate_tile=32)
            same(ref, got)


def test_autotuner_times_the_axes(mid):
    calc, ix, queries = mid
    nq, k = len(queries), 10
    dq = native.DeviceBuffer.from_array(queries)
    di = native.DeviceBuffer(nq * k * 4)
    dd = native.DeviceBuffer(nq * k * 4)
    dc = native.DeviceBuffer(nq * 4)
    t = calc.autotune(dq.ptr, nq, 128, k, di.ptr, dd.ptr, dc.ptr, k, repeats=2,
                      correction=1, entry_points=32)
    say("autotune: waves %d tile %d hash %d grid %d, %.3f ms (default %.3f), tried %d"
        % (t.waves_per_query, t.candidate_tile, t.hash_log2, t.grid_waves, t.ms,
           t.default_ms, t.tried))
    assert t.waves_per_query in (0, 1, 2, 4)
    assert t.candidate_tile in ((0,) if t.waves_per_query == 0 else TILES)
    assert t.rejected == 0
    assert t.tried >= 10  # the first stage and the nine (waves, tile) pairs
    ref = calc.search_batch(queries, 128, k, correction=1, entry_points=32)
    got = (di.download(np.int32, nq * k).reshape(nq, k),
           dd.download(np.float32, nq * k).reshape(nq, k), dc.download(np.int32, nq))
    same(ref, got)


def test_unsupported_combinations_are_refused(mid):
    calc, ix, queries = mid
    q = queries[:4]
    for kw in (dict(waves_per_query=3), dict(waves_per_query=2, candidate_tile=20),
               dict(waves_per_query=2, source_window=2),
               dict(waves_per_query=2, correction=0)):
        args = dict(correction=1, entry_points=32)
        args.update(kw)
        with pytest.raises(native.RoverError):
            calc.search_batch(q, 64, 10, **args)
