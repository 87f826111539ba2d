"""GPU parity: the HIP search path through the C-ABI against the reference's
golden outputs and the CPU oracle.  Integer/index results and float
distances are compared bit-for-bit (the kernel keeps the reference's
operation order)."""
import sys

import numpy as np
import pytest

import pyoracle
from rover_amd import index as rindex
from rover_amd import native
from rover_amd.search import Calculator, pure_pq_search_scalar

pytestmark = pytest.mark.gpu
PQ_SETS = ["pq_d128_m32", "pq_d96_m48", "pq_d64_ks16", "pq_d128_m32_r8"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pq_data(g):
    return rindex.PqIndexData(g["graph"], int(g["entry"]), g["codes"], g["codebook"],
                              g["res_codes"], g["res_qparams"], int(g["res_bits"]))


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def calc():
    c = Calculator()
    yield c
    c.close()


@pytest.mark.parametrize("name", PQ_SETS)
def test_pq_search_matches_reference(golden, calc, name):
    g = golden(name)
    say("case", name)
    calc.load_pq_index(pq_data(g), chunk_vectors=300)
    for mode, ep, expand, L in g["search_cases"]:
        key = "m%d_e%d_x%d_L%d" % (mode, ep, expand, L)
        ids, d, cnt, st = calc.search_batch(g["queries"], int(L), int(L),
                                            correction=int(mode), entry_points=int(ep),
                                            expand_count=int(expand), stats=True,
                                            hash_log2=14)
        for qi in range(len(g["queries"])):
            c = int(g[key + "_counts"][qi])
            assert cnt[qi] == c, (key, qi, cnt[qi])
            assert np.array_equal(ids[qi, :c], g[key + "_ids"][qi][:c]), (key, qi)
            assert np.array_equal(bits(d[qi, :c]), bits(g[key + "_dists"][qi][:c])), (key, qi)
            assert tuple(int(x) for x in st[qi]) == tuple(int(x) for x in g[key + "_stats"][qi]), \
                (key, qi, st[qi], g[key + "_stats"][qi])


@pytest.mark.parametrize("name", PQ_SETS)
def test_lossy_visited_hash_keeps_results(golden, calc, name):
    """A 16-slot visited hash with one probe drops most entries: nodes get
    re-scored, but ids, distances and the expansion count are unchanged."""
    g = golden(name)
    say("case", name, "lossy hash")
    calc.load_pq_index(pq_data(g))
    for mode, ep, expand, L in g["search_cases"][::3]:
        key = "m%d_e%d_x%d_L%d" % (mode, ep, expand, L)
        ids, d, cnt, st = calc.search_batch(g["queries"], int(L), int(L),
                                            correction=int(mode), entry_points=int(ep),
                                            expand_count=int(expand), stats=True,
                                            hash_log2=4, probe_limit=1)
        for qi in range(len(g["queries"])):
            c = int(g[key + "_counts"][qi])
            assert np.array_equal(ids[qi, :c], g[key + "_ids"][qi][:c])
            assert np.array_equal(bits(d[qi, :c]), bits(g[key + "_dists"][qi][:c]))
            assert int(st[qi, 2]) == int(g[key + "_stats"][qi][2])
            assert int(st[qi, 0]) >= int(g[key + "_stats"][qi][0])


@pytest.mark.parametrize("name", PQ_SETS)
def test_exact_visited_sets_match_reference(golden, calc, name):
    """Exact visited sets in HBM (ROVER_HASH_EXACT): every counter equals the
    reference's, as with a large LDS hash; launches of 3 and of the default
    number of waves (each wave reusing its set under new stamps) and a
    repeated launch give the same results."""
    g = golden(name)
    say("case", name, "exact visited sets")
    calc.load_pq_index(pq_data(g))
    for mode, ep, expand, L in g["search_cases"]:
        key = "m%d_e%d_x%d_L%d" % (mode, ep, expand, L)
        for grid in (0, 3, 0):
            ids, d, cnt, st = calc.search_batch(g["queries"], int(L), int(L),
                                                correction=int(mode), entry_points=int(ep),
                                                expand_count=int(expand), stats=True,
                                                hash_log2=native.HASH_EXACT, grid_waves=grid)
            for qi in range(len(g["queries"])):
                c = int(g[key + "_counts"][qi])
                assert cnt[qi] == c, (key, qi, cnt[qi])
                assert np.array_equal(ids[qi, :c], g[key + "_ids"][qi][:c]), (key, qi)
                assert np.array_equal(bits(d[qi, :c]), bits(g[key + "_dists"][qi][:c]))
                assert tuple(int(x) for x in st[qi]) == \
                    tuple(int(x) for x in g[key + "_stats"][qi]), (key, qi, grid)


@pytest.mark.parametrize("name", PQ_SETS)
def test_refine_matches_reference(golden, calc, name):
    g = golden(name)
    say("case", name, "refine")
    calc.load_pq_index(pq_data(g))
    mode, ep, expand, L = g["refine_source"]
    F, K = int(g["refine_F"]), int(g["refine_K"])
    ids, d, cnt = calc.search_batch(g["queries"], int(L), K, correction=0,
                                    refine_candidates=F, entry_points=1,
                                    expand_count=int(expand))
    assert np.all(cnt == K)
    assert np.array_equal(ids, g["refine_ids"])
    assert np.array_equal(bits(d), bits(g["refine_dists"]))


def test_sq_search_matches_reference(golden, calc):
    g = golden("pq_d128_m32")
    s = golden("sq_d128")
    dim = g["base"].shape[1]
    for b in (8, 4):
        say("case sq", b)
        calc.rover_configure(l1_bits=b)
        calc.retrieval_load(g["graph"], s["sq%d_codes" % b], dim, int(g["entry"]),
                            s["sq%d_scale" % b], s["sq%d_zero" % b])
        for ep in (1, 32):
            for L in (16, 64):
                key = "sq%d_e%d_L%d" % (b, ep, L)
                ids, d, cnt, st = calc.search_batch(g["queries"], L, L,
                                                    method=native.METHOD_SQ,
                                                    entry_points=ep, expand_count=32,
                                                    stats=True, hash_log2=14,
                                                    traversal=native.TRAVERSAL_BEST_FIRST)
                assert np.array_equal(ids, s[key + "_ids"]), key
                assert np.array_equal(bits(d), bits(s[key + "_dists"])), key
                assert np.array_equal(st.astype(np.int64), s[key + "_stats"].astype(np.int64))
                # the IDL op, one query at a time: the DSP search op's
                # distance-only traversal over its graph entry + LCG seeds
                # (the oracle's SQ traversal over the same seed list)
                oix = pyoracle.Index(g["graph"], sq_codes=s["sq%d_codes" % b], sq_bits=b,
                                     sq_scale=float(s["sq%d_scale" % b]),
                                     sq_zero=float(s["sq%d_zero" % b]), dim=dim)
                seeds = pyoracle.seed_ids(len(g["graph"]), ep, int(g["entry"]), 42)
                for qi in range(3):
                    i1, d1 = calc.retrieval_search(g["queries"][qi], ep, 32, L, 10)
                    oi, od, _ = pyoracle.sqmd_search(oix, g["queries"][qi], seeds, ep, 32, L, 10)
                    assert np.array_equal(i1, oi)
                    assert np.array_equal(bits(d1), bits(od))
        calc.retrieval_unload()


@pytest.mark.parametrize("name", PQ_SETS)
def test_library_dropin_matches_reference(golden, name):
    g = golden(name)
    say("case", name, "library call")
    n = g["graph"].shape[0]
    for mode, ep, expand, L in g["search_cases"][::2]:
        key = "m%d_e%d_x%d_L%d" % (mode, ep, expand, L)
        entries = pyoracle.seed_ids(n, int(ep), int(g["entry"]), 42)
        res = (g["res_codes"], g["res_qparams"], int(g["res_bits"])) if mode else None
        for qi in range(0, len(g["queries"]), 3):
            ids, d, st = pure_pq_search_scalar(g["graph"], g["codes"], g["codebook"],
                                               g["queries"][qi], entries, int(expand),
                                               int(L), int(L), chunk_vectors=256,
                                               residual=res)
            c = int(g[key + "_counts"][qi])
            assert len(ids) == c
            assert np.array_equal(ids, g[key + "_ids"][qi][:c])
            assert np.array_equal(bits(d), bits(g[key + "_dists"][qi][:c]))
            assert int(st[2]) == int(g[key + "_stats"][qi][2])


def test_error_statuses(golden, calc):
    g = golden("pq_d128_m32")
    n = g["graph"].shape[0]
    say("case errors")
    # entry id out of range -> -1, as the reference returns
    ids, _, _ = pure_pq_search_scalar(g["graph"], g["codes"], g["codebook"],
                                      g["queries"][0],
                                      np.array([int(g["entry"]), n], dtype=np.int32),
                                      32, 16, 16)
    assert ids is None
    # -1 in the middle of a graph row -> per-query status
    bad = g["graph"].copy()
    bad[int(g["entry"]), 3] = -1
    data = pq_data(g)
    data.graph = bad
    calc.load_pq_index(data)
    _, _, cnt = calc.search_batch(g["queries"][:2], 16, 16)
    assert np.all(cnt == -2)
    # k larger than L is a bad parameter
    calc.load_pq_index(pq_data(g))
    _, _, cnt = calc.search_batch(g["queries"][:2], np.array([16, 8], dtype=np.int32),
                                  np.array([10, 9], dtype=np.int32))
    assert cnt[0] == 10 and cnt[1] == -5


@pytest.mark.parametrize("name", PQ_SETS)
def test_device_encoders(golden, name):
    g = golden(name)
    say("case", name, "encoders")
    codes, res, qp = rindex.encode_on_device(g["base"], g["codebook"], int(g["res_bits"]))
    assert np.array_equal(codes, g["codes"])
    assert np.array_equal(res, g["res_codes"])
    assert np.array_equal(bits(qp), bits(g["res_qparams"]))


@pytest.fixture(scope="module")
def mid_index():
    from rover_amd import data
    base, queries = data.workload(12000, 128, 64, 8.0, 96)
    ix = rindex.build_pq_index(base, m=32, R=32, iters=6, train_max=12000)
    return base, queries, ix


def test_mid_scale_uneven_batch_vs_oracle(mid_index, calc):
    """12k x 128 index: per-query L and k in one launch, ONLINE and plain,
    against the oracle run at each query's own L and k."""
    base, queries, ix = mid_index
    say("case mid-scale uneven batch")
    calc.load_pq_index(ix)
    rng = np.random.default_rng(5)
    nq = len(queries)
    Ls = rng.integers(16, 300, size=nq).astype(np.int32)
    Ks = np.minimum(rng.choice([1, 10, 100], size=nq), Ls).astype(np.int32)
    oix = pyoracle.Index.from_pq(ix)
    entries = pyoracle.seed_ids(ix.n, 8, ix.entry, 42)
    for mode in (0, 1):
        ids, d, cnt, st = calc.search_batch(queries, Ls, Ks, correction=mode,
                                            entry_points=8, expand_count=32,
                                            stats=True, hash_log2=14)
        oids, od, ocnt, ost = pyoracle.search_batch(oix, mode, queries, entries, 32,
                                                    Ls, Ks)
        assert np.array_equal(cnt, ocnt)
        for qi in range(nq):
            c = cnt[qi]
            assert np.array_equal(ids[qi, :c], oids[qi, :c]), (mode, qi)
            assert np.array_equal(bits(d[qi, :c]), bits(od[qi, :c])), (mode, qi)
        assert np.array_equal(st.astype(np.int64), ost.astype(np.int64))


def test_mid_scale_refine_vs_oracle(mid_index, calc):
    base, queries, ix = mid_index
    say("case mid-scale refine")
    calc.load_pq_index(ix)
    oix = pyoracle.Index.from_pq(ix)
    entries = pyoracle.seed_ids(ix.n, 1, ix.entry, 42)
    for F in (20, 64, 150):
        ids, d, cnt = calc.search_batch(queries, 160, 10, refine_candidates=F,
                                        entry_points=1, expand_count=32)
        oids, od, ocnt, _ = pyoracle.search_batch(oix, 0, queries, entries, 32, 160, 10,
                                                  refine=F)
        assert np.array_equal(cnt, ocnt)
        assert np.array_equal(ids, oids)
        assert np.array_equal(bits(d), bits(od))


@pytest.mark.parametrize("name", PQ_SETS)
def test_decode_matches_reference(golden, calc, name):
    """rover_pure_pq_decode equals the reference's residual decoder
    (retrieval_pq_residual_query_decode) on the fixture's sampled ids, bit
    for bit."""
    g = golden(name)
    say("case", name, "decode")
    calc.load_pq_index(pq_data(g))
    ids = g["adc_sample"][:16]
    got = calc.pure_pq_decode(ids)
    assert np.array_equal(bits(got), bits(g["decoded"]))
    with pytest.raises(native.RoverError):
        calc.pure_pq_decode([len(g["graph"])])


def test_refine_text_trace(golden, calc, tmp_path):
    """The pure-PQ demo's refine trace written from GPU results: Q/C lines
    carry the search pool, V lines the decoded vectors of the refine
    candidates, O lines the refined top-k; every number equals the oracle's
    (search, residual decode, refine)."""
    import pyoracle
    g = golden("pq_d128_m32_r8")
    calc.load_pq_index(pq_data(g))
    n = len(g["graph"])
    L, F, K = 64, 20, 10
    ents = native.make_seed_ids(n, 32, int(g["entry"]))
    ids, d, cnt = calc.search_batch(g["queries"], L, L, correction=1, entry_points=32)
    rid, rd, rcnt = calc.search_batch(g["queries"], L, K, correction=1, entry_points=32,
                                      refine_candidates=F)
    path = tmp_path / "refine.trace"
    with native.Trace(path, native.TRACE_TEXT) as t:
        for qi in range(len(g["queries"])):
            c = int(cnt[qi])
            t.query(L, qi, ids[qi, :c], d[qi, :c], rid[qi, :K], rd[qi, :K], refine_budget=F,
                    decoded=calc.pure_pq_decode(ids[qi, :min(c, F)]))
    oix = pyoracle.Index.from_pq(rindex.PqIndexData(g["graph"], int(g["entry"]), g["codes"],
                                                    g["codebook"], g["res_codes"],
                                                    g["res_qparams"], int(g["res_bits"])))
    lines = path.read_text().splitlines()
    at = 0
    for qi, q in enumerate(g["queries"]):
        oi, od, _ = pyoracle.search(oix, pyoracle.SCORE_ONLINE, q, ents, 32, L, L)
        head = lines[at].split()
        assert head == ["Q", str(L), str(qi), str(len(oi)), str(F)]
        at += 1
        for i in range(len(oi)):
            tag, a, b = lines[at].split()
            assert tag == "C" and int(a) == oi[i] and float.fromhex(b) == float(od[i])
            at += 1
        for i in range(min(len(oi), F)):
            f = lines[at].split()
            want = pyoracle.residual_decode(oix, int(oi[i]))
            assert f[0] == "V" and int(f[1]) == oi[i]
            assert np.array_equal(np.array([float.fromhex(x) for x in f[2:]], np.float32), want)
            at += 1
        fi, fd = pyoracle.refine(oix, q, oi[:F], K)
        for i in range(K):
            tag, a, b = lines[at].split()
            assert tag == "O" and int(a) == fi[i] and float.fromhex(b) == float(fd[i])
            at += 1
    assert at == len(lines)
