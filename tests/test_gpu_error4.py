"""Third-level refine (error4) on the GPU through the C-ABI: the device
encoder against the reference builder's codes, the fused search + ERROR4
refine against the reference's refine outputs and the oracle, and the
loader's rejections.  Ids and distances are compared bit-for-bit."""
import hashlib
import os
import sys

import numpy as np
import pytest

import pyoracle
from rover_amd import artifacts, data, native
from rover_amd import index as rindex
from rover_amd.search import Calculator

pytestmark = pytest.mark.gpu
PQ_SETS = ["pq_d128_m32", "pq_d96_m48", "pq_d64_ks16", "pq_d128_m32_r8"]
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def e4(name):
    z = np.load(os.path.join(GOLDEN, name + "_error4.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


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


@pytest.mark.parametrize("name", PQ_SETS)
def test_device_error4_encoder_matches_reference_builder(golden, name):
    g, e = golden(name), e4(name)
    say("case", name, "error4 encode")
    codes, qp = rindex.encode_error4_on_device(g["base"], g["codes"], g["codebook"],
                                               g["res_codes"], g["res_qparams"],
                                               int(g["res_bits"]))
    assert np.array_equal(codes, e["e4_codes"])
    assert np.array_equal(bits(qp), bits(e["e4_qparams"]))


@pytest.mark.parametrize("name", PQ_SETS)
def test_search_refine_error4_matches_reference(golden, calc, name):
    g, e = golden(name), e4(name)
    say("case", name, "search + refine error4")
    calc.load_pq_index(pq_data(g))
    image = e["artifact"].tobytes()
    calc.pure_pq_error4_load(image, hashlib.sha256(image).digest())
    mode, ep, expand, L = (int(v) for v in g["refine_source"])
    F, K = int(g["refine_F"]), int(g["refine_K"])
    for rmode, key in ((native.REFINE_ERROR4, "refine_e4"),
                       (native.REFINE_CURRENT, "refine_current"), (0, "refine_current")):
        ids, d, cnt = calc.search_batch(g["queries"], L, K, correction=mode,
                                        refine_candidates=F, entry_points=ep,
                                        expand_count=expand, refine_mode=rmode)
        assert np.all(cnt == K), (rmode, cnt)
        assert np.array_equal(ids, e[key + "_ids"]), rmode
        assert np.array_equal(bits(d), bits(e[key + "_dists"])), rmode


def test_error4_loader_rejections(golden, calc):
    g, e = golden("pq_d64_ks16"), e4("pq_d64_ks16")
    say("case error4 loader rejections")
    calc.load_pq_index(pq_data(g))
    image = bytearray(e["artifact"].tobytes())
    q = g["queries"][:2]
    F, K = int(g["refine_F"]), int(g["refine_K"])

    def rc_of(fn):
        try:
            fn()
        except native.RoverError as err:
            return err.code
        return 0

    # ERROR4 before any artifact: the launch fails
    assert rc_of(lambda: calc.search_batch(q, 32, K, refine_candidates=F,
                                           refine_mode=native.REFINE_ERROR4)) == native.ROVER_EFAILED

    def load_rc(buf, sha=None):
        return rc_of(lambda: calc.pure_pq_error4_load(bytes(buf), sha))

    def mutated(off, val):
        b = bytearray(image)
        b[off:off + len(val)] = val
        return b

    assert load_rc(mutated(0, b"SQERR4V1")) == native.ROVER_EBADPARM
    assert load_rc(mutated(12, (7).to_bytes(4, "little"))) == native.ROVER_EBADPARM     # n
    assert load_rc(mutated(16, (32).to_bytes(4, "little"))) == native.ROVER_EBADPARM    # dim
    assert load_rc(mutated(28, (2048).to_bytes(4, "little"))) == native.ROVER_EBADPARM  # block
    assert load_rc(mutated(500, bytes([image[500] ^ 1]))) == native.ROVER_EBADPARM      # payload
    assert load_rc(image + b"\0") == native.ROVER_EBADPARM
    assert load_rc(image, b"\0" * 32) == native.ROVER_EBADPARM                          # file digest
    neg = mutated(256, np.float32(-1.0).tobytes())
    neg[56:88] = hashlib.sha256(bytes(neg[256:])).digest()
    assert load_rc(neg) == native.ROVER_EBADPARM                                        # scale < 0
    assert load_rc(image) == 0
    # refine_mode ERROR4 needs refine candidates; unknown modes are rejected
    assert rc_of(lambda: calc.search_batch(q, 32, K, refine_candidates=0,
                                           refine_mode=native.REFINE_ERROR4)) == native.ROVER_EBADPARM
    assert rc_of(lambda: calc.search_batch(q, 32, K, refine_candidates=F,
                                           refine_mode=3)) == native.ROVER_EBADPARM
    # the artifact is bound to "no id map": with a map its open fails in the
    # reference (retrieval_final_refine.c:167-180) and the launch fails here
    assert int(e["rc_e4_unbound_map"]) == -103
    n = g["graph"].shape[0]
    calc.pure_pq_original_ids(np.arange(n, dtype=np.int32)[::-1])
    assert rc_of(lambda: calc.search_batch(q, 32, K, refine_candidates=F,
                                           refine_mode=native.REFINE_ERROR4)) == native.ROVER_EBADPARM
    calc.pure_pq_original_ids(None)
    ids, d, cnt = calc.search_batch(q, 32, K, refine_candidates=F,
                                    refine_mode=native.REFINE_ERROR4)
    assert np.all(cnt == K)


def test_error4_bound_to_an_id_map(calc):
    """An artifact written with the id-map digest of a graph->original map
    refines with that map (ties broken by original id), as the oracle does."""
    say("case error4 with an id map")
    base, queries = data.workload(3000, 64, 24, 8.0, 16, base_seed=7, query_seed=8)
    idx = rindex.build_pq_index(base, 16, 16, ks=256, res_bits=8, iters=6)
    codes, qp = rindex.encode_error4_on_device(base, idx.codes, idx.codebook, idx.res_codes,
                                               idx.res_qparams, 8)
    perm = np.random.default_rng(3).permutation(3000).astype(np.int32)
    image = artifacts.error4_image(codes, qp, 64, 8, 8,
                                   maps_sha256=artifacts.maps_digest(perm))
    calc.load_pq_index(idx)
    calc.pure_pq_original_ids(perm)
    calc.pure_pq_error4_load(image)
    seeds = pyoracle.seed_ids(3000, 8, idx.entry, 42)
    ids, d, cnt = calc.search_batch(queries, 48, 10, correction=1, refine_candidates=24,
                                    entry_ids=seeds, expand_count=16,
                                    refine_mode=native.REFINE_ERROR4)
    oix = pyoracle.Index(idx.graph, idx.codes, idx.codebook, idx.res_codes, idx.res_qparams,
                         8, e4_codes=codes, e4_qparams=qp)
    for qi, q in enumerate(queries):
        pool, _, _ = pyoracle.search(oix, pyoracle.SCORE_ONLINE, q, seeds, 16, 48, 48)
        oi, od = pyoracle.refine(oix, q, pool[:24], 10, graph_to_original=perm)
        assert cnt[qi] == 10
        assert np.array_equal(ids[qi], oi), qi
        assert np.array_equal(bits(d[qi]), bits(od)), qi
    calc.pure_pq_original_ids(None)


def test_mid_scale_error4_refine_vs_oracle(calc):
    """20,000 x 128 vectors, M=32 x 256 with 4-bit residuals (the paired
    online scorer), online correction, L=64, refine of the top 32 with the
    error4 level, per-query entry rows; every query against the oracle."""
    say("case mid-scale error4: building")
    base, queries = data.workload(20000, 128, 64, 8.0, 200, base_seed=11, query_seed=12)
    idx = rindex.build_pq_index(base, 32, 32, ks=256, res_bits=4, iters=6)
    codes, qp = rindex.encode_error4_on_device(base, idx.codes, idx.codebook, idx.res_codes,
                                               idx.res_qparams, 4)
    calc.load_pq_index(idx)
    calc.pure_pq_error4_load(artifacts.error4_image(codes, qp, 128, 8, 4))
    rng = np.random.default_rng(4)
    rows = np.concatenate([np.full((200, 1), idx.entry, dtype=np.int32),
                           rng.integers(0, 20000, (200, 7), dtype=np.int32)], axis=1)
    say("case mid-scale error4: searching")
    ids, d, cnt, st = calc.search_batch(queries, 64, 10, correction=1, refine_candidates=32,
                                        entry_ids=rows, expand_count=32, stats=True,
                                        refine_mode=native.REFINE_ERROR4)
    oix = pyoracle.Index(idx.graph, idx.codes, idx.codebook, idx.res_codes, idx.res_qparams,
                         4, e4_codes=codes, e4_qparams=qp)
    oids, od, ocnt, ost = pyoracle.search_batch(oix, pyoracle.SCORE_ONLINE, queries, rows,
                                                32, 64, 10, refine=32)
    assert np.array_equal(cnt, ocnt)
    assert np.array_equal(ids, oids)
    assert np.array_equal(bits(d), bits(od))
    gt = data.ground_truth(base, queries, k=10)
    r4 = data.recall_at_k(ids, cnt, gt, 10)
    ids1, _, cnt1 = calc.search_batch(queries, 64, 10, correction=1, refine_candidates=32,
                                      entry_ids=rows, expand_count=32)
    r1 = data.recall_at_k(ids1, cnt1, gt, 10)
    say("mid-scale: recall@10 refine %.4f, refine + error4 %.4f" % (r1, r4))
    assert r4 >= r1
