"""The live SQ path pinned to the reference: the oracle's SQ traversal
(distance-only heaps) against sqmd_cpu_state_run_fake_dsp, the seed
generator against retrieval_sqmd_make_seed_ids, and the SQ residual
encoder and exact distance against what the mathematics fixes.

The sqmd_d128 fixture holds the reference's own outputs
(tests/golden/make_golden.py sqmd_case): SQ8 with 4-bit residuals and SQ4
with 8-bit residuals, correction off and on, several (entry points, expand,
L) settings, every query of the PQ fixture."""
import ctypes as C
import os

import numpy as np
import pytest

import pyoracle
from rover_amd import index as rindex

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def sqmd():
    g = load("pq_d128_m32")
    s = load("sq_d128")
    z = load("sqmd_d128")
    base = g["base_x16"].astype(np.float32) / np.float32(16.0)
    queries = g["queries_x16"].astype(np.float32) / np.float32(16.0)
    return g, s, z, base, queries


def oindex(g, s, z, b, dim):
    return pyoracle.Index(g["graph"], res_codes=z["sq%d_res_codes" % b],
                          res_qparams=z["sq%d_res_qparams" % b],
                          res_bits=int(z["sq%d_res_bits" % b]), sq_codes=s["sq%d_codes" % b],
                          sq_bits=b, sq_scale=float(s["sq%d_scale" % b]),
                          sq_zero=float(s["sq%d_zero" % b]), dim=dim)


def test_sqmd_traversal_matches_reference(sqmd):
    g, s, z, base, queries = sqmd
    n, dim = base.shape
    entry = int(g["entry"])
    for b, corr, ep, expand, L in (tuple(int(v) for v in c) for c in z["cases"]):
        ix = oindex(g, s, z, b, dim)
        seeds = pyoracle.sqmd_seed_ids(n, ep, entry, 42)
        assert np.array_equal(seeds, z["seeds_e%d" % ep])
        key = "sq%d_c%d_e%d_x%d_L%d" % (b, corr, ep, expand, L)
        for qi, q in enumerate(queries):
            ids, d, _ = pyoracle.sqmd_search(ix, q, seeds, ep, expand, L, L, corr)
            c = int(z[key + "_counts"][qi])
            assert len(ids) == c, (key, qi)
            assert np.array_equal(ids, z[key + "_ids"][qi][:c]), (key, qi)
            assert np.array_equal(bits(d), bits(z[key + "_dists"][qi][:c])), (key, qi)


def test_fixture_exercises_ties(sqmd):
    """SQ code distances are integers times scale^2, so equal distances are
    common; the fixture's uncorrected results must contain them (otherwise
    the distance-only heap order would be untested)."""
    g, s, z, base, queries = sqmd
    ties = 0
    for key in z:
        if key.endswith("_dists") and "_c0_" in key:
            d = z[key]
            ties += int((np.diff(d, axis=1) == 0).sum())
    assert ties > 100


def test_sqmd_seed_generator(sqmd):
    """The oracle and the library's host generator against the reference
    (fixture seeds, and the seeds fixture's n = 65536 lists)."""
    from rover_amd import native
    L = native.lib()
    g, s, z, base, queries = sqmd
    n = base.shape[0]
    for ep in (1, 8, 32):
        out = np.zeros(ep, dtype=np.int32)
        cnt = L.rover_sqmd_seed_ids(n, ep, int(g["entry"]), 42, out.ctypes.data, ep)
        assert np.array_equal(out[:cnt], z["seeds_e%d" % ep])
        assert np.array_equal(pyoracle.sqmd_seed_ids(n, ep, int(g["entry"]), 42),
                              z["seeds_e%d" % ep])
    sd = load("seeds")
    for seed in (42, 1, 12345):
        out = np.zeros(64, dtype=np.int32)
        cnt = L.rover_sqmd_seed_ids(65536, 64, -1, seed, out.ctypes.data, 64)
        assert np.array_equal(out[:cnt], sd["sqmd_seed%d" % seed])
    # more seeds than vectors: every id once, the entry first
    out = np.zeros(16, dtype=np.int32)
    cnt = L.rover_sqmd_seed_ids(10, 16, 3, 42, out.ctypes.data, 16)
    assert cnt == 10 and out[0] == 3 and sorted(out[:cnt]) == list(range(10))


def test_sq_residual_encoder_properties(sqmd):
    """demo_quantize_residual's contract: zero is the row minimum of the
    residual, the dequantized residual is within half a step of it, codes
    span the full range when the residual has one, and round-half-away
    holds at exact half steps."""
    g, s, z, base, queries = sqmd
    dim = base.shape[1]
    for b in (8, 4):
        codes, scale, zero = s["sq%d_codes" % b], s["sq%d_scale" % b], s["sq%d_zero" % b]
        rb = int(z["sq%d_res_bits" % b])
        res, qp = rindex.encode_sq_residual(base, codes, b, scale, zero, rb)
        assert np.array_equal(res, z["sq%d_res_codes" % b])
        deq = rindex.decode_sq(codes, b, scale, zero, dim).astype(np.float64)
        r = base.astype(np.float64) - deq
        levels = 15 if rb == 4 else 255
        q = np.zeros((len(base), dim), dtype=np.int64)
        if rb == 8:
            q[:] = res
        else:
            q[:, 0::2] = res[:, :(dim + 1) // 2] & 15
            q[:, 1::2] = res[:, :dim // 2] >> 4
        assert np.all(qp[:, 1] == (base - deq.astype(np.float32)).min(axis=1))
        rec = q * qp[:, 0:1].astype(np.float64) + qp[:, 1:2].astype(np.float64)
        assert np.all(np.abs(rec - r) <= qp[:, 0:1] * 0.5 * (1 + 1e-4) + 1e-7)
        assert np.all(q.max(axis=1) == levels)
        assert np.all(q.min(axis=1) == 0)
    # half steps round away from zero: residuals 0, 0.5, 1.5, 15 with scale 1
    x = np.array([[0.0, 0.5, 1.5, 15.0]], dtype=np.float32)
    l1 = np.zeros((1, 4), dtype=np.uint8)
    res, qp = rindex.encode_sq_residual(x, l1, 8, np.float32(1.0), np.float32(0.0), 4)
    assert qp[0, 0] == 1.0 and qp[0, 1] == 0.0
    assert list(res[0]) == [0 | (1 << 4), 2 | (15 << 4)]


def test_sq_residual_distance_closed_form(sqmd):
    """The exact SQ + residual distance against a float64 evaluation of the
    same reconstruction (f32 rounding of each step bounds the gap)."""
    g, s, z, base, queries = sqmd
    dim = base.shape[1]
    for b in (8, 4):
        ix = oindex(g, s, z, b, dim)
        codes = s["sq%d_codes" % b]
        deq = rindex.decode_sq(codes, b, s["sq%d_scale" % b], s["sq%d_zero" % b], dim)
        rb = int(z["sq%d_res_bits" % b])
        res, qp = z["sq%d_res_codes" % b], z["sq%d_res_qparams" % b]
        rq = np.zeros((len(base), dim), dtype=np.float64)
        if rb == 8:
            rq[:] = res
        else:
            rq[:, 0::2] = res[:, :(dim + 1) // 2] & 15
            rq[:, 1::2] = res[:, :dim // 2] >> 4
        rec = deq.astype(np.float64) + rq * qp[:, 0:1] + qp[:, 1:2]
        for q in queries[:3]:
            ref = ((q.astype(np.float64) - rec) ** 2).sum(axis=1)
            got = np.array([pyoracle.sq_residual_distance(ix, q, i) for i in range(0, 1500, 7)])
            assert np.allclose(got, ref[::7], rtol=2e-5, atol=1e-3)


def test_sqmd_reduces_to_exact_sort_on_complete_graph():
    """On a complete graph with L >= n the traversal visits every node and
    returns all of them in ascending distance order."""
    rng = np.random.default_rng(9)
    n, dim = 32, 16
    base = rng.normal(size=(n, dim)).astype(np.float32) * 4
    graph = np.array([[j for j in range(n) if j != i] + [-1] for i in range(n)],
                     dtype=np.int32)
    scale, zero = rindex.sq_params(base, 8)
    codes = rindex.encode_sq(base, 8, scale, zero)
    ix = pyoracle.Index(graph, sq_codes=codes, sq_bits=8, sq_scale=float(scale),
                        sq_zero=float(zero), dim=dim)
    q = rng.normal(size=dim).astype(np.float32)
    seeds = pyoracle.sqmd_seed_ids(n, 4, 0, 42)
    ids, d, st = pyoracle.sqmd_search(ix, q, seeds, 4, 32, n, n)
    assert sorted(ids.tolist()) == list(range(n))
    assert np.all(np.diff(d) >= 0)
    qq = pyoracle.sq_quantize_query(q, 8, float(scale), float(zero))
    want = np.array([pyoracle.sq_distance(qq, codes[i], dim, 8, float(scale)) for i in ids],
                    dtype=np.float32)
    assert np.array_equal(bits(d), bits(want))
