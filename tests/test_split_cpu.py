"""The SQ split distance pinned to the reference: the oracle's traversal
with split distances against the reference's own arithmetic and fake-DSP
turn loop (tests/golden/sqmd_split.npz, written by make_golden.py
split_case; and the live reference where it is built), the aligned query
quantization against round-half-away-from-zero on the refined grid, the
split distance against the squared distance it rewrites (in float64), the
bias against an exactly rounded sum, and the SQMDBIAS image layout and
checks."""
import math
import os
import struct

import numpy as np
import pytest

import pyoracle
from rover_amd import artifacts, data
from rover_amd import index as rindex

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def split():
    g, s, z = load("pq_d128_m32"), load("sq_d128"), load("sqmd_split")
    base = g["base_x16"].astype(np.float32) / np.float32(16.0)
    queries = g["queries_x16"].astype(np.float32) / np.float32(16.0)
    return g, s, z, base, queries


def oracle_index(g, s, z, rb):
    return pyoracle.Index(g["graph"], res_codes=z["r%d_res_codes" % rb],
                          res_qparams=z["r%d_res_qparams" % rb], res_bits=rb,
                          sq_codes=s["sq4_codes"], sq_bits=4,
                          sq_scale=float(s["sq4_scale"]), sq_zero=float(s["sq4_zero"]),
                          dim=128, split_bias=z["r%d_bias" % rb])


def test_oracle_matches_reference_fixture(split):
    g, s, z, base, queries = split
    n = g["graph"].shape[0]
    for rb, ep, expand, L in (tuple(int(v) for v in c) for c in z["cases"]):
        oix = oracle_index(g, s, z, rb)
        key = "r%d_e%d_x%d_L%d" % (rb, ep, expand, L)
        seeds = np.zeros(ep, dtype=np.int32)
        cnt = pyoracle.oracle().orc_sqmd_seed_ids(n, ep, int(z["entry"]), 42,
                                                  seeds.ctypes.data, ep)
        for qi, q in enumerate(queries):
            ids, d, _ = pyoracle.sqmd_search(oix, q, seeds[:cnt], ep, expand, L, L,
                                             correction=2)
            c = int(z[key + "_counts"][qi])
            assert ids is not None and len(ids) == c, (key, qi)
            assert np.array_equal(ids, z[key + "_ids"][qi][:c]), (key, qi)
            assert np.array_equal(bits(d), bits(z[key + "_dists"][qi][:c])), (key, qi)


@pytest.mark.skipif(not pyoracle.ref_available(), reason="compiled reference not built")
def test_oracle_matches_live_reference():
    """A second shape: 2000 points, 24 dims (odd packing boundaries), SQ4,
    8-bit residuals, exact kNN + RNG graph from the oracle's builder."""
    n, dim = 2000, 24
    base, queries = data.workload(n, dim, 16, 8.0, 6, base_seed=5, query_seed=6)
    scale, zero = rindex.sq_params(base, 4)
    codes = rindex.encode_sq(base, 4, scale, zero)
    res, rqp = rindex.encode_sq_residual(base, codes, 4, scale, zero, 8)
    bias = artifacts.split_bias(codes, scale, zero, res, rqp, 8, dim)
    graph = pyoracle.build_exact(np.ascontiguousarray(base), 16, 32, 7, 1)
    oix = pyoracle.Index(graph, res_codes=res, res_qparams=rqp, res_bits=8, sq_codes=codes,
                         sq_bits=4, sq_scale=float(scale), sq_zero=float(zero), dim=dim,
                         split_bias=bias)
    ref = pyoracle.reference()
    P = lambda a: a.ctypes.data  # noqa: E731
    for ep, expand, L in ((16, 16, 32), (1, 16, 48)):
        seeds = np.zeros(ep, dtype=np.int32)
        cnt = pyoracle.oracle().orc_sqmd_seed_ids(n, ep, 0, 42, seeds.ctypes.data, ep)
        for q in queries:
            ids = np.zeros(L, dtype=np.int32)
            d = np.zeros(L, dtype=np.float32)
            rc = ref.ref_sqmd_split_search(P(graph), 16, n, P(codes), dim, float(scale),
                                           float(zero), P(res), P(rqp), 8, P(bias), P(q), ep, 0,
                                           42, expand, L, 8, P(ids), P(d), L)
            oi, od, _ = pyoracle.sqmd_search(oix, q, seeds[:cnt], ep, expand, L, L,
                                             correction=2)
            assert rc == len(oi)
            assert np.array_equal(oi, ids[:rc])
            assert np.array_equal(bits(od), bits(d[:rc]))


def test_aligned_query_quantization():
    scale, zero = np.float32(3.0), np.float32(-10.0)
    step = float(scale) / 17.0
    # exact half steps round away from zero; the ends clip
    x = np.array([zero + step * 2.5, zero + step * 3.5, zero - step * 0.5, zero - 40.0,
                  zero + step * 300.0, zero + step * 254.49, zero], dtype=np.float32)
    qu8, sq = pyoracle.sq_query_aligned_u8(x, float(scale), float(zero))
    raw = (x.astype(np.float64) - float(zero)) / step
    want = np.clip(np.sign(raw) * np.floor(np.abs(raw) + 0.5), 0, 255).astype(np.uint8)
    assert np.array_equal(qu8, want)
    assert sq == int(want.sum())
    assert pyoracle.sq_query_aligned_u8(np.array([np.nan], np.float32), 1.0, 0.0) is None
    assert pyoracle.sq_query_aligned_u8(np.ones(4, np.float32), 0.0, 0.0) is None


def test_split_distance_is_the_refined_squared_distance(split):
    """coarse + bias - 2 <q, r> rewrites |q~ - (a + r)|^2, q~ the quantized
    query: equal to float64 evaluation within f32 rounding."""
    g, s, z, base, queries = split
    scale, zero = float(s["sq4_scale"]), float(s["sq4_zero"])
    codes = s["sq4_codes"]
    for rb in (4, 8):
        oix = oracle_index(g, s, z, rb)
        a_all = artifacts._unpack_u4(codes, 128).astype(np.float64) * scale + zero
        rc = z["r%d_res_codes" % rb]
        rc = artifacts._unpack_u4(rc, 128) if rb == 4 else rc
        qp = z["r%d_res_qparams" % rb].astype(np.float64)
        r_all = rc.astype(np.float64) * qp[:, :1] + qp[:, 1:]
        for q in queries[:4]:
            qu8, _ = pyoracle.sq_query_aligned_u8(q, scale, zero)
            qt = qu8.astype(np.float64) * (np.float32(scale) / 17.0) + zero
            for vid in range(0, codes.shape[0], 97):
                want = float(((qt - (a_all[vid] + r_all[vid])) ** 2).sum())
                got = pyoracle.sq_split_distance(oix, q, vid)
                assert abs(got - want) <= 2e-5 * max(want, 1.0) + 1e-3, (rb, vid, got, want)


def test_bias_matches_an_exactly_rounded_sum(split):
    g, s, z, base, queries = split
    scale, zero = float(s["sq4_scale"]), float(s["sq4_zero"])
    codes = s["sq4_codes"]
    for rb in (4, 8):
        bias = z["r%d_bias" % rb]
        rc = z["r%d_res_codes" % rb]
        rc = artifacts._unpack_u4(rc, 128) if rb == 4 else rc
        qp = z["r%d_res_qparams" % rb].astype(np.float64)
        a_all = artifacts._unpack_u4(codes, 128).astype(np.float64) * scale + zero
        off = 0
        for vid in range(0, codes.shape[0], 13):
            r = rc[vid].astype(np.float64) * qp[vid, 0] + qp[vid, 1]
            exact = np.float32(math.fsum(2.0 * a_all[vid] * r + r * r))
            ulps = abs(int(bits(bias[vid:vid + 1])[0]) - int(bits(np.array([exact]))[0]))
            assert ulps <= 1, (rb, vid)
            off += ulps
        assert off <= 2


def test_bias_image_layout_and_checks(split):
    g, s, z, base, queries = split
    bias = z["r4_bias"]
    fp = artifacts.source_fingerprint(b"\1" * 32, b"\2" * 32, b"\3" * 32)
    img = artifacts.split_bias_image(bias, 128, 4, fp)
    # the header is the reference's <8sIIIIIIQQ16s>: 64 bytes
    assert len(img) == 64 + 4 * bias.size
    magic, ver, hb, n, dim, lb, rb, pb, fnv, got_fp = struct.unpack_from("<8sIIIIIIQQ16s", img)
    assert (magic, ver, hb, n, dim, lb, rb, pb) == (b"SQMDBIAS", 1, 64, bias.size, 128, 4, 4,
                                                    4 * bias.size)
    assert got_fp == fp and fnv == artifacts.fnv1a64(img[64:])
    # FNV-1a64 known answers
    assert artifacts.fnv1a64(b"") == 0xcbf29ce484222325
    assert artifacts.fnv1a64(b"a") == 0xaf63dc4c8601ec8c
    hdr, back = artifacts.parse_split_bias(img)
    assert np.array_equal(bits(back), bits(bias)) and hdr["fingerprint"] == fp
    bad = bytearray(img)
    bad[64] ^= 1
    with pytest.raises(ValueError):
        artifacts.parse_split_bias(bytes(bad))
    with pytest.raises(ValueError):
        artifacts.parse_split_bias(b"SQMDBIAX" + img[8:])
    with pytest.raises(ValueError):
        artifacts.parse_split_bias(img[:-4])
