"""Third-level refine (error4): the oracle's encoder and ERROR4 refine, the
artifact reader/writer and the library's SHA-256, pinned to the reference.

The *_error4 fixtures hold what the reference's own compiled C produced on
the PQ fixtures: retrieval_build_error4's artifact (its residual decoder as
the reconstruction) and retrieval_final_refine in CURRENT and ERROR4 modes
on the fixture's refine candidates (tests/golden/make_golden.py).
"""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

import pyoracle
from rover_amd import artifacts

PQ_SETS = ["pq_d128_m32", "pq_d96_m48", "pq_d64_ks16", "pq_d128_m32_r8"]
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def e4(name):
    z = np.load(os.path.join(GOLDEN, name + "_error4.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def recon_all(g):
    ix = pyoracle.Index(g["graph"], g["codes"], g["codebook"], g["res_codes"],
                        g["res_qparams"], int(g["res_bits"]))
    return np.stack([pyoracle.residual_decode(ix, i) for i in range(len(g["base"]))])


@pytest.mark.parametrize("name", PQ_SETS)
def test_error4_encoder_matches_reference_builder(golden, name):
    g, e = golden(name), e4(name)
    codes, qp = pyoracle.error4_encode(g["base"], recon_all(g))
    assert np.array_equal(codes, e["e4_codes"])
    assert np.array_equal(bits(qp), bits(e["e4_qparams"]))


@pytest.mark.parametrize("name", PQ_SETS)
def test_error4_refine_matches_reference(golden, name):
    g, e = golden(name), e4(name)
    F, K = int(g["refine_F"]), int(g["refine_K"])
    mode, ep, expand, L = (int(v) for v in g["refine_source"])
    cand = g["m%d_e%d_x%d_L%d_ids" % (mode, ep, expand, L)][:, :F]
    plain = pyoracle.Index(g["graph"], g["codes"], g["codebook"], g["res_codes"],
                           g["res_qparams"], int(g["res_bits"]))
    third = pyoracle.Index(g["graph"], g["codes"], g["codebook"], g["res_codes"],
                           g["res_qparams"], int(g["res_bits"]), e4_codes=e["e4_codes"],
                           e4_qparams=e["e4_qparams"])
    changed = 0
    for qi, q in enumerate(g["queries"]):
        i1, d1 = pyoracle.refine(plain, q, cand[qi], K)
        assert np.array_equal(i1, e["refine_current_ids"][qi])
        assert np.array_equal(bits(d1), bits(e["refine_current_dists"][qi]))
        i2, d2 = pyoracle.refine(third, q, cand[qi], K)
        assert np.array_equal(i2, e["refine_e4_ids"][qi])
        assert np.array_equal(bits(d2), bits(e["refine_e4_dists"][qi]))
        changed += int(not np.array_equal(i1, i2))
    # the 4-bit-residual fixtures exercise a level that changes the ranking
    # (after 8-bit residuals the error left is too small to)
    assert changed > 0 or int(g["res_bits"]) == 8


@pytest.mark.parametrize("name", PQ_SETS)
def test_error4_level_bounds(golden, name):
    """What the quantizer fixes independently of the reference: each
    dequantized error is within half a step of the true error, zero is the
    row minimum, codes span 0..15 when the error has a range, and adding the
    level brings the reconstruction closer to the raw vector."""
    g, e = golden(name), e4(name)
    x = g["base"].astype(np.float64)
    rec = recon_all(g).astype(np.float64)
    err = x - rec
    dim = x.shape[1]
    codes = np.zeros((len(x), dim), dtype=np.int64)
    codes[:, 0::2] = e["e4_codes"][:, :(dim + 1) // 2] & 15
    codes[:, 1::2] = e["e4_codes"][:, :dim // 2] >> 4
    s = e["e4_qparams"][:, 0:1].astype(np.float64)
    z = e["e4_qparams"][:, 1:2].astype(np.float64)
    assert np.all(z[:, 0] == (g["base"] - recon_all(g)).min(axis=1))
    level = codes * s + z
    assert np.all(np.abs(level - err) <= s * 0.5 * (1 + 1e-4) + 1e-6)
    has_range = (err.max(axis=1) - err.min(axis=1)) > 0
    assert np.all(codes[has_range].min(axis=1) == 0)
    assert np.all(codes[has_range].max(axis=1) == 15)
    before = ((x - rec) ** 2).sum(axis=1)
    after = ((x - rec - level) ** 2).sum(axis=1)
    assert np.all(after <= before + 1e-9)
    assert after.mean() < 0.05 * before.mean()


def test_oracle_encoder_detects_planted_mistakes(golden):
    """A rounding-mode or nibble-order slip in an encoder fails the
    reference's codes: the same rows encoded with round-half-even or with the
    high nibble first differ from the fixture."""
    g, e = golden("pq_d128_m32"), e4("pq_d128_m32")
    err = (g["base"] - recon_all(g)).astype(np.float32)
    lo = err.min(axis=1, keepdims=True)
    scale = (np.maximum(err.max(axis=1, keepdims=True) - lo, np.float32(1e-12))
             / np.float32(15)).astype(np.float32)
    t = (err - lo) * (np.float32(1) / scale)
    half_even = np.clip(np.rint(t), 0, 15).astype(np.uint8)
    swapped = (half_even[:, 0::2] << 4) | half_even[:, 1::2]
    floor_half = np.clip(np.floor(t + np.float32(0.5)), 0, 15).astype(np.uint8)
    ref_pack = floor_half[:, 0::2] | (floor_half[:, 1::2] << 4)
    assert np.array_equal(ref_pack, e["e4_codes"])
    assert not np.array_equal(swapped, e["e4_codes"])


@pytest.mark.parametrize("name", PQ_SETS)
def test_artifact_reader_and_writer_match_reference_file(name):
    e = e4(name)
    image = e["artifact"].tobytes()
    art = artifacts.parse_error4(image)
    assert art.magic == b"PQERR4V1" and art.n == len(e["e4_qparams"])
    assert np.array_equal(art.codes, e["e4_codes"])
    assert np.array_equal(bits(art.qparams), bits(e["e4_qparams"]))
    assert bytes(e["header"]) == image[:256]
    # rewriting with the header's digests reproduces the reference file byte for byte
    again = artifacts.error4_image(art.codes, art.qparams, art.dim, art.l1_bits,
                                   art.residual_bits, art.base_sha256, art.identity_sha256,
                                   art.maps_sha256)
    assert again == image
    assert art.maps_sha256 == artifacts.maps_digest(None)


def test_artifact_reader_rejects_what_the_reference_rejects():
    image = bytearray(e4("pq_d64_ks16")["artifact"].tobytes())
    artifacts.parse_error4(bytes(image))

    def bad(mutate):
        b = bytearray(image)
        mutate(b)
        with pytest.raises(ValueError):
            artifacts.parse_error4(bytes(b))

    bad(lambda b: b.__setitem__(slice(0, 8), b"PQERR4V2"))
    bad(lambda b: b.__setitem__(slice(8, 12), (2).to_bytes(4, "little")))
    bad(lambda b: b.__setitem__(slice(28, 32), (2048).to_bytes(4, "little")))
    bad(lambda b: b.__setitem__(300, b[300] ^ 1))           # payload digest
    bad(lambda b: b.__setitem__(len(b) - 1, b[-1] ^ 1))     # payload digest
    bad(lambda b: b.extend(b"\0"))                          # size
    neg = bytearray(image)
    neg[256:260] = np.float32(-1.0).tobytes()
    body = hashlib.sha256(bytes(neg[256:])).digest()
    neg[56:88] = body
    with pytest.raises(ValueError):
        artifacts.parse_error4(bytes(neg))


def test_library_sha256_matches_hashlib():
    from rover_amd import native
    L = native.lib()
    rng = np.random.default_rng(5)
    for n in (0, 1, 3, 55, 56, 57, 63, 64, 65, 119, 120, 128, 1000, 4097):
        data = rng.integers(0, 256, n, dtype=np.uint8)
        out = (C.c_uint8 * 32)()
        L.rover_sha256(data.ctypes.data if n else None, n, out)
        assert bytes(out) == hashlib.sha256(data.tobytes()).digest(), n
    image = e4("pq_d96_m48")["artifact"]
    out = (C.c_uint8 * 32)()
    L.rover_sha256(image[256:].ctypes.data, image.size - 256, out)
    assert bytes(out) == image[56:88].tobytes()


def sq_refine_inputs(golden, b):
    g, s = golden("pq_d128_m32"), golden("sq_d128")
    z, r = golden("sqmd_d128"), golden("sq_refine_d128")
    return g, s, z, r


@pytest.mark.parametrize("b", [8, 4])
def test_sq_error4_artifact_matches_reference(golden, b):
    """The SQ third level: the product's host encoder (rover_amd.index.
    encode_error4 over decode_sq_residual) reproduces the codes and qparams
    of the SQERR4V1 artifact the reference built (retrieval_build_error4,
    pq = 0, over retrieval_final_refine_decode_sq), and the artifact writer
    reproduces the whole file from them and its digests."""
    from rover_amd import index as rindex
    g, s, z, r = sq_refine_inputs(golden, b)
    ref_art = r["sq%d_artifact" % b].tobytes()
    art = artifacts.parse_error4(ref_art)
    assert art.magic == b"SQERR4V1"
    base = g["base"]
    dim = base.shape[1]
    recon = rindex.decode_sq_residual(s["sq%d_codes" % b], b, s["sq%d_scale" % b],
                                      s["sq%d_zero" % b], z["sq%d_res_codes" % b],
                                      z["sq%d_res_qparams" % b], int(z["sq%d_res_bits" % b]),
                                      dim)
    # the reconstruction is the oracle's SQ decoder's, bit for bit
    oix = pyoracle.Index(g["graph"], sq_codes=s["sq%d_codes" % b], sq_bits=b,
                         sq_scale=float(s["sq%d_scale" % b]), sq_zero=float(s["sq%d_zero" % b]),
                         res_codes=z["sq%d_res_codes" % b],
                         res_qparams=z["sq%d_res_qparams" % b],
                         res_bits=int(z["sq%d_res_bits" % b]), dim=dim)
    for i in range(0, len(base), 97):
        assert np.array_equal(bits(pyoracle.sq_decode(oix, i)), bits(recon[i]))
    codes, qp = rindex.encode_error4(base, recon)
    assert np.array_equal(codes, art.codes)
    assert np.array_equal(bits(qp), bits(art.qparams))
    mine = artifacts.error4_image(codes, qp, dim, art.l1_bits, art.residual_bits,
                                  base_sha256=art.base_sha256,
                                  identity_sha256=art.identity_sha256, pq=False)
    assert mine == ref_art


@pytest.mark.parametrize("b", [8, 4])
@pytest.mark.parametrize("F", [20, 64])
def test_sq_refine_oracle_matches_reference(golden, b, F):
    """The oracle's refine with the SQ decoder (orc_refine without a PQ
    codebook) equals retrieval_final_refine with retrieval_final_refine_decode_sq
    on the SQMD traversal's outputs: CURRENT, CURRENT with a permuted
    original-id map (ties by original id), and ERROR4 with the reference's
    SQERR4V1 codes."""
    g, s, z, r = sq_refine_inputs(golden, b)
    dim = g["base"].shape[1]
    art = artifacts.parse_error4(r["sq%d_artifact" % b].tobytes())
    common = dict(sq_codes=s["sq%d_codes" % b], sq_bits=b,
                  sq_scale=float(s["sq%d_scale" % b]), sq_zero=float(s["sq%d_zero" % b]),
                  res_codes=z["sq%d_res_codes" % b], res_qparams=z["sq%d_res_qparams" % b],
                  res_bits=int(z["sq%d_res_bits" % b]), dim=dim)
    plain = pyoracle.Index(g["graph"], **common)
    third = pyoracle.Index(g["graph"], e4_codes=art.codes, e4_qparams=art.qparams, **common)
    cand = z["sq%d_c1_e32_x32_L64_ids" % b]
    counts = z["sq%d_c1_e32_x32_L64_counts" % b]
    changed = 0
    for qi, q in enumerate(g["queries"]):
        c = min(int(counts[qi]), F)
        for ix, tag, mp in ((plain, "current", None), (plain, "current_perm", r["perm"]),
                            (third, "error4", None)):
            i1, d1 = pyoracle.refine(ix, q, cand[qi][:c], 10, mp)
            key = "sq%d_F%d_%s" % (b, F, tag)
            assert np.array_equal(i1, r[key + "_ids"][qi]), (key, qi)
            assert np.array_equal(bits(d1), bits(r[key + "_dists"][qi])), (key, qi)
        changed += int(not np.array_equal(r["sq%d_F%d_current_ids" % (b, F)][qi],
                                          r["sq%d_F%d_error4_ids" % (b, F)][qi]))
    # the fixture exercises a third level that changes some rankings
    assert changed > 0
