"""On-disk index artifacts the search path reads.

PQERR4V1 / SQERR4V1 -- the third-level ("error4") refine artifact: the
4-bit quantized error left after the PQ + residual (or SQ + residual)
reconstruction, one (scale, zero) pair and (dim+1)/2 code bytes per vector.
Layout (read by retrieval_final_refine.c:128-160, written by
retrieval_build_artifacts.c:25-66, pq_refine_artifacts.py:47-65 and
generate_final_refine.py:134-145):

    0    magic[8]                  b"PQERR4V1" or b"SQERR4V1"
    8    u32 version = 1, n, dim, l1_bits, residual_bits, block = 4096,
         two format flags = 1
    40   u64 qparams_bytes = 8 n, code_bytes_total = blocks * 4096
    56   sha256 of everything after the header (the payload)
    88   sha256 of the three base files (codes, residual, residual qparams)
    184  sha256 of the identity file (PQ codebook, or raw vectors for SQ)
    216  sha256 of the id maps (graph->original then graph->storage, int32),
         or of nothing when there are none
    248  8 zero bytes
    256  f32 (scale, zero) [n][2]
    256 + 8 n   4096-byte blocks of 4096 // ((dim+1)/2) rows, zero-padded

Codes are low nibble first (even dimension in the low nibble).

SQMDBIAS -- the split-distance bias of the SQ traversal (written by
generate_sqmd_correction_bias.py:86-190, read by
retrieval_rpc_demo_support.c:209-299): a 64-byte header
<8sIIIIIIQQ16s> (magic, version 1, header bytes 64, n, dim, l1 bits 4,
residual bits 4 or 8, payload bytes 4 n, FNV-1a64 of the payload, the
first 16 bytes of SHA-256 over the SHA-256 digests of the l1, residual and
qparams files) and f32 bias [n], bias_v = sum_d (2 a_d r_d + r_d^2) with
a = SQ level (code * scale + zero) and r = residual level, in float64.
"""
import hashlib
import struct
from dataclasses import dataclass

import numpy as np

HEADER = struct.Struct("<8s8I2Q32s96s32s32s8s")
BLOCK = 4096
PQ_MAGIC, SQ_MAGIC = b"PQERR4V1", b"SQERR4V1"
assert HEADER.size == 256


def sha256(data):
    return hashlib.sha256(data).digest()


def maps_digest(graph_to_original=None, n=None):
    """Digest of the id maps an artifact is bound to (PQ: no storage map, so
    graph->storage is the identity)."""
    if graph_to_original is None:
        return sha256(b"")
    g = np.ascontiguousarray(graph_to_original, dtype="<i4")
    n = len(g) if n is None else n
    return sha256(g.tobytes() + np.arange(n, dtype="<i4").tobytes())


def error4_geometry(n, dim):
    """(code bytes per row, rows per block, blocks)."""
    cb = (dim + 1) // 2
    vpb = BLOCK // cb
    if vpb <= 0:
        raise ValueError("dim too large for 4096-byte blocks")
    return cb, vpb, (n + vpb - 1) // vpb


@dataclass
class Error4Artifact:
    magic: bytes
    n: int
    dim: int
    l1_bits: int
    residual_bits: int
    codes: np.ndarray      # [n][(dim+1)/2] uint8, rows in storage order
    qparams: np.ndarray    # [n][2] float32 (scale, zero)
    payload_sha256: bytes
    base_sha256: tuple     # three 32-byte digests
    identity_sha256: bytes
    maps_sha256: bytes


def error4_image(codes, qparams, dim, l1_bits, residual_bits, base_sha256=(b"\0" * 32,) * 3,
                 identity_sha256=b"\0" * 32, maps_sha256=None, pq=True):
    """The artifact file image (bytes) for flat codes [n][(dim+1)/2] and
    qparams [n][2]."""
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    qparams = np.ascontiguousarray(qparams, dtype="<f4")
    n = qparams.shape[0]
    cb, vpb, blocks = error4_geometry(n, dim)
    if codes.shape != (n, cb) or qparams.shape != (n, 2):
        raise ValueError("codes must be [n][(dim+1)/2] and qparams [n][2]")
    body = np.zeros(blocks * BLOCK, dtype=np.uint8)
    rows = np.arange(n)
    starts = (rows // vpb) * BLOCK + (rows % vpb) * cb
    body[starts[:, None] + np.arange(cb)] = codes
    payload = qparams.tobytes() + body.tobytes()
    if maps_sha256 is None:
        maps_sha256 = maps_digest(None)
    if len(base_sha256) != 3 or any(len(h) != 32 for h in base_sha256):
        raise ValueError("three 32-byte base digests expected")
    header = HEADER.pack(PQ_MAGIC if pq else SQ_MAGIC, 1, n, dim, l1_bits, residual_bits,
                         BLOCK, 1, 1, 8 * n, blocks * BLOCK, sha256(payload),
                         b"".join(base_sha256), identity_sha256, maps_sha256, b"\0" * 8)
    return header + payload


def parse_error4(image, verify=True):
    """Reads an artifact image with the checks of the reference's open
    (header fields, section sizes, payload digest, finite qparams with
    scale >= 0); raises ValueError on any mismatch."""
    buf = memoryview(bytes(image)) if not isinstance(image, (bytes, memoryview)) else memoryview(image)
    if len(buf) < HEADER.size:
        raise ValueError("artifact shorter than its header")
    (magic, version, n, dim, lb, rb, block, f1, f2, qb, cbytes, payload_sha, bases,
     identity, maps, _pad) = HEADER.unpack(bytes(buf[:HEADER.size]))
    if magic not in (PQ_MAGIC, SQ_MAGIC):
        raise ValueError("not an error4 artifact")
    if n <= 0 or not 0 < dim <= 8192:
        raise ValueError("invalid dimensions")
    cb, vpb, blocks = error4_geometry(n, dim)
    if (version != 1 or lb not in (4, 8) or rb not in (4, 8) or block != BLOCK or f1 != 1
            or f2 != 1 or qb != 8 * n or cbytes != blocks * BLOCK
            or len(buf) != HEADER.size + qb + cbytes):
        raise ValueError("error4 header or size mismatch")
    if verify and sha256(buf[HEADER.size:]) != payload_sha:
        raise ValueError("error4 payload digest mismatch")
    qp = np.frombuffer(buf, dtype="<f4", count=2 * n, offset=HEADER.size).reshape(n, 2)
    if not np.isfinite(qp).all() or (qp[:, 0] < 0).any():
        raise ValueError("invalid error4 quantization parameters")
    body = np.frombuffer(buf, dtype=np.uint8, count=cbytes, offset=HEADER.size + qb)
    rows = np.arange(n)
    starts = (rows // vpb) * BLOCK + (rows % vpb) * cb
    codes = body[starts[:, None] + np.arange(cb)]
    return Error4Artifact(magic, n, dim, lb, rb, codes, qp.astype(np.float32), payload_sha,
                          (bases[:32], bases[32:64], bases[64:]), identity, maps)


def file_sha256(path, offset=0):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        f.seek(offset)
        for chunk in iter(lambda: f.read(8 << 20), b""):
            h.update(chunk)
    return h.digest()


def check_error4_identity(art, base_paths, identity_path):
    """The file-binding checks the library leaves to the caller
    (retrieval_final_refine.c:150-166): the digests of the three base files
    and of the identity file (PQ codebook) must equal the header's."""
    for i, p in enumerate(base_paths):
        if file_sha256(p) != art.base_sha256[i]:
            raise ValueError("error4 artifact is bound to another base file: %s" % p)
    if file_sha256(identity_path) != art.identity_sha256:
        raise ValueError("error4 artifact is bound to another identity file: %s" % identity_path)


BIAS_HEADER = struct.Struct("<8sIIIIIIQQ16s")
BIAS_MAGIC = b"SQMDBIAS"
assert BIAS_HEADER.size == 64


def fnv1a64(data):
    """FNV-1a 64 over bytes (the bias payload checksum)."""
    v = 14695981039346656037
    for b in bytes(data):
        v = ((v ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return v


def _fnv1a64_np(payload):
    # the same recurrence over a NumPy byte array, 64-bit wrap-around
    v = np.uint64(14695981039346656037)
    prime = np.uint64(1099511628211)
    with np.errstate(over="ignore"):
        for b in np.frombuffer(payload, dtype=np.uint8):
            v = (v ^ np.uint64(b)) * prime
    return int(v)


def _unpack_u4(packed, dim):
    out = np.empty((packed.shape[0], dim), dtype=np.uint8)
    out[:, 0::2] = packed[:, : (dim + 1) // 2] & np.uint8(0x0F)
    if dim > 1:
        out[:, 1::2] = packed[:, : dim // 2] >> np.uint8(4)
    return out


def split_bias(sq_codes, sq_scale, sq_zero, res_codes, res_qparams, res_bits, dim,
               chunk=8192):
    """Per-vector split-distance bias from SQ4 codes [n][(dim+1)/2] and
    residual codes [n][(dim+1)/2 or dim] with (scale, zero) [n][2]: the
    float64 arithmetic of generate_sqmd_correction_bias.py:159-176 (NumPy
    row sums), cast to f32."""
    n = sq_codes.shape[0]
    if res_bits not in (4, 8):
        raise ValueError("split bias needs 4- or 8-bit residuals")
    qp = np.asarray(res_qparams, dtype=np.float32)
    if not np.isfinite(qp).all() or (qp[:, 0] < 0).any():
        raise ValueError("invalid residual quantization parameters")
    out = np.empty(n, dtype="<f4")
    for s0 in range(0, n, chunk):
        s1 = min(n, s0 + chunk)
        a = _unpack_u4(np.asarray(sq_codes[s0:s1]), dim).astype(np.float64) * float(sq_scale) \
            + float(sq_zero)
        rc = (_unpack_u4(np.asarray(res_codes[s0:s1]), dim) if res_bits == 4
              else np.asarray(res_codes[s0:s1])).astype(np.float64)
        scales = qp[s0:s1, 0].astype(np.float64)[:, None]
        zeros = qp[s0:s1, 1].astype(np.float64)[:, None]
        r = rc * scales + zeros
        out[s0:s1] = np.sum(2.0 * a * r + r * r, axis=1, dtype=np.float64).astype("<f4")
    if not np.isfinite(out).all():
        raise ValueError("non-finite split bias")
    return out


def source_fingerprint(l1_sha256, residual_sha256, qparams_sha256):
    """The bias's binding to its three source files (digests as bytes)."""
    return hashlib.sha256(bytes(l1_sha256) + bytes(residual_sha256) +
                          bytes(qparams_sha256)).digest()[:16]


def split_bias_image(bias, dim, res_bits, fingerprint=b"\0" * 16):
    """The SQMDBIAS file image of a bias array."""
    bias = np.ascontiguousarray(bias, dtype="<f4")
    payload = bias.tobytes()
    return BIAS_HEADER.pack(BIAS_MAGIC, 1, 64, bias.shape[0], dim, 4, res_bits, len(payload),
                            _fnv1a64_np(payload), bytes(fingerprint)) + payload


def parse_split_bias(image):
    """Header fields and the bias of an SQMDBIAS image, checked as the
    reference's loader checks them (fingerprint left to the caller)."""
    buf = bytes(image)
    if len(buf) < 64:
        raise ValueError("split bias image too short")
    magic, ver, hb, n, dim, lb, rb, pb, fnv, fp = BIAS_HEADER.unpack_from(buf, 0)
    if magic != BIAS_MAGIC or ver != 1 or hb != 64 or pb != 4 * n or len(buf) != 64 + pb:
        raise ValueError("bad split bias header")
    if _fnv1a64_np(buf[64:]) != fnv:
        raise ValueError("split bias payload checksum mismatch")
    bias = np.frombuffer(buf, dtype="<f4", count=n, offset=64)
    if not np.isfinite(bias).all():
        raise ValueError("non-finite split bias")
    return dict(n=n, dim=dim, l1_bits=lb, residual_bits=rb, fingerprint=fp), bias.copy()
