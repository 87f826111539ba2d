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
