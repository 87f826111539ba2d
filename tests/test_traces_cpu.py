"""The refine traces' writers (rover_trace_*, host code, no device): the
pure-PQ demo's text trace (retrieval_rpc_demo_pure_pq.c:615-645) and the SQ
pipeline's binary P15TRC01 trace (retrieval_rpc_demo_pipeline.c:3047-3050,
3276-3308), parsed back and compared with what was written."""
import ctypes
import struct

import numpy as np
import pytest

from rover_amd import native


def libc_hexfloat(v):
    """C's %a of a double, from the C library itself."""
    libc = ctypes.CDLL(None)
    buf = ctypes.create_string_buffer(64)
    libc.snprintf(buf, 64, b"%a", ctypes.c_double(v))
    return buf.value.decode()


def test_text_trace(tmp_path):
    rng = np.random.default_rng(3)
    ids = rng.integers(0, 10000, size=12).astype(np.int32)
    d = np.sort(rng.random(12).astype(np.float32) * 1e3)
    d[0] = 0.0
    dec = rng.normal(size=(5, 8)).astype(np.float32)
    oi, od = ids[:4][::-1].copy(), d[:4] * np.float32(0.5)
    path = tmp_path / "trace.txt"
    with native.Trace(path, native.TRACE_TEXT) as t:
        t.query(64, 7, ids, d, oi, od, refine_budget=5, decoded=dec)
        t.query(32, 8, ids[:3], d[:3], ids[:2], d[:2])
    lines = path.read_text().splitlines()
    assert lines[0] == "Q 64 7 12 5"
    for i in range(12):
        tag, a, b = lines[1 + i].split()
        assert tag == "C" and int(a) == ids[i] and float.fromhex(b) == float(d[i])
        assert b == libc_hexfloat(float(d[i]))
    for i in range(5):
        f = lines[13 + i].split()
        assert f[0] == "V" and int(f[1]) == ids[i]
        assert [float.fromhex(x) for x in f[2:]] == [float(v) for v in dec[i]]
    for i in range(4):
        tag, a, b = lines[18 + i].split()
        assert tag == "O" and int(a) == oi[i] and float.fromhex(b) == float(od[i])
    assert lines[22] == "Q 32 8 3 0" and len(lines) == 22 + 1 + 3 + 2


def test_p15_binary_trace(tmp_path):
    rng = np.random.default_rng(4)
    ids = rng.integers(0, 1000, size=9).astype(np.int32)
    d = rng.random(9).astype(np.float32)
    path = tmp_path / "trace.bin"
    with native.Trace(path, native.TRACE_P15) as t:
        t.query(64, 3, ids, d, ids[:2], d[:2], entry_points=32)
    raw = path.read_bytes()
    assert raw[:8] == b"P15TRC01"
    ep, pool, q, cnt = struct.unpack_from("<4I", raw, 8)
    assert (ep, pool, q, cnt) == (32, 64, 3, 9)
    o = 24
    assert np.array_equal(np.frombuffer(raw, "<i4", 9, o), ids)
    assert np.array_equal(np.frombuffer(raw, "<f4", 9, o + 36).view(np.uint32),
                          d.view(np.uint32))
    (n,) = struct.unpack_from("<I", raw, o + 72)
    assert n == 2
    assert np.array_equal(np.frombuffer(raw, "<i4", 2, o + 76), ids[:2])
    assert len(raw) == o + 76 + 16
    # the binary trace refuses to overwrite (the pipeline opens it "wbx")
    with pytest.raises(native.RoverError):
        native.Trace(path, native.TRACE_P15)
