"""The graph builders' stable radix sort (rv_radix_sort, 8-bit digits):
64-bit keys alone and with 32-bit values, full and partial bit ranges,
sizes around the 4096-key tile and the scan's block boundaries, many equal
keys; the output must equal NumPy's stable sort (keys ordered on the bit
range, equal keys in input order with their values), and the inputs stay
as they were."""
import ctypes as C
import sys

import numpy as np
import pytest

from rover_amd import native

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def radix(keys, vals, begin, end):
    L = native.lib()
    L.rv_radix_scratch_bytes.restype = C.c_size_t
    L.rv_radix_scratch_bytes.argtypes = [C.c_size_t, C.c_int]
    L.rv_radix_sort.restype = C.c_int
    L.rv_radix_sort.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                C.c_int, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p]
    n = len(keys)
    dk = native.DeviceBuffer.from_array(keys)
    dko = native.DeviceBuffer(max(n, 1) * 8)
    dv = native.DeviceBuffer.from_array(vals) if vals is not None else None
    dvo = native.DeviceBuffer(max(n, 1) * 4) if vals is not None else None
    sb = L.rv_radix_scratch_bytes(n, 1 if vals is not None else 0)
    scratch = native.DeviceBuffer(sb)
    rc = L.rv_radix_sort(dk.ptr, dko.ptr, dv.ptr if dv else None, dvo.ptr if dvo else None, n,
                         begin, end, scratch.ptr, sb, None)
    assert rc == 0, native.last_error()
    native.synchronize()
    out_k = dko.download(np.uint64, n)
    out_v = dvo.download(np.uint32, n) if dvo else None
    assert np.array_equal(dk.download(np.uint64, n), keys)  # input unchanged
    return out_k, out_v


@pytest.mark.parametrize("n", [1, 255, 4096, 4097, 70000, 4096 * 300 + 17])
@pytest.mark.parametrize("bits", [(0, 64), (0, 37), (8, 40)])
def test_radix_sort_matches_numpy_stable_sort(n, bits):
    rng = np.random.default_rng(n * 7 + bits[1])
    begin, end = bits
    keys = rng.integers(0, 2**63, size=n, dtype=np.uint64) | \
        (rng.integers(0, 2, size=n, dtype=np.uint64) << np.uint64(63))
    # many equal keys on the sorted bits: stability is visible in the values
    keys[rng.random(n) < 0.3] = keys[0]
    vals = np.arange(n, dtype=np.uint32) ^ np.uint32(0x5a5a5a5a)
    say("case radix n=%d bits=%s" % (n, bits))
    sel = (keys >> np.uint64(begin)) & np.uint64((1 << (end - begin)) - 1 if end - begin < 64
                                                 else 0xFFFFFFFFFFFFFFFF)
    order = np.argsort(sel, kind="stable")
    k, v = radix(keys, vals, begin, end)
    assert np.array_equal(k, keys[order])
    assert np.array_equal(v, vals[order])
    k2, _ = radix(keys, None, begin, end)
    assert np.array_equal(k2, keys[order])
