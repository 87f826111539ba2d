"""rover_search_batch_device_noalloc against rover_search_batch_device.

The allocation-free entry point keeps its work counter and seed list in a
fixed per-device array; its results must equal the allocating entry point's
bit for bit (ids, distance bits, counts, counters) on two calls in a row, the
second with a smaller batch, so nothing left in the array by the first call
leaks into the second."""
import sys

import numpy as np
import pytest

from rover_amd import data, native
from rover_amd import index as rindex
from rover_amd.search import Calculator

pytestmark = pytest.mark.gpu

K = 10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    say("case no-alloc index 8000 x 128")
    base, queries = data.workload(8000, 128, 32, 8.0, 400, base_seed=71, query_seed=72)
    ix = rindex.build_pq_index(base, m=32, R=32, res_bits=8, iters=6, train_max=8000,
                               n_router=64)
    calc = Calculator()
    calc.load_pq_index(ix, with_router=False)
    rows = rindex.route_on_host(queries, ix.router_centroids, ix.router_ids, 4)
    yield calc, ix, queries, rows
    calc.close()


def run(calc, q, L, noalloc, rows=None, correction=1, entry_points=32, refine=0):
    nq = q.shape[0]
    dq = native.DeviceBuffer.from_array(np.ascontiguousarray(q, dtype=np.float32))
    de = None if rows is None else native.DeviceBuffer.from_array(
        np.ascontiguousarray(rows, dtype=np.int32))
    ids = native.DeviceBuffer.from_array(np.full(nq * K, -7, dtype=np.int32))
    d = native.DeviceBuffer.from_array(np.zeros(nq * K, dtype=np.float32))
    cnt = native.DeviceBuffer.from_array(np.full(nq, -9, dtype=np.int32))
    st = native.DeviceBuffer.from_array(np.zeros(nq * 4, dtype=np.uint64))
    calc.search_batch_device(dq.ptr, nq, L, K, ids.ptr, d.ptr, cnt.ptr, K,
                             correction=correction, refine_candidates=refine,
                             entry_points=entry_points, d_stats=st.ptr,
                             d_entry_ids=None if de is None else de.ptr,
                             entry_count=0 if rows is None else rows.shape[1],
                             entry_stride=0 if rows is None else rows.shape[1],
                             noalloc=noalloc)
    native.synchronize()
    return (ids.download(np.int32, nq * K).reshape(nq, K),
            d.download(np.float32, nq * K).reshape(nq, K),
            cnt.download(np.int32, nq),
            st.download(np.uint64, nq * 4).reshape(nq, 4))


def same(a, b):
    ia, da, ca, sa = a
    ib, db, cb, sb = b
    assert np.array_equal(ca, cb)
    assert (ca > 0).all()
    assert np.array_equal(ia, ib)
    assert np.array_equal(da.view(np.uint32), db.view(np.uint32))
    assert np.array_equal(sa, sb)


@pytest.mark.parametrize("correction,refine", [(1, 0), (0, 20)])
def test_noalloc_equals_device_batch_lcg_seeds(setup, correction, refine):
    calc, ix, queries, _ = setup
    for nq, L, ep in ((300, 64, 32), (90, 40, 8)):
        say("case noalloc lcg nq=%d L=%d entry_points=%d correction=%d refine=%d"
            % (nq, L, ep, correction, refine))
        q = queries[:nq]
        ref = run(calc, q, L, False, correction=correction, entry_points=ep, refine=refine)
        got = run(calc, q, L, True, correction=correction, entry_points=ep, refine=refine)
        same(got, ref)


def test_noalloc_equals_device_batch_entry_rows(setup):
    calc, ix, queries, rows = setup
    for nq, L in ((400, 48), (37, 32)):
        say("case noalloc rows nq=%d L=%d" % (nq, L))
        ref = run(calc, queries[:nq], L, False, rows=rows[:nq])
        got = run(calc, queries[:nq], L, True, rows=rows[:nq])
        same(got, ref)


def test_noalloc_rejects_what_needs_allocation(setup):
    calc, ix, queries, _ = setup
    say("case noalloc rejections")
    with pytest.raises(native.RoverError) as e:
        run(calc, queries[:4], 32, True, entry_points=16385)
    assert e.value.code == native.ROVER_EBADPARM
    assert "entry points" in str(e.value)
    dq = native.DeviceBuffer.from_array(queries[:4])
    out = native.DeviceBuffer(4 * K * 4)
    with pytest.raises(native.RoverError) as e:
        calc.search_batch_device(dq.ptr, 4, 32, K, out.ptr, out.ptr, out.ptr, K,
                                 route_entries=4, noalloc=True)
    assert e.value.code == native.ROVER_EBADPARM
