"""Suspended and resumed searches (RoverSearchBatch.reorder_step): every
search still running after N expansions saves its pool, flags, counters and
learned-stop state and leaves; the saved searches then resume, ordered by
the predictor's stop probability.  Resuming must give exactly the
unsuspended search: ids, distance bits, counts and all four counters, for
the learned stop (predictor every 1, 4 or 12 expansions) and for the plain
search (the predictor only orders the resumption), per-query L and k, pools
of one entry, and suspension steps from 1 to past every search's end."""
import sys

import numpy as np
import pytest

from rover_amd import data, native
from rover_amd import index as rindex
from rover_amd.search import Calculator

pytestmark = pytest.mark.gpu
EX = native.HASH_EXACT


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def mid():
    say("case reorder: 12k x 128 index")
    base, queries = data.workload(12000, 128, 64, 8.0, 400)
    ix = rindex.build_pq_index(base, m=32, R=32, res_bits=8, iters=6,
                               train_max=12000, n_router=0)
    c = Calculator()
    c.load_pq_index(ix)
    P = native.lib().rover_mlp_param_count()
    rng = np.random.default_rng(31)
    params = (rng.normal(size=P) * 0.3).astype(np.float32)
    yield c, ix, queries, params
    c.close()


def set_model(calc, params, every, min_steps):
    mean = np.zeros(16, np.float32)
    istd = np.ones(16, np.float32)
    native.check(native.lib().rover_learned_stop_set(
        calc.h, params.ctypes.data, len(params), mean.ctypes.data, istd.ctypes.data,
        every, 10, min_steps), "rover_learned_stop_set")


def same(a, b):
    assert np.array_equal(a[2], b[2])
    for i in range(len(a[2])):
        c = int(a[2][i])
        assert np.array_equal(a[0][i, :c], b[0][i, :c]), i
        assert np.array_equal(bits(a[1][i, :c]), bits(b[1][i, :c])), i
    assert np.array_equal(a[3], b[3])


def test_reorder_learned_stop_is_exact(mid):
    calc, ix, queries, params = mid
    for every, min_steps in ((12, 12), (4, 4), (1, 1)):
        set_model(calc, params, every, min_steps)
        for thr in (0.5, 0.9):
            kw = dict(correction=1, entry_points=32, stats=True, learned_stop=1,
                      stop_threshold=thr, hash_log2=EX)
            ref = calc.search_batch(queries, 200, 10, **kw)
            for step in (1, 12, 48, 5000):
                say("case reorder learned every %d thr %.1f step %d" % (every, thr, step))
                same(calc.search_batch(queries, 200, 10, reorder_step=step, **kw), ref)


def test_reorder_plain_is_exact(mid):
    """Without the learned stop the predictor only orders the resumption."""
    calc, ix, queries, params = mid
    set_model(calc, params, 12, 12)
    rng = np.random.default_rng(37)
    for mode in (0, 1):
        for case in ("fixed", "mixed", "tiny"):
            if case == "fixed":
                Ls, ks = 256, 10
            elif case == "mixed":
                ks = rng.choice([1, 10, 100], size=len(queries)).astype(np.int32)
                Ls = np.array([rng.integers(k, 301) for k in ks], dtype=np.int32)
            else:
                Ls = rng.choice([1, 2, 5, 10], size=len(queries)).astype(np.int32)
                ks = Ls.copy()
            kw = dict(correction=mode, entry_points=32, stats=True, hash_log2=EX)
            ref = calc.search_batch(queries, Ls, ks, **kw)
            for step in (2, 24):
                say("case reorder plain mode %d %s step %d" % (mode, case, step))
                same(calc.search_batch(queries, Ls, ks, reorder_step=step, **kw), ref)


def test_reorder_device_batch_and_refine(mid):
    """The device entry point with the final refine; a repeated launch (the
    per-query sets take new stamps) gives the same results."""
    calc, ix, queries, params = mid
    set_model(calc, params, 12, 12)
    nq, k = len(queries), 10
    dq = native.DeviceBuffer.from_array(queries)
    di = native.DeviceBuffer(nq * k * 4)
    dd = native.DeviceBuffer(nq * k * 4)
    dc = native.DeviceBuffer(nq * 4)
    ref = calc.search_batch(queries, 128, k, correction=1, entry_points=32,
                            refine_candidates=20, learned_stop=1, stop_threshold=0.9,
                            hash_log2=EX)
    for rep in range(3):
        say("case reorder device batch, repeat", rep)
        calc.search_batch_device(dq.ptr, nq, 128, k, di.ptr, dd.ptr, dc.ptr, k,
                                 correction=1, entry_points=32, refine_candidates=20,
                                 learned_stop=1, stop_threshold=0.9, hash_log2=EX,
                                 reorder_step=24)
        native.synchronize()
        got = (di.download(np.int32, nq * k).reshape(nq, k),
               dd.download(np.float32, nq * k).reshape(nq, k), dc.download(np.int32, nq))
        assert np.array_equal(got[2], ref[2])
        assert np.array_equal(got[0], ref[0])
        assert np.array_equal(bits(got[1]), bits(ref[1]))


def test_reorder_needs_exact_sets_and_window_1(mid):
    calc, ix, queries, params = mid
    set_model(calc, params, 12, 12)
    q = queries[:4]
    for kw in (dict(hash_log2=12), dict(hash_log2=EX, source_window=2),
               dict(hash_log2=EX, waves_per_query=2)):
        with pytest.raises(native.RoverError):
            calc.search_batch(q, 64, 10, correction=1, entry_points=32, reorder_step=12, **kw)
    other = Calculator()
    try:
        other.load_pq_index(ix)
        with pytest.raises(native.RoverError):  # no predictor loaded
            other.search_batch(q, 64, 10, correction=1, entry_points=32, hash_log2=EX,
                               reorder_step=12)
    finally:
        other.close()
