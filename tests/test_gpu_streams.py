"""Concurrency on one session: eager batches on two streams of one handle, and
two handles training the stop predictor at the same time.

A session's work counter and scratch are shared by its launches; a launch on
another stream waits on the device for the previous one, so batches enqueued
back to back on two streams give what each gives alone.  Training keeps its
partial-gradient slab per call, so two trainings on two host threads give the
weights each gives alone, bit for bit."""
import ctypes as C
import sys
import threading

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
    say("case streams index 8000 x 128")
    base, queries = data.workload(8000, 128, 32, 8.0, 1200, base_seed=81, query_seed=82)
    ix = rindex.build_pq_index(base, m=32, R=32, res_bits=8, iters=6, train_max=8000,
                               n_router=64)
    return base, queries, ix


class Out:
    def __init__(self, q):
        self.nq = q.shape[0]
        self.dq = native.DeviceBuffer.from_array(np.ascontiguousarray(q, dtype=np.float32))
        self.ids = native.DeviceBuffer.from_array(np.full(self.nq * K, -7, dtype=np.int32))
        self.d = native.DeviceBuffer.from_array(np.zeros(self.nq * K, dtype=np.float32))
        self.cnt = native.DeviceBuffer.from_array(np.full(self.nq, -9, dtype=np.int32))

    def get(self):
        return (self.ids.download(np.int32, self.nq * K),
                self.d.download(np.float32, self.nq * K).view(np.uint32),
                self.cnt.download(np.int32, self.nq))


def launch(calc, o, L, stream):
    calc.search_batch_device(o.dq.ptr, o.nq, L, K, o.ids.ptr, o.d.ptr, o.cnt.ptr, K,
                             correction=1, entry_points=32, stream=stream)


def test_two_streams_one_handle(setup):
    base, queries, ix = setup
    qa, qb = queries[:700], queries[700:1100]
    with Calculator() as calc:
        calc.load_pq_index(ix, with_router=False)
        sa, sb = native.Stream(), native.Stream()
        say("case streams: references one at a time")
        ra, rb = Out(qa), Out(qb)
        launch(calc, ra, 160, sa.ptr)
        sa.synchronize()
        launch(calc, rb, 96, sb.ptr)
        sb.synchronize()
        ref_a, ref_b = ra.get(), rb.get()
        assert (ref_a[2] == K).all() and (ref_b[2] == K).all()
        for rep in range(3):
            say("case streams: back to back, round %d" % rep)
            oa, ob = Out(qa), Out(qb)
            launch(calc, oa, 160, sa.ptr)
            launch(calc, ob, 96, sb.ptr)
            launch(calc, oa, 160, sa.ptr)
            sb.synchronize()
            sa.synchronize()
            for got, ref in ((oa.get(), ref_a), (ob.get(), ref_b)):
                for g, r in zip(got, ref):
                    assert np.array_equal(g, r)


def _train_params(q):
    t = native.RoverLearnedStopTrain()
    t.queries = q.ctypes.data
    t.num_queries = q.shape[0]
    t.pool_size = 96
    t.correction = 1
    t.label_k = 10
    t.entry_points = 8
    t.expand_count = 32
    t.stop_every = 4
    t.min_steps = 4
    t.trace_cap = 128
    t.steps = 150
    t.batch = 2048
    t.lr = 3e-3
    t.seed = 91
    return t


def _weights(calc):
    P = native.lib().rover_mlp_param_count()
    m = [np.zeros(P, np.float32), np.zeros(16, np.float32), np.zeros(16, np.float32)]
    native.check(native.lib().rover_learned_stop_get(calc.h, m[0].ctypes.data, P,
                                                     m[1].ctypes.data, m[2].ctypes.data),
                 "rover_learned_stop_get")
    return m


def test_two_handles_train_concurrently(setup):
    base, queries, ix = setup
    tq = [np.ascontiguousarray(data.gaussian_mixture(800, 128, 32, 8.0, s, centers_seed=81))
          for s in (93, 94)]
    calcs = [Calculator(), Calculator()]
    try:
        for c in calcs:
            c.load_pq_index(ix, with_router=False)
        say("case training alone")
        alone = []
        for c, q in zip(calcs, tq):
            t = _train_params(q)
            native.check(native.lib().rover_learned_stop_train(c.h, C.byref(t)), "train")
            alone.append((_weights(c), t.final_loss))
        say("case training on two threads")
        errs = []
        ts = [_train_params(q) for q in tq]

        def work(c, t):
            rc = native.lib().rover_learned_stop_train(c.h, C.byref(t))
            if rc != native.ROVER_SUCCESS:
                errs.append((rc, native.last_error()))

        th = [threading.Thread(target=work, args=(c, t)) for c, t in zip(calcs, ts)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        assert not errs, errs
        for c, t, (w0, l0) in zip(calcs, ts, alone):
            for a, b in zip(_weights(c), w0):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
            assert t.final_loss == l0
    finally:
        for c in calcs:
            c.close()
