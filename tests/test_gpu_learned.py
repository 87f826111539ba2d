"""Learned early termination on the GPU: the training kernels' gradients
against the CPU reference (oracle/mlp_ref.py), the plain search as the
fallback, and what the trained predictor gains over the plain search."""
import ctypes as C
import sys

import numpy as np
import pytest

import mlp_ref
from rover_amd import data, native
from rover_amd import index as rindex
from rover_amd.search import Calculator

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_forward_backward_matches_cpu_reference():
    say("case mlp gradients")
    L = native.lib()
    rng = np.random.default_rng(7)
    n = 1000
    p = (rng.normal(size=mlp_ref.P) * 0.25).astype(np.float32)
    x = rng.normal(size=(n, 16)).astype(np.float32)
    y = (rng.random(n) > 0.4).astype(np.float32)
    w = (rng.random(n) + 0.25).astype(np.float32)
    w[::7] = 0.0
    assert L.rover_mlp_param_count() == mlp_ref.P
    dp, dx, dy, dw = (native.DeviceBuffer.from_array(a) for a in (p, x, y, w))
    dg = native.DeviceBuffer.from_array(np.zeros(mlp_ref.P, dtype=np.float32))
    dl = native.DeviceBuffer.from_array(np.zeros(1, dtype=np.float32))
    do = native.DeviceBuffer(n * 4)
    native.check(L.rover_mlp_forward_backward(dp.ptr, dx.ptr, dy.ptr, dw.ptr, n,
                                              dg.ptr, dl.ptr, do.ptr), "fwd_bwd")
    native.synchronize()
    g = dg.download(np.float32, mlp_ref.P).astype(np.float64)
    loss = float(dl.download(np.float32, 1)[0])
    out = do.download(np.float32, n).astype(np.float64)
    rl, rg, rout = mlp_ref.loss_and_grad(p.astype(np.float64), x, y, w)
    assert np.allclose(out, rout, rtol=1e-4, atol=1e-4)
    assert abs(loss - rl) <= 1e-4 * abs(rl)
    scale = np.abs(rg).max()
    assert np.abs(g - rg).max() <= 2e-4 * scale, np.abs(g - rg).max() / scale
    # a gradient that is zero in the reference is zero on the GPU (masks)
    zero = rg == 0
    assert np.all(g[zero] == 0)


def test_adam_step_matches_formula():
    say("case adam")
    L = native.lib()
    rng = np.random.default_rng(8)
    p = rng.normal(size=mlp_ref.P).astype(np.float32)
    g = rng.normal(size=mlp_ref.P).astype(np.float32)
    m1 = np.zeros(mlp_ref.P, dtype=np.float32)
    m2 = np.zeros(mlp_ref.P, dtype=np.float32)
    dp, dg, d1, d2 = (native.DeviceBuffer.from_array(a) for a in (p, g, m1, m2))
    lr, scale = 1e-3, 0.5
    native.check(L.rover_mlp_adam_step(dp.ptr, dg.ptr, d1.ptr, d2.ptr, 1, C.c_float(lr),
                                       C.c_float(scale)), "adam")
    native.synchronize()
    got = dp.download(np.float32, mlp_ref.P).astype(np.float64)
    gs = g.astype(np.float64) * scale
    a = 0.1 * gs
    v = 0.001 * gs * gs
    lr_t = lr * np.sqrt(1 - 0.999) / (1 - 0.9)
    expect = p - lr_t * a / (np.sqrt(v) + 1e-8)
    assert np.allclose(got, expect, rtol=1e-5, atol=1e-6)


@pytest.fixture(scope="module")
def trained():
    say("case learned-stop training")
    base, queries = data.workload(12000, 128, 64, 8.0, 300)
    train_q = data.gaussian_mixture(3000, 128, 64, 8.0, 44, centers_seed=42)
    ix = rindex.build_pq_index(base, m=32, R=32, res_bits=8, iters=6,
                               train_max=12000, n_router=128)
    gt = data.ground_truth(base, queries, k=10)
    calc = Calculator()
    calc.load_pq_index(ix)
    t = native.RoverLearnedStopTrain()
    t.queries = train_q.ctypes.data
    t.num_queries = train_q.shape[0]
    t.pool_size = 128
    t.correction = 1
    t.label_k = 10
    t.entry_points = 1
    t.route_entries = 4
    t.expand_count = 32
    t.stop_every = 2
    t.min_steps = 4
    t.trace_cap = 256
    t.steps = 600
    t.batch = 4096
    t.lr = 3e-3
    t.seed = 44
    native.check(native.lib().rover_learned_stop_train(calc.h, C.byref(t)), "train")
    say("trained: loss %.4f samples %d positive %.3f" % (t.final_loss, t.samples,
                                                         t.positive_frac))
    yield calc, ix, queries, gt, t
    calc.close()


def test_training_converges(trained):
    calc, ix, queries, gt, t = trained
    assert t.samples > 10000
    assert 0.05 < t.positive_frac < 0.95
    # better than the constant predictor's BCE at the same label balance
    p = t.positive_frac
    const = -(p * np.log(p) + (1 - p) * np.log(1 - p))
    assert t.final_loss < 0.8 * const, (t.final_loss, const)


def test_threshold_above_one_is_the_plain_search(trained):
    """With a threshold no probability reaches, the learned search falls back
    to the plain fixed-L search: identical ids, distances and expansions."""
    calc, ix, queries, gt, t = trained
    a = calc.search_batch(queries, 128, 10, stats=True, correction=1,
                          route_entries=4)
    b = calc.search_batch(queries, 128, 10, stats=True, learned_stop=1,
                          stop_threshold=1.5, correction=1, route_entries=4)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    assert np.array_equal(a[3], b[3])


def test_learned_stop_gain(trained):
    """The trained predictor cuts expansions at the same beam width while
    recall@10 stays within 0.02 of the plain search that judges it (online
    correction, 8-bit residuals, routed per-query entries)."""
    calc, ix, queries, gt, t = trained
    ids_p, _, cnt_p, st_p = calc.search_batch(queries, 128, 10, stats=True,
                                              correction=1, route_entries=4)
    ids_l, _, cnt_l, st_l = calc.search_batch(queries, 128, 10, stats=True,
                                              learned_stop=1, stop_threshold=0.9,
                                              correction=1, route_entries=4)
    rp = data.recall_at_k(ids_p, cnt_p, gt, 10)
    rl = data.recall_at_k(ids_l, cnt_l, gt, 10)
    ep = st_p[:, 2].mean()
    el = st_l[:, 2].mean()
    say("plain recall %.4f expanded %.1f | learned recall %.4f expanded %.1f"
        % (rp, ep, rl, el))
    assert rl >= rp - 0.02
    assert el <= 0.85 * ep


def test_training_is_reproducible(trained):
    """The same traces, seed and schedule give the same weights bit for bit
    (the gradient is summed over workgroups in a fixed order)."""
    calc, ix, queries, gt, t = trained
    P = native.lib().rover_mlp_param_count()

    def weights():
        m = [np.zeros(P, np.float32), np.zeros(16, np.float32), np.zeros(16, np.float32)]
        native.check(native.lib().rover_learned_stop_get(calc.h, m[0].ctypes.data, P,
                                                         m[1].ctypes.data, m[2].ctypes.data),
                     "rover_learned_stop_get")
        return m

    first = weights()
    t2 = native.RoverLearnedStopTrain()
    C.memmove(C.addressof(t2), C.addressof(t), C.sizeof(t))
    native.check(native.lib().rover_learned_stop_train(calc.h, C.byref(t2)), "train again")
    second = weights()
    for a, b in zip(first, second):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert t2.final_loss == t.final_loss


def test_fused_predictor_matches_cpu_forward(trained):
    """The predictor fused in the search kernel (grouped BF16 weights, FP32
    accumulation) against oracle/mlp_ref.py's float64 forward pass: the
    features the plain search records every stop_every expansions
    (rover_search_trace_device), normalized in float32 as the kernel does,
    give per record a probability; the first record at or after min_steps
    whose probability reaches 0.9 is where the learned search must stop
    (its expansion count), or it runs as the plain search does.  Queries
    with a probability within 1e-4 of 0.9 before the stop (float32 against
    float64 arithmetic differ by about 1e-6 there) are left out as
    ambiguous; every other query must stop exactly there."""
    calc, ix, queries, gt, t = trained
    say("case fused predictor vs mlp_ref")
    L, K, cap = 128, 10, 512
    P = native.lib().rover_mlp_param_count()
    params = np.zeros(P, np.float32)
    mean = np.zeros(16, np.float32)
    istd = np.zeros(16, np.float32)
    native.check(native.lib().rover_learned_stop_get(calc.h, params.ctypes.data, P,
                                                     mean.ctypes.data, istd.ctypes.data),
                 "rover_learned_stop_get")
    rows = calc.route_entries(queries, 4)
    nq = queries.shape[0]
    dq = native.DeviceBuffer.from_array(queries)
    de = native.DeviceBuffer.from_array(rows)
    dids, dd, dc = (native.DeviceBuffer(nq * K * 4), native.DeviceBuffer(nq * K * 4),
                    native.DeviceBuffer(nq * 4))
    dfeat = native.DeviceBuffer(nq * cap * 16 * 4)
    dsteps = native.DeviceBuffer(nq * cap * 4)
    dmeta = native.DeviceBuffer(nq * 18 * 4)
    b = calc._batch_struct(dq.ptr, nq, L, K, native.METHOD_PQ, 1, 0, 1, 32, dids.ptr,
                           dd.ptr, dc.ptr, K, None, 0, 0, 0, 0, 0.0, None, None,
                           native.ENTRY_POINT_SEED, de.ptr, 4, 4)
    native.check(native.lib().rover_search_trace_device(
        calc.h, C.byref(b), int(t.stop_every), 10, cap, dfeat.ptr, dsteps.ptr, dmeta.ptr,
        None), "rover_search_trace_device")
    native.synchronize()
    feats = dfeat.download(np.float32, nq * cap * 16).reshape(nq, cap, 16)
    steps = dsteps.download(np.int32, nq * cap).reshape(nq, cap)
    meta = dmeta.download(np.int32, nq * 18).reshape(nq, 18)
    _, _, _, st_plain = calc.search_batch(queries, L, K, stats=True, correction=1,
                                          entry_ids=rows)
    _, _, _, st_learn = calc.search_batch(queries, L, K, stats=True, correction=1,
                                          entry_ids=rows, learned_stop=1,
                                          stop_threshold=0.9)
    checked = early = 0
    for q in range(nq):
        recs = int(meta[q, 0])
        assert 0 < recs < cap
        assert np.all(steps[q, :recs] == t.stop_every * np.arange(1, recs + 1))
        x = ((feats[q, :recs] - mean).astype(np.float32) * istd).astype(np.float32)
        logit, _ = mlp_ref.forward(params.astype(np.float64), x)
        p = 1.0 / (1.0 + np.exp(-logit))
        live = steps[q, :recs] >= t.min_steps
        hit = np.nonzero(live & (p >= 0.9))[0]
        stop_at = int(steps[q, hit[0]]) if len(hit) else int(st_plain[q, 2])
        upto = hit[0] + 1 if len(hit) else recs
        if np.any(live[:upto] & (np.abs(p[:upto] - 0.9) < 1e-4)):
            continue
        checked += 1
        early += stop_at < int(st_plain[q, 2])
        assert int(st_learn[q, 2]) == stop_at, (q, int(st_learn[q, 2]), stop_at)
    say("fused predictor: %d of %d queries checked, %d stopped early" % (checked, nq, early))
    assert checked >= 0.9 * nq
    assert early >= 0.2 * checked


def test_training_ignores_stale_device_memory(trained):
    """The trace search writes only the records it keeps; the training must
    not read the rest.  Device memory freed with NaN bit patterns in it (as
    recycled allocations may hold) must not reach the weights."""
    calc, ix, queries, gt, t0 = trained
    say("case training over recycled NaN memory")
    q = np.ascontiguousarray(data.gaussian_mixture(800, 128, 64, 8.0, 45, centers_seed=42))
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
    t.steps = 60
    t.batch = 2048
    t.lr = 3e-3
    t.seed = 91
    nrec = q.shape[0] * t.trace_cap
    junk = [native.DeviceBuffer.from_array(np.full(nbytes, 0xFF, dtype=np.uint8))
            for nbytes in (nrec * 16 * 4, nrec * 4, nrec * 4, nrec * 4)]
    native.synchronize()
    for j in junk:
        j.close()
    native.check(native.lib().rover_learned_stop_train(calc.h, C.byref(t)), "train")
    P = native.lib().rover_mlp_param_count()
    m = [np.zeros(P, np.float32), np.zeros(16, np.float32), np.zeros(16, np.float32)]
    native.check(native.lib().rover_learned_stop_get(calc.h, m[0].ctypes.data, P,
                                                     m[1].ctypes.data, m[2].ctypes.data),
                 "rover_learned_stop_get")
    assert np.isfinite(t.final_loss)
    assert all(np.isfinite(a).all() for a in m)
