"""Pins the oracle's source-window search (orc_search_window) to the
reference's live pipeline (pure_pq_source_pipeline_step,
retrieval_sqmd_pure_pq.c:722-762).

Fixtures: tests/golden/<pq fixture>_w2.npz, written by make_golden.py from
the compiled reference's pipeline with two source turns in flight and its
DSP side run between steps.  Window 1 must reproduce the scalar search
(results and all four counters); window 2 must reproduce the reference's
pipeline bit for bit, counters included, and must differ from window 1 on
some queries (a dropped lookahead would pass the window-1 identity alone).
"""
import numpy as np
import pytest

import pyoracle

W2_SETS = ["pq_d128_m32", "pq_d128_m32_r8"]


def oindex(g):
    return pyoracle.Index(g["graph"], g["codes"], g["codebook"], g["res_codes"],
                          g["res_qparams"], int(g["res_bits"]))


@pytest.mark.parametrize("name", W2_SETS)
def test_window1_is_the_scalar_search(golden, name):
    g = golden(name)
    ix = oindex(g)
    n = g["graph"].shape[0]
    for mode, ep, expand, L in g["search_cases"]:
        key = "m%d_e%d_x%d_L%d" % (mode, ep, expand, L)
        entries = pyoracle.seed_ids(n, int(ep), int(g["entry"]), 42)
        for qi, q in enumerate(g["queries"]):
            ids, d, st, hw = pyoracle.search_window(ix, int(mode), q, entries, int(expand),
                                                    int(L), int(L), 1)
            c = int(g[key + "_counts"][qi])
            assert np.array_equal(ids, g[key + "_ids"][qi][:c]), key
            assert np.array_equal(d.view(np.uint32), g[key + "_dists"][qi][:c].view(np.uint32))
            assert tuple(st) == tuple(int(x) for x in g[key + "_stats"][qi]), key
            assert hw == 1


@pytest.mark.parametrize("name", W2_SETS)
def test_window2_matches_reference_pipeline(golden, name):
    g = golden(name)
    w = golden(name + "_w2")
    ix = oindex(g)
    n = g["graph"].shape[0]
    differs = 0
    for mode, ep, expand, L in w["search_cases"]:
        key = "m%d_e%d_x%d_L%d" % (mode, ep, expand, L)
        entries = pyoracle.seed_ids(n, int(ep), int(g["entry"]), 42)
        for qi, q in enumerate(g["queries"]):
            ids, d, st, hw = pyoracle.search_window(ix, int(mode), q, entries, int(expand),
                                                    int(L), int(L), 2)
            c = int(w[key + "_counts"][qi])
            assert len(ids) == c, key
            assert np.array_equal(ids, w[key + "_ids"][qi][:c]), key
            assert np.array_equal(d.view(np.uint32), w[key + "_dists"][qi][:c].view(np.uint32))
            turns, work, high, _steps = (int(x) for x in w[key + "_stats"][qi])
            assert st[2] == turns and hw == high, key
            # scored = seeds scored + work items; seeds = entries minus duplicates
            seeds = len(np.unique(entries))
            assert st[0] == seeds + work, key
            c1 = int(g[key + "_counts"][qi])
            differs += not np.array_equal(ids, g[key + "_ids"][qi][:c1])
    assert differs > 0


def test_window2_invariants_tiny_graph():
    """A path graph 0-1-2-...: with one seed at 0 and L >= n both windows
    visit every node and return all of them in (distance, id) order; the
    pipeline dispatches one turn per pool member (each is expanded exactly
    once) and, with one unexpanded member at a time on a path, never has two
    turns pending."""
    n, dim = 24, 4
    rng = np.random.default_rng(3)
    x = rng.normal(size=(n, dim)).astype(np.float32) * 3
    m, ks = 2, 16
    cb = np.zeros((m, ks, dim // m), dtype=np.float32)
    # codebook = the points' own sub-vectors for the first 16 ids, so ADC is
    # exact for those; the rest reuse them (distances still well defined)
    for s in range(m):
        cb[s, :, :] = x[:ks, s * 2:(s + 1) * 2]
    codes = np.zeros((n, 1), dtype=np.uint8)
    codes[:, 0] = (np.arange(n) % 16) | ((np.arange(n) % 16) << 4)
    graph = np.full((n, 2), -1, dtype=np.int32)
    for i in range(n):
        row = [j for j in (i - 1, i + 1) if 0 <= j < n]
        graph[i, :len(row)] = row
    ix = pyoracle.Index(graph, codes, cb)
    q = x[5]
    for window in (1, 2):
        ids, d, st, hw = pyoracle.search_window(ix, 0, q, np.array([0], np.int32), 2, n, n,
                                                window)
        assert sorted(ids.tolist()) == list(range(n))
        keys = list(zip(d.tolist(), ids.tolist()))
        assert keys == sorted(keys)
        assert st[2] == n and st[0] == n
        assert hw == 1


def test_window2_matches_reference_live():
    """Fresh inputs against the compiled reference's pipeline (both windows,
    both correction modes)."""
    ref = pyoracle.reference()
    if ref is None:
        pytest.skip("compiled reference absent")
    from rover_amd import index as rindex
    rng = np.random.default_rng(17)
    n, dim, m, ks, R = 900, 32, 8, 256, 14
    x = (rng.normal(size=(n, dim)) * 5 + 20).astype(np.float32)
    cb = rindex.train_pq(x, m, ks, iters=3)
    codes = rindex.encode_pq(x, cb)
    graph = rindex.build_graph(x, R)
    res = np.zeros((n, dim // 2), dtype=np.uint8)
    qp = np.zeros((n, 2), dtype=np.float32)
    h = ref.ref_residual_open(x.ctypes.data, codes.ctypes.data, cb.ctypes.data, n,
                              dim, m, ks, 4, res.ctypes.data, qp.ctypes.data)
    assert h
    try:
        ix = pyoracle.Index(graph, codes, cb, res, qp, 4)
        for t in range(6):
            q = (rng.normal(size=dim) * 5 + 20).astype(np.float32)
            entries = pyoracle.seed_ids(n, 8, 0, 42)
            for mode in (0, 1):
                for window in (1, 2):
                    L = 48
                    ids_r = np.zeros(L, dtype=np.int32)
                    d_r = np.zeros(L, dtype=np.float32)
                    st_r = np.zeros(4, dtype=np.uint64)
                    rc = ref.ref_pipeline_search(graph.ctypes.data, R, n, codes.ctypes.data,
                                                 128, cb.ctypes.data, dim, m, ks,
                                                 q.ctypes.data, entries.ctypes.data,
                                                 len(entries), R, L, L, window, 1, mode, h,
                                                 ids_r.ctypes.data, d_r.ctypes.data, L,
                                                 st_r.ctypes.data)
                    ids, d, st, hw = pyoracle.search_window(ix, mode, q, entries, R, L, L,
                                                            window)
                    assert rc == len(ids)
                    assert np.array_equal(ids, ids_r)
                    assert np.array_equal(d.view(np.uint32), d_r.view(np.uint32))
                    assert st[2] == st_r[0] and hw == st_r[2]
    finally:
        ref.ref_residual_close(h)
