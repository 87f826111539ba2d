"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Run in a container that has the Rover source tree (oracle/Makefile builds
oracle/_ref/librover_ref.so from it):

    python tests/golden/make_golden.py           # every fixture
    python tests/golden/make_golden.py error4    # only the *_error4 files
    python tests/golden/make_golden.py r8        # only pq_d128_m32_r8 (+ error4)
    python tests/golden/make_golden.py sq_refine # only sq_refine_d128
    python tests/golden/make_golden.py w2        # only the *_w2 files

Every expected output in the fixtures comes from the reference's own scalar
C: retrieval_pure_pq_search_scalar_with_policy, the live pipeline with two
source turns in flight (pure_pq_source_pipeline_step), retrieval_pq_build_lut,
retrieval_pq_adc_ids_batch_scalar, retrieval_pq_train/encode_batch, the
PQRES01 residual writer/reader and its evaluator/decoder,
retrieval_final_refine (CURRENT and ERROR4), retrieval_build_error4, the
live SQ traversal (sqmd_cpu_state_run_fake_dsp, retrieval_sqmd_make_seed_ids),
the exact-provider graph build (retrieval_search_build_graph_ex), the SQ quantize/distance kernels and
retrieval_sqmd_make_seed_ids.  Inputs are small synthetic indexes; base
vectors are multiples of 1/16 so they are stored exactly as uint16.
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import ctypes as C  # noqa: E402

import pyoracle  # noqa: E402
from rover_amd import artifacts, data, index  # noqa: E402


def _p(a):
    return a.ctypes.data


def make_base(n, dim, comps, sigma, seed, nq):
    base, queries = data.workload(n, dim, comps, sigma, nq, base_seed=seed,
                                  query_seed=seed + 1)
    base = (np.round(base * 16.0) / 16.0).astype(np.float32)
    queries = (np.round(queries * 16.0) / 16.0).astype(np.float32)
    return base, queries


def ref_index(R, base, m, ks, res_bits, iters=8):
    ref = pyoracle.reference()
    n, dim = base.shape
    cb = np.zeros((m, ks, dim // m), dtype=np.float32)
    assert ref.ref_pq_train(_p(base), n, dim, m, ks, iters, _p(cb)) == 0
    code_bytes = index.code_bytes(m, ks)
    codes = np.zeros((n, code_bytes), dtype=np.uint8)
    assert ref.ref_pq_encode(_p(base), n, dim, m, ks, _p(cb), _p(codes)) == 0
    rb = index.residual_code_bytes(dim, res_bits)
    res = np.zeros((n, rb), dtype=np.uint8)
    qp = np.zeros((n, 2), dtype=np.float32)
    handle = ref.ref_residual_open(_p(base), _p(codes), _p(cb), n, dim, m, ks,
                                   res_bits, _p(res), _p(qp))
    assert handle
    graph = index.build_graph(base, R)
    return cb, codes, res, qp, graph, handle


def ref_search(handle, graph, codes, cb, query, entries, expand, L, num_out,
               mode, chunk_vectors, sq=None):
    ref = pyoracle.reference()
    n, R = graph.shape
    m, ks, sub = cb.shape
    ids = np.zeros(num_out, dtype=np.int32)
    d = np.zeros(num_out, dtype=np.float32)
    st = np.zeros(4, dtype=np.uint64)
    sq_codes, sq_bits, sq_scale, sq_zero = sq if sq else (None, 8, 1.0, 0.0)
    rc = ref.ref_search(_p(graph), R, n, _p(codes), chunk_vectors, _p(cb),
                        m * sub, m, ks, _p(query), _p(entries), len(entries),
                        expand, L, num_out, mode, handle,
                        None if sq_codes is None else _p(sq_codes), sq_bits,
                        sq_scale, sq_zero, _p(ids), _p(d), num_out, _p(st))
    return rc, ids, d, st


def pq_case(name, n, dim, m, ks, R, res_bits, comps, nq, Ls, expands, seed):
    ref = pyoracle.reference()
    base, queries = make_base(n, dim, comps, 8.0, seed, nq)
    cb, codes, res, qp, graph, handle = ref_index(R, base, m, ks, res_bits)
    entry = index.medoid(base)
    out = dict(base_x16=(base * 16).astype(np.uint16),
               queries_x16=(queries * 16).astype(np.uint16),
               codebook=cb, codes=codes, res_codes=res, res_qparams=qp,
               res_bits=np.int32(res_bits), graph=graph, entry=np.int32(entry))
    # LUT and ADC for every query, ADC of a sample of ids
    sample = np.arange(0, n, max(1, n // 97), dtype=np.int32)
    luts, adcs, evals, decs = [], [], [], []
    for q in queries:
        lut = np.zeros(m * ks, dtype=np.float32)
        assert ref.ref_build_lut(_p(q), _p(cb), m, ks, dim // m, _p(lut)) == 0
        a = np.zeros(len(sample), dtype=np.float32)
        assert ref.ref_adc_ids(_p(codes), n, codes.shape[1], 300, _p(sample),
                               len(sample), _p(lut), m, ks, _p(a)) == 0
        e = np.zeros(len(sample), dtype=np.float32)
        for i, vid in enumerate(sample):
            v = np.zeros(1, dtype=np.float32)
            assert ref.ref_residual_evaluate(handle, _p(q), int(vid), _p(v)) == 0
            e[i] = v[0]
        if len(luts) < 2:
            luts.append(lut)
        adcs.append(a)
        evals.append(e)
    for vid in sample[:16]:
        v = np.zeros(dim, dtype=np.float32)
        assert ref.ref_residual_decode(handle, int(vid), _p(v)) == 0
        decs.append(v)
    out.update(adc_sample=sample, luts=np.stack(luts), adcs=np.stack(adcs),
               online_evals=np.stack(evals), decoded=np.stack(decs))
    # searches: (mode, entry_points, expand, L) -> full pool, stats
    cases = []
    results = {}
    for mode in (0, 1):
        for ep in (1, 32):
            entries = pyoracle.seed_ids(n, ep, entry, 42)
            for expand in expands:
                for L in Ls:
                    key = "m%d_e%d_x%d_L%d" % (mode, ep, expand, L)
                    ids_all, d_all, st_all, cnt_all = [], [], [], []
                    for q in queries:
                        rc, ids, d, st = ref_search(handle, graph, codes, cb, q,
                                                    entries, expand, L, L, mode,
                                                    300)
                        assert rc >= 0, (key, rc)
                        ids_all.append(ids)
                        d_all.append(d)
                        st_all.append(st)
                        cnt_all.append(rc)
                    results[key + "_ids"] = np.stack(ids_all)
                    results[key + "_dists"] = np.stack(d_all)
                    results[key + "_stats"] = np.stack(st_all)
                    results[key + "_counts"] = np.array(cnt_all, dtype=np.int32)
                    cases.append((mode, ep, expand, L))
    out.update(results)
    out["search_cases"] = np.array(cases, dtype=np.int32)
    # final refine of the first F pool members -> top k
    F, K = 20, 10
    ref_ids, ref_d = [], []
    for qi, q in enumerate(queries):
        key = "m0_e1_x%d_L%d" % (expands[0], Ls[-1])
        cand = np.ascontiguousarray(results[key + "_ids"][qi][:F])
        ids = cand.copy()
        d = np.zeros(F, dtype=np.float32)
        rc = ref.ref_refine(handle, n, dim, _p(q), _p(ids), _p(d), F, K, None)
        assert rc == K, rc
        ref_ids.append(ids[:K].copy())
        ref_d.append(d[:K].copy())
    out.update(refine_F=np.int32(F), refine_K=np.int32(K),
               refine_source=np.array([0, 1, expands[0], Ls[-1]], dtype=np.int32),
               refine_ids=np.stack(ref_ids), refine_dists=np.stack(ref_d))
    # error cases: -1 in the middle of a row; out-of-range entry
    bad = graph.copy()
    bad[entry, 3] = -1
    rc_bad_row, _, _, _ = ref_search(handle, bad, codes, cb, queries[0],
                                     np.array([entry], dtype=np.int32), R, 16, 16,
                                     0, 300)
    rc_bad_entry, _, _, _ = ref_search(handle, graph, codes, cb, queries[0],
                                       np.array([entry, n], dtype=np.int32), R,
                                       16, 16, 0, 300)
    out.update(rc_bad_row=np.int32(rc_bad_row), rc_bad_entry=np.int32(rc_bad_entry))
    ref.ref_residual_close(handle)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")
    return base, queries, graph, entry


def sq_case(name, base, queries, graph, entry):
    """SQ distances through the reference traversal's policy hook."""
    ref = pyoracle.reference()
    n, dim = base.shape
    out = {}
    dummy_cb = np.zeros((16, 256, dim // 16), dtype=np.float32)
    dummy_codes = np.zeros((n, 16), dtype=np.uint8)
    for bits in (8, 4):
        scale, zero = index.sq_params(base, bits)
        codes = index.encode_sq(base, bits, scale, zero)
        out["sq%d_scale" % bits] = np.float32(scale)
        out["sq%d_zero" % bits] = np.float32(zero)
        out["sq%d_codes" % bits] = codes
        # quantized queries and distances for a sample of ids
        qq_all, dist_all = [], []
        sample = np.arange(0, n, max(1, n // 61))
        for q in queries:
            qq = np.zeros(dim, dtype=np.uint8)
            assert ref.ref_quantize_query(_p(q), dim, bits, scale, zero, _p(qq)) == 0
            qq_all.append(qq)
            dist_all.append(np.array(
                [ref.ref_sq_distance(_p(qq), _p(np.ascontiguousarray(codes[i])), dim,
                                     bits, scale) for i in sample], dtype=np.float32))
        out["sq%d_qq" % bits] = np.stack(qq_all)
        out["sq%d_sample" % bits] = sample.astype(np.int32)
        out["sq%d_sample_dists" % bits] = np.stack(dist_all)
        for ep in (1, 32):
            entries = pyoracle.seed_ids(n, ep, entry, 42)
            for L in (16, 64):
                key = "sq%d_e%d_L%d" % (bits, ep, L)
                ids_all, d_all, st_all = [], [], []
                for q in queries:
                    rc, ids, d, st = ref_search(None, graph, dummy_codes, dummy_cb, q,
                                                entries, graph.shape[1], L, L, 2, n,
                                                sq=(codes, bits, scale, zero))
                    assert rc == L, (key, rc)
                    ids_all.append(ids)
                    d_all.append(d)
                    st_all.append(st)
                out[key + "_ids"] = np.stack(ids_all)
                out[key + "_dists"] = np.stack(d_all)
                out[key + "_stats"] = np.stack(st_all)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def seeds_case(name):
    """retrieval_sqmd_make_seed_ids with n = 65536: its id (s/65536)%n is the
    LCG state's high half, which is also what the pure-PQ mapping
    (s*n)>>32 gives for this n."""
    ref = pyoracle.reference()
    out = {}
    for seed in (42, 1, 12345):
        ids = np.zeros(64, dtype=np.int32)
        cnt = ref.ref_sqmd_seed_ids(65536, 64, -1, seed, _p(ids), 64)
        out["sqmd_seed%d" % seed] = ids[:cnt].copy()
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def sqmd_case(name, pq_name="pq_d128_m32", sq_name="sq_d128"):
    """The reference's live SQ traversal: retrieval_sqmd_make_seed_ids, seed
    distances and sqmd_cpu_state_run_fake_dsp (distance-only heaps), with SQ
    code distances (correction 0) or exact SQ + residual distances
    (correction 1), on the PQ fixture's base, graph and queries and the SQ
    fixture's codes.  Residual codes come from rover_amd.index's encoder
    (they are inputs here).  Writes <name>.npz."""
    ref = pyoracle.reference()
    g = dict(np.load(os.path.join(HERE, pq_name + ".npz"), allow_pickle=False))
    s = dict(np.load(os.path.join(HERE, sq_name + ".npz"), allow_pickle=False))
    base = g["base_x16"].astype(np.float32) / np.float32(16.0)
    queries = g["queries_x16"].astype(np.float32) / np.float32(16.0)
    graph, entry = g["graph"], int(g["entry"])
    n, dim = base.shape
    out = dict(entry=np.int32(entry))
    cases = []
    for bits, res_bits in ((8, 4), (4, 8)):
        codes = s["sq%d_codes" % bits]
        scale, zero = np.float32(s["sq%d_scale" % bits]), np.float32(s["sq%d_zero" % bits])
        res, rqp = index.encode_sq_residual(base, codes, bits, scale, zero, res_bits)
        out["sq%d_res_codes" % bits] = res
        out["sq%d_res_qparams" % bits] = rqp
        out["sq%d_res_bits" % bits] = np.int32(res_bits)
        for corr in (0, 1):
            for ep, expand, L in ((1, 32, 16), (32, 32, 16), (32, 32, 64), (8, 12, 40)):
                key = "sq%d_c%d_e%d_x%d_L%d" % (bits, corr, ep, expand, L)
                ids_all, d_all, cnt_all = [], [], []
                for q in queries:
                    res_out = []
                    for lanes in (8, 1):
                        ids = np.zeros(L, dtype=np.int32)
                        d = np.zeros(L, dtype=np.float32)
                        seeds = np.zeros(ep, dtype=np.int32)
                        rc = ref.ref_sqmd_search(_p(graph), graph.shape[1], n, _p(codes), dim,
                                                 bits, scale, zero, _p(res), _p(rqp), res_bits,
                                                 corr, _p(q), ep, entry, 42, expand, L, lanes,
                                                 _p(ids), _p(d), L, _p(seeds))
                        assert rc > 0, (key, rc)
                        res_out.append((rc, ids.copy(), d.copy(), seeds.copy()))
                    # the lane split does not change the result order
                    assert res_out[0][0] == res_out[1][0]
                    assert np.array_equal(res_out[0][1], res_out[1][1])
                    assert np.array_equal(res_out[0][2].view(np.uint32),
                                          res_out[1][2].view(np.uint32))
                    rc, ids, d, seeds = res_out[0]
                    ids_all.append(ids)
                    d_all.append(d)
                    cnt_all.append(rc)
                out[key + "_ids"] = np.stack(ids_all)
                out[key + "_dists"] = np.stack(d_all)
                out[key + "_counts"] = np.array(cnt_all, dtype=np.int32)
                out["seeds_e%d" % ep] = seeds
                cases.append((bits, corr, ep, expand, L))
    out["cases"] = np.array(cases, dtype=np.int32)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def split_case(name, pq_name="pq_d128_m32", sq_name="sq_d128"):
    """The SQ traversal ranked by the split distance, through the
    reference's own arithmetic (retrieval_quantize_query_l1_aligned_u8,
    retrieval_compute_u8_u4_l1_distance_scalar, the residual dot kernels)
    composed as the pipeline composes them, and its fake-DSP turn loop;
    SQ4 codes of the SQ fixture, 4- and 8-bit residuals from
    rover_amd.index's encoder and biases from rover_amd.artifacts.split_bias
    (inputs here).  Writes <name>.npz."""
    ref = pyoracle.reference()
    g = dict(np.load(os.path.join(HERE, pq_name + ".npz"), allow_pickle=False))
    s = dict(np.load(os.path.join(HERE, sq_name + ".npz"), allow_pickle=False))
    base = g["base_x16"].astype(np.float32) / np.float32(16.0)
    queries = g["queries_x16"].astype(np.float32) / np.float32(16.0)
    graph, entry = g["graph"], int(g["entry"])
    n, dim = base.shape
    codes = s["sq4_codes"]
    scale, zero = np.float32(s["sq4_scale"]), np.float32(s["sq4_zero"])
    out = dict(entry=np.int32(entry))
    cases = []
    for res_bits in (4, 8):
        res, rqp = index.encode_sq_residual(base, codes, 4, scale, zero, res_bits)
        bias = artifacts.split_bias(codes, scale, zero, res, rqp, res_bits, dim)
        out["r%d_res_codes" % res_bits] = res
        out["r%d_res_qparams" % res_bits] = rqp
        out["r%d_bias" % res_bits] = bias
        for ep, expand, L in ((1, 32, 16), (32, 32, 16), (32, 32, 64), (8, 12, 40)):
            key = "r%d_e%d_x%d_L%d" % (res_bits, ep, expand, L)
            ids_all, d_all, cnt_all = [], [], []
            for q in queries:
                ids = np.zeros(L, dtype=np.int32)
                d = np.zeros(L, dtype=np.float32)
                rc = ref.ref_sqmd_split_search(_p(graph), graph.shape[1], n, _p(codes), dim,
                                               scale, zero, _p(res), _p(rqp), res_bits,
                                               _p(bias), _p(q), ep, entry, 42, expand, L, 8,
                                               _p(ids), _p(d), L)
                assert rc > 0, (key, rc)
                ids_all.append(ids)
                d_all.append(d)
                cnt_all.append(rc)
            out[key + "_ids"] = np.stack(ids_all)
            out[key + "_dists"] = np.stack(d_all)
            out[key + "_counts"] = np.array(cnt_all, dtype=np.int32)
            cases.append((res_bits, ep, expand, L))
    out["cases"] = np.array(cases, dtype=np.int32)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def graph_case(name):
    """The reference's search-build with the exact provider
    (retrieval_search_build_graph_ex, one iteration) on small inputs:
    integer-coordinate data (distance ties) and a mixture, prune modes 1 and
    0.  Writes <name>.npz."""
    ref = pyoracle.reference()
    out, cases = {}, []
    specs = ((600, 16, 16, 32, 1, 4, True), (600, 30, 12, 24, 0, 4, True),
             (1000, 32, 24, 64, 1, 16, False), (400, 8, 8, 8, 1, 2, True))
    for ci, (n, dim, r, cand, prune, comps, integer) in enumerate(specs):
        base, _ = data.workload(n, dim, comps, 8.0, 1, base_seed=70 + ci, query_seed=80 + ci)
        base = np.ascontiguousarray(np.round(base) if integer else base, dtype=np.float32)
        g = np.zeros((n, r), dtype=np.int32)
        assert ref.ref_build_exact(_p(base), n, dim, r, cand, 42, prune, 4, _p(g)) == 0
        out["base%d" % ci] = base
        out["graph%d" % ci] = g
        cases.append((n, dim, r, cand, prune, 42))
    out["cases"] = np.array(cases, dtype=np.int64)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def rnn_case(name):
    """The reference's RNN-Descent (retrieval_rnn_build_graph_ex with four
    worker threads, then retrieval_rnn_export_topk) on small inputs:
    integer-coordinate data (distance ties) and mixtures, several (s, r, t1,
    t2).  Writes <name>.npz."""
    ref = pyoracle.reference()
    out, cases = {}, []
    specs = ((600, 16, 8, 16, 2, 3, 4, True), (900, 32, 16, 32, 3, 2, 8, False),
             (500, 24, 12, 24, 4, 4, 4, True), (300, 8, 16, 16, 2, 2, 2, False))
    for ci, (n, dim, s_, r, t1, t2, comps, integer) in enumerate(specs):
        base, _ = data.workload(n, dim, comps, 8.0, 1, base_seed=90 + ci, query_seed=95 + ci)
        base = np.ascontiguousarray(np.round(base) if integer else base, dtype=np.float32)
        g = np.zeros((n, r), dtype=np.int32)
        assert ref.ref_rnn_build(_p(base), n, dim, s_, r, t1, t2, 42, 4, r, _p(g)) == 0
        out["base%d" % ci] = base
        out["graph%d" % ci] = g
        cases.append((n, dim, s_, r, t1, t2, 42))
    out["cases"] = np.array(cases, dtype=np.int64)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def error4_case(name):
    """Third-level refine for an existing PQ fixture: the reference's
    retrieval_build_error4 writes the PQERR4V1 artifact (its residual decoder
    as the reconstruction), and retrieval_final_refine in ERROR4 and CURRENT
    modes refines the fixture's refine candidates.  Writes <name>_error4.npz."""
    ref = pyoracle.reference()
    g = dict(np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False))
    base = g["base_x16"].astype(np.float32) / np.float32(16.0)
    queries = g["queries_x16"].astype(np.float32) / np.float32(16.0)
    cb, codes = g["codebook"], g["codes"]
    m, ks, sub = cb.shape
    n, dim = base.shape
    bits = int(g["res_bits"])
    res = np.zeros_like(g["res_codes"])
    qp = np.zeros_like(g["res_qparams"])
    handle = ref.ref_residual_open(_p(base), _p(codes), _p(cb), n, dim, m, ks,
                                   bits, _p(res), _p(qp))
    assert handle
    assert np.array_equal(res, g["res_codes"]) and np.array_equal(qp, g["res_qparams"])
    header = np.zeros(256, dtype=np.uint8)
    e4_qp = np.zeros((n, 2), dtype=np.float32)
    e4_codes = np.zeros((n, (dim + 1) // 2), dtype=np.uint8)
    rc = ref.ref_error4_build(handle, _p(base), _p(codes), codes.shape[1], _p(cb),
                              cb.size, n, dim, _p(header), _p(e4_qp), _p(e4_codes))
    assert rc == 0, rc
    size = ref.ref_error4_file(handle, None, 0)
    artifact = np.zeros(size, dtype=np.uint8)
    assert ref.ref_error4_file(handle, _p(artifact), size) == size
    F, K = int(g["refine_F"]), int(g["refine_K"])
    mode, ep, expand, L = (int(v) for v in g["refine_source"])
    key = "m%d_e%d_x%d_L%d" % (mode, ep, expand, L)
    out_ids = {1: [], 2: []}
    out_d = {1: [], 2: []}
    for qi, q in enumerate(queries):
        for rmode in (1, 2):
            ids = np.ascontiguousarray(g[key + "_ids"][qi][:F]).copy()
            d = np.zeros(F, dtype=np.float32)
            rc = ref.ref_refine_mode(handle, rmode, n, dim, _p(q), _p(ids), _p(d),
                                     F, K, None)
            assert rc == K, (rmode, rc)
            out_ids[rmode].append(ids[:K].copy())
            out_d[rmode].append(d[:K].copy())
    # an ERROR4 open with a graph->original map the artifact is not bound to
    perm = np.ascontiguousarray(np.arange(n, dtype=np.int32)[::-1])
    ids = np.ascontiguousarray(g[key + "_ids"][0][:F]).copy()
    d = np.zeros(F, dtype=np.float32)
    rc_map = ref.ref_refine_mode(handle, 2, n, dim, _p(queries[0]), _p(ids), _p(d),
                                 F, K, _p(perm))
    ref.ref_residual_close(handle)
    assert np.array_equal(out_ids[1], g["refine_ids"])
    out = dict(artifact=artifact, header=header, e4_qparams=e4_qp, e4_codes=e4_codes,
               refine_current_ids=np.stack(out_ids[1]),
               refine_current_dists=np.stack(out_d[1]),
               refine_e4_ids=np.stack(out_ids[2]), refine_e4_dists=np.stack(out_d[2]),
               rc_e4_unbound_map=np.int32(rc_map))
    path = os.path.join(HERE, name + "_error4.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def sq_refine_case(name="sq_refine_d128", pq_name="pq_d128_m32", sq_name="sq_d128",
                   sqmd_name="sqmd_d128"):
    """The SQ final refine (SQERR4V1): for SQ8 with 4-bit residuals and SQ4
    with 8-bit residuals, the reference builds the error4 artifact over its
    SQ decoder (retrieval_build_error4, pq = 0) and refines the SQMD
    traversal's outputs with retrieval_final_refine in CURRENT mode (also
    with a permuted original-id map) and ERROR4 mode."""
    ref = pyoracle.reference()
    g = dict(np.load(os.path.join(HERE, pq_name + ".npz")))
    s = dict(np.load(os.path.join(HERE, sq_name + ".npz")))
    z = dict(np.load(os.path.join(HERE, sqmd_name + ".npz")))
    base = (g["base_x16"].astype(np.float32) / np.float32(16.0))
    queries = (g["queries_x16"].astype(np.float32) / np.float32(16.0))
    n, dim = base.shape
    rng = np.random.default_rng(5)
    perm = rng.permutation(n).astype(np.int32)
    out = dict(perm=perm)
    for b in (8, 4):
        codes = np.ascontiguousarray(s["sq%d_codes" % b])
        res = np.ascontiguousarray(z["sq%d_res_codes" % b])
        rqp = np.ascontiguousarray(z["sq%d_res_qparams" % b], dtype=np.float32)
        rb = int(z["sq%d_res_bits" % b])
        h = ref.ref_sq_refine_open(_p(base), _p(codes), b, float(s["sq%d_scale" % b]),
                                   float(s["sq%d_zero" % b]), _p(res), rb, _p(rqp), n, dim)
        assert h
        size = ref.ref_sq_refine_artifact(h, None, 0)
        art = np.zeros(size, dtype=np.uint8)
        assert ref.ref_sq_refine_artifact(h, _p(art), size) == size
        out["sq%d_artifact" % b] = art
        cand = z["sq%d_c1_e32_x32_L64_ids" % b]
        counts = z["sq%d_c1_e32_x32_L64_counts" % b]
        for F in (20, 64):
            for mode, tag, mp in ((1, "current", None), (1, "current_perm", perm),
                                  (2, "error4", None)):
                ids_all, d_all = [], []
                for qi, q in enumerate(queries):
                    c = min(int(counts[qi]), F)
                    ids = np.ascontiguousarray(cand[qi][:c], dtype=np.int32).copy()
                    d = np.zeros(c, dtype=np.float32)
                    rc = ref.ref_sq_refine(h, mode, _p(q), _p(ids), _p(d), c, 10,
                                           None if mp is None else _p(mp))
                    assert rc == min(10, c), (b, tag, rc)
                    ids_all.append(ids[:10])
                    d_all.append(d[:10])
                out["sq%d_F%d_%s_ids" % (b, F, tag)] = np.stack(ids_all)
                out["sq%d_F%d_%s_dists" % (b, F, tag)] = np.stack(d_all)
        ref.ref_sq_refine_close(h)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def pq_r8_case():
    """The measured scorer's shape: M = 32 subspaces of 4 dims, 256
    centroids, 8-bit residuals, 32 neighbors per expansion (the two-lane
    online scorer) and 6 (at most 8 unvisited neighbors per expansion: the
    grouped scorer)."""
    pq_case("pq_d128_m32_r8", n=2000, dim=128, m=32, ks=256, R=32, res_bits=8,
            comps=48, nq=12, Ls=(16, 64, 160), expands=(32, 6), seed=72)
    error4_case("pq_d128_m32_r8")


def pipeline_case(pq_name):
    """The live pipeline with two source turns in flight
    (pure_pq_source_pipeline_step, window 2) on a PQ fixture's index: whole
    sorted pools, counts and (turns dispatched, work items, high water,
    steps) for every search case of the fixture, written to
    <pq_name>_w2.npz.  The residual handle is reopened from the fixture's
    base vectors; its residual codes must equal the fixture's."""
    ref = pyoracle.reference()
    z = np.load(os.path.join(HERE, pq_name + ".npz"))
    base = z["base_x16"].astype(np.float32) / 16.0
    queries = z["queries_x16"].astype(np.float32) / 16.0
    cb, codes, graph = z["codebook"], z["codes"], z["graph"]
    res_bits = int(z["res_bits"])
    n, dim = base.shape
    m, ks, _ = cb.shape
    R = graph.shape[1]
    res = np.zeros_like(z["res_codes"])
    qp = np.zeros_like(z["res_qparams"])
    handle = ref.ref_residual_open(_p(base), _p(codes), _p(cb), n, dim, m, ks,
                                   res_bits, _p(res), _p(qp))
    assert handle
    assert np.array_equal(res, z["res_codes"]) and np.array_equal(qp, z["res_qparams"])
    entry = int(z["entry"])
    out = {}
    for mode, ep, expand, L in z["search_cases"]:
        mode, ep, expand, L = int(mode), int(ep), int(expand), int(L)
        entries = pyoracle.seed_ids(n, ep, entry, 42)
        key = "m%d_e%d_x%d_L%d" % (mode, ep, expand, L)
        ids_all, d_all, st_all, cnt_all = [], [], [], []
        for q in queries:
            ids = np.zeros(L, dtype=np.int32)
            d = np.zeros(L, dtype=np.float32)
            st = np.zeros(4, dtype=np.uint64)
            rc = ref.ref_pipeline_search(_p(graph), R, n, _p(codes), 300, _p(cb), dim,
                                         m, ks, _p(q), _p(entries), len(entries),
                                         expand, L, L, 2, 1, mode, handle, _p(ids),
                                         _p(d), L, _p(st))
            assert rc >= 0, (key, rc)
            ids_all.append(ids)
            d_all.append(d)
            st_all.append(st)
            cnt_all.append(rc)
        out[key + "_ids"] = np.stack(ids_all)
        out[key + "_dists"] = np.stack(d_all)
        out[key + "_stats"] = np.stack(st_all)
        out[key + "_counts"] = np.array(cnt_all, dtype=np.int32)
    out["search_cases"] = z["search_cases"]
    ref.ref_residual_close(handle)
    path = os.path.join(HERE, pq_name + "_w2.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"],
                   check=True)
    if not pyoracle.ref_available():
        sys.exit("reference build missing: oracle/_ref/librover_ref.so")
    if sys.argv[1:] == ["graph"]:
        graph_case("graph_build")
        return
    if sys.argv[1:] == ["sqmd"]:
        sqmd_case("sqmd_d128")
        return
    if sys.argv[1:] == ["split"]:
        split_case("sqmd_split")
        return
    if sys.argv[1:] == ["rnn"]:
        rnn_case("rnn_build")
        return
    if sys.argv[1:] == ["sq_refine"]:
        sq_refine_case()
        return
    if sys.argv[1:] == ["r8"]:
        pq_r8_case()
        return
    if sys.argv[1:] == ["w2"]:
        for name in ("pq_d128_m32", "pq_d128_m32_r8"):
            pipeline_case(name)
        return
    if sys.argv[1:] == ["error4"]:
        for name in ("pq_d128_m32", "pq_d96_m48", "pq_d64_ks16"):
            error4_case(name)
        return
    base, queries, graph, entry = pq_case(
        "pq_d128_m32", n=1500, dim=128, m=32, ks=256, R=32, res_bits=4,
        comps=48, nq=12, Ls=(16, 64, 160), expands=(32, 12), seed=42)
    sq_case("sq_d128", base, queries, graph, entry)
    pq_case("pq_d96_m48", n=1000, dim=96, m=48, ks=256, R=48, res_bits=4,
            comps=32, nq=8, Ls=(24, 96), expands=(48,), seed=52)
    pq_case("pq_d64_ks16", n=800, dim=64, m=32, ks=16, R=16, res_bits=8,
            comps=24, nq=8, Ls=(16, 64), expands=(16,), seed=62)
    seeds_case("seeds")
    for name in ("pq_d128_m32", "pq_d96_m48", "pq_d64_ks16"):
        error4_case(name)
    pq_r8_case()
    for name in ("pq_d128_m32", "pq_d128_m32_r8"):
        pipeline_case(name)
    sq_refine_case()
    sqmd_case("sqmd_d128")
    split_case("sqmd_split")
    graph_case("graph_build")
    rnn_case("rnn_build")


if __name__ == "__main__":
    main()
