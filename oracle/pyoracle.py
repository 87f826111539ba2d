"""ctypes wrapper of the CPU oracle (liboracle.so) and, where it was built,
the compiled reference (_ref/librover_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "librover_ref.so")

SCORE_ADC, SCORE_ONLINE, SCORE_SQ = 0, 1, 2


class OrcIndex(C.Structure):
    _fields_ = [("graph", C.c_void_p), ("degree", C.c_int),
                ("num_vectors", C.c_int), ("codes", C.c_void_p),
                ("codebook", C.c_void_p), ("dim", C.c_int), ("m", C.c_int),
                ("ks", C.c_int), ("subdim", C.c_int),
                ("res_codes", C.c_void_p), ("res_qparams", C.c_void_p),
                ("res_bits", C.c_int), ("sq_codes", C.c_void_p),
                ("sq_bits", C.c_int), ("sq_scale", C.c_float),
                ("sq_zero", C.c_float), ("e4_codes", C.c_void_p),
                ("e4_qparams", C.c_void_p), ("split_bias", C.c_void_p)]


class OrcStats(C.Structure):
    _fields_ = [("pq_scored_candidates", C.c_uint64),
                ("corrected_candidates", C.c_uint64),
                ("expanded_nodes", C.c_uint64),
                ("visited_dedup_skips", C.c_uint64)]


def _p(a):
    return None if a is None else a.ctypes.data


_orc = None
_ref = None


def oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_LIB):
            subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
        L = C.CDLL(ORACLE_LIB)
        P, I, F = C.c_void_p, C.c_int, C.c_float
        L.orc_build_lut.argtypes = [P, P, I, I, I, P]
        L.orc_adc.argtypes = [P, P, I, I]
        L.orc_adc.restype = F
        L.orc_residual_encode.argtypes = [P, P, P, I, I, I, I, P, P]
        L.orc_residual_evaluate.argtypes = [P, P, I, P]
        L.orc_residual_decode.argtypes = [P, I, P]
        L.orc_sq_decode.argtypes = [P, I, P]
        L.orc_error4_encode.argtypes = [P, P, I, P, P]
        L.orc_sq_range_params.argtypes = [P, C.c_long, I, P, P]
        L.orc_sq_encode.argtypes = [P, I, I, F, F, P]
        L.orc_sq_quantize_query.argtypes = [P, I, I, F, F, P]
        L.orc_sq_distance.argtypes = [P, P, I, I, F]
        L.orc_sq_distance.restype = F
        L.orc_seed_ids.argtypes = [I, I, I, C.c_uint32, P]
        L.orc_search.argtypes = [P, I, P, P, I, I, I, I, P, P, I, P]
        L.orc_refine.argtypes = [P, P, P, I, I, P, P, P]
        L.orc_recall_hits.argtypes = [P, I, P, I, I]
        L.orc_fingerprint.argtypes = [C.c_uint64, P, P, I]
        L.orc_fingerprint.restype = C.c_uint64
        L.orc_search_batch.argtypes = [P, I, P, I, P, I, I, I, P, P, I, I, I, P,
                                       P, P, P]
        L.orc_search_window.argtypes = [P, I, P, P, I, I, I, I, I, P, P, I, P, P]
        L.orc_search_batch_window.argtypes = [P, I, P, I, P, I, I, I, P, P, I, I, I,
                                              I, P, P, P, P]
        L.orc_build_exact.argtypes = [P, I, I, I, I, C.c_uint32, I, P]
        L.orc_rnn_build.argtypes = [P, I, I, I, I, I, I, C.c_uint32, I, P]
        L.orc_l2.argtypes = [P, I, I, I]
        L.orc_l2.restype = F
        L.orc_sqmd_search.argtypes = [P, I, P, P, I, I, I, I, P, P, I, P]
        L.orc_sqmd_seed_ids.argtypes = [I, I, I, C.c_uint32, P, I]
        L.orc_sq_residual_distance.argtypes = [P, P, I]
        L.orc_sq_residual_distance.restype = F
        L.orc_sq_query_aligned_u8.argtypes = [P, I, F, F, P, P]
        L.orc_sq_split_distance.argtypes = [P, P, C.c_uint64, I]
        L.orc_sq_split_distance.restype = F
        _orc = L
    return _orc


def ref_available():
    return os.path.exists(REF_LIB)


def reference():
    """The compiled reference scalar C, or None where it was not built."""
    global _ref
    if _ref is None and ref_available():
        L = C.CDLL(REF_LIB)
        P, I, F = C.c_void_p, C.c_int, C.c_float
        L.ref_build_lut.argtypes = [P, P, I, I, I, P]
        L.ref_adc_ids.argtypes = [P, I, I, I, P, I, P, I, I, P]
        L.ref_pq_train.argtypes = [P, I, I, I, I, I, P]
        L.ref_pq_encode.argtypes = [P, I, I, I, I, P, P]
        L.ref_sqmd_seed_ids.argtypes = [I, I, I, C.c_uint, P, I]
        L.ref_quantize_query.argtypes = [P, I, I, F, F, P]
        L.ref_sq_distance.argtypes = [P, P, I, I, F]
        L.ref_sq_distance.restype = F
        L.ref_residual_open.argtypes = [P, P, P, I, I, I, I, I, P, P]
        L.ref_residual_open.restype = P
        L.ref_residual_close.argtypes = [P]
        L.ref_residual_fork.argtypes = [P]
        L.ref_residual_fork.restype = P
        L.ref_sq_refine_open.argtypes = [P, P, I, F, F, P, I, P, I, I]
        L.ref_sq_refine_open.restype = P
        L.ref_sq_refine_artifact.argtypes = [P, P, C.c_long]
        L.ref_sq_refine_artifact.restype = C.c_long
        L.ref_sq_refine.argtypes = [P, I, P, P, P, I, I, P]
        L.ref_sq_refine_close.argtypes = [P]
        L.ref_residual_evaluate.argtypes = [P, P, I, P]
        L.ref_residual_decode.argtypes = [P, I, P]
        L.ref_search.argtypes = [P, I, I, P, I, P, I, I, I, P, P, I, I, I, I, I,
                                 P, P, I, F, F, P, P, I, P]
        L.ref_refine.argtypes = [P, I, I, P, P, P, I, I, P]
        L.ref_pipeline_search.argtypes = [P, I, I, P, I, P, I, I, I, P, P, I, I, I, I,
                                          I, I, I, P, P, P, I, P]
        L.ref_error4_build.argtypes = [P, P, P, I, P, I, I, I, P, P, P]
        L.ref_error4_file.argtypes = [P, P, C.c_long]
        L.ref_error4_file.restype = C.c_long
        L.ref_refine_mode.argtypes = [P, I, I, I, P, P, P, I, I, P]
        L.ref_pq_artifacts_write.argtypes = [C.c_char_p, P, I, I, I, I, I, I, I, I, C.c_uint,
                                             P, P, P, P]
        L.ref_pq_artifacts_load.argtypes = [C.c_char_p, P, P, P, P, P, C.c_size_t, P,
                                            C.c_size_t]
        L.ref_residual_prefix.argtypes = [P]
        L.ref_residual_prefix.restype = C.c_char_p
        L.ref_pq_residual_read.argtypes = [C.c_char_p, P, I, I, I, I, I, I, P, P, P, P, I,
                                           P, I, P]
        L.ref_results_write.argtypes = [C.c_char_p, I, C.c_char_p, I, P, P, P, I]
        L.ref_nsg_load.argtypes = [C.c_char_p, I, I, P, P]
        L.ref_build_exact.argtypes = [P, I, I, I, I, C.c_uint, I, I, P]
        L.ref_rnn_build.argtypes = [P, I, I, I, I, I, I, C.c_uint, I, I, P]
        L.ref_sqmd_search.argtypes = [P, I, I, P, I, I, F, F, P, P, I, I, P, I, I,
                                      C.c_uint, I, I, I, P, P, I, P]
        L.ref_sqmd_split_search.argtypes = [P, I, I, P, I, F, F, P, P, I, P, P, I, I,
                                            C.c_uint, I, I, I, P, P, I]
        _ref = L
    return _ref


class Index:
    """Keeps the numpy arrays alive behind an OrcIndex struct."""

    def __init__(self, graph, codes=None, codebook=None, res_codes=None,
                 res_qparams=None, res_bits=4, sq_codes=None, sq_bits=8,
                 sq_scale=1.0, sq_zero=0.0, dim=None, e4_codes=None,
                 e4_qparams=None, split_bias=None):
        self.graph = np.ascontiguousarray(graph, dtype=np.int32)
        self.codes = None if codes is None else np.ascontiguousarray(codes, dtype=np.uint8)
        self.codebook = None if codebook is None else np.ascontiguousarray(codebook, dtype=np.float32)
        self.res_codes = None if res_codes is None else np.ascontiguousarray(res_codes, dtype=np.uint8)
        self.res_qparams = None if res_qparams is None else np.ascontiguousarray(res_qparams, dtype=np.float32)
        self.sq_codes = None if sq_codes is None else np.ascontiguousarray(sq_codes, dtype=np.uint8)
        self.e4_codes = None if e4_codes is None else np.ascontiguousarray(e4_codes, dtype=np.uint8)
        self.e4_qparams = None if e4_qparams is None else np.ascontiguousarray(e4_qparams, dtype=np.float32)
        self.split_bias = None if split_bias is None else np.ascontiguousarray(split_bias, dtype=np.float32)
        s = OrcIndex()
        s.graph = _p(self.graph)
        s.degree = self.graph.shape[1]
        s.num_vectors = self.graph.shape[0]
        if self.codebook is not None:
            m, ks, sub = self.codebook.shape
            s.codes = _p(self.codes)
            s.codebook = _p(self.codebook)
            s.dim, s.m, s.ks, s.subdim = m * sub, m, ks, sub
        else:
            s.dim = dim
        s.res_codes = _p(self.res_codes)
        s.res_qparams = _p(self.res_qparams)
        s.res_bits = res_bits
        s.sq_codes = _p(self.sq_codes)
        s.sq_bits = sq_bits
        s.sq_scale = sq_scale
        s.sq_zero = sq_zero
        s.e4_codes = _p(self.e4_codes)
        s.e4_qparams = _p(self.e4_qparams)
        s.split_bias = _p(self.split_bias)
        self.struct = s
        self.dim = s.dim

    @classmethod
    def from_pq(cls, data):
        return cls(data.graph, data.codes, data.codebook, data.res_codes,
                   data.res_qparams, data.res_bits)


def search(index, scorer, query, entries, expand, L, num_out):
    q = np.ascontiguousarray(query, dtype=np.float32)
    e = np.ascontiguousarray(entries, dtype=np.int32)
    ids = np.zeros(max(num_out, 1), dtype=np.int32)
    d = np.zeros(max(num_out, 1), dtype=np.float32)
    st = OrcStats()
    n = oracle().orc_search(C.byref(index.struct), scorer, _p(q), _p(e), len(e),
                            expand, L, num_out, _p(ids), _p(d), num_out,
                            C.byref(st))
    stats = (st.pq_scored_candidates, st.corrected_candidates,
             st.expanded_nodes, st.visited_dedup_skips)
    if n < 0:
        return None, None, stats
    return ids[:n].copy(), d[:n].copy(), stats


def search_window(index, scorer, query, entries, expand, L, num_out, window):
    """The live pipeline with `window` source turns in flight; returns ids,
    distances, stats (scored, corrected, turns, dedup) and the high water."""
    q = np.ascontiguousarray(query, dtype=np.float32)
    e = np.ascontiguousarray(entries, dtype=np.int32)
    ids = np.zeros(max(num_out, 1), dtype=np.int32)
    d = np.zeros(max(num_out, 1), dtype=np.float32)
    st = OrcStats()
    hw = C.c_int(0)
    n = oracle().orc_search_window(C.byref(index.struct), scorer, _p(q), _p(e), len(e),
                                   expand, L, num_out, window, _p(ids), _p(d), num_out,
                                   C.byref(st), C.byref(hw))
    stats = (st.pq_scored_candidates, st.corrected_candidates,
             st.expanded_nodes, st.visited_dedup_skips)
    if n < 0:
        return None, None, stats, 0
    return ids[:n].copy(), d[:n].copy(), stats, hw.value


def search_batch(index, scorer, queries, entries, expand, L, k, refine=0,
                 threads=None, window=1):
    """Many queries on host threads; L and k scalars or per-query arrays;
    entries a shared list or an [nq, S] array of per-query rows; window > 1
    searches with that many source turns in flight (search_window).
    Returns ids [nq, stride], dists, counts, stats [nq, 4]."""
    q = np.ascontiguousarray(queries, dtype=np.float32)
    nq = q.shape[0]
    Ls = np.ascontiguousarray(np.broadcast_to(np.asarray(L, dtype=np.int32), (nq,)))
    ks = np.ascontiguousarray(np.broadcast_to(np.asarray(k, dtype=np.int32), (nq,)))
    stride = int(ks.max()) if refine else int(ks.max())
    e = np.ascontiguousarray(entries, dtype=np.int32)
    ids = np.full((nq, stride), -1, dtype=np.int32)
    d = np.zeros((nq, stride), dtype=np.float32)
    cnt = np.zeros(nq, dtype=np.int32)
    st = (OrcStats * max(nq, 1))()
    threads = threads or os.cpu_count() or 1
    estride = e.shape[1] if e.ndim == 2 else 0
    oracle().orc_search_batch_window(C.byref(index.struct), scorer, _p(q), nq, _p(e),
                                     e.shape[1] if estride else len(e), estride, expand,
                                     _p(Ls), _p(ks), stride, refine, window, threads,
                                     _p(ids), _p(d), _p(cnt), st)
    stats = np.array([(s.pq_scored_candidates, s.corrected_candidates,
                       s.expanded_nodes, s.visited_dedup_skips) for s in st[:nq]],
                     dtype=np.uint64).reshape(nq, 4)
    return ids, d, cnt, stats


def refine(index, query, ids, output_k, graph_to_original=None):
    q = np.ascontiguousarray(query, dtype=np.float32)
    i = np.ascontiguousarray(ids, dtype=np.int32)
    oi = np.zeros(max(len(i), 1), dtype=np.int32)
    od = np.zeros(max(len(i), 1), dtype=np.float32)
    g = None if graph_to_original is None else np.ascontiguousarray(graph_to_original, dtype=np.int32)
    n = oracle().orc_refine(C.byref(index.struct), _p(q), _p(i), len(i), output_k,
                            _p(g), _p(oi), _p(od))
    if n < 0:
        return None, None
    return oi[:n].copy(), od[:n].copy()


def build_lut(query, codebook):
    m, ks, sub = codebook.shape
    lut = np.zeros(m * ks, dtype=np.float32)
    q = np.ascontiguousarray(query, dtype=np.float32)
    cb = np.ascontiguousarray(codebook, dtype=np.float32)
    rc = oracle().orc_build_lut(_p(q), _p(cb), m, ks, sub, _p(lut))
    assert rc == 0
    return lut


def adc(code, lut, m, ks):
    c = np.ascontiguousarray(code, dtype=np.uint8)
    return oracle().orc_adc(_p(c), _p(lut), m, ks)


def residual_evaluate(index, query, vid):
    out = np.zeros(1, dtype=np.float32)
    q = np.ascontiguousarray(query, dtype=np.float32)
    rc = oracle().orc_residual_evaluate(C.byref(index.struct), _p(q), vid, _p(out))
    return None if rc else float(out[0])


def residual_decode(index, vid):
    out = np.zeros(index.dim, dtype=np.float32)
    rc = oracle().orc_residual_decode(C.byref(index.struct), vid, _p(out))
    return None if rc else out


def sq_decode(index, vid):
    """orc_sq_decode: the SQ index's SQ + residual reconstruction."""
    out = np.zeros(index.dim, dtype=np.float32)
    rc = oracle().orc_sq_decode(C.byref(index.struct), vid, _p(out))
    return None if rc else out


def residual_encode(vec, pq_code, codebook, bits):
    m, ks, sub = codebook.shape
    dim = m * sub
    rb = dim // 2 + dim % 2 if bits == 4 else dim
    out = np.zeros(rb, dtype=np.uint8)
    qp = np.zeros(2, dtype=np.float32)
    v = np.ascontiguousarray(vec, dtype=np.float32)
    c = np.ascontiguousarray(pq_code, dtype=np.uint8)
    cb = np.ascontiguousarray(codebook, dtype=np.float32)
    rc = oracle().orc_residual_encode(_p(v), _p(c), _p(cb), m, ks, sub, bits,
                                      _p(out), _p(qp))
    assert rc == 0
    return out, qp


def error4_encode(x, recon):
    """Per-row error4 codes [n][(dim+1)/2] and (scale, zero) [n][2] of the
    raw vectors x against their reconstructions recon."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    r = np.ascontiguousarray(recon, dtype=np.float32)
    n, dim = x.shape
    codes = np.zeros((n, (dim + 1) // 2), dtype=np.uint8)
    qp = np.zeros((n, 2), dtype=np.float32)
    for i in range(n):
        rc = oracle().orc_error4_encode(_p(x[i]), _p(r[i]), dim, _p(codes[i]), _p(qp[i]))
        assert rc == 0
    return codes, qp


def sq_quantize_query(query, bits, scale, zero):
    q = np.ascontiguousarray(query, dtype=np.float32)
    out = np.zeros(len(q), dtype=np.uint8)
    oracle().orc_sq_quantize_query(_p(q), len(q), bits, scale, zero, _p(out))
    return out


def sq_distance(qq, code, dim, bits, scale):
    a = np.ascontiguousarray(qq, dtype=np.uint8)
    b = np.ascontiguousarray(code, dtype=np.uint8)
    return oracle().orc_sq_distance(_p(a), _p(b), dim, bits, scale)


def sq_range_params(x, bits):
    a = np.ascontiguousarray(x, dtype=np.float32)
    s = np.zeros(1, dtype=np.float32)
    z = np.zeros(1, dtype=np.float32)
    oracle().orc_sq_range_params(_p(a), a.size, bits, _p(s), _p(z))
    return np.float32(s[0]), np.float32(z[0])


def sq_encode(x, bits, scale, zero):
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, dim = x.shape
    cb = (dim + 1) // 2 if bits == 4 else dim
    out = np.zeros((n, cb), dtype=np.uint8)
    for i in range(n):
        oracle().orc_sq_encode(_p(x[i]), dim, bits, scale, zero, _p(out[i]))
    return out


class OrcSqmdStats(C.Structure):
    _fields_ = [("work_items", C.c_uint64), ("dedup_skips", C.c_uint64),
                ("invalid_skips", C.c_uint64), ("applied", C.c_uint64),
                ("turns", C.c_uint64), ("empty_turns", C.c_uint64)]


def sqmd_search(index, query, seeds, entry_points, expand, L, num_out, correction=0):
    """The live SQ traversal (distance-only heaps) for one query; returns
    (ids, dists, stats dict) or (None, None, stats)."""
    q = np.ascontiguousarray(query, dtype=np.float32)
    e = np.ascontiguousarray(seeds, dtype=np.int32)
    ids = np.zeros(max(num_out, 1), dtype=np.int32)
    d = np.zeros(max(num_out, 1), dtype=np.float32)
    st = OrcSqmdStats()
    n = oracle().orc_sqmd_search(C.byref(index.struct), correction, _p(q), _p(e), len(e),
                                 entry_points, expand, L, _p(ids), _p(d), num_out,
                                 C.byref(st))
    stats = {f: getattr(st, f) for f, _ in OrcSqmdStats._fields_}
    if n < 0:
        return None, None, stats
    return ids[:n].copy(), d[:n].copy(), stats


def sqmd_seed_ids(n, entry_points, graph_entry, seed=42):
    out = np.zeros(max(entry_points, 1), dtype=np.int32)
    cnt = oracle().orc_sqmd_seed_ids(n, entry_points, graph_entry, seed, _p(out),
                                     len(out))
    return out[:max(cnt, 0)]


def sq_query_aligned_u8(query, scale, zero):
    """The query on the 17x-refined SQ grid and its sum, or None."""
    q = np.ascontiguousarray(query, dtype=np.float32)
    out = np.zeros(q.shape[0], dtype=np.uint8)
    sq = C.c_uint64()
    if oracle().orc_sq_query_aligned_u8(_p(q), q.shape[0], scale, zero, _p(out), C.byref(sq)):
        return None
    return out, sq.value


def sq_split_distance(index, query, vid):
    qu8, sq = sq_query_aligned_u8(query, index.struct.sq_scale, index.struct.sq_zero)
    return oracle().orc_sq_split_distance(C.byref(index.struct), _p(qu8), sq, vid)


def sq_residual_distance(index, query, vid):
    q = np.ascontiguousarray(query, dtype=np.float32)
    return oracle().orc_sq_residual_distance(C.byref(index.struct), _p(q), vid)


def build_exact(base, r, candidates, seed=42, prune_mode=1):
    """The exact-candidate graph build; int32 [n][r]."""
    x = np.ascontiguousarray(base, dtype=np.float32)
    n, dim = x.shape
    out = np.zeros((n, r), dtype=np.int32)
    rc = oracle().orc_build_exact(_p(x), n, dim, r, candidates, seed, prune_mode, _p(out))
    assert rc == 0
    return out


def rnn_build(base, s, r, t1, t2, seed=42, top_k=None):
    """RNN-Descent; int32 [n][top_k], -1 padded."""
    x = np.ascontiguousarray(base, dtype=np.float32)
    n, dim = x.shape
    top_k = r if top_k is None else top_k
    out = np.zeros((n, top_k), dtype=np.int32)
    rc = oracle().orc_rnn_build(_p(x), n, dim, s, r, t1, t2, seed, top_k, _p(out))
    assert rc == 0
    return out


def seed_ids(n, entry_count, graph_entry, seed=42):
    out = np.zeros(max(entry_count, 1), dtype=np.int32)
    cnt = oracle().orc_seed_ids(n, entry_count, graph_entry, seed, _p(out))
    return out[:cnt]


def recall_hits(result_ids, result_count, gt_row, gt_k, k):
    """orc_recall_hits: demo_recall_hits_at_k for one query."""
    r = np.ascontiguousarray(result_ids, dtype=np.int32)
    g = np.ascontiguousarray(gt_row, dtype=np.int32)
    return oracle().orc_recall_hits(_p(r), int(result_count), _p(g), int(gt_k), int(k))


def fingerprint(ids_rows, dist_rows, counts, h=14695981039346656037):
    for i in range(len(counts)):
        c = int(counts[i])
        ids = np.ascontiguousarray(ids_rows[i][:c], dtype=np.int32)
        d = np.ascontiguousarray(dist_rows[i][:c], dtype=np.float32)
        h = oracle().orc_fingerprint(h, _p(ids), _p(d), c)
    return h
