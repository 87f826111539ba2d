"""Python mirror of the reference's operator interface over librover.so.

`Calculator` follows the FastRPC `calculator` interface (calculator.idl):
open a session, configure it, load an index, search.  Method names keep the
IDL op names; errors raise RoverError carrying the AEEResult code, where the
reference's host code would see a non-zero AEEResult.
"""
import ctypes as C

import numpy as np

from . import native
from .native import check, ptr


class Calculator:
    def __init__(self, uri=None):
        self.h = C.c_uint64(0)
        check(native.lib().calculator_open(uri, C.byref(self.h)), "calculator_open")
        self.pq = None
        self.sq = None

    # --------------------------------------------------------- lifecycle
    def close(self):
        if self.h.value:
            native.lib().calculator_close(self.h)
            self.h = C.c_uint64(0)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -------------------------------------------------------- IDL ops
    def rover_configure(self, abi_version=native.ABI_VERSION, l1_bits=8):
        check(native.lib().calculator_rover_configure(self.h, abi_version, l1_bits),
              "rover_configure")

    def retrieval_pure_pq_load(self, codes, codebook, chunk_vectors=None,
                               uuid=None):
        """calculator.idl:119-131 with host chunks instead of fds."""
        codes = np.ascontiguousarray(codes, dtype=np.uint8)
        codebook = np.ascontiguousarray(codebook, dtype=np.float32)
        n, code_bytes = codes.shape
        m, ks, sub = codebook.shape
        cv = chunk_vectors or n
        chunks = [np.ascontiguousarray(codes[s:s + cv]) for s in range(0, n, cv)]
        ptrs = (C.c_void_p * len(chunks))(*[c.ctypes.data for c in chunks])
        counts = np.array([c.shape[0] for c in chunks], dtype=np.int32)
        u = None if uuid is None else np.ascontiguousarray(uuid, dtype=np.uint8)
        check(native.lib().calculator_retrieval_pure_pq_load_dma(
            self.h, ptrs, len(chunks), ptr(counts), len(counts), cv,
            ptr(codebook), codebook.size, n, m * sub, m, ks, sub, code_bytes,
            ptr(u), 0 if u is None else len(u)), "retrieval_pure_pq_load_dma")
        self.pq = dict(n=n, dim=m * sub, m=m, ks=ks)

    def retrieval_pure_pq_unload(self):
        check(native.lib().calculator_retrieval_pure_pq_unload_dma(self.h),
              "retrieval_pure_pq_unload_dma")
        self.pq = None

    def pure_pq_graph_load(self, graph, entry):
        graph = np.ascontiguousarray(graph, dtype=np.int32)
        check(native.lib().rover_pure_pq_graph_load(
            self.h, ptr(graph), graph.size, graph.shape[0], graph.shape[1],
            int(entry)), "rover_pure_pq_graph_load")

    def pure_pq_residual_load(self, res_codes, qparams, bits):
        res_codes = np.ascontiguousarray(res_codes, dtype=np.uint8)
        qparams = np.ascontiguousarray(qparams, dtype=np.float32)
        check(native.lib().rover_pure_pq_residual_load(
            self.h, ptr(res_codes), res_codes.size, ptr(qparams), qparams.size,
            bits), "rover_pure_pq_residual_load")

    def pure_pq_original_ids(self, mapping):
        m = None if mapping is None else np.ascontiguousarray(mapping, dtype=np.int32)
        check(native.lib().rover_pure_pq_original_ids(
            self.h, ptr(m), 0 if m is None else len(m)), "rover_pure_pq_original_ids")

    def pure_pq_error4_load(self, artifact, expected_sha256=None):
        """Loads a PQERR4V1 third-level refine artifact (the file image as
        bytes or a uint8 array); expected_sha256 (32 bytes) checks the whole
        file as RetrievalFinalRefineConfig.artifact_sha256 does."""
        a = np.frombuffer(artifact, dtype=np.uint8) if isinstance(artifact, (bytes, bytearray)) \
            else np.ascontiguousarray(artifact, dtype=np.uint8)
        e = None if expected_sha256 is None else np.frombuffer(bytes(expected_sha256), dtype=np.uint8)
        if e is not None and e.size != 32:
            raise ValueError("expected_sha256 must be 32 bytes")
        check(native.lib().rover_pure_pq_error4_load(self.h, ptr(a), a.size, ptr(e)),
              "rover_pure_pq_error4_load")

    def sq_error4_load(self, artifact, expected_sha256=None):
        """Loads an SQERR4V1 artifact (file image) for the SQ traversals'
        final refine in ERROR4 mode."""
        a = np.frombuffer(artifact, dtype=np.uint8) if isinstance(artifact, (bytes, bytearray)) \
            else np.ascontiguousarray(artifact, dtype=np.uint8)
        e = None if expected_sha256 is None else np.frombuffer(bytes(expected_sha256), dtype=np.uint8)
        if e is not None and e.size != 32:
            raise ValueError("expected_sha256 must be 32 bytes")
        check(native.lib().rover_sq_error4_load(self.h, ptr(a), a.size, ptr(e)),
              "rover_sq_error4_load")

    def pure_pq_decode(self, ids):
        """Reconstructed vectors [len(ids)][dim] (PQ + residual decode)."""
        i = np.ascontiguousarray(ids, dtype=np.int32)
        out = np.zeros((len(i), self.pq["dim"]), dtype=np.float32)
        check(native.lib().rover_pure_pq_decode(self.h, ptr(i), len(i), ptr(out)),
              "rover_pure_pq_decode")
        return out

    def sq_decode(self, ids):
        """Reconstructed vectors [len(ids)][dim] (SQ + residual decode)."""
        i = np.ascontiguousarray(ids, dtype=np.int32)
        out = np.zeros((len(i), self.sq["dim"]), dtype=np.float32)
        check(native.lib().rover_sq_decode(self.h, ptr(i), len(i), ptr(out)), "rover_sq_decode")
        return out

    def sq_error4_unload(self):
        check(native.lib().rover_sq_error4_unload(self.h), "rover_sq_error4_unload")

    def pure_pq_error4_unload(self):
        check(native.lib().rover_pure_pq_error4_unload(self.h), "rover_pure_pq_error4_unload")

    def load_pq_index(self, data, chunk_vectors=None, with_router=True):
        """Loads a rover_amd.index.PqIndexData (codes, graph, residuals and,
        with with_router, its entry router)."""
        self.retrieval_pure_pq_load(data.codes, data.codebook, chunk_vectors)
        self.pure_pq_graph_load(data.graph, data.entry)
        if data.res_codes is not None:
            self.pure_pq_residual_load(data.res_codes, data.res_qparams, data.res_bits)
        if with_router and getattr(data, "router_centroids", None) is not None:
            self.pure_pq_router_load(data.router_centroids, data.router_ids)

    def pure_pq_router_load(self, centroids, rep_ids):
        c = np.ascontiguousarray(centroids, dtype=np.float32)
        r = np.ascontiguousarray(rep_ids, dtype=np.int32)
        check(native.lib().rover_pure_pq_router_load(self.h, ptr(c), c.size, ptr(r),
                                                      len(r)), "rover_pure_pq_router_load")

    def route_entries(self, queries, S):
        """Per-query entry ids [nq, S] chosen by the router (host arrays)."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        dq = native.DeviceBuffer.from_array(q)
        dout = native.DeviceBuffer(q.shape[0] * S * 4)
        check(native.lib().rover_route_entries(self.h, dq.ptr, q.shape[0], S, dout.ptr,
                                               None), "rover_route_entries")
        native.synchronize()
        return dout.download(np.int32, q.shape[0] * S).reshape(q.shape[0], S)

    def retrieval_load(self, graph, l1_codes, dim, entry, scale, zero):
        """calculator.idl:36-45 (SQ index; bits from rover_configure)."""
        graph = np.ascontiguousarray(graph, dtype=np.int32)
        l1_codes = np.ascontiguousarray(l1_codes, dtype=np.uint8)
        check(native.lib().calculator_retrieval_load(
            self.h, ptr(graph), graph.size, ptr(l1_codes), l1_codes.size,
            graph.shape[0], dim, graph.shape[1], int(entry), float(scale),
            float(zero)), "retrieval_load")
        self.sq = dict(n=graph.shape[0], dim=dim)

    def retrieval_load_dma(self, graph, l1_codes, dim, entry, scale, zero,
                           chunk_vectors=None):
        """calculator.idl:59-70: the SQ index from chunked codes (the IDL's
        fds become chunk pointers); graph None is RETRIEVAL_DMA_NO_GRAPH_FD
        (give it later with sq_graph_load)."""
        l1_codes = np.ascontiguousarray(l1_codes, dtype=np.uint8)
        n = l1_codes.shape[0]
        cv = chunk_vectors or n
        chunks = [np.ascontiguousarray(l1_codes[s:s + cv]) for s in range(0, n, cv)]
        ptrs = (C.c_void_p * len(chunks))(*[c.ctypes.data for c in chunks])
        g = None if graph is None else np.ascontiguousarray(graph, dtype=np.int32)
        degree = g.shape[1] if g is not None else 1
        check(native.lib().calculator_retrieval_load_dma(
            self.h, ptr(g), ptrs, len(chunks), cv, n, dim, degree, int(entry),
            float(scale), float(zero)), "retrieval_load_dma")
        self.sq = dict(n=n, dim=dim)

    def retrieval_unload_dma(self):
        check(native.lib().calculator_retrieval_unload_dma(self.h), "retrieval_unload_dma")

    def sq_graph_load(self, graph, entry):
        g = np.ascontiguousarray(graph, dtype=np.int32)
        check(native.lib().rover_sq_graph_load(self.h, ptr(g), g.size, g.shape[0],
                                               g.shape[1], int(entry)), "rover_sq_graph_load")

    def sq_residual_load(self, res_codes, qparams, bits):
        """Per-vector SQ residuals for the corrected SQ traversal."""
        r = np.ascontiguousarray(res_codes, dtype=np.uint8)
        q = np.ascontiguousarray(qparams, dtype=np.float32)
        check(native.lib().rover_sq_residual_load(self.h, ptr(r), r.size, ptr(q), q.size,
                                                  int(bits)), "rover_sq_residual_load")

    def sq_split_bias_load(self, image, fingerprint=None):
        """The SQMDBIAS artifact (file image) for searches with correction=2;
        fingerprint: its 16-byte binding to the l1 / residual / qparams
        files, checked when given."""
        buf = np.frombuffer(bytes(image), dtype=np.uint8)
        fp = None if fingerprint is None else np.frombuffer(bytes(fingerprint), dtype=np.uint8)
        if fp is not None and fp.size != 16:
            raise ValueError("fingerprint must be 16 bytes")
        check(native.lib().rover_sq_split_bias_load(self.h, ptr(buf), buf.size,
                                                    None if fp is None else ptr(fp)),
              "rover_sq_split_bias_load")

    def retrieval_unload(self):
        check(native.lib().calculator_retrieval_unload(self.h), "retrieval_unload")
        self.sq = None

    def retrieval_search(self, query, entry_points, expand_count, pool_size,
                         num_out):
        """calculator.idl:46-55 -> (ids, dists)."""
        q = np.ascontiguousarray(query, dtype=np.float32)
        ids = np.zeros(num_out, dtype=np.int32)
        d = np.zeros(num_out, dtype=np.float32)
        cnt = C.c_int32(0)
        check(native.lib().calculator_retrieval_search(
            self.h, ptr(q), len(q), entry_points, expand_count, pool_size,
            num_out, ptr(ids), num_out, ptr(d), num_out, C.byref(cnt)),
            "retrieval_search")
        return ids[:cnt.value].copy(), d[:cnt.value].copy()

    # ---------------------------------------------------- batch extension
    def _batch_struct(self, queries, nq, L, k, method, correction, refine,
                      entry_points, expand_count, out_ids, out_dists,
                      out_counts, stride, out_stats, hash_log2, probe_limit,
                      grid_waves, learned_stop, stop_threshold, pool_sizes,
                      num_outs, seed, entry_ids=None, entry_count=0,
                      entry_stride=0, route_entries=0, refine_mode=0, traversal=0,
                      source_window=0, waves_per_query=0, candidate_tile=0,
                      reorder_step=0):
        b = native.RoverSearchBatch()
        b.reorder_step = int(reorder_step)
        b.source_window = int(source_window)
        b.waves_per_query = int(waves_per_query)
        b.candidate_tile = int(candidate_tile)
        b.refine_mode = int(refine_mode)
        b.traversal = int(traversal)
        b.entry_ids = entry_ids
        b.entry_count = int(entry_count)
        b.entry_stride = int(entry_stride)
        b.route_entries = int(route_entries)
        b.queries = queries
        b.num_queries = nq
        b.pool_sizes = pool_sizes
        b.num_outs = num_outs
        b.pool_size = int(L)
        b.num_out = int(k)
        b.method = method
        b.correction = int(correction)
        b.refine_candidates = int(refine)
        b.entry_points = int(entry_points)
        b.seed = seed
        b.expand_count = int(expand_count)
        b.out_ids = out_ids
        b.out_dists = out_dists
        b.out_counts = out_counts
        b.out_stride = int(stride)
        b.out_stats = out_stats
        b.hash_log2 = int(hash_log2)
        b.probe_limit = int(probe_limit)
        b.grid_waves = int(grid_waves)
        b.learned_stop = int(learned_stop)
        b.stop_threshold = float(stop_threshold)
        return b

    def search_batch(self, queries, pool_size, num_out, method=native.METHOD_PQ,
                     correction=0, refine_candidates=0, entry_points=1,
                     expand_count=32, stats=False, hash_log2=0, probe_limit=0,
                     grid_waves=0, learned_stop=0, stop_threshold=0.0,
                     seed=native.ENTRY_POINT_SEED, entry_ids=None,
                     route_entries=0, refine_mode=0, traversal=0, source_window=0,
                     waves_per_query=0, candidate_tile=0, reorder_step=0):
        """Host arrays in and out.  entry_ids: explicit seeds, a list shared
        by all queries or a [nq, S] array of per-query rows (otherwise graph
        entry + LCG seeds); route_entries=S: per-query seeds from the router.  pool_size / num_out: int or per-query
        arrays.  Returns ids [nq, k_max], dists, counts (negative = per-query
        error status) and, with stats=True, [nq, 4] counters."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = q.shape[0]
        Ls = None if np.isscalar(pool_size) else np.ascontiguousarray(pool_size, dtype=np.int32)
        Ks = None if np.isscalar(num_out) else np.ascontiguousarray(num_out, dtype=np.int32)
        Lmax = int(pool_size) if Ls is None else int(Ls.max())
        Kmax = int(num_out) if Ks is None else int(Ks.max())
        ids = np.full((nq, Kmax), -1, dtype=np.int32)
        d = np.zeros((nq, Kmax), dtype=np.float32)
        cnt = np.zeros(nq, dtype=np.int32)
        st = np.zeros((nq, 4), dtype=np.uint64) if stats else None
        e = None if entry_ids is None else np.ascontiguousarray(entry_ids, dtype=np.int32)
        stride = e.shape[1] if e is not None and e.ndim == 2 else 0
        b = self._batch_struct(ptr(q), nq, Lmax, Kmax, method, correction,
                               refine_candidates, entry_points, expand_count,
                               ptr(ids), ptr(d), ptr(cnt), Kmax, ptr(st),
                               hash_log2, probe_limit, grid_waves, learned_stop,
                               stop_threshold, ptr(Ls), ptr(Ks), seed, ptr(e),
                               0 if e is None else (stride or len(e)), stride,
                               route_entries, refine_mode, traversal, source_window,
                               waves_per_query, candidate_tile, reorder_step)
        check(native.lib().rover_search_batch(self.h, C.byref(b)), "rover_search_batch")
        return (ids, d, cnt, st) if stats else (ids, d, cnt)

    def search_batch_device(self, d_queries, nq, pool_size, num_out, d_ids,
                            d_dists, d_counts, stride, method=native.METHOD_PQ,
                            correction=0, refine_candidates=0, entry_points=1,
                            expand_count=32, d_stats=None, hash_log2=0,
                            probe_limit=0, grid_waves=0, learned_stop=0,
                            stop_threshold=0.0, d_pool_sizes=None,
                            d_num_outs=None, stream=None,
                            seed=native.ENTRY_POINT_SEED, d_entry_ids=None,
                            entry_count=0, entry_stride=0, route_entries=0,
                            refine_mode=0, traversal=0, noalloc=False, source_window=0,
                            waves_per_query=0, candidate_tile=0, reorder_step=0):
        """Device pointers in and out; asynchronous.  With per-query device
        arrays, pool_size must be their maximum.  noalloc=True calls
        rover_search_batch_device_noalloc (fixed per-device workspace)."""
        b = self._batch_struct(d_queries, nq, pool_size, num_out, method,
                               correction, refine_candidates, entry_points,
                               expand_count, d_ids, d_dists, d_counts, stride,
                               d_stats, hash_log2, probe_limit, grid_waves,
                               learned_stop, stop_threshold, d_pool_sizes,
                               d_num_outs, seed, d_entry_ids, entry_count,
                               entry_stride, route_entries, refine_mode, traversal,
                               source_window, waves_per_query, candidate_tile, reorder_step)
        fn = native.lib().rover_search_batch_device_noalloc if noalloc else \
            native.lib().rover_search_batch_device
        check(fn(self.h, C.byref(b), stream),
              "rover_search_batch_device" + ("_noalloc" if noalloc else ""))


    def search_graph_create(self, d_queries, capacity, pool_size, num_out, d_ids,
                            d_dists, d_counts, stride, stream, d_num_queries=None,
                            method=native.METHOD_PQ, correction=0, refine_candidates=0,
                            entry_points=1, expand_count=32, d_stats=None, hash_log2=0,
                            probe_limit=0, grid_waves=0, learned_stop=0,
                            stop_threshold=0.0, d_pool_sizes=None, d_num_outs=None,
                            seed=native.ENTRY_POINT_SEED, d_entry_ids=None,
                            entry_count=0, entry_stride=0, refine_mode=0,
                            source_window=0):
        """Captures a device batch in a HIP graph (rover_search_graph_create);
        returns the graph handle for search_graph_launch."""
        b = self._batch_struct(d_queries, capacity, pool_size, num_out, method,
                               correction, refine_candidates, entry_points,
                               expand_count, d_ids, d_dists, d_counts, stride,
                               d_stats, hash_log2, probe_limit, grid_waves,
                               learned_stop, stop_threshold, d_pool_sizes,
                               d_num_outs, seed, d_entry_ids, entry_count,
                               entry_stride, 0, refine_mode, 0, source_window)
        g = C.c_void_p()
        check(native.lib().rover_search_graph_create(self.h, C.byref(b), d_num_queries,
                                                     stream, C.byref(g)),
              "rover_search_graph_create")
        return g.value

    @staticmethod
    def search_graph_launch(graph, stream):
        check(native.lib().rover_search_graph_launch(graph, stream), "rover_search_graph_launch")

    @staticmethod
    def search_graph_destroy(graph):
        check(native.lib().rover_search_graph_destroy(graph), "rover_search_graph_destroy")

    def autotune(self, d_queries, nq, pool_size, num_out, d_ids, d_dists, d_counts,
                 stride, stream=None, repeats=3, method=native.METHOD_PQ, correction=0,
                 refine_candidates=0, entry_points=1, expand_count=32,
                 d_entry_ids=None, entry_count=0, entry_stride=0,
                 seed=native.ENTRY_POINT_SEED, learned_stop=0, stop_threshold=0.0,
                 refine_mode=0, source_window=0, reorder_step=0):
        """rover_autotune on a device batch; returns native.RoverTuning."""
        b = self._batch_struct(d_queries, nq, pool_size, num_out, method, correction,
                               refine_candidates, entry_points, expand_count, d_ids,
                               d_dists, d_counts, stride, None, 0, 0, 0, learned_stop,
                               stop_threshold, None, None, seed, d_entry_ids, entry_count,
                               entry_stride, 0, refine_mode, 0, source_window,
                               reorder_step=reorder_step)
        t = native.RoverTuning()
        check(native.lib().rover_autotune(self.h, C.byref(b), stream, repeats, C.byref(t)),
              "rover_autotune")
        return t


def pure_pq_search_scalar(graph, codes, codebook, query, entry_ids,
                          expand_count, pool_capacity, num_out,
                          chunk_vectors=None, residual=None):
    """retrieval_pure_pq_search_scalar[_with_policy] through the C-ABI.
    residual=(codes, qparams, bits) selects the ONLINE policy.
    Returns (ids, dists, stats) or (None, None, stats) on -1."""
    L = native.lib()
    graph = np.ascontiguousarray(graph, dtype=np.int32)
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    codebook = np.ascontiguousarray(codebook, dtype=np.float32)
    n = codes.shape[0]
    cv = chunk_vectors or n
    chunks = [np.ascontiguousarray(codes[s:s + cv]) for s in range(0, n, cv)]
    cptr = (C.c_void_p * len(chunks))(*[c.ctypes.data for c in chunks])
    counts = np.array([c.shape[0] for c in chunks], dtype=np.int32)
    m, ks, sub = codebook.shape
    view = native.PurePqSearchIndexView(
        ptr(graph), graph.shape[1], n, C.cast(cptr, C.c_void_p).value, ptr(counts),
        len(chunks), cv, ptr(codebook), m * sub, m, ks, sub)
    q = np.ascontiguousarray(query, dtype=np.float32)
    e = np.ascontiguousarray(entry_ids, dtype=np.int32)
    req = native.PurePqSearchRequest(ptr(q), ptr(e), len(e), expand_count,
                                     pool_capacity, num_out)
    ids = np.zeros(num_out, dtype=np.int32)
    d = np.zeros(num_out, dtype=np.float32)
    st = native.PurePqSearchStats()
    keep = (graph, codes, chunks, cptr, counts, codebook, q, e)
    if residual is not None:
        rc_, qp_, bits = residual
        rc_ = np.ascontiguousarray(rc_, dtype=np.uint8)
        qp_ = np.ascontiguousarray(qp_, dtype=np.float32)
        rv = native.RoverResidualView(ptr(rc_), ptr(qp_), bits)
        ev = C.cast(L.rover_pq_residual_evaluator, C.c_void_p).value
        pol = native.PurePqDistancePolicy(1, ev, C.addressof(rv))
        keep = keep + (rc_, qp_, rv, pol)
        cnt = L.retrieval_pure_pq_search_scalar_with_policy(
            C.byref(view), C.byref(req), C.byref(pol), ptr(ids), ptr(d), num_out,
            C.byref(st))
    else:
        cnt = L.retrieval_pure_pq_search_scalar(
            C.byref(view), C.byref(req), ptr(ids), ptr(d), num_out, C.byref(st))
    del keep
    stats = (st.pq_scored_candidates, st.corrected_candidates, st.expanded_nodes,
             st.visited_dedup_skips)
    if cnt < 0:
        return None, None, stats
    return ids[:cnt].copy(), d[:cnt].copy(), stats
