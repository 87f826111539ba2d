"""ctypes bindings to librover.so (include/rover_gpu.h).

The library is the product: there is no Python or CPU fallback for search.
Loading fails loudly when the library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "librover.so")

ROVER_SUCCESS = 0
ROVER_EFAILED = 1
ROVER_ENOMEMORY = 2
ROVER_EBADPARM = 14
METHOD_PQ = 0
HASH_EXACT = -1  # RoverSearchBatch.hash_log2: exact visited sets in HBM
METHOD_SQ = 1
ENTRY_POINT_SEED = 42
ABI_VERSION = 1


class RoverError(RuntimeError):
    """A failed op; .code is its AEEResult (ROVER_EBADPARM, ...)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class PurePqSearchIndexView(C.Structure):
    _fields_ = [("graph", C.c_void_p), ("graph_degree", C.c_int),
                ("num_vectors", C.c_int), ("code_chunks", C.c_void_p),
                ("chunk_vector_counts", C.c_void_p), ("chunk_count", C.c_int),
                ("chunk_vectors", C.c_int), ("codebook", C.c_void_p),
                ("dim", C.c_int), ("m", C.c_int), ("ks", C.c_int),
                ("subdim", C.c_int)]


class PurePqSearchRequest(C.Structure):
    _fields_ = [("query", C.c_void_p), ("entry_ids", C.c_void_p),
                ("entry_count", C.c_int), ("expand_count", C.c_int),
                ("pool_capacity", C.c_int), ("num_out", C.c_int)]


class PurePqSearchStats(C.Structure):
    _fields_ = [("pq_scored_candidates", C.c_uint64),
                ("corrected_candidates", C.c_uint64),
                ("expanded_nodes", C.c_uint64),
                ("visited_dedup_skips", C.c_uint64)]


class PurePqDistancePolicy(C.Structure):
    _fields_ = [("mode", C.c_int), ("evaluator", C.c_void_p),
                ("context", C.c_void_p)]


class RoverResidualView(C.Structure):
    _fields_ = [("codes", C.c_void_p), ("qparams", C.c_void_p),
                ("bits", C.c_int)]


class RoverSearchBatch(C.Structure):
    _fields_ = [("queries", C.c_void_p), ("num_queries", C.c_int),
                ("pool_sizes", C.c_void_p), ("num_outs", C.c_void_p),
                ("pool_size", C.c_int), ("num_out", C.c_int),
                ("method", C.c_int), ("correction", C.c_int),
                ("refine_candidates", C.c_int), ("entry_points", C.c_int),
                ("seed", C.c_uint32), ("entry_ids", C.c_void_p),
                ("entry_count", C.c_int), ("entry_stride", C.c_int),
                ("route_entries", C.c_int), ("expand_count", C.c_int),
                ("out_ids", C.c_void_p), ("out_dists", C.c_void_p),
                ("out_counts", C.c_void_p), ("out_stride", C.c_int),
                ("out_stats", C.c_void_p), ("hash_log2", C.c_int),
                ("probe_limit", C.c_int), ("grid_waves", C.c_int),
                ("learned_stop", C.c_int), ("stop_threshold", C.c_float),
                ("refine_mode", C.c_int), ("traversal", C.c_int),
                ("source_window", C.c_int), ("waves_per_query", C.c_int),
                ("candidate_tile", C.c_int), ("reorder_step", C.c_int)]


REFINE_CURRENT, REFINE_ERROR4 = 1, 2
TRAVERSAL_DEFAULT, TRAVERSAL_BEST_FIRST, TRAVERSAL_SQMD, TRAVERSAL_SQ_DSP = 0, 1, 2, 3


class RoverLearnedStopTrain(C.Structure):
    _fields_ = [("queries", C.c_void_p), ("num_queries", C.c_int),
                ("pool_size", C.c_int), ("correction", C.c_int),
                ("label_k", C.c_int), ("label_slack", C.c_int),
                ("entry_points", C.c_int), ("entry_ids", C.c_void_p),
                ("entry_count", C.c_int), ("entry_stride", C.c_int),
                ("route_entries", C.c_int),
                ("expand_count", C.c_int),
                ("stop_every", C.c_int), ("min_steps", C.c_int),
                ("trace_cap", C.c_int), ("steps", C.c_int), ("batch", C.c_int),
                ("lr", C.c_float), ("pos_weight", C.c_float),
                ("seed", C.c_uint32), ("final_loss", C.c_float),
                ("samples", C.c_long), ("positive_frac", C.c_float)]


class RoverTuning(C.Structure):
    _fields_ = [("hash_log2", C.c_int), ("probe_limit", C.c_int),
                ("grid_waves", C.c_int), ("ms", C.c_float), ("default_ms", C.c_float),
                ("tried", C.c_int), ("rejected", C.c_int),
                ("waves_per_query", C.c_int), ("candidate_tile", C.c_int)]


# (name, restype, argtypes); every symbol include/rover_gpu.h declares.
_H = C.c_uint64
_P = C.c_void_p
_I = C.c_int
_F = C.c_float
SIGNATURES = [
    ("calculator_open", _I, [C.c_char_p, C.POINTER(_H)]),
    ("calculator_close", _I, [_H]),
    ("calculator_rover_configure", _I, [_H, C.c_int32, C.c_int32]),
    ("calculator_retrieval_load", _I,
     [_H, _P, _I, _P, _I, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _F, _F]),
    ("calculator_retrieval_search", _I,
     [_H, _P, _I, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _I, _P, _I,
      _P]),
    ("calculator_retrieval_unload", _I, [_H]),
    ("calculator_retrieval_load_dma", _I,
     [_H, _P, _P, _I, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _F, _F]),
    ("calculator_retrieval_unload_dma", _I, [_H]),
    ("rover_sq_graph_load", _I, [_H, _P, C.c_int64, C.c_int32, C.c_int32, C.c_int32]),
    ("calculator_retrieval_pure_pq_load_dma", _I,
     [_H, _P, _I, _P, _I, C.c_int32, _P, _I, C.c_int32, C.c_int32, C.c_int32,
      C.c_int32, C.c_int32, C.c_int32, _P, _I]),
    ("calculator_retrieval_pure_pq_unload_dma", _I, [_H]),
    ("rover_pure_pq_graph_load", _I,
     [_H, _P, C.c_int64, C.c_int32, C.c_int32, C.c_int32]),
    ("rover_pure_pq_residual_load", _I,
     [_H, _P, C.c_int64, _P, C.c_int64, C.c_int32]),
    ("rover_pure_pq_original_ids", _I, [_H, _P, _I]),
    ("rover_sq_residual_load", _I, [_H, _P, C.c_int64, _P, C.c_int64, C.c_int32]),
    ("rover_sq_split_bias_load", _I, [_H, _P, C.c_int64, _P]),
    ("rover_pure_pq_error4_load", _I, [_H, _P, C.c_int64, _P]),
    ("rover_pure_pq_error4_unload", _I, [_H]),
    ("rover_sq_error4_load", _I, [_H, _P, C.c_int64, _P]),
    ("rover_sq_error4_unload", _I, [_H]),
    ("rover_pure_pq_decode", _I, [_H, _P, _I, _P]),
    ("rover_sq_decode", _I, [_H, _P, _I, _P]),
    ("rover_trace_open", _I, [C.c_char_p, _I, C.POINTER(_P)]),
    ("rover_trace_query", _I, [_P, _I, _I, _I, _I, _P, _P, _I, _P, _I, _I, _P, _P, _I]),
    ("rover_trace_close", _I, [_P]),
    ("rover_sha256", None, [_P, C.c_size_t, _P]),
    ("rover_pure_pq_router_load", _I, [_H, _P, _I, _P, _I]),
    ("rover_route_entries", _I, [_H, _P, _I, _I, _P, _P]),
    ("rover_pq_residual_evaluator", _I, [_P, _I, _F, _P]),
    ("retrieval_pure_pq_search_scalar_with_policy", _I,
     [_P, _P, _P, _P, _P, _I, _P]),
    ("retrieval_pure_pq_search_scalar", _I, [_P, _P, _P, _P, _I, _P]),
    ("rover_pure_pq_view_release", None, []),
    ("rover_search_batch", _I, [_H, _P]),
    ("rover_search_batch_device", _I, [_H, _P, _P]),
    ("rover_search_batch_device_noalloc", _I, [_H, _P, _P]),
    ("rover_search_graph_create", _I, [_H, _P, _P, _P, C.POINTER(_P)]),
    ("rover_search_graph_launch", _I, [_P, _P]),
    ("rover_search_graph_destroy", _I, [_P]),
    ("rover_autotune", _I, [_H, _P, _P, _I, _P]),
    ("rover_make_seed_ids", _I, [_I, _I, _I, C.c_uint32, _P]),
    ("rover_result_fingerprint", C.c_uint64, [C.c_uint64, _P, _P, _I]),
    ("rover_sqmd_seed_ids", _I, [_I, _I, _I, C.c_uint32, _P, _I]),
    ("rover_learned_stop_train", _I, [_H, _P]),
    ("rover_search_trace_device", _I, [_H, _P, _I, _I, _I, _P, _P, _P, _P]),
    ("rover_learned_stop_set", _I, [_H, _P, _I, _P, _P, _I, _I, _I]),
    ("rover_learned_stop_get", _I, [_H, _P, _I, _P, _P]),
    ("rover_mlp_param_count", _I, []),
    ("rover_mlp_forward_backward", _I, [_P, _P, _P, _P, _I, _P, _P, _P]),
    ("rover_mlp_adam_step", _I, [_P, _P, _P, _P, _I, _F, _F]),
    ("rover_build_pq_encode", _I, [_P, _I, _I, _I, _I, _P, _P]),
    ("rover_build_residual_encode", _I, [_P, _P, _I, _I, _I, _I, _P, _I, _P, _P]),
    ("rover_build_graph_exact", _I, [_P, _I, _I, _I, _I, C.c_uint32, _I, _P]),
    ("rover_build_graph_exact_clustered", _I,
     [_P, _I, _I, _I, _I, C.c_uint32, _I, _P, _P, _I, _P, _P]),
    ("rover_build_graph_rnn", _I, [_P, _I, _I, _I, _I, _I, _I, C.c_uint32, _I, _P]),
    ("rover_build_error4_encode", _I,
     [_P, _P, _P, _P, _I, _I, _I, _I, _P, _I, _P, _P]),
    ("rover_device_alloc", _I, [C.POINTER(_P), C.c_size_t]),
    ("rover_device_free", _I, [_P]),
    ("rover_host_alloc", _I, [C.POINTER(_P), C.c_size_t]),
    ("rover_host_free", _I, [_P]),
    ("rover_device_upload", _I, [_P, _P, C.c_size_t]),
    ("rover_device_download", _I, [_P, _P, C.c_size_t]),
    ("rover_device_synchronize", _I, []),
    ("rover_device_copy", _I, [_P, _P, C.c_size_t]),
    ("rover_device_count", _I, []),
    ("rover_last_error", C.c_char_p, []),
    ("rover_event_create", _I, [C.POINTER(_P)]),
    ("rover_event_record", _I, [_P, _P]),
    ("rover_event_elapsed_ms", _I, [_P, _P, C.POINTER(_F)]),
    ("rover_event_destroy", _I, [_P]),
    ("rover_stream_create", _I, [C.POINTER(_P)]),
    ("rover_stream_destroy", _I, [_P]),
    ("rover_stream_synchronize", _I, [_P]),
]

_lib = None


def lib():
    """Loads librover.so once; raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RoverError(
                "librover.so not built (%s): run __graft_entry__.build()"
                % LIB_PATH)
        # One HIP runtime per process: librover.so needs libamdhip64.so.7
        # by name, and a library of that name already loaded (torch's, when
        # torch is imported first) serves it.  Loading librover.so first
        # would map the system runtime, and torch would then load its own
        # copy beside it; importing torch here keeps torch's the only one.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            if not hasattr(L, name):
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    msg = lib().rover_last_error()
    return msg.decode() if msg else ""


def check(rc, what):
    if rc != ROVER_SUCCESS:
        raise RoverError("%s failed (rc=%d): %s" % (what, rc, last_error()), rc)


def ptr(a):
    """Address of a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data


class DeviceBuffer:
    """HBM allocation owned by Python, freed on close/garbage collection."""

    def __init__(self, nbytes):
        p = C.c_void_p()
        check(lib().rover_device_alloc(C.byref(p), max(int(nbytes), 16)),
              "rover_device_alloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes)
        if a.nbytes:
            check(lib().rover_device_upload(buf.ptr, ptr(a), a.nbytes),
                  "rover_device_upload")
        return buf

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().rover_device_download(ptr(out), self.ptr, out.nbytes),
                  "rover_device_download")
        return out

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().rover_device_upload(self.ptr, ptr(a), a.nbytes),
              "rover_device_upload")

    def close(self):
        if self.ptr is not None and _lib is not None:
            _lib.rover_device_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PinnedArray:
    """A numpy array over page-locked host memory (rover_host_alloc)."""

    def __init__(self, shape, dtype):
        dt = np.dtype(dtype)
        n = int(np.prod(shape)) * dt.itemsize
        p = C.c_void_p()
        check(lib().rover_host_alloc(C.byref(p), max(n, 16)), "rover_host_alloc")
        self.ptr = p.value
        buf = (C.c_char * max(n, 16)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)

    def close(self):
        if self.ptr is not None and _lib is not None:
            self.array = None
            _lib.rover_host_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_count():
    return lib().rover_device_count()


def synchronize():
    check(lib().rover_device_synchronize(), "rover_device_synchronize")


def make_seed_ids(n, entry_points, graph_entry, seed=ENTRY_POINT_SEED):
    out = np.zeros(max(entry_points, 1), dtype=np.int32)
    cnt = lib().rover_make_seed_ids(n, entry_points, graph_entry, seed, ptr(out))
    return out[:cnt]


FINGERPRINT_BASIS = 14695981039346656037


def result_fingerprint(ids, dists, h=FINGERPRINT_BASIS):
    """rover_result_fingerprint over one result list (ids, distances)."""
    i = np.ascontiguousarray(ids, dtype=np.int32)
    d = np.ascontiguousarray(dists, dtype=np.float32)
    assert i.shape == d.shape
    return lib().rover_result_fingerprint(h, ptr(i), ptr(d), int(i.size))


def results_fingerprint(ids_rows, dist_rows, counts, h=FINGERPRINT_BASIS):
    """The demo's fingerprint chained over queries: row q contributes its
    first counts[q] results."""
    for q in range(len(counts)):
        c = int(counts[q])
        if c > 0:
            h = result_fingerprint(ids_rows[q, :c], dist_rows[q, :c], h)
    return h


TRACE_TEXT, TRACE_P15 = 1, 2


class Trace:
    """A refine trace file (rover_trace_open): TRACE_TEXT, the pure-PQ demo's
    text trace, or TRACE_P15, the SQ pipeline's binary trace."""

    def __init__(self, path, fmt):
        p = C.c_void_p()
        check(lib().rover_trace_open(str(path).encode(), fmt, C.byref(p)), "rover_trace_open")
        self.ptr = p.value

    def query(self, pool_size, query_index, ids, dists, out_ids, out_dists, refine_budget=0,
              entry_points=0, decoded=None):
        i = np.ascontiguousarray(ids, dtype=np.int32)
        d = np.ascontiguousarray(dists, dtype=np.float32)
        oi = np.ascontiguousarray(out_ids, dtype=np.int32)
        od = np.ascontiguousarray(out_dists, dtype=np.float32)
        v = None if decoded is None else np.ascontiguousarray(decoded, dtype=np.float32)
        check(lib().rover_trace_query(
            self.ptr, int(entry_points), int(pool_size), int(query_index), int(refine_budget),
            ptr(i), ptr(d), len(i), ptr(v), 0 if v is None else v.shape[0],
            0 if v is None else v.shape[1], ptr(oi), ptr(od), len(oi)), "rover_trace_query")

    def close(self):
        if self.ptr is not None:
            check(lib().rover_trace_close(self.ptr), "rover_trace_close")
        self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Event:
    def __init__(self):
        p = C.c_void_p()
        check(lib().rover_event_create(C.byref(p)), "rover_event_create")
        self.ptr = p.value

    def record(self, stream=None):
        check(lib().rover_event_record(self.ptr, stream), "rover_event_record")

    def elapsed_ms(self, stop):
        ms = C.c_float()
        check(lib().rover_event_elapsed_ms(self.ptr, stop.ptr, C.byref(ms)),
              "rover_event_elapsed_ms")
        return ms.value

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.rover_event_destroy(self.ptr)
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = C.c_void_p()
        check(lib().rover_stream_create(C.byref(p)), "rover_stream_create")
        self.ptr = p.value

    def synchronize(self):
        check(lib().rover_stream_synchronize(self.ptr), "rover_stream_synchronize")

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.rover_stream_destroy(self.ptr)
        except Exception:
            pass
