"""The reference demo's per-setting metrics file (metrics.csv).

One row per (entry points, beam width) setting with the columns and number
formats of retrieval_rpc_demo_pure_pq.c:453-462 (header) and 704-728 (rows):
recalls in percent with 8 decimals, integer microseconds, the search-pool
fingerprint as 16 hex digits.  Columns that describe the DSP session (ADC
microseconds, polls), the residual block cache and the source pipeline are
written as the GPU path has them: no DSP time, residuals resident in HBM
(no cache traffic), one source in flight per query.
"""

COLUMNS = (
    "dataset", "layout", "code_bits", "ks", "m", "code_bytes", "compression_x",
    "entry_points", "pool_size", "queries", "recall_at_1", "recall_at_10",
    "recall_at_100", "wall_us", "work_items", "dsp_adc_us", "dsp_wait_polls",
    "correction_mode", "residual_bits", "corrected_candidates",
    "residual_cache_hits", "residual_cache_misses", "residual_cache_errors",
    "artifact_prefix", "window", "source_jobs", "source_retired", "source_high_water",
    "top_k", "refine_mode", "refine_factor", "refine_budget", "refine_candidates",
    "refine_us", "refine_blocks", "refine_bytes", "refine_hits",
    "search_pool_fingerprint", "total_wall_us",
)

_FLOAT8 = {"recall_at_1", "recall_at_10", "recall_at_100"}


def header():
    return ",".join(COLUMNS) + "\n"


def row(**fields):
    """One CSV line; every column must be given.  Recalls are percentages,
    the fingerprint an integer."""
    missing = [c for c in COLUMNS if c not in fields]
    if missing:
        raise ValueError("metrics row lacks %s" % ", ".join(missing))
    out = []
    for c in COLUMNS:
        v = fields[c]
        if c in _FLOAT8:
            out.append("%.8f" % float(v))
        elif c == "search_pool_fingerprint":
            out.append("%016x" % int(v))
        elif isinstance(v, str):
            out.append(v)
        else:
            out.append("%d" % int(v))
    return ",".join(out) + "\n"


def recall_percent(result_ids, counts, gt, k):
    """100 * hits / (queries * k) with demo_recall_hits_at_k's hits
    (retrieval_rpc_demo_support.c:1709-1728)."""
    import numpy as np
    hits = 0
    for q in range(result_ids.shape[0]):
        c = min(int(counts[q]), k)
        if c > 0:
            hits += int(np.isin(result_ids[q, :c], gt[q, :k]).sum())
    return 100.0 * hits / (result_ids.shape[0] * float(k))
