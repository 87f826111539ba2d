"""Benchmark: QPS at recall@10 >= 0.95 on the 1M x 128 PQ index, one MI355X,
learned early termination against the plain fixed-L search.

Workload: N=1,000,000 base vectors, d=128, from a 1024-component Gaussian
mixture (centres U[0,100]^d, sigma 8, clipped to [0,255]); PQ M=32 x 256
(8-bit codes) plus per-vector residual codes; degree-32 graph from the
reference's exact search-build (exact 64-NN, random initial rows, RNG
pruning with fill), built on the device; 10,000 queries from the same
mixture.  The index is built on the GPU at start-up and the queries are
resident in HBM.

Entries: the reference's seeds, the medoid plus 31 ids from its LCG
(entry_points 32), the same list for every query and passed as entry rows.
The routed leg repeats the plain / learned comparison with per-query rows:
the representative nodes of the query's 4 nearest of 4096 k-means
centroids, chosen on the host before the timed region.

One "step" is one batched search of all queries.  For each mode (online
residual correction; ADC ranking + final refine) the beam width L is the
smallest of the sweep reaching recall@10 >= 0.95 against exact float32
ground truth; the plain search uses the faster mode.  The learned stop
predictor is trained on traces of that mode's plain search of 200,000
training queries at L=512, once per label slack (label 1 once all but
`slack` of the final top-10 are found), each swept the same way; the slack
reaching the target fastest is reported.  The L grid is refined by
bisection to the smallest L reaching the target.  `value` and
`ms_per_step` are end to end: each step uploads the queries and entry rows
from page-locked host memory, searches and downloads the results
(rover_search_batch); `config.device_resident` has the same searches with
everything already in HBM, whose kernel time the roofline uses.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import ctypes as C
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from rover_amd import artifacts, data, index, metrics, native  # noqa: E402
from rover_amd.search import Calculator  # noqa: E402

HBM_PEAK_GBS = 8000.0
T0 = time.time()


def log(*a):
    print("[bench %7.1fs]" % (time.time() - T0), *a, file=sys.stderr, flush=True)


def algorithmic_bytes(stats, R, code_bytes, res_bytes, dim, k, refine_F,
                      codebook_bytes, online):
    """Bytes one launch must move: graph rows of expanded nodes (4R each),
    codes of scored candidates, their residual codes and qparams in online
    mode, codes+residual+qparams of refined candidates, the query vector
    and the output; the codebook once."""
    E = stats[:, 2].astype(np.float64)
    Cq = stats[:, 0].astype(np.float64)
    b = 4.0 * R * E + code_bytes * Cq
    if online:
        b += (res_bytes + 8) * Cq
    b += (code_bytes + res_bytes + 8) * refine_F + 4 * dim + 8 * k
    return float(b.sum() + codebook_bytes)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform string."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def reference_leg(pyoracle, ix, base, queries, ents, gids, gd, L, k, args, threads):
    """The reference's own C search (retrieval_pure_pq_search_scalar_with_
    policy with its residual evaluator, compiled from its sources) over a
    bounded sample of the queries on `threads` host threads, each with its
    own residual reader; the share of queries whose ids and distance bits
    equal the GPU's, and whether the reference's residual encoder reproduces
    the index's residual codes."""
    import threading
    ref = pyoracle.reference()
    n_base = ix.graph.shape[0]
    res_r = np.zeros_like(ix.res_codes)
    qp_r = np.zeros_like(ix.res_qparams)
    xb = np.ascontiguousarray(base, dtype=np.float32)
    codes = np.ascontiguousarray(ix.codes)
    cb = np.ascontiguousarray(ix.codebook, dtype=np.float32)
    graph = np.ascontiguousarray(ix.graph, dtype=np.int32)
    h = ref.ref_residual_open(xb.ctypes.data, codes.ctypes.data, cb.ctypes.data, n_base,
                              args.dim, args.m, 256, ix.res_bits, res_r.ctypes.data,
                              qp_r.ctypes.data)
    if not h:
        return None
    forks = []
    try:
        enc_same = bool(np.array_equal(res_r, ix.res_codes) and
                        np.array_equal(qp_r.view(np.uint32), ix.res_qparams.view(np.uint32)))
        forks = [ref.ref_residual_fork(h) for _ in range(threads)]
        if not all(forks):
            return None
        nq = queries.shape[0]
        done = [0] * threads
        same = [0] * threads
        deadline = time.time() + args.cpu_seconds

        def work(t):
            ids_r = np.zeros(k, dtype=np.int32)
            d_r = np.zeros(k, dtype=np.float32)
            st_r = np.zeros(4, dtype=np.uint64)
            for qi in range(t, nq, threads):
                if time.time() >= deadline:
                    break
                q = np.ascontiguousarray(queries[qi], dtype=np.float32)
                e = np.ascontiguousarray(ents[qi], dtype=np.int32)
                rc = ref.ref_search(graph.ctypes.data, graph.shape[1], n_base,
                                    codes.ctypes.data, 16384, cb.ctypes.data, args.dim, args.m,
                                    256, q.ctypes.data, e.ctypes.data, len(e), args.R, L, k, 1,
                                    forks[t], None, 8, C.c_float(1.0), C.c_float(0.0),
                                    ids_r.ctypes.data, d_r.ctypes.data, k, st_r.ctypes.data)
                same[t] += int(rc == k and np.array_equal(ids_r, gids[qi]) and
                               np.array_equal(d_r.view(np.uint32), gd[qi].view(np.uint32)))
                done[t] += 1

        t1 = time.time()
        th = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        el = time.time() - t1
        total = sum(done)
        return {"value": round(total / el, 1), "unit": "queries/s", "cores": threads,
                "kind": "reference",
                "sample": "%d of the queries (query i on thread i mod %d), plain online "
                          "search at L=%d (%.1f s)" % (total, threads, L, el),
                "gpu_results_identical_frac": sum(same) / max(total, 1),
                "residual_encoder_identical": enc_same}
    finally:
        for f in forks:
            if f:
                ref.ref_residual_close(f)
        ref.ref_residual_close(h)


def batching_leg(calc, queries, ents, S, mode, refine_of, args, cap=4096, kmax=100,
                 lmax=512, refine_mode=0):
    """Uneven batches through one captured HIP graph: batch size uniform in
    1..4096, per-query k from {1, 10, 100} and beam width uniform in
    max(16, k)..512 (k <= L is required), queries drawn from the query set.
    Each batch uploads its queries, entry rows, L and k and the live count,
    replays the graph and downloads the results; wall latency includes those
    transfers, device latency is the replay alone (HIP events)."""
    rng = np.random.default_rng(2024)
    dim = queries.shape[1]
    dq = native.DeviceBuffer(cap * dim * 4)
    de = native.DeviceBuffer(cap * S * 4)
    dL = native.DeviceBuffer(cap * 4)
    dK = native.DeviceBuffer(cap * 4)
    ids = native.DeviceBuffer(cap * kmax * 4)
    dd = native.DeviceBuffer(cap * kmax * 4)
    cnt = native.DeviceBuffer(cap * 4)
    live = native.DeviceBuffer(4)
    stream = native.Stream()
    live.upload(np.array([cap], dtype=np.int32))
    dL.upload(np.full(cap, lmax, dtype=np.int32))
    dK.upload(np.full(cap, 10, dtype=np.int32))
    dq.upload(np.ascontiguousarray(queries[:cap], dtype=np.float32))
    de.upload(np.ascontiguousarray(ents[:cap], dtype=np.int32))
    g = calc.search_graph_create(dq.ptr, cap, lmax, kmax, ids.ptr, dd.ptr, cnt.ptr, kmax,
                                 stream.ptr, d_num_queries=live.ptr,
                                 correction=1 if mode == "online" else 0,
                                 refine_candidates=refine_of(mode, lmax), refine_mode=refine_mode,
                                 expand_count=args.R, d_pool_sizes=dL.ptr, d_num_outs=dK.ptr,
                                 d_entry_ids=de.ptr, entry_count=S, entry_stride=S)
    e0, e1 = native.Event(), native.Event()
    wall, dev, total = [], [], 0
    try:
        calc.search_graph_launch(g, stream.ptr)  # warm-up
        stream.synchronize()
        log("batching: graph captured (capacity %d, L <= %d, k <= %d); %d batches"
            % (cap, lmax, kmax, args.batches))
        for _ in range(args.batches):
            n = int(rng.integers(1, cap + 1))
            sel = rng.integers(0, len(queries), n)
            kq = rng.choice(np.array([1, 10, 100], dtype=np.int32), n)
            lq = (np.maximum(16, kq) + rng.integers(0, lmax + 1 - np.maximum(16, kq))
                  ).astype(np.int32)
            q = np.ascontiguousarray(queries[sel], dtype=np.float32)
            r = np.ascontiguousarray(ents[sel], dtype=np.int32)
            t0 = time.perf_counter()
            dq.upload(q)
            de.upload(r)
            dL.upload(lq)
            dK.upload(kq)
            live.upload(np.array([n], dtype=np.int32))
            e0.record(stream.ptr)
            calc.search_graph_launch(g, stream.ptr)
            e1.record(stream.ptr)
            stream.synchronize()
            c = cnt.download(np.int32, n)
            ids.download(np.int32, n * kmax)
            dd.download(np.float32, n * kmax)
            wall.append((time.perf_counter() - t0) * 1e3)
            dev.append(e0.elapsed_ms(e1))
            total += n
            assert np.array_equal(c, kq), "batch search failed: %s" % np.unique(c)
    finally:
        calc.search_graph_destroy(g)
    wall, dev = np.array(wall), np.array(dev)
    out = dict(batches=args.batches, queries=total,
               batch_size="uniform 1..%d" % cap, L="uniform max(16,k)..%d" % lmax,
               k="uniform {1,10,100}", capture="one HIP graph, live count in HBM",
               wall_p50_ms=round(float(np.percentile(wall, 50)), 3),
               wall_p99_ms=round(float(np.percentile(wall, 99)), 3),
               device_p50_ms=round(float(np.percentile(dev, 50)), 3),
               device_p99_ms=round(float(np.percentile(dev, 99)), 3),
               wall_qps=round(float(total / (wall.sum() / 1e3)), 1),
               device_qps=round(float(total / (dev.sum() / 1e3)), 1))
    log("batching:", out)
    return out


def sq_leg_run(args, n=100_000, nq=1000, L=64, k=10, ep=32):
    """The small SQ workload: 100k x 128 mixture (256 components, sigma 8),
    SQ8 codes with 4-bit per-vector residuals, degree-32 graph, entry =
    medoid, 1k queries, L=64, k=10.  Times the reference's SQ traversal
    (distance-only heaps; plain and with the exact residual correction) and
    the best-first SQ search over the same index, HIP events around single
    launches with queries in HBM; the oracle's SQ traversal on the host for
    comparison.  The graph is built on the device by the reference's
    exact-candidate search-build (exact 64-NN, RNG prune to 32).  The same
    base with SQ4 codes and 4-bit residuals runs the traversal plain,
    corrected and with the split distance (SQMDBIAS bias)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    base, queries = data.workload(n, 128, 256, 8.0, nq, base_seed=7, query_seed=8)
    t_build = time.time()
    graph = index.build_graph_exact(base, 32, 64, 42, 1)  # exact 64-NN + RNG prune
    t_build = time.time() - t_build
    entry = index.medoid(base)
    scale, zero = index.sq_params(base, 8)
    codes = index.encode_sq(base, 8, scale, zero)
    res, rqp = index.encode_sq_residual(base, codes, 8, scale, zero, 4)
    gt = data.ground_truth(base, queries, k=k)
    calc = Calculator()
    calc.rover_configure(l1_bits=8)
    calc.retrieval_load(graph, codes, 128, entry, scale, zero)
    calc.sq_residual_load(res, rqp, 4)
    dq = native.DeviceBuffer.from_array(queries)
    d_ids, d_d = native.DeviceBuffer(nq * k * 4), native.DeviceBuffer(nq * k * 4)
    d_cnt, d_st = native.DeviceBuffer(nq * 4), native.DeviceBuffer(nq * 32)
    out = {"workload": "100k x 128 mixture (256 comps, sigma 8), SQ8 + 4-bit residuals, "
                       "R=32 (exact 64-NN + RNG prune, built on the device), entry medoid, "
                       "%d seeds, %d queries, L=%d, k=%d" % (ep, nq, L, k),
           "graph_build_s": round(t_build, 3)}
    oix = pyoracle.Index(graph, res_codes=res, res_qparams=rqp, res_bits=4, sq_codes=codes,
                         sq_bits=8, sq_scale=float(scale), sq_zero=float(zero), dim=128)
    seeds = pyoracle.sqmd_seed_ids(n, ep, entry, 42)
    graph = np.ascontiguousarray(graph, dtype=np.int32)

    def reference_sq(search, gpu_ids, budget_s=5.0):
        """The reference's own SQ traversal (its fake-DSP turn loop, compiled)
        on one host thread over the first queries for about budget_s, and the
        share whose top-k ids equal the GPU's."""
        if not pyoracle.ref_available():
            return {}
        oid = np.zeros(L, dtype=np.int32)
        od = np.zeros(L, dtype=np.float32)
        done = same_r = 0
        t1 = time.time()
        while done < nq and time.time() - t1 < budget_s:
            q = np.ascontiguousarray(queries[done], dtype=np.float32)
            rc = search(q, oid, od)
            same_r += int(rc >= k and np.array_equal(oid[:k], gpu_ids[done]))
            done += 1
        el_r = time.time() - t1
        return dict(reference_qps_1thread=round(done / el_r, 1), reference_sample=done,
                    reference_ids_identical_frac=same_r / max(done, 1))

    for name, trav, corr in (("sqmd", native.TRAVERSAL_SQMD, 0),
                             ("sqmd_corrected", native.TRAVERSAL_SQMD, 1),
                             ("best_first", native.TRAVERSAL_BEST_FIRST, 0)):
        def launch():
            calc.search_batch_device(dq.ptr, nq, L, k, d_ids.ptr, d_d.ptr, d_cnt.ptr, k,
                                     method=native.METHOD_SQ, correction=corr,
                                     entry_points=ep, expand_count=32, d_stats=d_st.ptr,
                                     traversal=trav)
        launch()
        native.synchronize()
        ids = d_ids.download(np.int32, nq * k).reshape(nq, k)
        cnt = d_cnt.download(np.int32, nq)
        st = d_st.download(np.uint64, nq * 4).reshape(nq, 4)
        assert (cnt == k).all(), (name, np.unique(cnt))
        for _ in range(args.warmup):
            launch()
        e0, e1 = native.Event(), native.Event()
        e0.record()
        for _ in range(args.steps):
            launch()
        e1.record()
        ms = e0.elapsed_ms(e1) / args.steps
        row = dict(recall10=round(data.recall_at_k(ids, cnt, gt, k), 4),
                   qps=round(nq / (ms / 1e3), 1), ms_per_batch=round(ms, 4),
                   mean_expanded=round(float(st[:, 2].mean()), 2))
        if trav == native.TRAVERSAL_SQMD:
            # the oracle's SQ traversal on one host thread, and the identity check
            t1 = time.time()
            same = 0
            for qi in range(nq):
                oi, od, _ = pyoracle.sqmd_search(oix, queries[qi], seeds, ep, 32, L, k, corr)
                same += int(np.array_equal(oi, ids[qi]))
            el = time.time() - t1
            row.update(cpu_oracle_qps_1thread=round(nq / el, 1),
                       gpu_ids_identical_frac=same / nq)
            row.update(reference_sq(lambda q, oid, od: pyoracle.reference().ref_sqmd_search(
                graph.ctypes.data, graph.shape[1], n, codes.ctypes.data, 128, 8,
                C.c_float(scale), C.c_float(zero), res.ctypes.data, rqp.ctypes.data, 4, corr,
                q.ctypes.data, ep, int(entry), 42, 32, L, 8, oid.ctypes.data, od.ctypes.data, L,
                None), ids))
        log("sq leg %s: %s" % (name, row))
        out[name] = row
    # graph construction: the device builders on the whole base, the
    # compiled reference's builders (16 host threads) on a 20k-point sample
    builds = {"device_exact_s": round(t_build, 3)}
    t1 = time.time()
    rnn = index.build_graph_rnn(base, 16, 32, 4, 4, 42)
    builds["device_rnn_s"] = round(time.time() - t1, 3)
    builds["rnn_mean_degree"] = round(float((rnn >= 0).sum(1).mean()), 2)
    if pyoracle.ref_available():
        sample = np.ascontiguousarray(base[:20_000])
        ns = sample.shape[0]
        ref = pyoracle.reference()
        g = np.zeros((ns, 32), dtype=np.int32)
        t1 = time.time()
        rc = ref.ref_build_exact(sample.ctypes.data, ns, 128, 32, 64, 42, 1, 16, g.ctypes.data)
        builds["reference_exact_20k_s_16threads"] = round(time.time() - t1, 3) if rc == 0 else None
        t1 = time.time()
        rc = ref.ref_rnn_build(sample.ctypes.data, ns, 128, 16, 32, 4, 4, 42, 16, 32,
                               g.ctypes.data)
        builds["reference_rnn_20k_s_16threads"] = round(time.time() - t1, 3) if rc == 0 else None
        t1 = time.time()
        gd = index.build_graph_exact(sample, 32, 64, 42, 1)
        builds["device_exact_20k_s"] = round(time.time() - t1, 3)
        t1 = time.time()
        gr = index.build_graph_rnn(sample, 16, 32, 4, 4, 42)
        builds["device_rnn_20k_s"] = round(time.time() - t1, 3)
        builds["rnn_20k_identical_to_reference"] = bool(np.array_equal(gr, g))
        del gd
    log("graph builds: %s" % builds)
    out["graph_builds"] = builds
    # the split distance needs SQ4 codes: the same base and graph with SQ4
    # codes, 4-bit residuals of the SQ4 reconstruction and their SQMDBIAS
    scale4, zero4 = index.sq_params(base, 4)
    codes4 = index.encode_sq(base, 4, scale4, zero4)
    res4, rqp4 = index.encode_sq_residual(base, codes4, 4, scale4, zero4, 4)
    bias4 = artifacts.split_bias(codes4, scale4, zero4, res4, rqp4, 4, 128)
    calc.rover_configure(l1_bits=4)
    calc.retrieval_load(graph, codes4, 128, entry, scale4, zero4)
    calc.sq_residual_load(res4, rqp4, 4)
    calc.sq_split_bias_load(artifacts.split_bias_image(bias4, 128, 4))
    oix4 = pyoracle.Index(graph, res_codes=res4, res_qparams=rqp4, res_bits=4,
                          sq_codes=codes4, sq_bits=4, sq_scale=float(scale4),
                          sq_zero=float(zero4), dim=128, split_bias=bias4)
    for name, corr in (("sq4_sqmd", 0), ("sq4_sqmd_corrected", 1), ("sq4_sqmd_split", 2)):
        def launch():
            calc.search_batch_device(dq.ptr, nq, L, k, d_ids.ptr, d_d.ptr, d_cnt.ptr, k,
                                     method=native.METHOD_SQ, correction=corr,
                                     entry_points=ep, expand_count=32, d_stats=d_st.ptr,
                                     traversal=native.TRAVERSAL_SQMD)
        launch()
        native.synchronize()
        ids = d_ids.download(np.int32, nq * k).reshape(nq, k)
        cnt = d_cnt.download(np.int32, nq)
        st = d_st.download(np.uint64, nq * 4).reshape(nq, 4)
        assert (cnt == k).all(), (name, np.unique(cnt))
        for _ in range(args.warmup):
            launch()
        e0, e1 = native.Event(), native.Event()
        e0.record()
        for _ in range(args.steps):
            launch()
        e1.record()
        ms = e0.elapsed_ms(e1) / args.steps
        same = 0
        for qi in range(nq):
            oi, od, _ = pyoracle.sqmd_search(oix4, queries[qi], seeds, ep, 32, L, k, corr)
            same += int(np.array_equal(oi, ids[qi]))
        row = dict(recall10=round(data.recall_at_k(ids, cnt, gt, k), 4),
                   qps=round(nq / (ms / 1e3), 1), ms_per_batch=round(ms, 4),
                   mean_expanded=round(float(st[:, 2].mean()), 2),
                   gpu_ids_identical_frac=same / nq)
        if corr == 2:
            row.update(reference_sq(lambda q, oid, od: pyoracle.reference().ref_sqmd_split_search(
                graph.ctypes.data, graph.shape[1], n, codes4.ctypes.data, 128,
                C.c_float(scale4), C.c_float(zero4), res4.ctypes.data, rqp4.ctypes.data, 4,
                bias4.ctypes.data, q.ctypes.data, ep, int(entry), 42, 32, L, 8, oid.ctypes.data,
                od.ctypes.data, L), ids))
        else:
            row.update(reference_sq(lambda q, oid, od: pyoracle.reference().ref_sqmd_search(
                graph.ctypes.data, graph.shape[1], n, codes4.ctypes.data, 128, 4,
                C.c_float(scale4), C.c_float(zero4), res4.ctypes.data, rqp4.ctypes.data, 4,
                corr, q.ctypes.data, ep, int(entry), 42, 32, L, 8, oid.ctypes.data,
                od.ctypes.data, L, None), ids))
        log("sq leg %s: %s" % (name, row))
        out[name] = row
    calc.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--m", type=int, default=32)
    ap.add_argument("--R", type=int, default=32)
    ap.add_argument("--components", type=int, default=1024)
    ap.add_argument("--res-bits", type=int, default=8)
    ap.add_argument("--router", type=int, default=4096,
                    help="entry-router centroids (k-means over the base)")
    ap.add_argument("--router-seeding", choices=["stride", "kmeans++"], default="stride",
                    help="initial router centroids: uniform-stride sample rows or D^2 "
                         "(k-means++) seeding")
    ap.add_argument("--route", type=int, default=4,
                    help="entries per query: representatives of its nearest centroids")
    ap.add_argument("--sigma", type=float, default=8.0,
                    help="standard deviation of the mixture's components")
    ap.add_argument("--build-clusters", type=int, default=0,
                    help="k-means clusters for --graph clustered (default n / 1024)")
    ap.add_argument("--uneven-L", default="",
                    help="lo,hi: also search every query once with its own beam width "
                         "uniform in [lo, hi], plain and learned stop, in one launch")
    ap.add_argument("--graph", choices=["backlinks", "reference", "clustered"],
                    default="reference",
                    help="backlinks: exact 64-NN + RNG prune + reverse edges (torch); "
                         "reference: the reference's exact search-build on the device "
                         "(random initial rows, RNG prune with fill)")
    ap.add_argument("--entry", choices=["routed", "medoid"], default="medoid",
                    help="routed: per-query rows from the k-means router; medoid: "
                         "the reference's seeds (medoid + LCG ids, --entry-points)")
    ap.add_argument("--entry-points", type=int, default=32)
    ap.add_argument("--no-window2-leg", dest="window2_leg", action="store_false",
                    help="skip the source-window-2 (live pipeline) plain / learned leg")
    ap.add_argument("--no-routed-leg", action="store_true",
                    help="skip the plain / learned comparison with routed entry rows")
    ap.add_argument("--Ls", default="10,12,16,20,24,32,48,64,96,128,192,256,384,512,768,1024",
                    help="beam widths swept per mode (each >= k)")
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--refine", type=int, default=64,
                    help="refine candidates F of the ADC mode")
    ap.add_argument("--modes", default="online,adc")
    ap.add_argument("--only-sq", action="store_true",
                    help="run only the SQ leg (graph build and SQ searches) and print its "
                         "JSON; a profiling aid, not the benchmark line")
    ap.add_argument("--no-sq", action="store_true",
                    help="skip the SQ leg (100k x 128 SQ8, L=64: the reference's SQ "
                         "traversal, plain and corrected, and the best-first SQ search)")
    ap.add_argument("--final-refine", type=int, default=0,
                    help="F > 0: every search of the sweeps and timings re-ranks its top "
                         "F pool members in FP64 (--final-refine-mode) before the top k")
    ap.add_argument("--final-refine-mode", choices=["current", "error4"], default="error4")
    ap.add_argument("--refine-factor", type=int, default=2,
                    help="refine leg: refine the top factor*k of the pool, with and "
                         "without the error4 level (0 skips the leg)")
    ap.add_argument("--target", type=float, default=0.95)
    ap.add_argument("--no-bisect", dest="bisect", action="store_false",
                    help="take the smallest L of the --Ls grid reaching the target, "
                         "without the bisection between grid points")
    ap.add_argument("--metrics-csv", default="",
                    help="also write the reference demo's metrics.csv rows "
                         "(retrieval_rpc_demo_pure_pq.c:453-462, 704-728) here")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the default launch configuration instead of the tuned one")
    ap.add_argument("--plain-tuning", default="",
                    help="H,P,G[,W,T[,R]]: visited hash (-1 exact), probe limit, grid "
                         "(waves, or workgroups with W), waves per query, candidate "
                         "tile and reorder step for every plain launch, in place of the "
                         "autotuner and the reorder timing")
    ap.add_argument("--reorder-steps", default="0,16,24,32,48",
                    help="suspension steps timed for each headline launch "
                         "(RoverSearchBatch.reorder_step; 0 = one pass); the fastest is "
                         "kept (empty: no reordering)")
    ap.add_argument("--learned-tuning", default="",
                    help="H,P,G as --plain-tuning, for the learned-stop launches")
    ap.add_argument("--batches", type=int, default=48,
                    help="uneven batches replayed through one captured HIP graph "
                         "(0: skip)")
    ap.add_argument("--train-queries", type=int, default=200_000)
    ap.add_argument("--train-L", type=int, default=512)
    ap.add_argument("--label-k", type=int, default=10)
    ap.add_argument("--label-slacks", default="0,1",
                    help="label slacks to train (label 1 once all but `slack` of the "
                         "final top-K are found); the best at the target is kept")
    ap.add_argument("--stop-every", type=int, default=16)
    ap.add_argument("--stop-threshold", type=float, default=0.9,
                    help="stop probability at which the learned stop ends a search")
    ap.add_argument("--trace-cap", type=int, default=18,
                    help="trace records per training query (steps up to "
                         "trace-cap * stop-every)")
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--save-index", default="",
                    help="write the built index, queries and ground truth (.npz)")
    ap.add_argument("--load-index", default="",
                    help="reuse an index written by --save-index (no rebuild)")
    ap.add_argument("--dump-query-stats", default="",
                    help="write the per-query counters of the timed plain and headline "
                         "searches to this .npz")
    ap.add_argument("--save-model", default="",
                    help="write the chosen learned-stop model (.npz)")
    ap.add_argument("--load-model", default="",
                    help="use a model written by --save-model instead of training")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="also run one plain copy kernel of this many MiB (a "
                         "known-traffic reference for profiler byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: the search runs on a single MI355X "
              "(queries are independent; multi-GPU is data parallel over queries)",
              file=sys.stderr)
        sys.exit(2)
    if args.only_sq:
        print(json.dumps({"sq": sq_leg_run(args)}), flush=True)
        return
    sweep_Ls = [int(v) for v in args.Ls.split(",")]
    if min(sweep_Ls) < args.k:
        ap.error("every beam width in --Ls must be >= k (%d)" % args.k)

    base = None  # kept for the reference's residual encoder (reference_cpu)
    graph_build = None
    if args.load_index:
        z = np.load(args.load_index, allow_pickle=False)
        ix = index.PqIndexData(z["graph"], int(z["entry"]), z["codes"], z["codebook"],
                               z["res_codes"], z["res_qparams"], int(z["res_bits"]),
                               router_centroids=z["router_centroids"],
                               router_ids=z["router_ids"])
        queries, gt = z["queries"], z["gt"]
        e4 = (z["e4_codes"], z["e4_qparams"]) if "e4_codes" in z.files else None
        assert queries.shape[0] == args.nq and ix.res_bits == args.res_bits
        log("index loaded from", args.load_index)
    else:
        import torch
        dev = torch.device("cuda:0")
        log("device", torch.cuda.get_device_name(0))
        base, queries = data.workload(args.n, args.dim, args.components, args.sigma, args.nq)
        log("data", base.shape, queries.shape)
        graph = None
        if args.graph == "reference":
            tg = time.time()
            graph = index.build_graph_exact(base, args.R, 2 * args.R, 42, 1)
            graph_build = dict(build_s=round(time.time() - tg, 2))
            log("reference search-build graph on the device: %.2f s" % (time.time() - tg))
        elif args.graph == "clustered":
            tg = time.time()
            if args.build_clusters <= 0:
                args.build_clusters = max(1, args.n // 1024)
            assign, cent = index.cluster_for_build(base, args.build_clusters)
            tc = time.time() - tg
            st = []
            graph = index.build_graph_exact_clustered(base, args.R, 2 * args.R, 42, 1,
                                                      assign=assign, centroids=cent, stats=st)
            graph_build = dict(clusters=args.build_clusters, kmeans_s=round(tc, 2),
                               build_s=round(time.time() - tg - tc, 2), node_blocks=st[0],
                               cluster_pairs=st[1], block_cluster_pairs=st[2],
                               blocks_with_pairs=st[3])
            log("reference search-build graph, clustered exact candidates: %s" % graph_build)
            del assign, cent
        ix = index.build_pq_index(base, m=args.m, R=args.R, res_bits=args.res_bits,
                                  iters=20, train_max=100000, device=dev,
                                  n_router=args.router, graph=graph,
                                  router_seeding=args.router_seeding)
        log("index built: graph", ix.graph.shape, "medoid", ix.entry, "router",
            ix.router_centroids.shape)
        gt = data.ground_truth(base, queries, k=100, device=dev)
        log("ground truth done")
        e4 = None
        if args.refine_factor > 0 or (args.final_refine > 0 and
                                      args.final_refine_mode == "error4"):
            e4 = index.encode_error4_on_device(base, ix.codes, ix.codebook, ix.res_codes,
                                               ix.res_qparams, ix.res_bits)
            log("error4 codes encoded on the device")
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        if args.save_index:
            np.savez(args.save_index, graph=ix.graph, entry=np.int32(ix.entry),
                     codes=ix.codes, codebook=ix.codebook, res_codes=ix.res_codes,
                     res_qparams=ix.res_qparams, res_bits=np.int32(ix.res_bits),
                     router_centroids=ix.router_centroids, router_ids=ix.router_ids,
                     queries=queries, gt=gt,
                     **({} if e4 is None else dict(e4_codes=e4[0], e4_qparams=e4[1])))
            log("index saved to", args.save_index)

    if args.calibrate_copy_mb:
        nb = args.calibrate_copy_mb << 20
        a, b2 = native.DeviceBuffer(nb), native.DeviceBuffer(nb)
        native.check(native.lib().rover_device_copy(b2.ptr, a.ptr, nb), "copy")
        native.synchronize()
        log("calibration copy: %d MiB read + %d MiB written" % (args.calibrate_copy_mb,
                                                                args.calibrate_copy_mb))
        a.close()
        b2.close()

    if args.entry == "medoid":
        seeds = native.make_seed_ids(len(ix.graph), args.entry_points, ix.entry)
        S = len(seeds)

        def entry_rows(q):
            return np.ascontiguousarray(np.tile(seeds, (len(q), 1)), dtype=np.int32)
    else:
        S = args.route

        def entry_rows(q):
            return index.route_on_host(q, ix.router_centroids, ix.router_ids, S)
    ents = entry_rows(queries)
    log("entry rows chosen on the host: %d queries x %d ids" % ents.shape)
    calc = Calculator()
    calc.load_pq_index(ix, with_router=False)
    log("index loaded into HBM")
    nq, k = args.nq, args.k
    dq = native.DeviceBuffer.from_array(queries)
    d_ent = native.DeviceBuffer.from_array(ents)
    # the entry rows the search launches take (the routed leg swaps them)
    es = dict(ptr=d_ent.ptr, S=S)
    d_ids = native.DeviceBuffer(nq * k * 4)
    d_d = native.DeviceBuffer(nq * k * 4)
    d_cnt = native.DeviceBuffer(nq * 4)
    d_st = native.DeviceBuffer(nq * 32)

    final_rmode = 0
    if args.final_refine > 0:
        final_rmode = native.REFINE_ERROR4 if args.final_refine_mode == "error4" else \
            native.REFINE_CURRENT
        if final_rmode == native.REFINE_ERROR4:
            calc.pure_pq_error4_load(artifacts.error4_image(e4[0], e4[1], args.dim, 8,
                                                            args.res_bits))
        log("final refine: top %d re-ranked in FP64 (%s)" % (args.final_refine,
                                                              args.final_refine_mode))

    def refine_of(mode, L):
        if args.final_refine > 0:
            return min(args.final_refine, L)
        return 0 if mode == "online" else min(args.refine, L)

    # (mode, L, learned, window) -> (hash_log2, probe_limit, grid_waves,
    # waves_per_query, candidate_tile, reorder_step)
    tuned = {}
    # a learned-stop predictor is loaded in the session (the reorder timing
    # needs one to order the resumed searches)
    have_model = {"set": False}
    # source window of every launch below (1: the scalar search; 2: the live
    # pipeline's two turns in flight, the window-2 leg)
    win = {"w": 1}
    # fixed launch configurations (profiler runs): every L of the mode
    fixed = {lrn: (tuple(int(v) for v in spec.split(",")) + (0, 0, 0))[:6]
             for lrn, spec in ((False, args.plain_tuning), (True, args.learned_tuning)) if spec}

    def run(mode, L, stats=False, learned=False, refine=None, refine_mode=0, exact_hash=False):
        h2, pl, gw, wq, tl, rs = tuned.get((mode, L, learned, win["w"]),
                                           fixed.get(learned, (0, 0, 0, 0, 0, 0)))
        if exact_hash:
            # exact visited sets: the counters are the distinct nodes scored
            # and expanded (what the oracle's stats count)
            h2, pl = native.HASH_EXACT, 0
        calc.search_batch_device(dq.ptr, nq, L, k, d_ids.ptr, d_d.ptr, d_cnt.ptr, k,
                                 correction=1 if mode == "online" else 0,
                                 refine_candidates=refine_of(mode, L) if refine is None
                                 else min(refine, L),
                                 refine_mode=final_rmode if refine is None else refine_mode,
                                 expand_count=args.R,
                                 d_stats=d_st.ptr if stats else None,
                                 learned_stop=1 if learned else 0,
                                 stop_threshold=args.stop_threshold, d_entry_ids=es["ptr"],
                                 entry_count=es["S"], entry_stride=es["S"], hash_log2=h2,
                                 probe_limit=pl, grid_waves=gw, source_window=win["w"],
                                 waves_per_query=wq, candidate_tile=tl, reorder_step=rs)

    def autotune(mode, L, learned):
        t = calc.autotune(dq.ptr, nq, L, k, d_ids.ptr, d_d.ptr, d_cnt.ptr, k, repeats=3,
                          correction=1 if mode == "online" else 0,
                          refine_candidates=refine_of(mode, L), expand_count=args.R,
                          d_entry_ids=es["ptr"], entry_count=es["S"], entry_stride=es["S"],
                          learned_stop=1 if learned else 0,
                          stop_threshold=args.stop_threshold, refine_mode=final_rmode,
                          source_window=win["w"])
        tuned[(mode, L, learned, win["w"])] = (t.hash_log2, t.probe_limit, t.grid_waves,
                                              t.waves_per_query, t.candidate_tile, 0)
        info = dict(hash_log2=t.hash_log2, probe_limit=t.probe_limit,
                    grid_waves=t.grid_waves, waves_per_query=t.waves_per_query,
                    candidate_tile=t.candidate_tile, ms=round(t.ms, 4),
                    default_ms=round(t.default_ms, 4), tried=t.tried, rejected=t.rejected)
        log("autotuned %s%s L=%d: %s" % (mode, "+learned" if learned else "", L, info))
        return info

    def choose_reorder(mode, L, learned):
        """Times the tuned launch suspended after each of --reorder-steps
        expansions (0: one pass) and keeps the fastest.  The searches resume
        ordered by the learned-stop predictor's probability (for the plain
        search the predictor only orders them), so the results are the same
        for every step; needs exact visited sets and a predictor."""
        key = (mode, L, learned, win["w"])
        cfg = tuned.get(key)
        steps = [int(v) for v in args.reorder_steps.split(",") if v.strip()]
        if not cfg or cfg[0] != native.HASH_EXACT or cfg[3] or not have_model["set"] or \
                not any(steps):
            return None
        table = {}
        for rs in steps:
            tuned[key] = cfg[:5] + (rs,)
            for _ in range(3):
                run(mode, L, learned=learned)
            native.synchronize()
            e0, e1 = native.Event(), native.Event()
            e0.record()
            for _ in range(5):
                run(mode, L, learned=learned)
            e1.record()
            native.synchronize()
            table[rs] = round(e0.elapsed_ms(e1) / 5, 4)
        best = min(table, key=table.get)
        tuned[key] = cfg[:5] + (best,)
        log("reorder %s%s L=%d: ms by suspension step %s; kept %d" %
            (mode, "+learned" if learned else "", L, table, best))
        return dict(reorder_step=best, ms_by_step={str(k): v for k, v in table.items()})

    def probe(mode, L, learned, rows, bisect=False):
        run(mode, L, stats=True, learned=learned)
        native.synchronize()
        ids = d_ids.download(np.int32, nq * k).reshape(nq, k)
        cnt = d_cnt.download(np.int32, nq)
        st = d_st.download(np.uint64, nq * 4).reshape(nq, 4)
        assert (cnt == k).all(), "search failed: %s" % np.unique(cnt)
        rec = data.recall_at_k(ids, cnt, gt, k)
        e0, e1 = native.Event(), native.Event()
        e0.record()
        run(mode, L, learned=learned)
        e1.record()
        ms = e0.elapsed_ms(e1)
        rows.append(dict(L=L, recall10=round(rec, 4), qps=round(nq / (ms / 1e3), 1),
                         mean_expanded=round(float(st[:, 2].mean()), 2),
                         mean_scored=round(float(st[:, 0].mean()), 2),
                         **({"bisect": True} if bisect else {})))
        log("%s%s L=%d recall@10=%.4f qps=%.0f expanded=%.1f scored=%.1f%s" %
            (mode, "+learned" if learned else "", L, rec, nq / (ms / 1e3),
             st[:, 2].mean(), st[:, 0].mean(), " (bisection)" if bisect else ""))
        return dict(L=L, stats=st.copy(), qps=nq / (ms / 1e3), recall=rec)

    def sweep(mode, learned):
        """The grid of --Ls up to the first L reaching the target, then a
        bisection over the integers between the last grid L below the
        target and that one: the smallest L whose recall@10 reaches it
        (recall grows with L)."""
        rows, chosen, lo = [], None, max(k, 1) - 1
        log("sweep %s%s over L = %s" % (mode, "+learned" if learned else "", args.Ls))
        for L in sweep_Ls:
            r = probe(mode, L, learned, rows)
            if r["recall"] >= args.target:
                chosen = r
                break
            lo = L
        if chosen is not None and args.bisect:
            hi = chosen["L"]
            while hi - lo > 1:
                mid = (lo + hi) // 2
                r = probe(mode, mid, learned, rows, bisect=True)
                if r["recall"] >= args.target:
                    hi, chosen = mid, r
                else:
                    lo = mid
        return rows, chosen

    def timed(mode, L, learned):
        for _ in range(args.warmup):
            run(mode, L, learned=learned)
        native.synchronize()
        ev = [native.Event() for _ in range(args.steps + 1)]
        ev[0].record()
        for i in range(args.steps):
            run(mode, L, learned=learned)
            ev[i + 1].record()
        native.synchronize()
        return sum(ev[i].elapsed_ms(ev[i + 1]) for i in range(args.steps)) / args.steps

    # end-to-end timing through the host-buffer entry point: page-locked
    # query, entry-row and result buffers; each step uploads the queries and
    # entry rows, searches and downloads the results (rover_search_batch
    # returns when the results are in host memory)
    pin_q = native.PinnedArray(queries.shape, np.float32)
    pin_q.array[:] = queries
    pin_e = native.PinnedArray(ents.shape, np.int32)
    pin_e.array[:] = ents
    pin_ids = native.PinnedArray((nq, k), np.int32)
    pin_d = native.PinnedArray((nq, k), np.float32)
    pin_cnt = native.PinnedArray((nq,), np.int32)

    def timed_host(configs):
        """End-to-end steps of each (mode, L, learned) configuration, the
        configurations interleaved step by step so drift on the host or the
        device affects them alike.  Returns per configuration the mean ms
        per step and whether the results equal the device-resident run's."""
        structs = []
        for mode, L, learned in configs:
            h2, pl, gw, wq, tl, rs = tuned.get((mode, L, learned, 1),
                                               fixed.get(learned, (0, 0, 0, 0, 0, 0)))
            structs.append(calc._batch_struct(
                pin_q.ptr, nq, L, k, native.METHOD_PQ, 1 if mode == "online" else 0,
                refine_of(mode, L), 1, args.R, pin_ids.ptr, pin_d.ptr, pin_cnt.ptr, k, None,
                h2, pl, gw, 1 if learned else 0, args.stop_threshold, None, None,
                native.ENTRY_POINT_SEED,
                pin_e.ptr, S, S, 0, final_rmode, 0, 0, wq, tl, reorder_step=rs))
        times = [[] for _ in configs]
        gc.collect()
        gc.disable()  # no collector pauses inside the timed calls
        for step in range(args.warmup + args.steps):
            for i, b in enumerate(structs):
                t0 = time.perf_counter()
                native.check(native.lib().rover_search_batch(calc.h, C.byref(b)),
                             "rover_search_batch")
                if step >= args.warmup:
                    times[i].append(time.perf_counter() - t0)
        gc.enable()
        out = []
        for (mode, L, learned), b, t in zip(configs, structs, times):
            native.check(native.lib().rover_search_batch(calc.h, C.byref(b)),
                         "rover_search_batch")
            host_ids, host_d = pin_ids.array.ravel().copy(), pin_d.array.ravel().copy()
            run(mode, L, learned=learned)
            native.synchronize()
            same = bool(np.array_equal(host_ids, d_ids.download(np.int32, nq * k)) and
                        np.array_equal(host_d.view(np.uint32),
                                       d_d.download(np.uint32, nq * k)))
            log("end to end %s%s L=%d: steps %s ms" % (mode, "+learned" if learned else "", L,
                                                       " ".join("%.3f" % (1e3 * x) for x in t)))
            out.append((1e3 * sum(t) / len(t), same, 1e3 * float(np.median(t))))
        return out

    # ---- plain search: best mode at the target recall
    plain = {}
    for mode in args.modes.split(","):
        rows, chosen = sweep(mode, False)
        plain[mode] = dict(sweep=rows, chosen=chosen)
    reached = [(v["chosen"]["qps"], m) for m, v in plain.items() if v["chosen"]]
    if reached:
        mode = max(reached)[1]
    else:
        mode = args.modes.split(",")[0]
        last = plain[mode]["sweep"][-1]
        run(mode, last["L"], stats=True)
        native.synchronize()
        plain[mode]["chosen"] = dict(L=last["L"], recall=last["recall10"],
                                     qps=last["qps"], stats=d_st.download(
                                         np.uint64, nq * 4).reshape(nq, 4))
        log("no L reached the target recall; using mode %s at its largest L" % mode)
    L_plain = plain[mode]["chosen"]["L"]

    # ---- learned stop on the same mode
    learned = None
    if not args.no_learned:
        if not args.load_model:
            train_q = data.gaussian_mixture(args.train_queries, args.dim, args.components,
                                            args.sigma, 44, centers_seed=42)
            # the mixture's centres depend only on centers_seed, as for the base
            train_ents = entry_rows(train_q)
            log("training: %d queries at L=%d, entry rows chosen"
                % (len(train_q), args.train_L))
        P = native.lib().rover_mlp_param_count()

        def set_model(m):
            native.check(native.lib().rover_learned_stop_set(
                calc.h, m[0].ctypes.data, P, m[1].ctypes.data, m[2].ctypes.data,
                args.stop_every, args.label_k, args.stop_every), "rover_learned_stop_set")
            have_model["set"] = True

        def train_model(slack, tq, tents, tS):
            t = native.RoverLearnedStopTrain()
            t.queries = tq.ctypes.data
            t.num_queries = tq.shape[0]
            t.pool_size = args.train_L
            t.correction = 1 if mode == "online" else 0
            t.label_k = args.label_k
            t.label_slack = slack
            t.entry_points = 1
            t.entry_ids = tents.ctypes.data
            t.entry_count = tS
            t.entry_stride = tS
            t.expand_count = args.R
            t.stop_every = args.stop_every
            t.min_steps = args.stop_every
            t.trace_cap = args.trace_cap
            t.steps = args.train_steps
            t.batch = 8192
            t.lr = 3e-3
            t.seed = 44
            tt = time.time()
            native.check(native.lib().rover_learned_stop_train(calc.h, C.byref(t)),
                         "rover_learned_stop_train")
            training = dict(train_queries=len(tq), train_L=args.train_L,
                            mode=mode, label_k=args.label_k, label_slack=slack,
                            stop_every=args.stop_every, trace_cap=args.trace_cap,
                            samples=int(t.samples), positive_frac=round(t.positive_frac, 4),
                            final_loss=round(t.final_loss, 5),
                            seconds=round(time.time() - tt, 2), threshold=args.stop_threshold)
            log("learned stop trained:", training)
            model = [np.zeros(P, np.float32), np.zeros(16, np.float32),
                     np.zeros(16, np.float32)]
            native.check(native.lib().rover_learned_stop_get(
                calc.h, model[0].ctypes.data, P, model[1].ctypes.data,
                model[2].ctypes.data), "rover_learned_stop_get")
            return model, training

        tried = []
        if args.load_model:
            z = np.load(args.load_model, allow_pickle=False)
            model = [z["params"], z["mean"], z["inv_std"]]
            training = json.loads(str(z["training"]))
            set_model(model)
            log("learned stop model loaded from", args.load_model)
            rows, chosen = sweep(mode, True)
            tried.append(dict(sweep=rows, chosen=chosen, training=training, model=model))
        for slack in ([] if args.load_model else
                      [int(v) for v in args.label_slacks.split(",")]):
            model, training = train_model(slack, train_q, train_ents, S)
            rows, chosen = sweep(mode, True)
            tried.append(dict(sweep=rows, chosen=chosen, training=training, model=model))
        # the slack whose sweep reaches the target fastest (else the best recall)
        hit = [v for v in tried if v["chosen"]]
        if hit:
            learned = max(hit, key=lambda v: v["chosen"]["qps"])
        else:
            learned = max(tried, key=lambda v: v["sweep"][-1]["recall10"])
        main_model = learned["model"]
        set_model(main_model)
        if args.save_model:
            m = learned["model"]
            np.savez(args.save_model, params=m[0], mean=m[1], inv_std=m[2],
                     training=json.dumps(learned["training"]))
        if learned["chosen"] is None:
            last = learned["sweep"][-1]
            run(mode, last["L"], stats=True, learned=True)
            native.synchronize()
            learned["chosen"] = dict(L=last["L"], recall=last["recall10"], qps=last["qps"],
                                     stats=d_st.download(np.uint64, nq * 4).reshape(nq, 4))
        log("learned stop: label slack %d chosen" % learned["training"]["label_slack"])
        others = [dict(label_slack=v["training"]["label_slack"], sweep=v["sweep"])
                  for v in tried if v is not learned]
        learned = dict(sweep=learned["sweep"], chosen=learned["chosen"],
                       training=learned["training"], other_slacks=others)

    # ---- timed region: K steps of the full batch.  The headline is end to
    # end (queries and entry rows uploaded, results downloaded); the
    # device-resident kernel time is measured too and sets the roofline.
    log("timing plain %s at L=%d: %d warmup + %d steps" % (mode, L_plain, args.warmup,
                                                            args.steps))
    tuning = {"plain": autotune(mode, L_plain, False)
              if not (args.no_autotune or False in fixed) else None}
    if tuning["plain"]:
        tuning["plain"]["reorder"] = choose_reorder(mode, L_plain, False)
    ms_plain_dev = timed(mode, L_plain, False)
    if learned:
        L_l = learned["chosen"]["L"]
        tuning["learned"] = autotune(mode, L_l, True) \
            if not (args.no_autotune or True in fixed) else None
        if tuning["learned"]:
            tuning["learned"]["reorder"] = choose_reorder(mode, L_l, True)
        ms_l_dev = timed(mode, L_l, True)
        (ms_plain, same_plain, med_plain), (ms_l, same_l, med_l) = timed_host(
            [(mode, L_plain, False), (mode, L_l, True)])
    else:
        ((ms_plain, same_plain, med_plain),) = timed_host([(mode, L_plain, False)])
    qps_plain = nq / (ms_plain / 1e3)
    log("timed plain %s: L=%d end to end %.3f ms/step %.0f QPS; device-resident %.3f ms"
        % (mode, L_plain, ms_plain, qps_plain, ms_plain_dev))
    device_resident = {"plain": dict(L=L_plain, ms_per_step=round(ms_plain_dev, 4),
                                     qps=round(nq / (ms_plain_dev / 1e3), 1))}
    median_ms = {"plain": round(med_plain, 4)}
    host_same = {"plain": same_plain}
    if learned:
        host_same["learned"] = same_l
        median_ms["learned"] = round(med_l, 4)
        qps_l = nq / (ms_l / 1e3)
        device_resident["learned"] = dict(L=L_l, ms_per_step=round(ms_l_dev, 4),
                                          qps=round(nq / (ms_l_dev / 1e3), 1),
                                          speedup_vs_plain=round(ms_plain_dev / ms_l_dev, 3))
        log("timed learned %s: L=%d end to end %.3f ms/step %.0f QPS (%.2fx plain); "
            "device-resident %.3f ms" % (mode, L_l, ms_l, qps_l, qps_l / qps_plain, ms_l_dev))
        ms_per_step, qps, L, label, kernel_ms = ms_l, qps_l, L_l, "learned-stop", ms_l_dev
    else:
        ms_per_step, qps, L, label, kernel_ms = ms_plain, qps_plain, L_plain, "plain", ms_plain_dev
    host_qps = qps_plain
    # counters for the algorithmic bytes: the timed launch's own counters
    # include nodes re-scored after the tuned (small) visited hash dropped
    # them; a near-exact hash gives the distinct counts (same ids)
    run(mode, L, stats=True, learned=label == "learned-stop", exact_hash=True)
    native.synchronize()
    stats = d_st.download(np.uint64, nq * 4).reshape(nq, 4)
    run(mode, L_plain, stats=True, exact_hash=True)
    native.synchronize()
    plain_exact_stats = d_st.download(np.uint64, nq * 4).reshape(nq, 4)
    if args.dump_query_stats:
        # per-query counters (scored, corrected, expanded, dedup) of the two
        # timed searches, for scheduling studies
        np.savez(args.dump_query_stats, plain=plain_exact_stats, headline=stats,
                 plain_L=L_plain, headline_L=L, headline_search=label)
        log("per-query counters written to", args.dump_query_stats)

    # ---- window-2 leg: the live pipeline with two source turns in flight
    # (pure_pq_source_pipeline_step, SQMD_MAX_INFLIGHT_SOURCES=2): each
    # source is chosen before the previous turn's candidates are merged, so
    # the merge overlaps the next neighbor-row load.  Its results are the
    # reference pipeline's, not the scalar search's; plain and learned stop
    # (the same model) are swept to the target and timed device-resident.
    window2 = None
    if args.window2_leg:
        win["w"] = 2
        window2 = {"what": "source window 2 (the live pipeline's two turns in flight), "
                           "the same seeds and model; device-resident launches"}
        w_rows, wch = sweep(mode, False)
        if wch:
            if not args.no_autotune:
                window2["plain_tuning"] = autotune(mode, wch["L"], False)
            ms_wp = timed(mode, wch["L"], False)
            window2["plain"] = dict(L=wch["L"], recall10=round(wch["recall"], 4),
                                    qps=round(nq / (ms_wp / 1e3), 1),
                                    ms_per_step=round(ms_wp, 4),
                                    mean_expanded=round(float(wch["stats"][:, 2].mean()), 2),
                                    vs_window1_plain=round(ms_plain_dev / ms_wp, 3))
            log("window 2, timed plain: %s" % window2["plain"])
            window2["plain_sweep"] = w_rows
            if learned:
                l_rows, wlch = sweep(mode, True)
                if wlch:
                    if not args.no_autotune:
                        window2["learned_tuning"] = autotune(mode, wlch["L"], True)
                    ms_wl = timed(mode, wlch["L"], True)
                    window2["learned"] = dict(
                        L=wlch["L"], recall10=round(wlch["recall"], 4),
                        qps=round(nq / (ms_wl / 1e3), 1), ms_per_step=round(ms_wl, 4),
                        mean_expanded=round(float(wlch["stats"][:, 2].mean()), 2),
                        speedup_vs_window2_plain=round(ms_wp / ms_wl, 3),
                        speedup_vs_window1_plain=round(ms_plain_dev / ms_wl, 3))
                    log("window 2, timed learned: %s" % window2["learned"])
                window2["learned_sweep"] = l_rows
            if not args.no_cpu:
                # the plain window-2 results against the oracle's restatement
                # of the pipeline (orc_search_window) on a sample of queries
                sys.path.insert(0, os.path.join(ROOT, "oracle"))
                import pyoracle
                ns = min(nq, 2000)
                run(mode, wch["L"], stats=True, exact_hash=True)
                native.synchronize()
                gids = d_ids.download(np.int32, nq * k).reshape(nq, k)[:ns]
                gd = d_d.download(np.float32, nq * k).reshape(nq, k)[:ns]
                gst = d_st.download(np.uint64, nq * 4).reshape(nq, 4)[:ns]
                oix_w = pyoracle.Index.from_pq(ix)
                sc_w = pyoracle.SCORE_ONLINE if mode == "online" else pyoracle.SCORE_ADC
                oi, od, oc, ost = pyoracle.search_batch(oix_w, sc_w, queries[:ns], ents[:ns],
                                                        args.R, wch["L"], k,
                                                        refine=refine_of(mode, wch["L"]),
                                                        threads=16, window=2)
                window2["oracle_identical_frac"] = float(
                    ((gids == oi).all(axis=1) &
                     (gd.view(np.uint32) == od.view(np.uint32)).all(axis=1)).mean())
                window2["oracle_stats_identical_frac"] = float(
                    (gst.astype(np.int64) == ost.astype(np.int64)).all(axis=1).mean())
                window2["oracle_sample"] = ns
                log("window 2: GPU equals the oracle's pipeline on %.4f of %d queries "
                    "(counters %.4f)" % (window2["oracle_identical_frac"], ns,
                                         window2["oracle_stats_identical_frac"]))
        win["w"] = 1

    # ---- routed leg: the same plain / learned comparison with per-query entry
    # rows from the k-means router (representatives of the centroids nearest
    # the query) in place of the shared seed list
    routed = None
    if not args.no_routed_leg:
        rS = args.route
        rents = index.route_on_host(queries, ix.router_centroids, ix.router_ids, rS)
        d_rent = native.DeviceBuffer.from_array(rents)
        saved_tuned = dict(tuned)
        tuned.clear()
        es.update(ptr=d_rent.ptr, S=rS)
        routed = {"entries": "per query: representatives of its %d nearest of %d k-means "
                             "centroids" % (rS, len(ix.router_ids))}
        log("routed leg: %s" % routed["entries"])
        _, pch = sweep(mode, False)
        if pch:
            if not args.no_autotune:
                autotune(mode, pch["L"], False)
            ms_rp = timed(mode, pch["L"], False)
            routed["plain"] = dict(L=pch["L"], recall10=round(pch["recall"], 4),
                                   qps=round(nq / (ms_rp / 1e3), 1),
                                   ms_per_step=round(ms_rp, 4))
            log("routed leg, timed plain: %s" % routed["plain"])
            if learned and not args.load_model:
                tr_ents = index.route_on_host(train_q, ix.router_centroids, ix.router_ids, rS)
                rmodel, rtraining = train_model(0, train_q, tr_ents, rS)
                set_model(rmodel)
                _, lch = sweep(mode, True)
                if lch:
                    if not args.no_autotune:
                        autotune(mode, lch["L"], True)
                    ms_rl = timed(mode, lch["L"], True)
                    routed["learned"] = dict(
                        L=lch["L"], recall10=round(lch["recall"], 4),
                        qps=round(nq / (ms_rl / 1e3), 1), ms_per_step=round(ms_rl, 4),
                        speedup_vs_plain=round(ms_rp / ms_rl, 3), label_slack=0)
                    log("routed leg, timed learned: %s" % routed["learned"])
                set_model(main_model)
        es.update(ptr=d_ent.ptr, S=S)
        tuned.clear()
        tuned.update(saved_tuned)
        d_rent.close()

    # ---- uneven beam widths: every query once with its own L, plain and
    # learned stop, in one launch (per-query pool sizes in HBM)
    uneven = None
    if args.uneven_L:
        lo_L, hi_L = (int(v) for v in args.uneven_L.split(","))
        lo_L = max(lo_L, k)
        Lq = np.random.default_rng(4096).integers(lo_L, hi_L + 1, size=nq).astype(np.int32)
        d_L = native.DeviceBuffer.from_array(Lq)
        uneven = {"L": "per query uniform in [%d, %d], one launch" % (lo_L, hi_L),
                  "mean_L": round(float(Lq.mean()), 1)}

        def go(lrn):
            calc.search_batch_device(dq.ptr, nq, hi_L, k, d_ids.ptr, d_d.ptr, d_cnt.ptr, k,
                                     correction=1 if mode == "online" else 0,
                                     refine_candidates=refine_of(mode, lo_L),
                                     refine_mode=final_rmode,
                                     expand_count=args.R, learned_stop=1 if lrn else 0,
                                     stop_threshold=args.stop_threshold, d_pool_sizes=d_L.ptr,
                                     d_entry_ids=es["ptr"], entry_count=es["S"],
                                     entry_stride=es["S"])

        for name, lrn in (("plain", False), ("learned", True)) if learned else \
                (("plain", False),):
            go(lrn)
            native.synchronize()
            cnt = d_cnt.download(np.int32, nq)
            assert (cnt == k).all(), "uneven-L search failed: %s" % np.unique(cnt)
            rec = data.recall_at_k(d_ids.download(np.int32, nq * k).reshape(nq, k), cnt, gt, k)
            for _ in range(args.warmup):
                go(lrn)
            e0, e1 = native.Event(), native.Event()
            e0.record()
            for _ in range(args.steps):
                go(lrn)
            e1.record()
            native.synchronize()
            ms = e0.elapsed_ms(e1) / args.steps
            uneven[name] = dict(recall10=round(rec, 4), qps=round(nq / (ms / 1e3), 1),
                                ms_per_step=round(ms, 4))
            log("uneven L %s: %s" % (name, uneven[name]))
        if learned:
            uneven["learned_speedup_vs_plain"] = round(
                uneven["plain"]["ms_per_step"] / uneven["learned"]["ms_per_step"], 3)
        d_L.close()

    # ---- refine leg: the top refine_factor*k pool members re-ranked in FP64
    # from PQ + residual (CURRENT) and with the error4 level (ERROR4)
    refine_leg = None
    if args.refine_factor > 0 and e4 is not None:
        calc.pure_pq_error4_load(artifacts.error4_image(e4[0], e4[1], args.dim, 8,
                                                        args.res_bits))
        F = args.refine_factor * k
        refine_leg = {"refine_candidates": F, "mode": mode}
        for name, rmode in (("current", native.REFINE_CURRENT),
                            ("error4", native.REFINE_ERROR4)):
            rows, hit = [], None
            for Lr in sweep_Ls:
                if Lr < F:
                    continue
                run(mode, Lr, refine=F, refine_mode=rmode)
                native.synchronize()
                ids = d_ids.download(np.int32, nq * k).reshape(nq, k)
                cnt = d_cnt.download(np.int32, nq)
                assert (cnt == k).all(), "refine search failed: %s" % np.unique(cnt)
                rec = data.recall_at_k(ids, cnt, gt, k)
                for _ in range(args.warmup):
                    run(mode, Lr, refine=F, refine_mode=rmode)
                e0, e1 = native.Event(), native.Event()
                e0.record()
                for _ in range(args.steps):
                    run(mode, Lr, refine=F, refine_mode=rmode)
                e1.record()
                ms = e0.elapsed_ms(e1) / args.steps
                rows.append(dict(L=Lr, recall10=round(rec, 4), qps=round(nq / (ms / 1e3), 1)))
                log("refine %s F=%d L=%d recall@10=%.4f qps=%.0f" % (name, F, Lr, rec,
                                                                    nq / (ms / 1e3)))
                if rec >= args.target:
                    hit = rows[-1]
                    break
            refine_leg[name] = dict(sweep=rows, at_target=hit)

    sq_leg = None
    if not args.no_sq:
        sq_leg = sq_leg_run(args)

    batching = None
    if args.batches > 0:
        batching = batching_leg(calc, queries, ents, S, mode, refine_of, args,
                                refine_mode=final_rmode)

    # restore the main batch's results for the checks below
    run(mode, L_plain)
    native.synchronize()

    code_bytes = index.code_bytes(args.m, 256)
    res_bytes = index.residual_code_bytes(args.dim, args.res_bits)
    algo = algorithmic_bytes(stats, args.R, code_bytes, res_bytes, args.dim, k,
                             refine_of(mode, L), ix.codebook.nbytes, mode == "online")
    achieved = algo / (kernel_ms / 1e3) / 1e9
    # online scoring on 32 x 4-dim PQ with <= 32 neighbors runs the two-lane
    # instantiation (scorer 3); ADC ranking is scorer 0
    scorer = (3 if args.m * 4 == args.dim and args.m == 32 and args.R <= 32 else 1) \
        if mode == "online" else 0
    tl_ = tuned.get((mode, L, label == "learned-stop", 1),
                    fixed.get(label == "learned-stop", (0, 0, 0, 0, 0, 0)))
    if tl_[3]:
        kernel = "rv_search_mw_kernel<%d, %d, %d>" % (args.res_bits,
                                                     1 if label == "learned-stop" else 0,
                                                     tl_[3])
    elif tl_[5]:
        # suspended and resumed: two dispatches per step (the searches up to
        # the suspension step, then the resumed ones) and the ordering kernel
        kernel = "rv_search_kernel<%d, 3, 1>" % scorer
    else:
        kernel = "rv_search_kernel<%d, %d, 1>" % (scorer, 1 if label == "learned-stop" else 0)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": None, "kernel": kernel,
                "algorithmic_bytes_per_launch": int(algo),
                "kernel_ms": round(kernel_ms, 4),
                "counters": "distinct scored / expanded nodes (exact visited sets); "
                            "the plain search's equal the oracle's PurePqSearchStats"}
    if tl_[5] and not tl_[3]:
        roofline["launch"] = ("per step: %s twice (searches up to %d expansions, then the "
                              "suspended ones resumed) and rv_susp_order_kernel; kernel_ms "
                              "and traffic are per step" % (kernel, tl_[5]))
    # HBM bytes per launch come from the profiler's FETCH_SIZE / WRITE_SIZE
    # passes over this same workload (calibrated against a copy kernel),
    # recorded next to the profile summary; used when kernel and L match
    rec_path = os.path.join(ROOT, "profiles", "rv_search_kernel_online_1M_L%d.json" % L)
    if os.path.exists(rec_path) and args.n == 1_000_000 and args.nq == 10_000:
        rec = json.load(open(rec_path))
        per = rec.get("traffic_bytes_per_launch", {}).get(kernel)
        if per and rec.get("L") == L:
            roofline["traffic"] = int(per)
            roofline["traffic_source"] = os.path.relpath(rec_path, ROOT)

    cpu = None
    reference_cpu = None
    fingerprint = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        oix = pyoracle.Index.from_pq(ix)
        scorer = pyoracle.SCORE_ONLINE if mode == "online" else pyoracle.SCORE_ADC
        # every host thread this process may run on, up to the 16 a
        # one-GPU share of the host provides
        avail = (len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
                 else (os.cpu_count() or 1))
        threads = max(1, min(16, avail))
        host = dict(cpu_model=cpu_model(), threads_used=threads, threads_available=avail,
                    cpus_visible=os.cpu_count())
        log("cpu baseline: the oracle on %d host threads (%s)" % (threads, host["cpu_model"]))
        # whole passes over the query set until about cpu_seconds of work
        passes, searched = 0, 0
        t1 = time.time()
        while True:
            oids, od, ocnt, ost = pyoracle.search_batch(oix, scorer, queries, ents, args.R,
                                                        L_plain, k,
                                                        refine=refine_of(mode, L_plain),
                                                        threads=threads)
            passes += 1
            searched += nq
            el = time.time() - t1
            if el >= args.cpu_seconds or passes >= 200:
                break
            if passes % 10 == 0:
                log("cpu baseline: pass %d, %.1f s" % (passes, el))
        run(mode, L_plain)
        native.synchronize()
        gids = d_ids.download(np.int32, nq * k).reshape(nq, k)
        gd = d_d.download(np.float32, nq * k).reshape(nq, k)
        same = float(((gids == oids).all(axis=1) &
                      (gd.view(np.uint32) == od.view(np.uint32)).all(axis=1)).mean())
        # the plain search's distinct scored / corrected / expanded / dedup
        # counters equal the oracle's PurePqSearchStats
        stats_same = float((plain_exact_stats.astype(np.int64) ==
                            ost.astype(np.int64)).all(axis=1).mean())
        cpu = {"value": round(searched / el, 1), "unit": "queries/s", "cores": threads,
               "kind": "port",
               "sample": "%d passes over the %d queries (%d searches), plain %s search "
                         "at L=%d, one query per host thread (%.1f s)"
                         % (passes, nq, searched, mode, L_plain, el),
               "host": host,
               "gpu_results_identical_frac": same,
               "gpu_stats_identical_frac": stats_same}
        log("cpu baseline: %.0f QPS on %d threads; GPU ids+distances identical on %.4f, "
            "counters on %.4f" % (searched / el, threads, same, stats_same))

        # the reference demo's result fingerprint (rpc_demo_pure_pq.c:387-400,
        # 556, 614): FNV-1a 64 over every query's whole sorted pool, plain
        # search, on the GPU (the library's fingerprint) and in the oracle
        fids = native.DeviceBuffer(nq * L_plain * 4)
        fd = native.DeviceBuffer(nq * L_plain * 4)
        fcnt = native.DeviceBuffer(nq * 4)
        calc.search_batch_device(dq.ptr, nq, L_plain, L_plain, fids.ptr, fd.ptr, fcnt.ptr,
                                 L_plain, correction=1 if mode == "online" else 0,
                                 expand_count=args.R, d_entry_ids=d_ent.ptr, entry_count=S,
                                 entry_stride=S)
        native.synchronize()
        g_fp = native.results_fingerprint(fids.download(np.int32, nq * L_plain).reshape(nq, -1),
                                          fd.download(np.float32, nq * L_plain).reshape(nq, -1),
                                          fcnt.download(np.int32, nq))
        fi, fdd, fc, _ = pyoracle.search_batch(oix, scorer, queries, ents, args.R, L_plain,
                                               L_plain, threads=threads)
        o_fp = pyoracle.fingerprint(fi, fdd, fc)
        fingerprint = {"search_pool_fingerprint": "%016x" % g_fp,
                       "oracle": "%016x" % o_fp, "equal": g_fp == o_fp,
                       "what": "plain %s search, L=%d, %d queries, whole sorted pools"
                               % (mode, L_plain, nq)}
        log("result fingerprint: GPU %016x oracle %016x" % (g_fp, o_fp))
        fids.close()
        fd.close()
        fcnt.close()

        # the reference's own C search (retrieval_pure_pq_search_scalar_with_
        # policy with its residual evaluator, compiled from its sources) on
        # the same host threads, one residual reader per thread, over a
        # bounded sample of the same queries
        if pyoracle.ref_available() and base is not None and mode == "online":
            reference_cpu = reference_leg(pyoracle, ix, base, queries, ents, gids, gd,
                                          L_plain, k, args, threads)
            if reference_cpu:
                reference_cpu["host"] = host
                log("reference cpu: %s" % reference_cpu)

    if args.metrics_csv:
        # one row per timed setting, as the reference demo writes them
        lines = [metrics.header()]
        settings = [("plain", L_plain, False, ms_plain)]
        if label == "learned-stop":
            settings.append(("learned-stop", L, True, ms_per_step))
        for name, Lx, lrn, ms in settings:
            kk = min(100, Lx)
            mids = native.DeviceBuffer(nq * Lx * 4)
            md = native.DeviceBuffer(nq * Lx * 4)
            mcnt = native.DeviceBuffer(nq * 4)
            mst = native.DeviceBuffer(nq * 32)
            calc.search_batch_device(dq.ptr, nq, Lx, Lx, mids.ptr, md.ptr, mcnt.ptr, Lx,
                                     correction=1 if mode == "online" else 0,
                                     expand_count=args.R, d_stats=mst.ptr,
                                     learned_stop=1 if lrn else 0,
                                     stop_threshold=args.stop_threshold,
                                     d_entry_ids=d_ent.ptr, entry_count=S, entry_stride=S,
                                     hash_log2=native.HASH_EXACT)
            native.synchronize()
            mi = mids.download(np.int32, nq * Lx).reshape(nq, Lx)
            mdd = md.download(np.float32, nq * Lx).reshape(nq, Lx)
            mc = mcnt.download(np.int32, nq)
            ms_ = mst.download(np.uint64, nq * 4).reshape(nq, 4)
            total_us = int(round(ms * 1e3))
            lines.append(metrics.row(
                dataset="mixture-%dx%d" % (len(ix.graph), args.dim), layout="original",
                code_bits=8, ks=256, m=args.m, code_bytes=index.code_bytes(args.m, 256),
                compression_x=args.dim * 4 // index.code_bytes(args.m, 256),
                entry_points=S, pool_size=Lx, queries=nq,
                recall_at_1=metrics.recall_percent(mi, mc, gt, 1),
                recall_at_10=metrics.recall_percent(mi, mc, gt, 10),
                recall_at_100=metrics.recall_percent(mi, mc, gt, kk),
                wall_us=total_us // nq, work_items=int(ms_[:, 0].sum()) // nq, dsp_adc_us=0,
                dsp_wait_polls=0, correction_mode="online" if mode == "online" else "off",
                residual_bits=args.res_bits if mode == "online" else 0,
                corrected_candidates=int(ms_[:, 1].sum()), residual_cache_hits=0,
                residual_cache_misses=0, residual_cache_errors=0,
                artifact_prefix=name, window=1, source_jobs=int(ms_[:, 2].sum()),
                source_retired=int(ms_[:, 2].sum()), source_high_water=1, top_k=k,
                refine_mode=0, refine_factor=0, refine_budget=0, refine_candidates=0,
                refine_us=0, refine_blocks=0, refine_bytes=0, refine_hits=0,
                search_pool_fingerprint=native.results_fingerprint(mi, mdd, mc),
                total_wall_us=total_us))
            for b_ in (mids, md, mcnt, mst):
                b_.close()
        with open(args.metrics_csv, "w") as f:
            f.write("".join(lines))
        log("metrics.csv rows written to", args.metrics_csv)

    def brief(v):
        if not v:
            return None
        c = v["chosen"]
        return {"L": c["L"] if c else None,
                "recall10": round(float(c["recall"]), 4) if c else None,
                "reached_target": c is not None, "sweep": v["sweep"]}

    out = {
        "metric": "QPS @ recall@10=0.95, 1M×128 PQ, learned-stop vs plain, 1×MI355X",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "%s x %d Gaussian mixture (%d comps, sigma %g); PQ %dx256 + "
                        "%d-bit residuals; R=%d graph (%s); %d queries, k=%d"
                        % ("%gM" % (len(ix.graph) / 1e6), args.dim, args.components,
                           args.sigma, args.m, args.res_bits, args.R,
                           "exact %d-NN, RNG prune, backlinks" % (2 * args.R)
                           if args.graph == "backlinks" else
                           "the reference's exact search-build: %d-NN%s, random initial "
                           "rows, RNG prune with fill"
                           % (2 * args.R, " over %d k-means clusters, certified"
                              % args.build_clusters if args.graph == "clustered" else ""),
                           nq, k),
            "graph_build": graph_build,
            "search": label, "mode": mode, "L": L,
            "refine_candidates": refine_of(mode, L),
            "entries": ("per query: representatives of its %d nearest of %d k-means "
                        "centroids, passed as entry_ids rows" % (S, len(ix.router_ids))
                        if args.entry == "routed" else
                        "the reference's seeds: medoid + %d LCG ids" % (S - 1)),
            "timing": "end to end per step: page-locked host queries and entry rows "
                      "uploaded, search, results downloaded (rover_search_batch); "
                      "device_resident: the same search with everything in HBM, "
                      "HIP events on the launch stream (the roofline's time)",
            "plain_qps": round(qps_plain, 1), "plain_L": L_plain,
            "plain_ms_per_step": round(ms_plain, 4),
            "learned_speedup_vs_plain": round(qps / qps_plain, 3) if learned else None,
            "device_resident": device_resident,
            "host_results_identical": host_same,
            "end_to_end_median_ms": median_ms,
            "result_fingerprint": fingerprint,
            "plain_modes": {m: brief(v) for m, v in plain.items()},
            "learned": None if not learned else dict(
                brief(learned), training=learned["training"],
                other_slacks=learned["other_slacks"]),
            "refine": refine_leg,
            "sq": sq_leg,
            "reference_cpu": reference_cpu,
            "routed": routed,
            "window2": window2,
            "uneven_L": uneven,
            "launch_tuning": tuning,
            "uneven_batches": batching,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    calc.close()


if __name__ == "__main__":
    main()
