"""Index building: PQ codebooks and codes, 4/8-bit residuals, global SQ
codes, and the degree-R RNG-pruned graph with a medoid entry point.

Encoders are bit-identical to the reference's (checked against the compiled
reference by the tests):
  * PQ encode: argmin over k of the sequential f32 sub-L2, first minimum
    wins (retrieval_pq_core.c:226-268);
  * residual: per-vector min/max of (x - centroid), scale=(max-min)/levels,
    code=clamp(roundf((r-min)/scale)) (retrieval_pq_residual.c:314-385);
  * SQ: one global scale/zero over all values
    (retrieval_rpc_demo_support.c:573-700).
Codebook training follows the reference trainer's shape (uniform-stride
initial centroids, Lloyd iterations; retrieval_pq_core.c:96-191) but its
floating-point summation order is not the reference's, so codebooks are
not bit-identical; search parity is unaffected because the GPU and the
oracle read the same artifact.

Graph construction follows retrieval_search_build_core.c: exact 2R nearest
neighbours sorted by (distance, id) (:314-361), RNG occlusion pruning
`d(kept, c) < d(u, c)` (:904-930), fill from the remaining candidates in
order (:932-947), rows sorted by (distance, id).
"""
import ctypes as C

import numpy as np


# ------------------------------------------------------------------ PQ

def _sub_l2_f32(x, c):
    """Sequential f32 sum over the last axis of (x - c)^2, x [..., s],
    c [..., s] broadcastable (retrieval_pq_core.c:16-25)."""
    diff = (x - c).astype(np.float32)
    acc = np.zeros(diff.shape[:-1], dtype=np.float32)
    for d in range(diff.shape[-1]):
        acc = (acc + diff[..., d] * diff[..., d]).astype(np.float32)
    return acc


def uniform_stride_sample(n, count):
    """Row indices k*(n-1)/(count-1) (the reference's uniform-stride
    strategy, retrieval_pq_core.c:140-149 applied to training rows)."""
    if count >= n:
        return np.arange(n)
    return (np.arange(count, dtype=np.int64) * (n - 1)) // max(count - 1, 1)


def _segment_means(torch, pts, assign, k, cent):
    """Lloyd's centroid update on the device without floating-point atomics
    (so a build is reproducible run to run): points stably sorted by their
    cluster, float64 prefix sums differenced at the cluster boundaries;
    empty clusters keep their centroid."""
    order = torch.argsort(assign, stable=True)
    cs = torch.cumsum(pts[order].double(), 0)
    cs = torch.cat([torch.zeros_like(cs[:1]), cs])
    counts = torch.bincount(assign, minlength=k)
    ends = torch.cumsum(counts, 0)
    sums = cs[ends] - cs[ends - counts]
    nz = counts > 0
    cent[nz] = (sums[nz] / counts[nz, None].double()).to(cent.dtype)


def train_pq(x, m, ks=256, iters=20, train_max=100000, device=None):
    """k-means codebook float32 [m, ks, dim/m]."""
    n, dim = x.shape
    sub = dim // m
    tr = x[uniform_stride_sample(n, min(n, train_max))]
    cnt = tr.shape[0]
    init = (np.arange(ks, dtype=np.int64) * (cnt - 1)) // max(ks - 1, 1)
    if device is not None:
        import torch
        t = torch.from_numpy(np.ascontiguousarray(tr)).to(device)
        out = np.empty((m, ks, sub), dtype=np.float32)
        for s in range(m):
            xs = t[:, s * sub:(s + 1) * sub].contiguous()
            cent = xs[torch.from_numpy(init).to(device)].clone()
            for _ in range(iters):
                d = (xs * xs).sum(1, keepdim=True) - 2.0 * xs @ cent.T + \
                    (cent * cent).sum(1)[None, :]
                a = torch.argmin(d, dim=1)
                _segment_means(torch, xs, a, ks, cent)
            out[s] = cent.cpu().numpy()
        return out
    out = np.empty((m, ks, sub), dtype=np.float32)
    for s in range(m):
        xs = np.ascontiguousarray(tr[:, s * sub:(s + 1) * sub])
        cent = xs[init].copy()
        for _ in range(iters):
            d = (xs * xs).sum(1)[:, None] - 2.0 * xs @ cent.T + \
                (cent * cent).sum(1)[None, :]
            a = np.argmin(d, axis=1)
            sums = np.zeros_like(cent, dtype=np.float64)
            np.add.at(sums, a, xs)
            counts = np.bincount(a, minlength=ks)
            nz = counts > 0
            cent[nz] = (sums[nz] / counts[nz, None]).astype(np.float32)
        out[s] = cent
    return out


def code_bytes(m, ks):
    return m // 2 + m % 2 if ks <= 16 else m


def encode_pq(x, codebook, chunk=4096):
    """uint8 [n, code_bytes]; 4-bit codes put even subspaces in the low
    nibble (retrieval_pq_core.c:54-85)."""
    m, ks, sub = codebook.shape
    n = x.shape[0]
    idx = np.empty((n, m), dtype=np.uint8)
    for s0 in range(0, n, chunk):
        xs = x[s0:s0 + chunk]
        for s in range(m):
            d = _sub_l2_f32(xs[:, None, s * sub:(s + 1) * sub], codebook[s][None])
            idx[s0:s0 + chunk, s] = np.argmin(d, axis=1)
    if ks > 16:
        return idx
    out = np.zeros((n, code_bytes(m, ks)), dtype=np.uint8)
    for s in range(m):
        if s & 1:
            out[:, s >> 1] |= (idx[:, s] << 4).astype(np.uint8)
        else:
            out[:, s >> 1] |= idx[:, s]
    return out


def pq_indices(codes, m, ks):
    if ks > 16:
        return codes[:, :m].astype(np.int64)
    lo = codes & 0x0F
    hi = codes >> 4
    out = np.empty((codes.shape[0], m), dtype=np.int64)
    out[:, 0::2] = lo[:, : (m + 1) // 2]
    out[:, 1::2] = hi[:, : m // 2]
    return out


def reconstruct_pq(codes, codebook):
    m, ks, sub = codebook.shape
    idx = pq_indices(codes, m, ks)
    parts = [codebook[s][idx[:, s]] for s in range(m)]
    return np.concatenate(parts, axis=1).astype(np.float32)


def _roundf_nonneg_or_clamp(v, levels):
    """clamp(roundf(v), 0, levels) for float32 v; roundf rounds half away
    from zero, and anything below zero clamps to 0."""
    r = np.floor(v.astype(np.float64) + 0.5)
    return np.clip(r, 0, levels).astype(np.int64)


def residual_code_bytes(dim, bits):
    return dim // 2 + dim % 2 if bits == 4 else dim


def encode_residual(x, codes, codebook, bits=4):
    """(codes uint8 [n, res_code_bytes], qparams float32 [n, 2])."""
    levels = 15 if bits == 4 else 255
    recon = reconstruct_pq(codes, codebook)
    r = (x.astype(np.float32) - recon).astype(np.float32)
    lo = r.min(axis=1)
    hi = r.max(axis=1)
    scale = ((hi - lo) / np.float32(levels)).astype(np.float32)
    scale = np.where(scale < np.float32(1e-12), np.float32(1e-12), scale).astype(np.float32)
    inv = (np.float32(1.0) / scale).astype(np.float32)
    v = ((r - lo[:, None]).astype(np.float32) * inv[:, None]).astype(np.float32)
    q = _roundf_nonneg_or_clamp(v, levels).astype(np.uint8)
    n, dim = x.shape
    if bits == 8:
        out = q
    else:
        out = np.zeros((n, residual_code_bytes(dim, 4)), dtype=np.uint8)
        out[:, : (dim + 1) // 2] |= q[:, 0::2]
        out[:, : dim // 2] |= (q[:, 1::2] << 4).astype(np.uint8)
    qp = np.stack([scale, lo], axis=1).astype(np.float32)
    return np.ascontiguousarray(out), np.ascontiguousarray(qp)


def sq_params(x, bits):
    """Global (scale, zero) over all values (demo_set_qparams_from_range)."""
    lo = np.float32(x.min())
    hi = np.float32(x.max())
    rng = np.float32(hi - lo)
    if rng < np.float32(1e-12):
        rng = np.float32(1e-12)
    return np.float32(rng / np.float32((1 << bits) - 1)), lo


def encode_sq(x, bits, scale, zero):
    levels = (1 << bits) - 1
    inv = np.float32(np.float32(1.0) / np.float32(scale))
    v = ((x.astype(np.float32) - np.float32(zero)).astype(np.float32) * inv).astype(np.float32)
    q = _roundf_nonneg_or_clamp(v, levels).astype(np.uint8)
    if bits == 8:
        return np.ascontiguousarray(q)
    n, dim = x.shape
    out = np.zeros((n, (dim + 1) // 2), dtype=np.uint8)
    out[:, : (dim + 1) // 2] |= q[:, 0::2]
    out[:, : dim // 2] |= (q[:, 1::2] << 4).astype(np.uint8)
    return out


def decode_sq_residual(l1_codes, l1_bits, scale, zero, res_codes, res_qparams, res_bits, dim):
    """The SQ index's reconstruction, float32 [n][dim]: per dimension
    ((l1 * scale) + zero) + ((res * rscale) + rzero), each step rounded in
    float32 (retrieval_final_refine_decode_sq, retrieval_final_refine.c:
    46-67)."""
    a = decode_sq(l1_codes, l1_bits, scale, zero, dim)
    if res_bits == 8:
        r = res_codes[:, :dim].astype(np.float32)
    else:
        r = np.zeros((res_codes.shape[0], dim), dtype=np.float32)
        r[:, 0::2] = res_codes[:, :(dim + 1) // 2] & 15
        r[:, 1::2] = res_codes[:, :dim // 2] >> 4
    qp = np.asarray(res_qparams, dtype=np.float32)
    level = ((r * qp[:, :1]).astype(np.float32) + qp[:, 1:2]).astype(np.float32)
    return (a + level).astype(np.float32)


def encode_error4(x, recon):
    """Third-level (error4) codes of the error x - recon, as
    retrieval_build_error4 forms them (retrieval_build_artifacts.c:25-66),
    every step a float32 rounding: e = x - recon; lo, hi = min, max of e over
    the row; scale = max(hi - lo, 1e-12) / 15; inv = 1 / scale;
    code = clamp(floor((e - lo) * inv + 0.5), 0, 15).  Returns codes
    [n][(dim+1)/2] (low nibble = even dimension) and (scale, lo) [n][2]."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    e = (x - np.asarray(recon, dtype=np.float32)).astype(np.float32)
    if not np.isfinite(e).all():
        raise ValueError("non-finite reconstruction error")
    lo = e.min(axis=1).astype(np.float32)
    hi = e.max(axis=1).astype(np.float32)
    span = np.maximum((hi - lo).astype(np.float32), np.float32(1e-12)).astype(np.float32)
    scale = (span / np.float32(15.0)).astype(np.float32)
    inv = (np.float32(1.0) / scale).astype(np.float32)
    t = ((e - lo[:, None]).astype(np.float32) * inv[:, None]).astype(np.float32)
    t = (t + np.float32(0.5)).astype(np.float32)
    code = np.clip(np.floor(t), 0, 15).astype(np.uint8)
    n, dim = x.shape
    out = np.zeros((n, (dim + 1) // 2), dtype=np.uint8)
    out[:, :(dim + 1) // 2] |= code[:, 0::2]
    out[:, :dim // 2] |= (code[:, 1::2] << 4).astype(np.uint8)
    return out, np.stack([scale, lo], axis=1).astype(np.float32)


def decode_sq(codes, bits, scale, zero, dim):
    """code * scale + zero per dimension in float32, two roundings
    (demo_quantize_l1_vector's dequantized vector)."""
    if bits == 8:
        q = codes.astype(np.float32)
    else:
        q = np.zeros((codes.shape[0], dim), dtype=np.float32)
        q[:, 0::2] = codes[:, :(dim + 1) // 2] & 15
        q[:, 1::2] = codes[:, :dim // 2] >> 4
    return ((q * np.float32(scale)).astype(np.float32) + np.float32(zero)).astype(np.float32)


def encode_sq_residual(x, l1_codes, l1_bits, scale, zero, residual_bits):
    """Per-vector residual of the SQ reconstruction
    (retrieval_rpc_demo_support.c:843-886 demo_quantize_residual over
    r = x - deq, :664-700): rmin / rmax of r, range = max(rmax - rmin, 1e-12),
    scale = range / levels, zero = rmin, code = clamp(roundf((r - zero) *
    (1 / scale)), 0, levels).  Returns (codes [n][code bytes], qparams
    [n][2] float32)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, dim = x.shape
    levels = 15 if residual_bits == 4 else 255
    r = (x - decode_sq(l1_codes, l1_bits, scale, zero, dim)).astype(np.float32)
    rmin = r.min(axis=1)
    rng = (r.max(axis=1) - rmin).astype(np.float32)
    rng = np.where(rng < np.float32(1e-12), np.float32(1e-12), rng).astype(np.float32)
    rs = (rng / np.float32(levels)).astype(np.float32)
    inv = (np.float32(1.0) / rs).astype(np.float32)
    v = ((r - rmin[:, None]).astype(np.float32) * inv[:, None]).astype(np.float32)
    q = _roundf_nonneg_or_clamp(v, levels).astype(np.uint8)
    if residual_bits == 8:
        codes = q
    else:
        codes = np.zeros((n, (dim + 1) // 2), dtype=np.uint8)
        codes[:, :(dim + 1) // 2] |= q[:, 0::2]
        codes[:, :dim // 2] |= (q[:, 1::2] << 4).astype(np.uint8)
    return np.ascontiguousarray(codes), np.stack([rs, rmin], axis=1).astype(np.float32)


# ---------------------------------------------------------------- graph

def _prune_rows(du, dpair, R, valid=None):
    """RNG occlusion over candidates sorted by (distance, id).
    du [B, C] distance u->c, dpair [B, C, C] distance c_i->c_j, valid [B, C]
    marks real candidates (default all).  Returns bool keep mask [B, C] with
    min(R, #valid) True per row: occlusion-survivors first
    (retrieval_search_build_core.c:904-930), then filled from the remaining
    candidates in order (:932-947)."""
    B, Cn = du.shape
    if valid is None:
        valid = np.ones((B, Cn), dtype=bool)
    keep = np.zeros((B, Cn), dtype=bool)
    count = np.zeros(B, dtype=np.int64)
    for i in range(Cn):
        occl = (keep[:, :i] & (dpair[:, :i, i] < du[:, i:i + 1])).any(axis=1)
        acc = (~occl) & (count < R) & valid[:, i]
        keep[:, i] = acc
        count += acc
    need = R - count
    rest = ~keep & valid
    rank = np.cumsum(rest, axis=1)
    keep |= rest & (rank <= need[:, None])
    return keep


def reverse_candidates(graph, x, R):
    """For each node v, up to R nodes u with an edge u -> v, closest first
    (the backlink lists of the reference's construction)."""
    n, deg = graph.shape
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    dst = graph.ravel().astype(np.int64)
    ok = dst >= 0
    src, dst = src[ok], dst[ok]
    diff = x[src].astype(np.float32) - x[dst].astype(np.float32)
    dist = np.einsum("ij,ij->i", diff, diff)
    order = np.lexsort((src, dist, dst))
    src, dst = src[order], dst[order]
    start = np.searchsorted(dst, np.arange(n))
    rank = np.arange(len(dst)) - start[dst]
    keep = rank < R
    rev = np.full((n, R), -1, dtype=np.int32)
    rev[dst[keep], rank[keep]] = src[keep]
    return rev


def _prune_candidate_rows(x, u_ids, cand, R):
    """RNG-prunes candidate rows (may hold -1 and duplicates) for nodes
    u_ids; returns int32 [len(u_ids), R] rows sorted by (distance, id)."""
    B, Cn = cand.shape
    xf = x
    cand = cand.astype(np.int64)
    valid = (cand >= 0) & (cand != u_ids[:, None])
    # drop duplicates: sort by id, mark repeats
    o = np.argsort(np.where(valid, cand, np.iinfo(np.int64).max), axis=1, kind="stable")
    cs = np.take_along_axis(cand, o, 1)
    vs = np.take_along_axis(valid, o, 1)
    dup = np.zeros_like(vs)
    dup[:, 1:] = cs[:, 1:] == cs[:, :-1]
    vs &= ~dup
    safe = np.where(vs, cs, 0)
    diff = xf[safe] - xf[u_ids][:, None, :]
    du = np.einsum("bcd,bcd->bc", diff, diff).astype(np.float32)
    du = np.where(vs, du, np.float32(np.inf))
    order = np.lexsort((np.where(vs, cs, np.iinfo(np.int64).max), du), axis=1)
    cs = np.take_along_axis(safe, order, 1)
    du = np.take_along_axis(du, order, 1)
    vs = np.take_along_axis(vs, order, 1)
    pts = xf[cs]
    g = np.einsum("bid,bjd->bij", pts, pts)
    sq = np.einsum("bid,bid->bi", pts, pts)
    dpair = sq[:, :, None] + sq[:, None, :] - 2.0 * g
    keep = _prune_rows(du, dpair, R, vs)
    rows = np.full((B, R), -1, dtype=np.int32)
    for i in range(B):
        r = cs[i][keep[i]][:R]
        rows[i, :len(r)] = r
    return rows


def build_graph(x, R, candidates=None, device=None, batch=None, backlinks=True):
    """int32 [n, R] neighbour rows (distance, id order), -1 padded when fewer
    candidates exist.  backlinks: after the kNN+RNG rows are built, each node
    re-prunes its row together with up to R nodes that link to it, so that
    nodes the occlusion rule left without in-edges become reachable."""
    n, dim = x.shape
    Cn = min(n - 1, candidates or 2 * R)
    if device is not None:
        g = _build_graph_torch(x, R, Cn, device, batch or 8192)
        if backlinks:
            g = _backlinks_torch(x, g, R, device, batch or 8192)
        return g
    graph = np.full((n, R), -1, dtype=np.int32)
    xf = x.astype(np.float32)
    norms = (xf * xf).sum(1)
    step = batch or 256
    for s in range(0, n, step):
        e = min(n, s + step)
        d = norms[s:e, None] - 2.0 * xf[s:e] @ xf.T + norms[None, :]
        d[np.arange(e - s), np.arange(s, e)] = np.inf
        kk = min(n - 1, Cn + 8)
        cand = np.argpartition(d, kk - 1, axis=1)[:, :kk]
        # exact f32 distances, sorted by (distance, id)
        diff = xf[cand] - xf[s:e, None, :]
        du = np.einsum("bcd,bcd->bc", diff, diff).astype(np.float32)
        order = np.lexsort((cand, du), axis=1)
        cand = np.take_along_axis(cand, order, 1)[:, :Cn]
        du = np.take_along_axis(du, order, 1)[:, :Cn]
        pts = xf[cand]
        g = np.einsum("bid,bjd->bij", pts, pts)
        sq = np.einsum("bid,bid->bi", pts, pts)
        dpair = sq[:, :, None] + sq[:, None, :] - 2.0 * g
        keep = _prune_rows(du, dpair, R)
        for i in range(e - s):
            row = cand[i][keep[i]][:R]
            graph[s + i, :len(row)] = row
    if not backlinks:
        return graph
    rev = reverse_candidates(graph, xf, R)
    out = np.full((n, R), -1, dtype=np.int32)
    for s in range(0, n, step):
        e = min(n, s + step)
        cand = np.concatenate([graph[s:e], rev[s:e]], axis=1)
        out[s:e] = _prune_candidate_rows(xf, np.arange(s, e), cand, R)
    return out


def _prune_torch(xt, u_ids, cand, R):
    """GPU counterpart of _prune_candidate_rows: cand [B, C] int64 (may hold
    -1 and duplicates), returns int32 rows [B, R] (distance, id order)."""
    import torch
    B, Cn = cand.shape
    dev = xt.device
    big = torch.iinfo(torch.int64).max
    valid = (cand >= 0) & (cand != u_ids[:, None])
    key = torch.where(valid, cand, torch.full_like(cand, big))
    key, _ = torch.sort(key, dim=1)
    dup = torch.zeros_like(valid)
    dup[:, 1:] = key[:, 1:] == key[:, :-1]
    valid = (key != big) & ~dup
    cand = torch.where(valid, key, torch.zeros_like(key))
    pts = xt[cand]
    diff = pts - xt[u_ids][:, None, :]
    du = (diff * diff).sum(-1)
    du = torch.where(valid, du, torch.full_like(du, float("inf")))
    # (distance, id) order: ids are already ascending, so a stable sort by
    # distance breaks ties by id
    o = torch.argsort(du, dim=1, stable=True)
    cand = torch.gather(cand, 1, o)
    du = torch.gather(du, 1, o)
    valid = torch.gather(valid, 1, o)
    pts = xt[cand].contiguous()
    sq = (pts * pts).sum(-1)
    g = torch.bmm(pts, pts.transpose(1, 2).contiguous())
    dpair = sq[:, :, None] + sq[:, None, :] - 2.0 * g
    keep = torch.zeros((B, Cn), dtype=torch.bool, device=dev)
    count = torch.zeros(B, dtype=torch.int64, device=dev)
    for i in range(Cn):
        if i:
            occl = (keep[:, :i] & (dpair[:, :i, i] < du[:, i:i + 1])).any(dim=1)
        else:
            occl = torch.zeros(B, dtype=torch.bool, device=dev)
        acc = (~occl) & (count < R) & valid[:, i]
        keep[:, i] = acc
        count += acc.to(torch.int64)
    need = R - count
    rest = ~keep & valid
    rank = torch.cumsum(rest.to(torch.int64), dim=1)
    keep |= rest & (rank <= need[:, None])
    pos = torch.cumsum(keep.to(torch.int64), dim=1) - 1
    rows = torch.full((B, R), -1, dtype=torch.int32, device=dev)
    sel = keep & (pos < R)
    bi = torch.arange(B, device=dev)[:, None].expand(B, Cn)
    rows[bi[sel], pos[sel]] = cand[sel].to(torch.int32)
    return rows


def _build_graph_torch(x, R, Cn, device, batch):
    import torch
    n, dim = x.shape
    xt = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(device)
    norms = (xt * xt).sum(1)
    graph = torch.full((n, R), -1, dtype=torch.int32, device=device)
    kk = min(n - 1, Cn + 8)
    for s in range(0, n, batch):
        e = min(n, s + batch)
        xb = xt[s:e]
        d = norms[s:e, None] - 2.0 * (xb @ xt.T) + norms[None, :]
        d[torch.arange(e - s, device=device), torch.arange(s, e, device=device)] = float("inf")
        cand = torch.topk(d, kk, dim=1, largest=False).indices
        del d
        diff = xt[cand] - xb[:, None, :]
        du = (diff * diff).sum(-1)
        o1 = torch.argsort(cand, dim=1, stable=True)
        cand = torch.gather(cand, 1, o1)
        du = torch.gather(du, 1, o1)
        o2 = torch.argsort(du, dim=1, stable=True)
        cand = torch.gather(cand, 1, o2)[:, :Cn].contiguous()
        graph[s:e] = _prune_torch(xt, torch.arange(s, e, device=device), cand, R)
    return graph.cpu().numpy()


def _backlinks_torch(x, graph, R, device, batch):
    """Each node re-prunes its row together with up to R closest nodes that
    link to it (GPU)."""
    import torch
    n, deg = graph.shape
    xt = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(device)
    g = torch.from_numpy(graph).to(device).to(torch.int64)
    src = torch.arange(n, device=device).repeat_interleave(deg)
    dst = g.reshape(-1)
    ok = dst >= 0
    src, dst = src[ok], dst[ok]
    dist = torch.empty(src.shape[0], device=device)
    step = 1 << 22
    for s in range(0, src.shape[0], step):
        a, b = src[s:s + step], dst[s:s + step]
        df = xt[a] - xt[b]
        dist[s:s + step] = (df * df).sum(-1)
    # order by (dst, dist, src) with stable sorts from the last key
    o = torch.argsort(src, stable=True)
    src, dst, dist = src[o], dst[o], dist[o]
    o = torch.argsort(dist, stable=True)
    src, dst, dist = src[o], dst[o], dist[o]
    o = torch.argsort(dst, stable=True)
    src, dst = src[o], dst[o]
    start = torch.searchsorted(dst, torch.arange(n, device=device))
    rank = torch.arange(dst.shape[0], device=device) - start[dst]
    keep = rank < R
    rev = torch.full((n, R), -1, dtype=torch.int64, device=device)
    rev[dst[keep], rank[keep]] = src[keep]
    out = torch.full((n, R), -1, dtype=torch.int32, device=device)
    for s in range(0, n, batch):
        e = min(n, s + batch)
        cand = torch.cat([g[s:e], rev[s:e]], dim=1)
        out[s:e] = _prune_torch(xt, torch.arange(s, e, device=device), cand, R)
    return out.cpu().numpy()


def medoid(x):
    """Entry point: the base vector nearest the centroid (the standard
    stand-in for the medoid, which is quadratic to compute exactly)."""
    n = x.shape[0]
    step = 1 << 18
    if n <= 1 << 22:
        c = x.astype(np.float64).mean(0)
    else:  # a float64 copy of a large set is avoided; the mean is summed by chunks
        c = sum(x[s:s + step].astype(np.float64).sum(0) for s in range(0, n, step)) / n
    d = np.concatenate([((x[s:s + step].astype(np.float64) - c) ** 2).sum(1)
                        for s in range(0, n, step)])
    return int(np.argmin(d))


def entry_list(x, count, entry, device=None, iters=10, sample=None):
    """Shared seed list: the medoid, then the base vector nearest each of
    `count` k-means centroids (deduplicated, in centroid order).

    A degree-R graph built from exact nearest neighbours of well-separated
    clusters has no edges between clusters; seeding every query with one
    node per k-means cell lets the best-first search start in the query's
    cluster.  The list is passed as the request's entry ids, exactly like the
    reference's seed list (PurePqSearchRequest.entry_ids)."""
    _, reps = kmeans_representatives(x, count, device, iters, sample)
    out, seen = [int(entry)], {int(entry)}
    for r in reps:
        if int(r) not in seen:
            seen.add(int(r))
            out.append(int(r))
    return np.array(out, dtype=np.int32)


def kmeans_representatives(x, count, device=None, iters=10, sample=None, seeding="stride"):
    """(centroids float32 [count, dim], representative ids int32 [count]):
    Lloyd k-means on a uniform-stride sample, then for every centroid the
    base vector nearest to it.  Used by the entry router: each query starts
    from the representatives of its nearest centroids.  seeding "stride":
    the initial centroids are uniform-stride sample rows; "kmeans++" (device
    only): D^2 seeding, which puts a centroid in every well-separated group
    of the data."""
    n = x.shape[0]
    count = min(count, n)
    rows = uniform_stride_sample(n, min(n, sample or max(64 * count, 20000)))
    tr = np.ascontiguousarray(x[rows], dtype=np.float32)
    init = (np.arange(count, dtype=np.int64) * (len(tr) - 1)) // max(count - 1, 1)
    if seeding not in ("stride", "kmeans++") or (seeding != "stride" and device is None):
        raise ValueError("seeding: 'stride', or 'kmeans++' with a device")
    if device is not None:
        import torch
        t = torch.from_numpy(tr).to(device)
        cent = t[torch.from_numpy(init).to(device)].clone() if seeding == "stride" else \
            _kmeanspp_seeds(torch, t, count, 11)
        for _ in range(iters):
            d = (t * t).sum(1, keepdim=True) - 2.0 * t @ cent.T + (cent * cent).sum(1)[None]
            a = torch.argmin(d, dim=1)
            _segment_means(torch, t, a, count, cent)
        xt = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(device)
        xn = (xt * xt).sum(1)
        best = torch.full((count,), float("inf"), device=device)
        arg = torch.zeros(count, dtype=torch.int64, device=device)
        step = 1 << 18
        for s in range(0, n, step):
            d = xn[None, s:s + step] - 2.0 * cent @ xt[s:s + step].T
            v, i = torch.min(d, dim=1)
            upd = v < best
            best = torch.where(upd, v, best)
            arg = torch.where(upd, i + s, arg)
        reps = arg.cpu().numpy()
        cent = cent.cpu().numpy()
    else:
        cent = tr[init].copy()
        for _ in range(iters):
            d = (tr * tr).sum(1)[:, None] - 2.0 * tr @ cent.T + (cent * cent).sum(1)[None]
            a = np.argmin(d, axis=1)
            sums = np.zeros_like(cent, dtype=np.float64)
            np.add.at(sums, a, tr)
            cnt = np.bincount(a, minlength=count)
            nz = cnt > 0
            cent[nz] = (sums[nz] / cnt[nz, None]).astype(np.float32)
        xf = np.ascontiguousarray(x, dtype=np.float32)
        d = (xf * xf).sum(1)[None, :] - 2.0 * cent @ xf.T
        reps = np.argmin(d, axis=1)
    return (np.ascontiguousarray(cent, dtype=np.float32),
            np.ascontiguousarray(reps, dtype=np.int32))


def route_on_host(queries, centroids, rep_ids, S, chunk=8192):
    """Caller-side entry selection: for every query the representative ids
    of its S nearest centroids (nearest first), as an int32 [nq, S] array of
    per-query entry rows for a search request's entry_ids.  Distances are
    ||q||^2 - 2 q.c + ||c||^2 in float32 BLAS (near ties may order
    differently from the device router's sequential sums; both paths take
    whatever rows the caller passes)."""
    q = np.ascontiguousarray(queries, dtype=np.float32)
    c = np.ascontiguousarray(centroids, dtype=np.float32)
    S = int(S)
    assert 0 < S <= c.shape[0]
    cn = (c * c).sum(1)
    out = np.empty((q.shape[0], S), dtype=np.int32)
    cols = np.arange(S)
    for s in range(0, q.shape[0], chunk):
        d = cn[None, :] - 2.0 * (q[s:s + chunk] @ c.T)
        part = np.argpartition(d, S - 1, axis=1)[:, :S] if S < c.shape[0] else \
            np.tile(np.arange(c.shape[0]), (d.shape[0], 1))
        pd = np.take_along_axis(d, part, axis=1)
        order = np.lexsort((part, pd), axis=1)
        best = np.take_along_axis(part, order, axis=1)
        out[s:s + chunk] = rep_ids[best[:, cols]]
    return out


class PqIndexData:
    """Arrays of a pure-PQ index (all host numpy)."""

    def __init__(self, graph, entry, codes, codebook, res_codes=None,
                 res_qparams=None, res_bits=4, entries=None,
                 router_centroids=None, router_ids=None):
        self.graph = np.ascontiguousarray(graph, dtype=np.int32)
        self.entry = int(entry)
        # entry router: centroids and their representative node ids
        self.router_centroids = None if router_centroids is None else \
            np.ascontiguousarray(router_centroids, dtype=np.float32)
        self.router_ids = None if router_ids is None else \
            np.ascontiguousarray(router_ids, dtype=np.int32)
        # shared seed list (medoid first); None -> the graph entry alone
        self.entries = None if entries is None else np.ascontiguousarray(entries, dtype=np.int32)
        self.codes = np.ascontiguousarray(codes, dtype=np.uint8)
        self.codebook = np.ascontiguousarray(codebook, dtype=np.float32)
        self.res_codes = None if res_codes is None else np.ascontiguousarray(res_codes)
        self.res_qparams = None if res_qparams is None else np.ascontiguousarray(res_qparams)
        self.res_bits = res_bits

    @property
    def n(self):
        return self.graph.shape[0]

    @property
    def degree(self):
        return self.graph.shape[1]

    @property
    def m(self):
        return self.codebook.shape[0]

    @property
    def ks(self):
        return self.codebook.shape[1]

    @property
    def dim(self):
        return self.codebook.shape[0] * self.codebook.shape[2]


def build_pq_index(x, m, R, ks=256, res_bits=4, iters=20, train_max=100000,
                   device=None, graph=None, n_entries=0, n_router=0, router_seeding="stride"):
    """n_entries > 0 adds a shared seed list of the medoid plus n_entries
    k-means representatives (see entry_list); n_router > 0 adds an entry
    router of n_router centroids (see kmeans_representatives)."""
    codebook = train_pq(x, m, ks, iters=iters, train_max=train_max, device=device)
    if device is not None and x.shape[0] > 100000:
        codes, res, qp = encode_on_device(x, codebook, res_bits)
    else:
        codes = encode_pq(x, codebook)
        res, qp = encode_residual(x, codes, codebook, res_bits)
    if graph is None:
        graph = build_graph(x, R, device=device)
    entry = medoid(x)
    entries = entry_list(x, n_entries, entry, device=device) if n_entries > 0 else None
    rc, ri = kmeans_representatives(x, n_router, device, seeding=router_seeding) \
        if n_router > 0 else (None, None)
    return PqIndexData(graph, entry, codes, codebook, res, qp, res_bits, entries, rc, ri)


def encode_on_device(x, codebook, res_bits):
    """PQ codes and residuals with the library's HIP encoders."""
    from . import native
    L = native.lib()
    m, ks, sub = codebook.shape
    n, dim = x.shape
    cb = code_bytes(m, ks)
    rb = residual_code_bytes(dim, res_bits)
    dx = native.DeviceBuffer.from_array(np.ascontiguousarray(x, dtype=np.float32))
    dcb = native.DeviceBuffer.from_array(codebook)
    dcodes = native.DeviceBuffer(n * cb)
    dres = native.DeviceBuffer(n * rb)
    dqp = native.DeviceBuffer(n * 8)
    native.check(L.rover_build_pq_encode(dx.ptr, n, dim, m, ks, dcb.ptr, dcodes.ptr),
                 "rover_build_pq_encode")
    native.check(L.rover_build_residual_encode(dx.ptr, dcodes.ptr, n, dim, m, ks,
                                               dcb.ptr, res_bits, dres.ptr, dqp.ptr),
                 "rover_build_residual_encode")
    native.synchronize()
    codes = dcodes.download(np.uint8, n * cb).reshape(n, cb)
    res = dres.download(np.uint8, n * rb).reshape(n, rb)
    qp = dqp.download(np.float32, n * 2).reshape(n, 2)
    return codes, res, qp


def encode_error4_on_device(x, codes, codebook, res_codes, res_qparams, res_bits):
    """Third-level refine codes with the library's HIP encoder
    (rover_build_error4_encode): returns ([n][(dim+1)/2] uint8 codes,
    [n][2] float32 (scale, zero)).  Raises if any row's error is not finite."""
    from . import native
    L = native.lib()
    m, ks, sub = codebook.shape
    n, dim = x.shape
    e4b = (dim + 1) // 2
    if (m * sub != dim or codes.shape != (n, code_bytes(m, ks))
            or res_codes.shape != (n, residual_code_bytes(dim, res_bits))
            or res_qparams.shape != (n, 2)):
        raise ValueError("error4 encode: inconsistent index shapes")
    dx = native.DeviceBuffer.from_array(np.ascontiguousarray(x, dtype=np.float32))
    dcodes = native.DeviceBuffer.from_array(np.ascontiguousarray(codes, dtype=np.uint8))
    dcb = native.DeviceBuffer.from_array(np.ascontiguousarray(codebook, dtype=np.float32))
    dres = native.DeviceBuffer.from_array(np.ascontiguousarray(res_codes, dtype=np.uint8))
    dqp = native.DeviceBuffer.from_array(np.ascontiguousarray(res_qparams, dtype=np.float32))
    de4 = native.DeviceBuffer(n * e4b)
    de4qp = native.DeviceBuffer(n * 8)
    native.check(L.rover_build_error4_encode(dx.ptr, dcodes.ptr, dres.ptr, dqp.ptr, n, dim,
                                             m, ks, dcb.ptr, res_bits, de4.ptr, de4qp.ptr),
                 "rover_build_error4_encode")
    e4 = de4.download(np.uint8, n * e4b).reshape(n, e4b)
    e4qp = de4qp.download(np.float32, n * 2).reshape(n, 2)
    if not np.isfinite(e4qp).all():
        raise ValueError("non-finite reconstruction error")
    return e4, e4qp


def build_graph_rnn(x, s=16, R=32, t1=4, t2=8, seed=42, top_k=None):
    """RNN-Descent on the device (rover_build_graph_rnn): int32 [n, top_k]
    (top_k defaults to R), rows by (distance, id), -1 padded."""
    from . import native
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, dim = x.shape
    top_k = R if top_k is None else top_k
    dx = native.DeviceBuffer.from_array(x)
    dg = native.DeviceBuffer(n * top_k * 4)
    native.check(native.lib().rover_build_graph_rnn(dx.ptr, n, dim, s, R, t1, t2, seed, top_k,
                                                    dg.ptr), "rover_build_graph_rnn")
    return dg.download(np.int32, n * top_k).reshape(n, top_k)


def build_graph_exact(x, R, candidates=64, seed=42, prune_mode=1):
    """The reference's exact-candidate graph build on the device
    (rover_build_graph_exact): int32 [n, R], rows by (distance, id)."""
    from . import native
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, dim = x.shape
    dx = native.DeviceBuffer.from_array(x)
    dg = native.DeviceBuffer(n * R * 4)
    native.check(native.lib().rover_build_graph_exact(dx.ptr, n, dim, R, candidates, seed,
                                                      prune_mode, dg.ptr),
                 "rover_build_graph_exact")
    return dg.download(np.int32, n * R).reshape(n, R)


def _kmeanspp_seeds(torch, tr, count, seed):
    """k-means++ seeding on the device rows tr: the first centre uniform, each
    next one drawn with probability proportional to the squared distance to
    the nearest centre so far (numpy's generator from `seed` draws the
    uniforms, so the seeds are reproducible)."""
    rng = np.random.default_rng(seed)
    pick = [int(rng.integers(tr.shape[0]))]
    d2 = ((tr - tr[pick[0]]) ** 2).sum(1)
    u = rng.random(count)
    for i in range(1, count):
        cs = torch.cumsum(d2.double(), 0)
        j = int(torch.searchsorted(cs, cs[-1:] * float(u[i])).item())
        j = min(j, tr.shape[0] - 1)
        pick.append(j)
        d2 = torch.minimum(d2, ((tr - tr[j]) ** 2).sum(1))
    return tr[torch.tensor(pick, device=tr.device)].clone()


def cluster_for_build(x, clusters, iters=8, sample=262144, seed=7, device="cuda",
                      chunk=65536):
    """A k-means clustering of x for build_graph_exact_clustered: k-means++
    seeding (each next centre drawn with probability proportional to the
    squared distance to the nearest centre so far) and Lloyd's iterations on
    a uniform-stride sample (float32 on the device), then every row assigned
    to its nearest centroid.  Returns (assign int32 [n], centroids float32
    [clusters, dim]).  Any clustering gives the same graph; a tight one
    leaves fewer clusters to search per node (seeding by squared distance
    puts a centre in every well-separated group, where uniform seeding
    leaves about a third of them to share a cluster with a neighbour)."""
    import torch
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, dim = x.shape
    clusters = max(1, min(int(clusters), n))
    tr = torch.from_numpy(x[uniform_stride_sample(n, min(n, max(sample, 64 * clusters)))]
                          ).to(device)
    cent = _kmeanspp_seeds(torch, tr, clusters, seed)

    def nearest(pts):
        d = (cent * cent).sum(1)[None, :] - 2.0 * pts @ cent.T
        return torch.argmin(d, dim=1)

    for _ in range(iters):
        a = torch.cat([nearest(tr[i:i + chunk]) for i in range(0, tr.shape[0], chunk)])
        _segment_means(torch, tr, a, clusters, cent)
    assign = np.empty(n, dtype=np.int32)
    for i in range(0, n, chunk):
        pts = torch.from_numpy(x[i:i + chunk]).to(device)
        assign[i:i + chunk] = nearest(pts).to(torch.int32).cpu().numpy()
    return assign, cent.cpu().numpy()


def build_graph_exact_clustered(x, R, candidates=64, seed=42, prune_mode=1, clusters=None,
                                assign=None, centroids=None, stats=None):
    """build_graph_exact's graph, bit for bit, computed over a clustering
    (rover_build_graph_exact_clustered): own-cluster candidates plus every
    cluster a triangle-inequality bound does not rule out.  assign/centroids
    default to cluster_for_build(x, clusters or n / 1024).  stats (a list)
    receives [node blocks (16 members of one cluster), cluster pairs screened
    in, (block, cluster) pairs searched, blocks with such pairs]."""
    from . import native
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, dim = x.shape
    if assign is None:
        k = clusters if clusters is not None else max(1, n // 1024)
        assign, centroids = cluster_for_build(x, k)
    assign = np.ascontiguousarray(assign, dtype=np.int32)
    centroids = np.ascontiguousarray(centroids, dtype=np.float32)
    if assign.shape != (n,) or centroids.ndim != 2 or centroids.shape[1] != dim:
        raise ValueError("assign must be [n] and centroids [clusters, dim]")
    dx = native.DeviceBuffer.from_array(x)
    da = native.DeviceBuffer.from_array(assign)
    dc = native.DeviceBuffer.from_array(centroids)
    dg = native.DeviceBuffer(n * R * 4)
    st = (C.c_int64 * 4)()
    native.check(native.lib().rover_build_graph_exact_clustered(
        dx.ptr, n, dim, R, candidates, seed, prune_mode, da.ptr, dc.ptr, centroids.shape[0],
        dg.ptr, st), "rover_build_graph_exact_clustered")
    if stats is not None:
        stats[:] = list(st)
    return dg.download(np.int32, n * R).reshape(n, R)
