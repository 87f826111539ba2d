"""Synthetic workloads and ground truth.

Gaussian-mixture base and query sets: K centres drawn U[0,100]^d, each
point picks a centre uniformly and adds N(0, sigma^2 I), clipped to
[0, 255], float32.  Base, queries and training queries use separate seeds.
Ground truth is exact top-k by squared L2 with the (distance, id)
tie-break, candidates from a float32 GEMM re-ranked in float64.
"""
import numpy as np


def gaussian_mixture(n, dim, components, sigma, seed, centers_seed=None):
    """Returns float32 [n, dim].  The centres depend only on centers_seed
    (default: seed), so base and query sets share one mixture."""
    crng = np.random.default_rng(seed if centers_seed is None else centers_seed)
    centers = crng.uniform(0.0, 100.0, size=(components, dim)).astype(np.float32)
    rng = np.random.default_rng(seed + 7919 if centers_seed is not None else seed + 1)
    out = np.empty((n, dim), dtype=np.float32)
    chunk = 1 << 16
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        comp = rng.integers(0, components, size=e - s)
        noise = rng.standard_normal(size=(e - s, dim), dtype=np.float32)
        x = centers[comp] + np.float32(sigma) * noise
        np.clip(x, 0.0, 255.0, out=x)
        out[s:e] = x
    return out


def workload(n, dim, components, sigma, nq, base_seed=42, query_seed=43):
    """Base and query sets drawn from the same mixture."""
    base = gaussian_mixture(n, dim, components, sigma, base_seed,
                            centers_seed=base_seed)
    queries = gaussian_mixture(nq, dim, components, sigma, query_seed,
                               centers_seed=base_seed)
    return base, queries


def _exact_rerank(base, q, cand, k):
    """Exact (float64) re-rank of candidate ids, tie-break by id."""
    diff = base[cand].astype(np.float64) - q[None, :].astype(np.float64)
    d = np.einsum("ij,ij->i", diff, diff)
    order = np.lexsort((cand, d))
    return cand[order[:k]]


def ground_truth(base, queries, k=100, device=None, margin=32, verify=16):
    """Exact top-k ids per query, int32 [nq, k]: float32 ||x||^2 - 2 q.x
    screens the k + margin best per query, which are re-ranked exactly
    (float64, ties by id).  On a device the base is screened in slices of
    2^20 rows, 512 queries at a time (small products and top-k calls), and
    `verify` queries spread over the set are checked against a float64 scan
    of the whole base."""
    n = base.shape[0]
    kk = min(n, k + margin)
    if device is not None:
        import torch
        xb = torch.from_numpy(base).to(device)
        nb = (xb * xb).sum(1)
        out = np.empty((queries.shape[0], k), dtype=np.int32)
        cols, step = 1 << 20, 512
        for s in range(0, queries.shape[0], step):
            e = min(queries.shape[0], s + step)
            xq = torch.from_numpy(queries[s:e]).to(device)
            vals, idx = [], []
            for c0 in range(0, n, cols):
                c1 = min(n, c0 + cols)
                d = nb[None, c0:c1] - 2.0 * (xq @ xb[c0:c1].T)
                v, i = torch.topk(d, min(kk, c1 - c0), dim=1, largest=False)
                vals.append(v)
                idx.append(i + c0)
            v, i = torch.cat(vals, 1), torch.cat(idx, 1)
            sel = torch.topk(v, kk, dim=1, largest=False).indices
            cand = torch.gather(i, 1, sel).cpu().numpy()
            for j in range(e - s):
                out[s + j] = _exact_rerank(base, queries[s + j], cand[j], k)
        for qi in np.linspace(0, queries.shape[0] - 1, min(verify, queries.shape[0])).astype(int):
            d = ((xb.double() - torch.from_numpy(queries[qi]).to(device).double()) ** 2).sum(1) \
                if n <= (1 << 22) else torch.cat([
                    ((xb[c0:c0 + (1 << 22)].double() -
                      torch.from_numpy(queries[qi]).to(device).double()) ** 2).sum(1)
                    for c0 in range(0, n, 1 << 22)])
            dk = torch.topk(d, k, largest=False).values[-1].item()
            got = ((base[out[qi]].astype(np.float64) - queries[qi].astype(np.float64)) ** 2
                   ).sum(1)
            if got.max() > dk * (1 + 1e-9) + 1e-9:
                raise RuntimeError("ground truth check failed for query %d" % qi)
        return out
    nb = (base.astype(np.float32) ** 2).sum(1)
    out = np.empty((queries.shape[0], k), dtype=np.int32)
    step = max(1, min(256, (1 << 26) // max(n, 1)))
    for s in range(0, queries.shape[0], step):
        e = min(queries.shape[0], s + step)
        d = nb[None, :] - 2.0 * (queries[s:e] @ base.T)
        cand = np.argpartition(d, kk - 1, axis=1)[:, :kk] if kk < n else \
            np.tile(np.arange(n), (e - s, 1))
        for i in range(e - s):
            out[s + i] = _exact_rerank(base, queries[s + i], cand[i], k)
    return out


def recall_at_k(result_ids, counts, gt, k):
    """Mean recall@k with demo_recall_hits_at_k's definition
    (retrieval_rpc_demo_support.c:1709-1728): |{i<k: res[i] in gt[:k]}|/k."""
    hits = 0
    for q in range(result_ids.shape[0]):
        c = min(int(counts[q]), k)
        if c <= 0:
            continue
        hits += np.isin(result_ids[q, :c], gt[q, :k]).sum()
    return hits / (result_ids.shape[0] * float(k))


def write_fvecs(path, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    with open(path, "wb") as f:
        dim = np.int32(x.shape[1])
        for row in x:
            f.write(dim.tobytes())
            f.write(row.tobytes())


def read_fvecs(path):
    raw = np.fromfile(path, dtype=np.int32)
    dim = int(raw[0])
    return raw.reshape(-1, dim + 1)[:, 1:].view(np.float32).copy()


def write_ivecs(path, x):
    x = np.ascontiguousarray(x, dtype=np.int32)
    with open(path, "wb") as f:
        dim = np.int32(x.shape[1])
        for row in x:
            f.write(dim.tobytes())
            f.write(row.tobytes())


def read_ivecs(path):
    raw = np.fromfile(path, dtype=np.int32)
    dim = int(raw[0])
    return raw.reshape(-1, dim + 1)[:, 1:].copy()
