"""RNN-Descent pinned to the reference: the oracle's restatement of
retrieval_rnn_build_graph_ex + retrieval_rnn_export_topk against graphs the
compiled reference wrote (tests/golden/rnn_build.npz; integer-coordinate
inputs make distance ties), against the live reference (single- and
multi-threaded, which must agree: each pass edits only its own rows and
applies insertions in sorted order), and against what the construction
fixes: rows ascending by (distance, id), no self edges or duplicates,
-1 only as padding, and on points along a line the path graph the
relative-neighborhood rule converges to."""
import numpy as np
import pytest

import pyoracle
from rover_amd import data


def l2(x, a, b):
    return pyoracle.oracle().orc_l2(x.ctypes.data, x.shape[1], int(a), int(b))


def test_oracle_matches_reference_fixture(golden):
    z = golden("rnn_build")
    for ci, (n, dim, s, r, t1, t2, seed) in enumerate(z["cases"]):
        g = pyoracle.rnn_build(z["base%d" % ci], int(s), int(r), int(t1), int(t2), int(seed))
        assert np.array_equal(g, z["graph%d" % ci]), ci


@pytest.mark.skipif(not pyoracle.ref_available(), reason="compiled reference not built")
@pytest.mark.parametrize("spec", [(700, 20, 10, 20, 3, 2, 5), (1200, 64, 16, 32, 2, 4, 9),
                                  (400, 12, 24, 24, 3, 3, 13)])
def test_oracle_matches_live_reference(spec):
    n, dim, s, r, t1, t2, seed = spec
    base, _ = data.workload(n, dim, 12, 8.0, 1, base_seed=seed, query_seed=seed + 1)
    base = np.ascontiguousarray(np.round(base * 2) / 2, dtype=np.float32)
    want = pyoracle.rnn_build(base, s, r, t1, t2, seed)
    for threads in (1, 3):
        got = np.zeros((n, r), dtype=np.int32)
        assert pyoracle.reference().ref_rnn_build(base.ctypes.data, n, dim, s, r, t1, t2, seed,
                                                  threads, r, got.ctypes.data) == 0
        assert np.array_equal(got, want), threads


def test_construction_invariants(golden):
    z = golden("rnn_build")
    for ci, (n, dim, s, r, t1, t2, seed) in enumerate(z["cases"]):
        x, g = z["base%d" % ci], z["graph%d" % ci]
        for u in range(0, int(n), 7):
            row = g[u][g[u] >= 0]
            # -1 only as a suffix
            assert (g[u][len(row):] == -1).all()
            assert len(set(row.tolist())) == len(row) and u not in row
            keys = [(l2(x, u, v), int(v)) for v in row]
            assert keys == sorted(keys)


@pytest.mark.parametrize("perm_seed", [1, 4])
def test_points_on_a_line_give_the_path_graph(perm_seed):
    """On points along a line the relative-neighborhood rule keeps exactly
    each point's nearest neighbor on either side (a farther point on the
    same side is occluded by the nearer one, one across is not), so a
    converged build is the path graph.  60 points in shuffled order, s = r =
    8, four rounds of eight passes converge."""
    n = 60
    perm = np.random.RandomState(perm_seed).permutation(n)
    x = np.ascontiguousarray(np.arange(n, dtype=np.float32)[perm][:, None])
    g = pyoracle.rnn_build(x, 8, 8, 4, 8, 42)
    node_at = np.argsort(perm)
    for k in range(n):
        want = {int(node_at[q]) for q in (perm[k] - 1, perm[k] + 1) if 0 <= q < n}
        assert set(int(v) for v in g[k] if v >= 0) == want, k
