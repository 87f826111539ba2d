"""Graph construction on the GPU (rover_build_graph_exact) against the
reference's outputs and the oracle, row for row; and the 100k-node build
(the small SQ workload's graph) timed and checked for its invariants."""
import sys
import time

import numpy as np
import pytest

import pyoracle
from rover_amd import data
from rover_amd import index as rindex

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_device_build_matches_reference_fixture(golden):
    z = golden("graph_build")
    for ci, (n, dim, r, cand, prune, seed) in enumerate(z["cases"]):
        say("case graph fixture", ci, int(n), int(dim), int(r), int(cand), int(prune))
        g = rindex.build_graph_exact(z["base%d" % ci], int(r), int(cand), int(seed), int(prune))
        assert np.array_equal(g, z["graph%d" % ci]), ci


@pytest.mark.parametrize("spec", [(5000, 128, 32, 64, 1), (3000, 96, 48, 64, 1),
                                  (4000, 30, 16, 48, 0), (2000, 200, 32, 64, 1),
                                  (1500, 131, 16, 32, 1), (777, 260, 24, 40, 0),
                                  (3000, 96, 48, 96, 1), (2500, 64, 32, 128, 1),
                                  (1800, 100, 64, 100, 0)])
def test_device_build_matches_oracle(spec):
    n, dim, r, cand, prune = spec
    say("case graph oracle", spec)
    base, _ = data.workload(n, dim, 32, 8.0, 1, base_seed=n + dim, query_seed=1)
    g = rindex.build_graph_exact(base, r, cand, 42, prune)
    want = pyoracle.build_exact(base, r, cand, 42, prune)
    assert np.array_equal(g, want)


def test_device_build_100k():
    say("case graph 100k")
    base, _ = data.workload(100_000, 128, 256, 8.0, 1)
    t = time.time()
    g = rindex.build_graph_exact(base, 32, 64, 42, 1)
    el = time.time() - t
    assert g.shape == (100_000, 32) and (g >= 0).all() and (g < 100_000).all()
    srt = np.sort(g, axis=1)
    assert not np.any(srt[:, 1:] == srt[:, :-1])
    assert not np.any(g == np.arange(100_000)[:, None])
    # spot rows against the oracle's exact distances: ascending (distance, id)
    x = np.ascontiguousarray(base, dtype=np.float32)
    for u in range(0, 100_000, 9973):
        d = [pyoracle.oracle().orc_l2(x.ctypes.data, 128, u, int(v)) for v in g[u]]
        keys = list(zip(d, g[u].tolist()))
        assert keys == sorted(keys)
        # nothing occludes the nearest neighbor, so it heads the row
        near = np.argsort(((x.astype(np.float64) - x[u]) ** 2).sum(axis=1))[:6]
        best = min((pyoracle.oracle().orc_l2(x.ctypes.data, 128, u, int(v)), int(v))
                   for v in near if v != u)
        assert best[1] == g[u][0]
    say("graph 100k x 128, R=32, 64 exact candidates: %.2f s" % el)
