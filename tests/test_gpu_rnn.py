"""RNN-Descent on the GPU (rover_build_graph_rnn) against the reference's
outputs and the oracle, row for row, including a multi-round build with
reverse edges and rows filled to r; and a 100k-node build timed and checked
for its invariants."""
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


def test_device_rnn_matches_reference_fixture(golden):
    z = golden("rnn_build")
    for ci, (n, dim, s, r, t1, t2, seed) in enumerate(z["cases"]):
        say("case rnn fixture", ci)
        g = rindex.build_graph_rnn(z["base%d" % ci], int(s), int(r), int(t1), int(t2), int(seed))
        assert np.array_equal(g, z["graph%d" % ci]), ci


@pytest.mark.parametrize("spec", [(5000, 128, 16, 32, 2, 4), (3000, 96, 24, 48, 3, 3),
                                  (2000, 30, 8, 16, 4, 4), (1500, 40, 64, 64, 2, 2)])
def test_device_rnn_matches_oracle(spec):
    n, dim, s, r, t1, t2 = spec
    say("case rnn oracle", spec)
    base, _ = data.workload(n, dim, 32, 8.0, 1, base_seed=n + dim, query_seed=1)
    g = rindex.build_graph_rnn(base, s, r, t1, t2, 42)
    want = pyoracle.rnn_build(base, s, r, t1, t2, 42)
    assert np.array_equal(g, want)


def test_device_rnn_top_k_and_rejections():
    base, _ = data.workload(800, 16, 8, 8.0, 1, base_seed=3, query_seed=4)
    g8 = rindex.build_graph_rnn(base, 8, 16, 2, 2, 7, top_k=8)
    g16 = rindex.build_graph_rnn(base, 8, 16, 2, 2, 7)
    assert np.array_equal(g8, g16[:, :8])
    for bad in (dict(s=0), dict(s=20, R=16), dict(R=65, s=8), dict(t1=0), dict(top_k=17)):
        kw = dict(s=8, R=16, t1=1, t2=1, seed=1)
        kw.update(bad)
        with pytest.raises(Exception):
            rindex.build_graph_rnn(base, **kw)


def test_device_rnn_100k():
    say("case rnn 100k")
    base, _ = data.workload(100_000, 128, 256, 8.0, 1)
    t = time.time()
    g = rindex.build_graph_rnn(base, 16, 32, 4, 4, 42)
    el = time.time() - t
    n = 100_000
    valid = g >= 0
    assert g.shape == (n, 32) and (g < n).all()
    assert not np.any(g == np.arange(n)[:, None])
    srt = np.sort(np.where(valid, g, -np.arange(1, 33)[None, :]), axis=1)
    assert not np.any(srt[:, 1:] == srt[:, :-1])
    x = np.ascontiguousarray(base, dtype=np.float32)
    for u in range(0, n, 9973):
        row = g[u][g[u] >= 0]
        assert (g[u][len(row):] == -1).all()
        d = [pyoracle.oracle().orc_l2(x.ctypes.data, 128, u, int(v)) for v in row]
        keys = list(zip(d, row.tolist()))
        assert keys == sorted(keys)
    say("rnn 100k x 128, s=16, R=32, t1=4, t2=4: %.2f s, mean degree %.1f"
        % (el, valid.sum(1).mean()))
