"""Bundles loaded into the library through rover_amd.bundle and searched on
the GPU: a PQ bundle with its error4 artifact gives the reference's ERROR4
refine results, an SQ bundle gives the reference's SQ traversal results."""
import os
import sys

import numpy as np
import pytest

from rover_amd import bundle, native
from rover_amd import index as rindex
from rover_amd.search import Calculator

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def test_pq_bundle_error4_search(tmp_path):
    g, e = load("pq_d96_m48"), load("pq_d96_m48_error4")
    queries = g["queries_x16"].astype(np.float32) / np.float32(16.0)
    bundle.write_pq_bundle(tmp_path / "b", g["graph"], int(g["entry"]), g["codebook"],
                           g["codes"], g["res_codes"], g["res_qparams"], int(g["res_bits"]),
                           error4=e["artifact"].tobytes())
    b = bundle.open_bundle(tmp_path / "b")
    say("case pq bundle")
    with Calculator() as calc:
        b.load(calc)
        mode, ep, expand, L = (int(v) for v in g["refine_source"])
        F, K = int(g["refine_F"]), int(g["refine_K"])
        ids, d, cnt = calc.search_batch(queries, L, K, correction=mode, refine_candidates=F,
                                        entry_points=ep, expand_count=expand,
                                        refine_mode=native.REFINE_ERROR4)
        assert np.all(cnt == K)
        assert np.array_equal(ids, e["refine_e4_ids"])
        assert np.array_equal(bits(d), bits(e["refine_e4_dists"]))


def test_sq_bundle_sqmd_search(tmp_path):
    g, s, z = load("pq_d128_m32"), load("sq_d128"), load("sqmd_d128")
    queries = g["queries_x16"].astype(np.float32) / np.float32(16.0)
    dim = queries.shape[1]
    bundle.write_sq_bundle(tmp_path / "s", g["graph"], int(g["entry"]), s["sq8_codes"], 8,
                           s["sq8_scale"], s["sq8_zero"], z["sq8_res_codes"],
                           z["sq8_res_qparams"], int(z["sq8_res_bits"]), dim)
    b = bundle.open_bundle(tmp_path / "s")
    say("case sq bundle")
    with Calculator() as calc:
        b.load(calc)
        for corr in (0, 1):
            key = "sq8_c%d_e32_x32_L64" % corr
            ids, d, cnt = calc.search_batch(queries, 64, 64, method=native.METHOD_SQ,
                                            correction=corr, entry_points=32, expand_count=32)
            assert np.array_equal(cnt, z[key + "_counts"])
            for qi in range(len(queries)):
                c = cnt[qi]
                assert np.array_equal(ids[qi, :c], z[key + "_ids"][qi][:c])
                assert np.array_equal(bits(d[qi, :c]), bits(z[key + "_dists"][qi][:c]))
