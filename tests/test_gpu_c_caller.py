"""A C program that includes only rover_gpu.h and links librover.so makes
the reference's library search call (retrieval_pure_pq_search_scalar and
_with_policy with the residual evaluator) and gets the compiled reference's
golden results: ids, distance bits and stats (tests/c/search_once.c)."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from rover_amd import native

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("c") / "search_once")
    libdir = os.path.dirname(native.LIB_PATH)
    subprocess.run(["gcc", "-std=gnu11", "-O2", os.path.join(ROOT, "tests", "c", "search_once.c"),
                    "-I", os.path.join(ROOT, "include"), "-L", libdir, "-lrover",
                    "-Wl,-rpath," + libdir, "-o", out], check=True)
    return out


@pytest.mark.parametrize("mode", [0, 1])
def test_c_caller_gets_reference_results(golden, exe, tmp_path, mode):
    g = golden("pq_d128_m32_r8")
    n, degree = g["graph"].shape
    m, ks, sub = g["codebook"].shape
    dim = m * sub
    ep, expand, L = 32, 32, 64
    key = "m%d_e%d_x%d_L%d" % (mode, ep, expand, L)
    entries = native.make_seed_ids(n, ep, int(g["entry"]))
    for qi in range(3):
        say("case C caller mode=%d query=%d" % (mode, qi))
        path = tmp_path / ("in%d_%d.bin" % (mode, qi))
        with open(path, "wb") as f:
            f.write(struct.pack("<13i", n, degree, dim, m, ks, g["codes"].shape[1], len(entries),
                                expand, L, L, mode, int(g["res_bits"]), 300))
            for a, t in ((g["graph"], "<i4"), (g["codes"], "u1"), (g["codebook"], "<f4"),
                         (g["queries"][qi], "<f4"), (entries, "<i4")):
                f.write(np.ascontiguousarray(a, dtype=t).tobytes())
            if mode:
                f.write(np.ascontiguousarray(g["res_codes"], dtype="u1").tobytes())
                f.write(np.ascontiguousarray(g["res_qparams"], dtype="<f4").tobytes())
        r = subprocess.run([exe, str(path)], stdout=subprocess.PIPE, text=True, timeout=120)
        assert r.returncode == 0, r.stdout
        lines = r.stdout.split("\n")
        head = lines[0].split()
        count = int(head[1])
        assert count == int(g[key + "_counts"][qi])
        assert [int(v) for v in head[3:7]] == [int(v) for v in g[key + "_stats"][qi]]
        got = [l.split() for l in lines[1:1 + count]]
        assert [int(a) for a, _ in got] == [int(v) for v in g[key + "_ids"][qi][:count]]
        want_bits = np.ascontiguousarray(g[key + "_dists"][qi][:count],
                                         dtype=np.float32).view(np.uint32)
        assert [int(b, 16) for _, b in got] == [int(v) for v in want_bits]
