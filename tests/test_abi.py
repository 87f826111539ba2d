"""The C-ABI library loads and exports every function include/rover_gpu.h
declares (no compute, no device needed)."""
import ctypes
import os
import re

from rover_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "rover_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"typedef[^;]*;", "", text)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "sizeof", "PurePqDistanceEvaluatorFn"}
    return sorted(set(n for n in names if n not in skip and not n.isupper()))


def test_library_exports_header():
    lib = native.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    bound = {s[0] for s in native.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_seed_ids_host_only():
    # host-side helper, no device involved
    ids = native.make_seed_ids(1000, 4, 5, 42)
    assert ids[0] == 5 and len(ids) == 4


REFERENCE_TREE = os.environ.get("ROVER_REFERENCE_TREE", "/root/reference")
REF_HEADER = os.path.join(REFERENCE_TREE, "src", "dsp_demo", "retrieval_sqmd_pure_pq.h")


def test_struct_layout_matches_reference_header(tmp_path):
    """The library-search structs of rover_gpu.h have the layout of the
    reference's own retrieval_sqmd_pure_pq.h (sizes, every field offset, the
    correction-mode values), checked by compiling tests/c/layout_probe.c
    against each header and comparing.  Needs a Rover source tree."""
    import pytest
    import subprocess
    if not os.path.exists(REF_HEADER):
        pytest.skip("no Rover source tree at %s" % REFERENCE_TREE)
    c = os.path.join(ROOT, "tests", "c")
    objs = []
    for name, flags in (("ref", ["-DLAYOUT_REFERENCE", "-I", os.path.join(REFERENCE_TREE, "src"),
                                 "-I", os.path.join(REFERENCE_TREE, "config")]),
                        ("ours", ["-I", os.path.join(ROOT, "include")])):
        o = str(tmp_path / (name + ".o"))
        subprocess.run(["gcc", "-std=gnu11", "-c", os.path.join(c, "layout_probe.c"), "-o", o]
                       + flags, check=True)
        objs.append(o)
    exe = str(tmp_path / "layout")
    subprocess.run(["gcc", os.path.join(c, "layout_main.c")] + objs + ["-o", exe], check=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 0, r.stdout
    assert "equal" in r.stdout
