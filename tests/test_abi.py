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
