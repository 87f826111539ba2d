import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def pytest_collection_modifyitems(session, config, items):
    """torch ships its own HIP runtime next to the system one that
    librover.so links; in one process torch's must initialize first (some
    GPU tests build graphs with torch after the library has run).  Touch
    torch's device once before any GPU test when GPU tests are selected."""
    if "not gpu" in (config.getoption("markexpr") or ""):
        return
    if any(item.get_closest_marker("gpu") for item in items):
        try:
            import torch
            if torch.cuda.is_available():
                torch.zeros(1, device="cuda:0")
        except Exception:
            pass


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    out = {k: z[k] for k in z.files}
    if "base_x16" in out:
        out["base"] = out["base_x16"].astype(np.float32) / np.float32(16.0)
        out["queries"] = out["queries_x16"].astype(np.float32) / np.float32(16.0)
    return out


@pytest.fixture(scope="session")
def golden():
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = load_golden(name)
        return cache[name]
    return get
