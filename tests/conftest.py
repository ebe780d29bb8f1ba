import os
import sys

import pytest
# torch first: importing it loads its bundled HIP runtime, and libalma loaded
# afterwards binds to that same runtime (one HIP runtime per process), so
# tests that hand torch device tensors to the library share one device view.
import torch  # noqa: F401

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "protein-structure-analysis_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: longer CPU test")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def ref_module():
    """The reference module built from its own sources into oracle/_ref
    (golden-vector harness); skipped where it was not built."""
    import glob
    import importlib.machinery
    import importlib.util
    hits = glob.glob(os.path.join(ROOT, "oracle", "_ref", "alma_ref*.so"))
    if not hits:
        pytest.skip("reference harness not built (oracle/_ref)")
    loader = importlib.machinery.ExtensionFileLoader("alma_ref", hits[0])
    spec = importlib.util.spec_from_loader("alma_ref", loader)
    mod = importlib.util.module_from_spec(spec)
    loader.exec_module(mod)
    return mod
