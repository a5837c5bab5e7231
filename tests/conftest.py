import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "gpu-path-delay-coverage_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")

# torch ships its own HIP runtime; the ranking tests pass torch tensors to the
# library, so torch is loaded before the library in every test process
try:
    import torch  # noqa: F401
except ImportError:
    pass
