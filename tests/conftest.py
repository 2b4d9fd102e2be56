"""Test configuration: import paths and the `gpu` marker.

Tests marked `gpu` need a HIP device (an MI355X); everything else runs on a
CPU-only host.  The oracle (oracle/) is the checker; the product is
ramses-cpp_amd/librhydro.so, reached through its C-ABI.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("ramses-cpp_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls the C-ABI library")
