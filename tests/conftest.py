"""Shared test setup: import paths and the `gpu` marker.

Tests marked `gpu` need a HIP device (an MI355X); they call the product
library through its C ABI.  Everything else runs on the CPU."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "gdistger-re_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls the C ABI")
