"""The test-only torch.autograd.Function over the loss kernels, in a child
process (see tests/autograd_check.py)."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_autograd_function_in_subprocess():
    here = os.path.dirname(os.path.abspath(__file__))
    r = subprocess.run([sys.executable, os.path.join(here, "autograd_check.py")], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0 and "OK" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
