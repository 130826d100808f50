import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def _ensure_built():
    """Build the oracle (gcc) and the product library (hipcc) when absent."""
    if not os.path.exists(os.path.join(ROOT, "oracle", "libcfd_oracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    if not os.path.exists(os.path.join(ROOT, "cfd_amd", "libcfd_amd.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "cfd_amd", "csrc")], check=True)
    drv = os.path.join(ROOT, "cfd_amd", "fortran", "littleharsh_gpu")
    if not os.path.exists(drv) and os.path.exists("/opt/rocm/bin/amdflang"):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "cfd_amd", "fortran")], check=False)


_ensure_built()
