import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def _ensure_built():
    # Tests use what build() produced; they build only when nothing is there.
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if not os.path.exists(os.path.join(ROOT, "beep_amd", "libbeep.so")):
        subprocess.check_call(["make", "-s", "-j8", "-C", os.path.join(ROOT, "beep_amd", "csrc")])


_ensure_built()


def pytest_collection_modifyitems(config, items):
    # GPU tests run only where a HIP device is visible; CPU runs select
    # -m "not gpu" explicitly anyway.
    pass
