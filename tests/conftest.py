import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tblis-strassen_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running (full-size configurations)")


def pytest_runtest_logstart(nodeid, location):
    # MISTRA_TEST_TRACE=1: name every test on stdout as it starts (a run that ends abruptly still names it)
    if os.environ.get("MISTRA_TEST_TRACE") == "1":
        print(f"\n[start] {nodeid}", flush=True)
