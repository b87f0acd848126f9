"""Test configuration.

Markers:
  gpu  -- needs an MI355X (gfx950) and the built libecsg_hip.so; these are the
          parity tests proper and call the product through its C ABI.
Everything else runs on the CPU: the oracle (CPU restatement) against the
reference's golden vectors and mathematical invariants, the host logic, and
that the C-ABI library loads and exports its symbols.
  slow -- full-length CPU runs of the golden flows (about 3 minutes each);
          skipped unless ECSG_RUN_SLOW=1.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and libecsg_hip.so")
    config.addinivalue_line("markers", "slow: long CPU runs (minutes)")


def pytest_collection_modifyitems(config, items):
    have_gpu = os.path.exists("/dev/kfd")
    run_slow = os.environ.get("ECSG_RUN_SLOW", "") not in ("", "0")
    for it in items:
        if "gpu" in it.keywords and not have_gpu:
            it.add_marker(pytest.mark.skip(reason="no GPU device (/dev/kfd) on this machine"))
        if "slow" in it.keywords and not run_slow:
            it.add_marker(pytest.mark.skip(reason="long CPU run; set ECSG_RUN_SLOW=1"))
