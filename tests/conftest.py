import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: long CPU check, enabled with BPMAX_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("BPMAX_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow; set BPMAX_SLOW=1")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)
