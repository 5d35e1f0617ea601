import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _ensure_built():
    """Build the oracle (and the HIP library) only when nothing is built."""
    import subprocess
    if not os.path.exists(os.path.join(REPO, "oracle", "liboracle.so")):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle")], check=True)


_ensure_built()


@pytest.fixture(scope="session")
def golden():
    import json
    import glob
    out = {}
    for f in sorted(glob.glob(os.path.join(REPO, "tests", "golden", "*.json"))):
        with open(f) as fh:
            fx = json.load(fh)
        if "case" in fx:  # the full-size fixtures are loaded by name
            out[fx["case"]] = fx
    return out
