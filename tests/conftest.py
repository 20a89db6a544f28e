import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "multall-open-18.3_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_state(case):
    z = np.load(os.path.join(GOLDEN, f"{case}_setup.npz"), allow_pickle=False)
    st = {k: z[k] for k in z.files if k != "__scalars__"}
    st["scalars"] = json.loads(str(z["__scalars__"]))
    return st


def load_ref(case, n):
    z = np.load(os.path.join(GOLDEN, f"{case}_ref{n}.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


@pytest.fixture
def state_loader():
    return load_state


@pytest.fixture
def ref_loader():
    return load_ref
