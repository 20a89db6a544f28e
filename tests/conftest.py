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

# The long comparison runs 100 steps, except on the supersonic cascade: there
# a one-ulp change of any initial value grows to ~1e-3 of the solution
# between steps 20 and 30 (a sonic switch flips at one node), so it is
# compared after 20 steps (tests/golden/make_golden.py STEPS_OF).
LONG_STEPS = {"tiny3_sup": 20}


def long_steps(case):
    return LONG_STEPS.get(case, 100)


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
