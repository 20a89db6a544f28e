"""Steps queued as captured HIP graphs (the default) against the same steps
launched one by one (USE_GRAPHS = 0), bit for bit.

The graph path replays each captured chunk of steps many times, with the
library's own eager work (uploads, the set-up viscous call, read-backs) and
host synchronisations between the calls; every comparison is with the
launch-by-launch result of the same run.  The step state (step count, stop
tests, ramps) lives on the device, so the chunks are replayed unchanged from
step to step: the comparisons cover the ramp phase (NCHANGE), the every-5-step
checks, the viscous refreshes, printout steps, a stop request and a run that
ends by convergence.
"""
import numpy as np
import pytest

from conftest import load_state

VARS = ["RO", "ROVX", "ROVR", "RORVT", "ROE", "DRO", "DROVX", "DROVR", "DRORVT", "DROE", "STEP",
        "XFORCE", "RFORCE", "TFORCE", "QSOURCE"]


def _run(case, graphs, exact, pieces, nsteps=100, stop_after=None):
    from multall_hip.runtime import Solver
    st = load_state(case)
    st["scalars"]["NSTEPS_MAX"] = nsteps
    g = Solver(st, exact_order=exact)
    g.set_scalar("USE_GRAPHS", 1 if graphs else 0)
    g.start()
    out = []
    for i, n in enumerate(pieces):
        if stop_after is not None and i == stop_after:
            g.request_stop()
        out.append(g.step(n))
        g.sync()
    return g, out


@pytest.mark.gpu
@pytest.mark.parametrize("case,exact", [("tiny3_sa", False), ("tiny3_ml", True), ("tiny3_nout", False),
                                        ("tiny3_sss", False)])
def test_graph_replays_equal_launches(case, exact):
    pieces = [7, 13, 1, 4, 25, 60]
    a, ra = _run(case, True, exact, pieces)
    b, rb = _run(case, False, exact, [100])
    assert ra[-1] == rb[-1] and ra[-1]["nstep"] == 100 and ra[-1]["ifend"] == 1
    for v in VARS:
        x, y = a.get(v), b.get(v)
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (case, v)


@pytest.mark.gpu
def test_stop_request_ends_at_the_next_tenth_step():
    """mh_request_stop after 13 steps: the run ends at step 20 (the reference
    reads `stopit` on steps divisible by 10), in both forms."""
    a, ra = _run("tiny3_ml", True, False, [13, 100], nsteps=100, stop_after=1)
    b, rb = _run("tiny3_ml", False, False, [13, 100], nsteps=100, stop_after=1)
    assert ra[-1]["nstep"] == 20 and ra[-1]["ifend"] == 1, ra
    assert ra == rb
    for v in VARS:
        assert np.array_equal(a.get(v).view(np.uint32), b.get(v).view(np.uint32)), v


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["tiny3_sa", "tiny3_sa2", "tiny3_nl", "tiny3_ml"])
def test_concurrent_equations_equal_one_stream(case):
    """Fast mode with the equations on three streams (LANES = 1, the default:
    energy and the turbulence transport beside continuity, then the three
    momentum equations side by side) against all of them in order on one
    stream (LANES = 0), 100 steps, bit for bit."""
    from multall_hip.runtime import Solver
    runs = []
    for lanes in (1, 0):
        st = load_state(case)
        st["scalars"]["NSTEPS_MAX"] = 100
        g = Solver(st, exact_order=False)
        g.set_scalar("LANES", lanes)
        g.start()
        runs.append((g, g.step(100)))
    (a, ra), (b, rb) = runs
    assert ra == rb and ra["nstep"] == 100
    names = VARS + (["TRANS_DYN_VIS", "TRANS_KVIS"] if load_state(case)["scalars"]["ILOS"] >= 200 else [])
    for v in names:
        assert np.array_equal(a.get(v).view(np.uint32), b.get(v).view(np.uint32)), (case, v)
