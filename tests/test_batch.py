"""mh_step_batch: several machines stepped concurrently, each on its own HIP
stream and host thread, give bit for bit what each gives stepped alone."""
import numpy as np
import pytest

from conftest import load_state

CASES = ["tiny3_ml", "tiny3_sa", "tiny1_inv", "tiny3_nl"]


@pytest.mark.gpu
def test_batch_equals_sequential():
    from multall_hip.runtime import Solver, step_batch
    states = [load_state(c) for c in CASES]
    for st in states:
        st["scalars"]["NSTEPS_MAX"] = 100

    def fresh():
        out = []
        for st in states:
            g = Solver(st)
            g.start()
            out.append(g)
        return out

    alone = fresh()
    r_alone = [g.step(37) for g in alone]
    together = fresh()
    r_together = step_batch(together, 37)
    assert r_alone == r_together
    for c, a, b in zip(CASES, alone, together):
        for v in ("RO", "ROVX", "ROVR", "RORVT", "ROE", "XFORCE", "QSOURCE"):
            x, y = a.get(v), b.get(v)
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (c, v)


@pytest.mark.gpu
def test_batch_rejects_a_context_twice():
    from multall_hip.runtime import MhError, Solver, step_batch
    st = load_state("tiny1_inv")
    st["scalars"]["NSTEPS_MAX"] = 10
    g = Solver(st)
    g.start()
    with pytest.raises(MhError):
        step_batch([g, g], 1)
