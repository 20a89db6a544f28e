"""The tuned tile shapes: every shape of the damping/corner-distribution
kernel (2, 4 or 8 K lines per workgroup) and of SMOOTH_VAR's register form
(8 or 16 wavefronts per J plane) gives the same bits as the fixed shapes, in
both reduction modes, on decks with one and with four pitchwise chunks
(IM - 1 <= 64 and IM = 72), and the tuner's own choice does too."""
import numpy as np
import pytest

from conftest import load_state

VARS = ["RO", "ROVX", "ROVR", "RORVT", "ROE", "DRO", "DROE", "STEP"]


def _run(case, exact, steps, **scalars):
    from multall_hip.runtime import Solver
    st = load_state(case)
    st["scalars"]["NSTEPS_MAX"] = steps
    g = Solver(st, exact_order=exact)
    for k, v in scalars.items():
        g.set_scalar(k, v)
    g.start()
    r = g.step(steps + 10)
    out = {v: g.get(v).copy() for v in VARS}
    g.close()
    return r, out


@pytest.mark.gpu
@pytest.mark.parametrize("case,exact", [("tiny3_sa", False), ("tiny3_ml", True), ("tiny3_wide", False),
                                        ("tiny3_sss", False)])
def test_tile_shapes_give_the_same_bits(case, exact):
    r0, base = _run(case, exact, 30, TUNE=0)
    runs = [dict(SCAT_TK=tk, SMOOTH_WAVES=nw) for tk in (2, 4, 8) for nw in (8, 16)] + [dict(TUNE=1)]
    for sc in runs:
        r, out = _run(case, exact, 30, **sc)
        assert r == r0, (case, sc)
        for v in VARS:
            assert np.array_equal(out[v].view(np.uint32), base[v].view(np.uint32)), (case, sc, v)


@pytest.mark.gpu
def test_bad_forced_shapes_are_refused():
    from multall_hip.runtime import MhError, Solver
    st = load_state("tiny3_ml")
    g = Solver(st)
    g.set_scalar("SCAT_TK", 3)
    with pytest.raises(MhError):
        g.start()
    g.set_scalar("SCAT_TK", 0)
    g.set_scalar("SMOOTH_WAVES", 12)
    with pytest.raises(MhError):
        g.start()
    g.close()
