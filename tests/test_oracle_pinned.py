"""The C restatement (oracle/) reproduces the compiled reference bit-for-bit.

Golden vectors: tests/golden/<case>_ref<N>.npz, written by make_golden.py from
the reference solver itself (amdflang -O3, FP32) after N time steps of the
same deck; inputs are the reference's own post-SETUP COMMON state.
"""
import numpy as np
import pytest

import orc
from conftest import load_ref, load_state, long_steps

VARS = ["RO", "ROVX", "ROVR", "RORVT", "ROE", "DRO", "DROVX", "DROVR", "DRORVT", "DROE", "STEP",
        "XFORCE", "RFORCE", "TFORCE", "QSOURCE"]

# inviscid decks; mixing-length (LOSS) decks covering the chord-Reynolds and
# power-law laminar viscosity, smooth and rough walls, a rotor tip gap, the
# fixed wall y+, a shear-free hub, exit smoothing, IM = 72, the SSS scheme
# with residual smoothing (ITIMST = -3) and mass-flow forcing (IN_FLOW = 2
# and 3), a deck whose inter-row gaps the reader re-grids (ISHIFT = 2, new
# trailing-edge cusps, input grid angles); the inlet and exit condition
# options (IPOUT = 1, -1, 3; IN_PRESS = -3, 1, 4; IN_VTAN = 1, 2; IN_VR = -3,
# 0; inlet yaw and pitch angles), an inviscid deck whose flow relative to the
# second row is supersonic at the first mixing plane (NEW_MIXPLAN's
# Prandtl-Meyer sweeps) and a deck with a printout step inside the run (the
# viscous refresh at a printout step); NEW_LOSS decks
# (tip gap; power-law viscosity, rough walls, shear-free hub and transition
# points with a threshold that leaves part of the flow laminar); and the
# Spalart-Allmaras decks below
CASES = ["tiny1_inv", "tiny3_inv", "tiny3_ml", "tiny3_ml2", "tiny1_mlw", "tiny3_sa", "tiny3_sa2", "tiny3_exit", "tiny3_wide",
         "tiny3_nl", "tiny3_nl2", "tiny3_sss", "tiny3_flow2", "tiny3_flow3", "tiny3_ml3", "tiny3_nl3", "tiny3_sa3",
         "tiny3_gap", "tiny3_bc1", "tiny3_bc2", "tiny2_bc3", "tiny3_sup", "tiny3_nout", "tiny3_cusp2", "tiny2_rep"]
# Spalart-Allmaras decks (rotor tip gap; power-law viscosity, rough walls,
# shear-free casing and the mixing-length relaxation source) also pin the
# turbulence state
SA_VARS = ["TRANS_DYN_VIS", "DEL_DYNVIS", "TRANS_KVIS", "T_SOURCE", "Y_PLUS"]
# decks with the Shih wall functions also pin the smoothed wall pressure
# gradient, y+ and the wall functions' own y+ state
WALL_VARS = ["DPDS_CELL", "YPLUS_I1", "YPLUS_IM", "YPLUS_K1", "YPLUS_KM"]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("nsteps", [1, 100])
def test_oracle_matches_reference_bitwise(case, nsteps):
    nsteps = long_steps(case) if nsteps == 100 else nsteps
    st = load_state(case)
    st["scalars"]["NSTEPS_MAX"] = nsteps
    o = orc.Oracle(st)
    o.start()
    assert o.step(nsteps + 10) == nsteps          # the run ends at NSTEPS_MAX like the reference
    ref = load_ref(case, nsteps)
    for v in VARS + [v for v in SA_VARS + WALL_VARS if v in ref]:
        a, b = o[v], ref[v]
        bad = np.count_nonzero(a.view(np.uint32) != b.view(np.uint32))
        assert bad == 0, f"{case} {nsteps} steps: {v} differs in {bad} of {a.size} values"


def test_oracle_freestream_preserved():
    """Uniform state on closed cells gives zero flux residual: with the inlet
    and exit conditions matching the uniform state, one step leaves U unchanged
    to rounding (area closure, mf:6707-6720)."""
    st = load_state("tiny1_inv")
    # the closure identity: sum of the six face areas of each cell vanishes
    aqx, abx, asx = st["AQX"], st["ABX"], st["ASX"]
    aqr, abr, asr = st["AQR"], st["ABR"], st["ASR"]
    km, jm, im = aqx.shape
    xa = aqx[:-1, 1:, :-1] - aqx[:-1, :-1, :-1] + abx[:-1, 1:, :-1] - abx[:-1, 1:, 1:] \
        + asx[:-1, 1:, :-1] - asx[1:, 1:, :-1]
    ra = aqr[:-1, 1:, :-1] - aqr[:-1, :-1, :-1] + abr[:-1, 1:, :-1] - abr[:-1, 1:, 1:] \
        + asr[:-1, 1:, :-1] - asr[1:, 1:, :-1]
    scale = np.abs(aqx).max()
    assert np.abs(xa).max() < 1e-5 * scale
    # radial closure leaves exactly -8*VOLOR (mf:6694-6700)
    assert np.allclose(-0.125 * ra, st["VOLOR"][:-1, 1:, :-1], rtol=1e-5, atol=1e-6 * np.abs(ra).max())


@pytest.mark.parametrize("case", ["tiny3_inv", "tiny3_ml", "tiny3_sa", "tiny3_wide"])
def test_hundred_step_ulp_spread(case):
    """How far 100 FP32 steps of the reference's arithmetic (the oracle) carry
    a one-ulp change of a single initial value: the noise floor below which a
    100-step difference from the reference means nothing.  The GPU parity
    tests' 100-step bound (2.5e-5 of each variable's scale) must stay at least
    twice this spread."""
    from test_gpu_parity import TOL_100, _scales
    st = load_state(case)
    st["scalars"]["NSTEPS_MAX"] = 100
    o = orc.Oracle(st)
    o.start()
    assert o.step(110) == 100
    base = {v: o[v].copy() for v in ["RO", "ROVX", "ROVR", "RORVT", "ROE"]}
    sc = _scales(base)
    spread = 0.0
    for field, (k, j, i) in (("RO", (4, 30, 5)), ("ROE", (2, 10, 3)), ("ROVX", (6, 50, 7))):
        st2 = {n: (np.array(v) if n != "scalars" else dict(v)) for n, v in st.items()}
        a = st2[field]
        k, j, i = min(k, a.shape[0] - 1), min(j, a.shape[1] - 1), min(i, a.shape[2] - 1)
        a[k, j, i] = np.nextafter(a[k, j, i], np.float32(np.inf))
        o2 = orc.Oracle(st2)
        o2.start()
        o2.step(110)
        for v in base:
            spread = max(spread, float(np.abs(o2[v].astype(np.float64) - base[v]).max() / sc[v]))
    print(f"{case}: one-ulp spread after 100 steps {spread:.2e}")
    assert 1e-7 < spread <= TOL_100 / 2, spread

