"""GPU parity: the HIP library against the oracle and the reference.

- exact-order mode, 1 step: every field bit-identical to the C restatement
  except where a transcendental (powf in the inlet condition and the Mach
  limiter, asinf/atan2f/sinf/cosf in NEW_MIXPLAN) enters, which device and
  host libm round differently; those cases get a 2e-6 relative bound.
- both modes, 1 and 100 steps against the reference's own output:
  max|gpu-ref| / scale, where scale is max|ref| of the variable for RO and
  ROE, of the momentum magnitude |rho V| for ROVX and ROVR, and of |RORVT|
  for RORVT (a near-zero component such as the radial momentum is measured
  against the momentum it is part of, not against itself).  After 1 step the
  bound is 1e-5 (observed ~1e-7).  After 100 steps it is TOL_100 = 2.5e-5:
  FP32 time marching carries a single-ulp change of one initial value to
  2.5e-6 .. 1.1e-5 of the scale after 100 steps of the reference's own
  arithmetic (test_oracle_pinned.py::test_hundred_step_ulp_spread measures it
  with the oracle and holds it below TOL_100 / 2), so a few ulps of
  difference anywhere (transcendentals, or the fast mode's FP64 sums) cannot
  be told apart from the reference beyond that.
"""
import numpy as np
import pytest

import orc
from conftest import load_ref, load_state, long_steps

VARS = ["RO", "ROVX", "ROVR", "RORVT", "ROE", "DRO", "DROVX", "DROVR", "DRORVT", "DROE", "STEP"]
FORCES = ["XFORCE", "RFORCE", "TFORCE", "QSOURCE"]
TOL = 1e-5
TOL_100 = 2.5e-5
# inviscid decks; the last one's flow relative to the second row is
# supersonic at the first mixing plane, so NEW_MIXPLAN's Prandtl-Meyer angle
# sweeps run
INVISCID = ["tiny1_inv", "tiny3_inv", "tiny3_sup"]
# mixing-length (LOSS) decks: chord-Reynolds and power-law viscosity, smooth
# and rough walls, rotor tip gap, fixed wall y+ with a shear-free hub, exit
# smoothing on the last three J planes, IM = 72, the SSS scheme with
# residual smoothing (ITIMST = -3), mass-flow forcing towards the mean flow
# (IN_FLOW = 2) and towards a fixed flow (IN_FLOW = 3), the Shih wall
# functions without the pressure-gradient term (YPLUSWALL = -6), re-gridded
# gaps, the inlet and exit condition options (IPOUT = 1, -1, 3; IN_PRESS =
# -3, 1, 4; IN_VTAN = 1, 2; IN_VR = -3, 0; inlet yaw and pitch) and printout
# steps inside the run (viscous refreshes at NOUT)
MIXING_LENGTH = ["tiny3_ml", "tiny3_ml2", "tiny1_mlw", "tiny3_exit", "tiny3_wide", "tiny3_sss", "tiny3_flow2",
                 "tiny3_flow3", "tiny3_ml3", "tiny3_gap", "tiny3_bc1", "tiny3_bc2", "tiny2_bc3", "tiny3_nout",
                 "tiny3_cusp2", "tiny2_rep"]
# Spalart-Allmaras decks: rotor tip gap; power-law viscosity, rough walls,
# shear-free casing, mixing-length relaxation source; the Rotor 37 test
# deck's options (Shih wall functions with the pressure-gradient term,
# YPLUSWALL = -11, and the helicity and pressure-gradient source factors)
# with transition points
SA = ["tiny3_sa", "tiny3_sa2", "tiny3_sa3"]
SA_VARS = ["TRANS_DYN_VIS", "TRANS_KVIS", "T_SOURCE"]
# NEW_LOSS decks: rotor tip gap; power-law viscosity, rough walls, shear-free
# hub, transition points and a threshold that leaves part of the flow
# laminar; the Shih wall functions with the pressure-gradient term
NEW_LOSS = ["tiny3_nl", "tiny3_nl2", "tiny3_nl3"]
# the wall functions' smoothed pressure gradient and y+ state
WALL_VARS = ["DPDS_CELL", "YPLUS_I1", "YPLUS_IM", "YPLUS_K1", "YPLUS_KM"]
WALLFUN = ["tiny3_ml3", "tiny3_nl3", "tiny3_sa3"]
VISCOUS = MIXING_LENGTH + NEW_LOSS + SA


def _solver(st, exact):
    from multall_hip.runtime import Solver
    s = Solver(st, exact_order=exact)
    s.start()
    return s


def _nerr(a, b, scale=None):
    a = np.asarray(a).reshape(np.shape(b))   # the library returns non-node arrays flat
    scale = np.abs(b).max() if scale is None else scale
    return float(np.abs(a.astype(np.float64) - b).max() / max(scale, 1e-30))


def _scales(ref):
    rovt = ref["RORVT"]  # |rho V| bound: components are comparable in size
    mom = np.sqrt(ref["ROVX"].astype(np.float64) ** 2 + ref["ROVR"].astype(np.float64) ** 2).max()
    return {"RO": np.abs(ref["RO"]).max(), "ROE": np.abs(ref["ROE"]).max(), "ROVX": mom, "ROVR": mom,
            "RORVT": np.abs(rovt).max()}


@pytest.mark.gpu
@pytest.mark.parametrize("case", INVISCID + VISCOUS)
def test_one_step_vs_oracle_exact_order(case):
    st = load_state(case)
    st["scalars"]["NSTEPS_MAX"] = 1
    o = orc.Oracle(st)
    o.start()
    o.step(1)
    g = _solver(st, exact=True)
    g.step(1)
    for v in VARS:
        a, b = g.get(v), o[v]
        ndiff = int(np.count_nonzero(a != b))
        err = _nerr(a, b)
        print(f"{case} {v}: {ndiff} differing values, normalized max err {err:.2e}")
        assert err <= 2e-6, (v, err)
    # the step ends the run, so LOSS has refreshed the forces (logf and powf
    # differ by ulps between device and host libm)
    for v in FORCES + (SA_VARS if case in SA else []) + (WALL_VARS if case in WALLFUN else []):
        err = _nerr(g.get(v), o[v])
        print(f"{case} {v}: normalized max err {err:.2e}")
        assert err <= 1e-5, (v, err)


@pytest.mark.gpu
@pytest.mark.parametrize("case", VISCOUS)
def test_loss_call_vs_oracle(case):
    """One LOSS / NEW_LOSS / SPAL_LOSS call from the post-SETUP state
    (RF_VIS = 1, FMIXUP = 1, NSTEP = 1, as SETUP makes it): forces, y+, the
    laminar viscosity, (NEW_LOSS, SA) the viscosity ratio after the
    transition test and (SA) the turbulent viscosity and its source."""
    st = load_state(case)
    st["scalars"]["NSTEPS_MAX"] = 100
    o = orc.Oracle(st)
    for k, v in (("RF_VIS", 1.0), ("FMIXUP", 1.0), ("NSTEP", 1)):
        o.lib.orc_scalar(o.h, k.encode(), float(v))
    if case in SA:
        o.lib.orc_spal_loss(o.h)
    elif case in NEW_LOSS:
        o.lib.orc_new_loss(o.h)
    else:
        o.lib.orc_loss(o.h)
    from multall_hip.runtime import Solver
    g = Solver(st, exact_order=True)
    g.loss_init()
    vis = o.scalar("VISLAM")
    if st["scalars"]["REYNO"] >= 0.0001:
        assert abs(g.scalar("VISLAM") - vis) <= 1e-6 * abs(vis), (g.scalar("VISLAM"), vis)
    extra = SA_VARS + ["VISC_RAT", "VISC_LAM"] if case in SA else ["VISC_RAT", "VISC_LAM"] if case in NEW_LOSS else []
    extra = extra + (WALL_VARS if case in WALLFUN else [])
    for v in FORCES + ["Y_PLUS"] + extra:
        b = o[v]
        a = g.get(v).reshape(b.shape)
        err = _nerr(a, b)
        print(f"{case} {v}: {int(np.count_nonzero(a != b))} differing values, normalized max err {err:.2e}")
        assert err <= 1e-5, (v, err)


@pytest.mark.gpu
@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("case", INVISCID + VISCOUS)
@pytest.mark.parametrize("nsteps", [1, 100])
def test_vs_reference(case, nsteps, exact):
    nsteps = long_steps(case) if nsteps == 100 else nsteps
    st = load_state(case)
    st["scalars"]["NSTEPS_MAX"] = nsteps
    g = _solver(st, exact=exact)
    r = g.step(nsteps + 10)
    assert r["nstep"] == nsteps and r["ifend"] == 1
    ref = load_ref(case, nsteps)
    sc = _scales(ref)
    for v in ["RO", "ROVX", "ROVR", "RORVT", "ROE"]:
        err = _nerr(g.get(v), ref[v], sc[v])
        print(f"{case} {nsteps} exact={exact} {v}: scaled max err {err:.2e}")
        assert err <= (TOL if nsteps == 1 else TOL_100), (v, err)
    for v in [v for v in SA_VARS if v in ref]:
        # T_SOURCE is production minus destruction of the turbulent
        # viscosity: the cancellation amplifies the state's ~1e-6 rounding
        # differences (device and host logf/powf), so it gets 1e-4
        tol = 1e-4 if v == "T_SOURCE" else TOL
        err = _nerr(g.get(v), ref[v])
        print(f"{case} {nsteps} exact={exact} {v}: normalized max err {err:.2e}")
        assert err <= tol, (v, err)
    if nsteps > 1:
        # SUMFLX's per-J integrals against what the reference's EFICOOL wrote
        # at the end of the same run (FP64 here, FP32 sequential there).
        # The entropy function
        # CP*ln(T/TO1) - R*ln(p/PO1) is a small difference of two terms that
        # are ~100x larger in a nearly loss-free (inviscid) run, so the
        # state's few-1e-6 rounding differences reach its integral amplified
        # by that factor: 2e-4 there, 5e-5 for the other integrals.
        integ = g.integrals()
        names = ["BLADE_FLOW", "SUMPO", "SUMTO", "SUMRVT", "SUM_ENTPY"]
        for q, name in enumerate(names):
            err = _nerr(integ[q], ref["SUMFLX"][q])
            print(f"{case} exact={exact} {name}: normalized max err {err:.2e}")
            assert err <= (2e-4 if name == "SUM_ENTPY" else 5e-5), (name, err)
    for v in ["DRO", "DROVX", "DROVR", "DRORVT", "DROE", "STEP"]:
        print(f"{case} {nsteps} exact={exact} {v}: normalized max err {_nerr(g.get(v), ref[v]):.2e} (reported)")


@pytest.mark.gpu
@pytest.mark.parametrize("case", INVISCID + VISCOUS)
def test_host_path_vs_reference(case):
    """The whole drop-in path: the deck read and set up by the Fortran host,
    the SETUP-time LOSS call on the GPU (forces against the reference's
    post-SETUP state), then 100 steps (the supersonic cascade 20) against
    the reference."""
    import os
    from conftest import GOLDEN
    from multall_hip import host
    from multall_hip.runtime import Solver
    st = host.load_deck(os.path.join(GOLDEN, "decks", f"{case}.dat"))
    n = long_steps(case)
    assert st["scalars"]["NSTEPS_MAX"] == n
    g = Solver(st, exact_order=False)
    g.loss_init()
    ref0 = load_state(case)
    for v in FORCES:
        err = _nerr(g.get(v), ref0[v])
        print(f"{case} set-up {v}: normalized max err {err:.2e}")
        assert err <= 1e-5, (v, err)
    g.start()
    r = g.step(110)
    assert r["nstep"] == n and r["ifend"] == 1
    ref = load_ref(case, n)
    sc = _scales(ref)
    for v in ["RO", "ROVX", "ROVR", "RORVT", "ROE"]:
        err = _nerr(g.get(v), ref[v], sc[v])
        print(f"{case} host path {n} steps {v}: scaled max err {err:.2e}")
        assert err <= TOL_100, (v, err)
