"""mh_eficool, EFICOOL's stage and machine performance (uncooled branch,
mf:11253-11548), against what the reference printed at the end of the same
run (tests/golden/make_eficool.py): the three-row synthetic decks (rows in
stages 1, 1, 2: two stage records and the whole machine's) and the single
turbine stator row, 100 steps from the host's set-up in the exact-order mode
(the state bit for bit the reference's FP32 state).

The library forms the figures in FP64 from its FP64 SUMFLX integrals; the
reference forms them in its own precision from its sums.  Each figure is
held to its FP64 build's value within 1e-4 (relative for pressures,
temperatures, flows, the pressure ratio and the work; absolute for the
efficiencies and reactions), or within twice the distance between the
reference's FP32 and FP64 builds where that is larger (the FP32 build's
own rounding of sums over a plane).  An efficiency is a ratio over the
stagnation-temperature change DELTO, which the synthetic decks' stages
leave small after 100 steps (0.6 K for the whole three-row machine): a
change of 1e-6 of the mass-averaged temperatures, FP32 rounding, moves it by
|eta| x 1e-6 x TO / |DELTO|, and twice that is also allowed.  The planes, the
stage numbering and the compressor/turbine choice must be the reference's
exactly.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

CASES = ["tiny3_ml", "tiny3_sa", "config0"]
RELATIVE = ["po_in", "po_out", "ps_in", "ps_out", "to_in", "to_out", "flow_in", "flow_out", "work_kw"]
ABSOLUTE = ["eta_tt", "eta_ts", "eta_poly", "reaction", "reaction_hub", "reaction_mid", "reaction_tip"]
TOL = 1e-4


def compare(got: list, ref64: list, ref32: list, tol_rel: float, tol_abs: float, what: str) -> None:
    """Stage-by-stage comparison of mh_eficool records with the reference's
    printed ones (FP64 build, FP32 build for the spread)."""
    assert [r["stage"] for r in got] == [r["stage"] for r in ref64], (what, got, ref64)
    for g, a, b in zip(got, ref64, ref32):
        for k in ("j1", "j2m1", "jmix"):
            assert g[k] == a[k], (what, g["stage"], k, g[k], a[k])
        assert bool(g["compressor"]) == (not a["turbine"]), (what, g["stage"])
        pr = 1.0 / g["po_ratio"] if a["turbine"] else g["po_ratio"]
        checks = [("po_ratio", pr, a["po_ratio_printed"], b["po_ratio_printed"], True)]
        checks += [(k, g[k], a[k], b[k], True) for k in RELATIVE if k in a]
        checks += [(k, g[k], a[k], b[k], False) for k in ABSOLUTE if k in a]
        # an efficiency's sensitivity to a 1e-6 relative change of the stagnation temperatures
        cond = 1e-6 * a["to_in"] / max(abs(a["to_out"] - a["to_in"]), 1e-30)
        for k, x, y, z, rel in checks:
            scale = abs(y) if rel else 1.0
            err, spread = abs(x - y) / scale, abs(z - y) / scale
            bound = max(tol_rel if rel else tol_abs, 2.0 * spread)
            if k.startswith("eta"):
                bound = max(bound, 2.0 * abs(y) * cond)
            print(f"{what} stage {g['stage']} {k}: library {x:.8g} reference {y:.8g} error {err:.2e} "
                  f"(FP32 build {spread:.2e}, bound {bound:.1e})")
            assert err <= bound, (what, g["stage"], k, x, y)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_eficool_vs_reference_printout(case):
    from multall_hip import host
    from multall_hip.runtime import Solver
    with open(os.path.join(GOLDEN, f"eficool_{case}.json")) as f:
        fx = json.load(f)
    st = host.load_deck(os.path.join(GOLDEN, fx["deck"]))
    g = Solver(st, exact_order=True)
    g.loss_init()
    g.start()
    r = g.step(fx["steps"] + 10)
    assert r["nstep"] == fx["steps"] and r["ifend"] == 1
    got = g.eficool()
    g.close()
    compare(got, fx["fp64"], fx["fp32"], TOL, TOL, case)


@pytest.mark.gpu
def test_eficool_default_stages_and_capacity():
    """Without NSTAGE the reference's default numbering (rows 1 and 2 in
    stage 1, row 3 in stage 2; mf:704-706) gives the same records as the
    deck's own; too small an output array is refused and says how many
    records there are."""
    import ctypes
    from multall_hip import host
    from multall_hip.runtime import MhPerf, Solver
    st = host.load_deck(os.path.join(GOLDEN, "decks", "tiny3_ml.dat"))
    assert list(st["NSTAGE"]) == [1, 1, 2]
    a = Solver(st)
    a.loss_init()
    a.start()
    a.step(20)
    ra = a.eficool()
    b = Solver({k: v for k, v in st.items() if k != "NSTAGE"})
    b.loss_init()
    b.start()
    b.step(20)
    rb = b.eficool()
    # (the machine record's reactions are NaN: compare as text)
    assert json.dumps(ra) == json.dumps(rb) and [x["stage"] for x in ra] == [1, 2, 0]
    out = (MhPerf * 2)()
    n = ctypes.c_int(0)
    rc = a.L.mh_eficool(a.h, out, 2, ctypes.byref(n))
    assert rc != 0 and n.value == 3
    for x in ra:
        for k, v in x.items():
            assert np.isfinite(v) or k.startswith("reaction"), (k, v)
    a.close()
    b.close()
