"""Converged accuracy against the FP64 build of the reference.

The library is run to convergence (the reference's own stop test) on the
reference's Rotor 37 deck (NASA transonic compressor rotor, 45 x 150 x 49
nodes, 36 blades), as it ships (Spalart-Allmaras with its helicity and
pressure-gradient source factors and the Shih wall functions with the
pressure-gradient term) and with plain Spalart-Allmaras and the log-law
walls, and on two of the reference's one-stage machines as they ship (the
axial compressor and the axial turbine: rotor and stator across a mixing
plane, gaps re-gridded by the reader, mixing-length model, 37 x 222 x 37 and
46 x 282 x 46 nodes) and on its six-stage compressor (twelve rows, eleven
mixing planes, 37 x 1282 x 37 nodes; reference on its long build).  Mass flow, total pressure ratio and total-to-total
isentropic efficiency are compared with the FP64 build of the reference
(-fdefault-real-8) run on the same deck: within 0.05 % for mass flow and
pressure ratio and 0.05 points for efficiency.

Fixtures: tests/golden/converged_<case>.json hold the per-J SUMFLX integrals
(BLADE_FLOW, SUMPO, SUMTO, SUMRVT, SUM_ENTPY) that EFICOOL wrote at the end of
the FP64 and FP32 reference runs (mf:10748-10755), made by
tests/golden/make_converged.py.  The quantities are formed from those
integrals the way EFICOOL mass-averages them (mf:10761-10770).
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

CASES = ["r37samods", "r37sa", "axialcompr", "axialturb", "sixstage"]
BOUND_REL = 5e-4     # 0.05 % (mass flow, pressure ratio)
BOUND_ETA = 5e-4     # 0.05 points of efficiency


def quantities(integrals, ga):
    """Exit mass flow, mass-averaged total pressure and total temperature
    ratios (exit / inlet) and the total-to-total isentropic efficiency of a
    compressor or a turbine, from the per-J integrals."""
    flow, spo, sto = (np.asarray(integrals[q], np.float64) for q in range(3))
    p0r = (spo[-1] / flow[-1]) / (spo[0] / flow[0])
    t0r = (sto[-1] / flow[-1]) / (sto[0] / flow[0])
    if p0r > 1.0:   # compressor
        eta = (p0r ** ((ga - 1.0) / ga) - 1.0) / (t0r - 1.0)
    else:           # turbine
        eta = (1.0 - t0r) / (1.0 - p0r ** ((ga - 1.0) / ga))
    return {"mass_flow": float(flow[-1]), "p0_ratio": float(p0r), "t0_ratio": float(t0r), "eta": float(eta)}


def _fixture(case):
    with open(os.path.join(GOLDEN, f"converged_{case}.json")) as f:
        return json.load(f)


def _ga(case):
    from multall_hip import host
    st = host.load_deck(os.path.join(GOLDEN, _fixture(case)["deck"]))
    return st, float(st["scalars"]["GA"])


def _check(q, ref, label):
    for k in ("mass_flow", "p0_ratio"):
        err = abs(q[k] - ref[k]) / abs(ref[k])
        print(f"{label} {k}: {q[k]:.7g} vs FP64 {ref[k]:.7g}, relative error {err:.2e}")
        assert err <= BOUND_REL, (label, k, err)
    err = abs(q["eta"] - ref["eta"])
    print(f"{label} efficiency: {q['eta']:.6f} vs FP64 {ref['eta']:.6f}, error {100 * err:.4f} points")
    assert err <= BOUND_ETA, (label, err)


@pytest.mark.parametrize("case", CASES)
def test_fp32_reference_within_bounds_of_fp64(case):
    """The bounds are met by the reference's own FP32 build: FP32 arithmetic
    itself is not the limit at convergence."""
    d = _fixture(case)
    _, ga = _ga(case)
    q64 = quantities(d["fp64"]["integrals"], ga)
    q32 = quantities(d["fp32"]["integrals"], ga)
    assert abs(q64["p0_ratio"] - 1.0) > 0.05 and 0.5 < q64["eta"] < 1.0, q64
    _check(q32, q64, "reference FP32")


@pytest.mark.gpu
@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("case", CASES)
def test_converged_vs_fp64_reference(case, mixed):
    """Plain FP32 and the mixed form (SSS histories stored in BF16, FP32
    arithmetic): both within the bounds of the FP64 reference."""
    from multall_hip.runtime import Solver
    d = _fixture(case)
    st, ga = _ga(case)
    g = Solver(st, exact_order=False)
    g.set_scalar("MIXED", 1 if mixed else 0)
    g.loss_init()
    g.start()
    r = g.step(int(st["scalars"]["NSTEPS_MAX"]) + 10)
    # the run ends by the reference's convergence test, found on the device
    assert r["ifend"] == 1 and r["nstep"] < int(st["scalars"]["NSTEPS_MAX"])
    print(f"GPU stopped at step {r['nstep']} (EAVG {r['eavg']:.3g}, ECONT {r['econt']:.3g}); "
          f"reference FP64 at {d['fp64']['nstep']}, FP32 at {d['fp32']['nstep']}")
    q = quantities(g.integrals()[:5], ga)
    q64 = quantities(d["fp64"]["integrals"], ga)
    label = "GPU (mixed form)" if mixed else "GPU (fast mode)"
    _check(q, q64, label)
    # EFICOOL's stage (and machine) figures, mh_eficool
    perf = g.eficool()
    if d["fp64"].get("eficool"):
        from test_gpu_eficool import compare
        compare(perf, d["fp64"]["eficool"], d["fp32"]["eficool"], BOUND_REL, BOUND_ETA, f"{case} {label}")
    else:
        # no printout kept: each record's planes in the FP64 reference's integrals
        f64 = np.asarray(d["fp64"]["integrals"], np.float64)
        for rec in perf:
            j1, j2 = rec["j1"] - 1, rec["j2m1"] - 1
            pr = (f64[1][j2] / f64[0][j2]) / (f64[1][j1] / f64[0][j1])
            tr = (f64[2][j2] / f64[0][j2]) / (f64[2][j1] / f64[0][j1])
            x = (ga - 1.0) / ga
            eta = (pr ** x - 1.0) / (tr - 1.0) if tr > 1.0 else (tr - 1.0) / (pr ** x - 1.0)
            for name, a, b in (("flow_in", rec["flow_in"], f64[0][j1]), ("flow_out", rec["flow_out"], f64[0][j2]),
                               ("po_ratio", rec["po_ratio"], pr), ("to ratio", rec["to_out"] / rec["to_in"], tr)):
                print(f"{case} {label} stage {rec['stage']} {name}: {a:.7g} vs FP64 {b:.7g}")
                assert abs(a - b) <= BOUND_REL * abs(b), (rec["stage"], name, a, b)
            assert abs(rec["eta_tt"] - eta) <= BOUND_ETA, (rec["stage"], rec["eta_tt"], eta)
        assert [r["stage"] for r in perf] == list(range(1, len(perf))) + [0]
