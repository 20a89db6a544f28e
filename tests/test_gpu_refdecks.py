"""The HIP library on the reference's own test decks against the oracle.

Each deck under tests/golden/decks/ref/ (the reference's multall-test-cases,
verbatim: one- to three-row axial and radial machines and the twelve-row
six-stage compressor, inter-row gaps re-gridded by the reader, new
trailing-edge cusps, a KM = 5 inviscid shock case, Rotor 37) is set up by the
Fortran host, whose set-up equals the reference's bit for bit
(test_host_setup.py; the six-stage compressor to the last bits of its last
wake line).  The library's SETUP-time
viscous call (mh_loss_init) forms the forces and turbulence state; the
oracle makes the same call itself (LOSS, NEW_LOSS or SPAL_LOSS with
RF_VIS = 1, FMIXUP = 1, NSTEP = 1, as SETUP does, mf:7346-7357) and the two
are compared within 1e-5 of each array's largest magnitude.  Then, to test
the step apart from the viscous call, the oracle takes the library's
post-call state and one time step in the reference's statement order
(exact-order mode) is compared: conserved variables, changes and time steps
within 2e-6 of each variable's largest magnitude (the transcendentals device
and host libm round differently), the forces that the run's closing viscous
call refreshes within 1e-5.
"""
import os

import numpy as np
import pytest

import orc
from conftest import GOLDEN
from test_gpu_parity import FORCES, VARS, _nerr

DECK_DIR = os.path.join(GOLDEN, "decks", "ref")
DECKS = sorted(os.listdir(DECK_DIR)) if os.path.isdir(DECK_DIR) else []
# what the SETUP-time viscous call writes (per model), carried into the oracle
LOSS_OUT = FORCES + ["Y_PLUS", "VISC_RAT", "VISC_LAM", "TRANS_DYN_VIS", "DEL_DYNVIS", "TRANS_KVIS",
                     "T_SOURCE", "DPDS_CELL", "YPLUS_I1", "YPLUS_IM", "YPLUS_K1", "YPLUS_KM"]


@pytest.mark.gpu
@pytest.mark.parametrize("deck", DECKS)
def test_reference_deck_one_step_vs_oracle(deck):
    from multall_hip import host
    from multall_hip.runtime import Solver
    st = host.load_deck(os.path.join(DECK_DIR, deck))
    sc = st["scalars"]
    sc["NSTEPS_MAX"] = 1
    print(f"{deck}: IM={sc['IM']} JM={sc['JM']} KM={sc['KM']} ILOS={sc['ILOS']} NROWS={sc['NROWS']}", flush=True)
    g = Solver(st, exact_order=True)
    ost = orc.with_workspace(dict(st))
    if sc["ILOS"] != 0:
        g.loss_init()
        own = orc.Oracle(ost)
        for k, v in (("RF_VIS", 1.0), ("FMIXUP", 1.0), ("NSTEP", 1)):
            own.lib.orc_scalar(own.h, k.encode(), float(v))
        if sc["ILOS"] >= 200:
            own.lib.orc_spal_loss(own.h)
        elif sc["ILOS"] >= 100:
            own.lib.orc_new_loss(own.h)
        else:
            own.lib.orc_loss(own.h)
        outs = FORCES + (["TRANS_DYN_VIS", "TRANS_KVIS"] if sc["ILOS"] >= 200 else [])
        for v in outs:
            err = _nerr(g.get(v), own[v])
            print(f"{deck} set-up call {v}: normalized max err {err:.2e}")
            assert err <= 1e-5, (deck, v, err)
        for n in LOSS_OUT:
            if n in ost:
                ost[n] = g.get(n).reshape(np.shape(ost[n])).astype(np.float32)
    o = orc.Oracle(ost)
    o.start()
    assert o.step(5) == 1
    g.start()
    assert g.step(5)["nstep"] == 1
    for v in VARS:
        a, b = g.get(v), o[v]
        err = _nerr(a, b)
        print(f"{deck} {v}: {int(np.count_nonzero(a != b))} differing values, normalized max err {err:.2e}")
        assert err <= 2e-6, (deck, v, err)
    if sc["ILOS"] != 0:
        for v in FORCES:
            err = _nerr(g.get(v), o[v])
            print(f"{deck} {v}: normalized max err {err:.2e}")
            assert err <= 1e-5, (deck, v, err)
