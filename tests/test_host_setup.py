"""The Fortran host's deck reader and grid set-up (host/mh_host.f90)
reproduce the reference's post-SETUP COMMON state bit for bit.

Golden: tests/golden/<case>_setup.npz, the reference's own COMMON dump after
SETUP for the deck tests/golden/decks/<case>.dat.  The reference's SETUP also
makes the first LOSS call (forces, y+, VISLAM); the host leaves that to the
library (mh_loss_init), so those arrays are compared on the GPU instead.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_state

CASES = ["tiny1_inv", "tiny3_inv", "tiny3_ml", "tiny3_ml2", "tiny1_mlw", "tiny3_sa", "tiny3_sa2", "tiny3_exit", "tiny3_wide",
         "tiny3_nl", "tiny3_nl2", "tiny3_sss", "tiny3_flow2", "tiny3_flow3", "tiny3_ml3", "tiny3_nl3", "tiny3_sa3",
         "tiny3_gap", "tiny3_bc1", "tiny3_bc2", "tiny2_bc3", "tiny3_sup", "tiny3_nout", "tiny3_cusp2", "tiny2_rep"]
LOSS_STATE = {"XFORCE", "RFORCE", "TFORCE", "QSOURCE", "Y_PLUS", "VISC_RAT", "VISC_LAM", "TRANS_DYN_VIS",
              "TRANS_KVIS", "T_SOURCE"}


@pytest.mark.parametrize("case", CASES)
def test_host_setup_matches_reference(case):
    from multall_hip import host
    st = host.load_deck(os.path.join(GOLDEN, "decks", f"{case}.dat"))
    ref = load_state(case)
    compared = 0
    skip = set(LOSS_STATE)
    if st["scalars"]["ILOS"] == 0:
        skip |= {"DWALLSQ", "XLENGTH", "FST_RAT"}   # the host forms wall distances only for viscous runs
    for name, a in st.items():
        if name == "scalars" or name in skip or name not in ref:
            continue
        b = np.asarray(ref[name])
        a = np.asarray(a).ravel()[: b.size] if a.size != b.size else np.asarray(a).ravel()
        b = b.ravel()[: a.size]
        if a.dtype == np.float32:
            bad = np.count_nonzero(a.view(np.uint32) != b.astype(np.float32).view(np.uint32))
        else:
            bad = np.count_nonzero(a != b)
        assert bad == 0, f"{case}: {name} differs in {bad} of {a.size} values"
        compared += 1
    assert compared >= 60
    for k, v in st["scalars"].items():
        if k in ref["scalars"] and k not in ("NSTEPS_MAX", "VISLAM", "FTCOND"):
            assert np.float32(v) == np.float32(ref["scalars"][k]), (k, v, ref["scalars"][k])


def test_reload_keeps_no_state_of_the_previous_deck():
    """A deck read after another one sets up exactly as in a fresh process:
    options a deck leaves unread (the mixing-plane parameters of a one-row
    machine, IFMIX = 0; the Spalart-Allmaras factors of a mixing-length run)
    keep the reference's fresh-COMMON zero rather than the previous deck's
    values.  Rotor 37 (IFMIX = 0) after a three-row Spalart-Allmaras deck."""
    from multall_hip import host
    r37 = os.path.join(GOLDEN, "decks", "converged_r37sa.dat")
    fresh = host.load_deck(r37)
    host.load_deck(os.path.join(GOLDEN, "decks", "tiny3_sa2.dat"))
    again = host.load_deck(r37)
    assert fresh["scalars"] == again["scalars"]
    for name, a in fresh.items():
        if name != "scalars":
            assert np.array_equal(a, again[name]), name


REF_DIGESTS = os.path.join(GOLDEN, "refdeck_setup_digests.json")
# the reference's own test decks that fit its FP32 build: one- to three-row
# axial and radial machines re-gridded between rows (ISHIFT = 2) with and
# without new trailing-edge cusps (IF_CUSP = 1: the radial inflow turbine and
# Harrison's cascade), a KM = 5 inviscid shock case and Rotor 37 (ISHIFT = 0)
REF_DECKS = sorted(json.load(open(REF_DIGESTS))) if os.path.exists(REF_DIGESTS) else []


@pytest.mark.parametrize("deck", REF_DECKS)
def test_host_setup_of_reference_decks(deck):
    """The host's set-up of the reference's own test decks equals the
    reference's post-SETUP COMMON state bit for bit, array by array (SHA-256
    of the FP32/INT32 values; golden digests written by
    tests/golden/make_refdeck_digests.py from the reference itself)."""
    import hashlib
    from multall_hip import host
    want = json.load(open(REF_DIGESTS))[deck]
    st = host.load_deck(os.path.join(GOLDEN, "decks", "ref", deck))
    for name, w in want["arrays"].items():
        a = np.ascontiguousarray(st[name])
        assert list(a.shape) == w["shape"] and str(a.dtype) == w["dtype"], name
        got = hashlib.sha256(a.tobytes()).hexdigest()
        assert got == w["sha256"], f"{deck}: {name} differs (sum {a.astype(np.float64).sum()!r} vs {w['sum']!r})"
    for k, v in want["scalars"].items():
        assert np.float32(st["scalars"][k]) == np.float32(v), (deck, k, st["scalars"][k], v)
    assert len(want["arrays"]) >= 100


CFG_DIGESTS = os.path.join(GOLDEN, "config_setup_digests.json")
# the benchmark workloads: the single stator row, the 1.5-stage turbine and
# the first and last speedline points (tests/golden/make_configs.py)
CFG_DECKS = sorted(json.load(open(CFG_DIGESTS))) if os.path.exists(CFG_DIGESTS) else []


@pytest.mark.parametrize("deck", CFG_DECKS)
def test_host_setup_of_config_decks(deck):
    """The host's set-up of the benchmark decks equals the reference's
    post-SETUP state bit for bit (SHA-256 of every array the library takes),
    so the GPU tests on these decks start where the reference starts."""
    import hashlib
    from multall_hip import host
    want = json.load(open(CFG_DIGESTS))[deck]
    st = host.load_deck(os.path.join(GOLDEN, "decks", f"{deck}.dat"))
    for name, w in want["arrays"].items():
        a = np.ascontiguousarray(st[name])
        assert list(a.shape) == w["shape"] and str(a.dtype) == w["dtype"], name
        got = hashlib.sha256(a.tobytes()).hexdigest()
        assert got == w["sha256"], f"{deck}: {name} differs (sum {a.astype(np.float64).sum()!r} vs {w['sum']!r})"
    for k, v in want["scalars"].items():
        assert np.float32(st["scalars"][k]) == np.float32(v), (deck, k, st["scalars"][k], v)
    assert len(want["arrays"]) >= 80


@pytest.mark.parametrize("deck", ["tiny3_ml", "tiny3_gap"])
def test_grid_out_equals_the_reference_file(deck, tmp_path):
    """The plotting grid file the host writes (grid_out, mf:6585-6599) is,
    byte for byte, the file the reference's SETUP wrote for the same deck
    (tests/golden/make_outputs.py); tiny3_gap has its gaps re-gridded."""
    from multall_hip import host
    host.load_deck(os.path.join(GOLDEN, "decks", f"{deck}.dat"))
    host.write_grid_out(str(tmp_path / "grid_out"))
    got = (tmp_path / "grid_out").read_bytes()
    want = open(os.path.join(GOLDEN, f"{deck}.grid_out"), "rb").read()
    assert len(got) == len(want) and got == want
