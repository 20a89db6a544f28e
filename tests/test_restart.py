"""Starting from a restart file (IF_RESTART = 1): the host reads the
reference's own plotting/restart file flow_out (written by the reference at
the end of a 40-step run, mf:5951-5976) in place of the initial guess
(mf:6919-6999), and the library continues from it.

Golden (tests/golden/make_restart.py): the restart file, the restart deck,
the reference's post-SETUP state of its restart run and its state 40 steps
later.  The host's set-up must equal the reference's bit for bit (the
secondary variables formed from the file, NCHANGE = 100, FACMIXUP = 1, the
shifted exit pressure profile); the library's 40 steps in the exact-order
mode must be within the 100-step bound of the reference's."""
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN, load_state
from test_host_setup import LOSS_STATE

CASES = ["tiny3_ml", "tiny3_sa"]
TOL = 2.5e-5


def _fortran_records(path):
    raw = open(path, "rb").read()
    recs, pos = [], 0
    while pos < len(raw):
        n = int(np.frombuffer(raw, np.int32, 1, pos)[0])
        recs.append(raw[pos + 4: pos + 4 + n])
        pos += 8 + n
    return recs


def _load_restart(case, tmp_path):
    """The restart deck read and set up in a directory holding the
    reference's restart file (the host reads ./flow_out, as the reference's
    unit 7)."""
    from multall_hip import host
    shutil.copy(os.path.join(GOLDEN, f"restart_{case}.flow_out"), tmp_path / "flow_out")
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        return host.load_deck(os.path.join(GOLDEN, "decks", f"restart_{case}.dat"))
    finally:
        os.chdir(cwd)


@pytest.mark.parametrize("case", CASES)
def test_restart_setup_matches_reference(case, tmp_path):
    st = _load_restart(case, tmp_path)
    ref = load_state(f"restart_{case}")
    compared = 0
    for name, a in st.items():
        if name == "scalars" or name in LOSS_STATE or name not in ref:
            continue
        a, b = np.asarray(a).ravel(), np.asarray(ref[name]).ravel()
        n = min(a.size, b.size)
        a, b = a[:n], b[:n]
        bad = np.count_nonzero(a.view(np.uint32) != b.astype(np.float32).view(np.uint32)) \
            if a.dtype == np.float32 else np.count_nonzero(a != b)
        assert bad == 0, f"{case}: {name} differs in {bad} of {n} values"
        compared += 1
    assert compared >= 60
    sc = st["scalars"]
    assert sc["NCHANGE"] == 100 and sc["IF_RESTART"] == 1
    for k, v in sc.items():
        if k in ref["scalars"] and k not in ("NSTEPS_MAX", "VISLAM", "FTCOND"):
            assert np.float32(v) == np.float32(ref["scalars"][k]), (k, v, ref["scalars"][k])
    # the conserved variables and the SA variable are the file's records
    recs = _fortran_records(os.path.join(GOLDEN, f"restart_{case}.flow_out"))
    shape = st["RO"].shape
    for i, name in ((1, "RO"), (2, "ROVX"), (3, "ROVR"), (5, "ROE"), (7, "TRANS_DYN_VIS")):
        want = np.frombuffer(recs[i], np.float32).reshape(shape)
        assert np.array_equal(st[name], want), name


def test_restart_without_a_file_is_an_error(tmp_path):
    from multall_hip import host
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        with pytest.raises(ValueError, match="flow_out"):
            host.load_deck(os.path.join(GOLDEN, "decks", "restart_tiny3_ml.dat"))
    finally:
        os.chdir(cwd)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_restart_run_vs_reference(case, tmp_path):
    from multall_hip.runtime import Solver
    st = _load_restart(case, tmp_path)
    ref = np.load(os.path.join(GOLDEN, f"restart_{case}_ref40.npz"), allow_pickle=False)
    g = Solver(st, exact_order=True)
    g.loss_init()
    g.start()
    r = g.step(50)
    assert r["nstep"] == int(ref["NSTEP"]) and r["ifend"] == 1
    for v in ("RO", "ROVX", "ROVR", "RORVT", "ROE") + (("TRANS_DYN_VIS",) if st["scalars"]["ILOS"] >= 200 else ()):
        a, b = g.get(v).astype(np.float64), ref[v].astype(np.float64)
        err = float(np.abs(a - b).max() / np.abs(b).max())
        print(f"{case} restart {v}: {err:.2e}")
        assert err <= TOL, (v, err)
    g.close()
