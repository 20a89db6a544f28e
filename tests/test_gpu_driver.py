"""The Fortran driver program (bin/multall_hip: the reference's main program
with LOOP replaced by the library, the deck on standard input) on a
three-row mixing-length deck: it reads and sets up the deck (the wall
distances formed on the device by mh_set_xlength), marches the deck's 100
steps in chunks of 5 and writes flow_out in the reference's record layout
(multall-open-18.3.f:5951-5976), grid_out, stage.log and mixbconds.  The density, energy and angular momentum
it writes match the reference's own state after 100 steps within the
100-step bound of the parity tests.  The angular-momentum record is ROVT as
the reference leaves it: refreshed from RORVT at the start of the last step
(mf:3833) and then not updated by that step's TSTEP, so it lags the marched
RORVT by one step away from the inlet and the limiters; the test checks both
that it matches the reference's ROVT and that it differs from RORVT/R."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, load_ref
from test_gpu_parity import TOL_100, _scales


def _records(path):
    """Fortran sequential unformatted records (4-byte length markers)."""
    raw = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(raw):
        n = int(np.frombuffer(raw, np.int32, 1, pos)[0])
        out.append(raw[pos + 4: pos + 4 + n])
        pos += 8 + n
    return out


@pytest.mark.gpu
def test_driver_program_marches_a_deck(tmp_path):
    exe = os.path.join(ROOT, "multall-open-18.3_amd", "bin", "multall_hip")
    with open(os.path.join(GOLDEN, "decks", "tiny3_ml.dat")) as f:
        deck = f.read()
    p = subprocess.run([exe], input=deck, cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-2000:]
    rec = _records(tmp_path / "flow_out")
    assert int(np.frombuffer(rec[0], np.int32)[0]) == 100
    ref = load_ref("tiny3_ml", 100)
    shape = ref["RO"].shape
    got = {"RO": np.frombuffer(rec[1], np.float32).reshape(shape),
           "ROE": np.frombuffer(rec[5], np.float32).reshape(shape),
           "ROVT": np.frombuffer(rec[4], np.float32).reshape(shape)}
    sc = _scales(ref)
    sc["ROVT"] = float(np.abs(ref["ROVT"]).max())
    for v, a in got.items():
        err = float(np.abs(a.astype(np.float64) - ref[v]).max() / sc[v])
        print(f"driver {v}: normalized max err {err:.2e} after 100 steps")
        assert err <= TOL_100, (v, err)
    # the stale record: one step behind the marched angular momentum
    r = np.load(os.path.join(GOLDEN, "tiny3_ml_setup.npz"))["R"]      # (KM, JM)
    marched = ref["RORVT"].astype(np.float64) / r[:, :, None]
    lag = float(np.abs(marched - ref["ROVT"])[:, 2:-2, :].max() / sc["ROVT"])
    print(f"ROVT lags RORVT/R by {lag:.2e} of its scale")
    assert lag > 10 * TOL_100
    # grid_out as the reference's SETUP wrote it, byte for byte
    assert (tmp_path / "grid_out").read_bytes() == open(os.path.join(GOLDEN, "tiny3_ml.grid_out"), "rb").read()
    # stage.log: one line per check in the reference's format (mf:5437-5438),
    # with its step, EMAX location and figures
    mine = (tmp_path / "stage.log").read_text().splitlines()
    want = open(os.path.join(GOLDEN, "tiny3_ml.stage_log")).read().splitlines()
    assert len(mine) == len(want), (len(mine), len(want))

    def fields(line):
        head, at = line.split("AT")
        step = int(head.split("STEP")[1])
        nums = [float(x) for x in head.split("ECONT")[1].replace(" FLOW", " ").split("STEP")[0].split()]
        return step, nums, [int(x) for x in at.split()]
    same_at = 0
    for a, b in zip(mine, want):
        sa, na, la = fields(a)
        sb, nb, lb = fields(b)
        assert sa == sb
        for x, y in zip(na, nb):
            assert abs(x - y) <= 2e-3 * max(abs(y), 1e-3), (a, b)
        same_at += la == lb
    print(f"stage.log: {len(mine)} lines, EMAX location equal on {same_at}")
    assert same_at >= len(want) - 2
    # mixbconds: MIX_BCONDS's profiles at the inlet and every JMIX in the
    # reference's layout; each number within twice its printed resolution
    # or 1e-4 of its value
    tm = (tmp_path / "mixbconds").read_text().split()
    tw = open(os.path.join(GOLDEN, "tiny3_ml.mixbconds")).read().split()
    assert len(tm) == len(tw), (len(tm), len(tw))
    worst = 0.0
    for a, b in zip(tm, tw):
        try:
            y = float(b)
        except ValueError:
            assert a == b, (a, b)
            continue
        res = 10.0 ** -(len(b.split(".")[1]) if "." in b else 0)
        err = abs(float(a) - y)
        worst = max(worst, err / (2 * res + 1e-4 * abs(y)))
        assert err <= 2 * res + 1e-4 * abs(y), (a, b)
    print(f"mixbconds: {len(tw)} tokens, worst error {worst:.2f} of its bound")
