"""Restart fixtures: the reference run from its own restart file.

For the three-row cascades with the mixing-length and Spalart-Allmaras
models, the reference (FP32 build from oracle/build_ref.sh, this container
only) runs the deck for 40 steps and writes its plotting/restart file
flow_out at the end (OUTPUT, mf:5951-5976); the same deck with IF_RESTART = 1
then runs 40 more steps from that file (mf:6919-6999: the file's flow in
place of the initial guess, NCHANGE = 100, FACMIXUP = 1).  Stored per case:

  restart_<case>.flow_out     the restart file (Fortran unformatted records)
  decks/restart_<case>.dat    the restart deck (IF_RESTART = 1, NSTEPS_MAX = 40)
  restart_<case>_setup.npz    the reference's post-SETUP state of the restart run
  restart_<case>_ref40.npz    its state after the 40 steps

Usage: python tests/golden/make_restart.py [case ...]
"""
from __future__ import annotations

import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "multall-open-18.3_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

CASES = ["tiny3_ml", "tiny3_sa"]
STEPS = 40
VARS = ["RO", "ROVX", "ROVR", "RORVT", "ROE", "DRO", "DROVX", "DROVR", "DRORVT", "DROE", "STEP"]
SA_VARS = ["TRANS_DYN_VIS", "DEL_DYNVIS"]


def set_value(text: str, label: str, col: int, value: str) -> str:
    """The deck with column `col` of the value line under the comment line
    that starts with `label` replaced."""
    lines = text.split("\n")
    i = next(n for n, l in enumerate(lines) if l.strip().upper().startswith(label))
    v = lines[i + 1].split()
    v[col] = value
    lines[i + 1] = "  " + "  ".join(v)
    return "\n".join(lines)


def restart_deck(text: str, restart: int, nsteps: int) -> str:
    return set_value(set_value(text, "NSTEPS_MAX", 0, str(nsteps)), "IF_RESTART", 0, str(restart))


def make(case: str) -> None:
    import make_golden
    import orc
    import refdump
    base = open(os.path.join(HERE, "decks", f"{case}.dat")).read()
    wd = tempfile.mkdtemp(prefix="mhrst_")
    try:
        refdump.run_reference(restart_deck(base, 0, STEPS), "fp32", workdir=wd)
        shutil.copy(os.path.join(wd, "flow_out"), os.path.join(HERE, f"restart_{case}.flow_out"))
        for f in os.listdir(wd):
            if f != "flow_out":
                os.remove(os.path.join(wd, f))
        deck = restart_deck(base, 1, STEPS)
        setup, final, _ = refdump.run_reference(deck, "fp32", workdir=wd)
    finally:
        shutil.rmtree(wd, ignore_errors=True)
    with open(os.path.join(HERE, "decks", f"restart_{case}.dat"), "w") as f:
        f.write(deck)
    st = orc.from_dump(setup)
    st["scalars"].pop("NSTEPS_MAX")
    st.pop("NOUT", None)
    make_golden.save_state(os.path.join(HERE, f"restart_{case}_setup.npz"), st)
    names = VARS + (SA_VARS if st["scalars"]["ILOS"] >= 200 else [])
    out = {k: final.node(k).astype(np.float32) for k in names}
    out["NSTEP"] = np.array(int(final["NSTEP"]))
    np.savez_compressed(os.path.join(HERE, f"restart_{case}_ref{STEPS}.npz"), **out)
    print(f"{case}: restart run of {STEPS} steps after {STEPS}; NCHANGE {st['scalars']['NCHANGE']}, "
          f"FACMIXUP {st['scalars'].get('FACMIXUP')}", flush=True)


if __name__ == "__main__":
    for c in (sys.argv[1:] or CASES):
        make(c)
