"""The reference's own output files for the driver and host tests.

Runs the reference built by oracle/build_ref.sh (FP32, this container only)
on decks under tests/golden/decks/ and keeps the files it writes besides
its state:

  <case>.grid_out    the plotting grid file SETUP writes (mf:6585-6599)
  <case>.stage_log   stage.log, the convergence lines of every check (mf:5437-5438)
  <case>.mixbconds   the mass-averaged profiles MIX_BCONDS writes at the end
                     of the run (mf:5445, 11906-12080)

Usage: python tests/golden/make_outputs.py [case ...]
"""
from __future__ import annotations

import os
import shutil
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

CASES = ["tiny3_ml", "tiny3_gap"]


def make(case: str) -> None:
    import refdump
    txt = open(os.path.join(HERE, "decks", f"{case}.dat")).read()
    wd = tempfile.mkdtemp(prefix="mhout_")
    try:
        refdump.run_reference(txt, "fp32", workdir=wd)
        shutil.copy(os.path.join(wd, "grid_out"), os.path.join(HERE, f"{case}.grid_out"))
        shutil.copy(os.path.join(wd, "stage.log"), os.path.join(HERE, f"{case}.stage_log"))
        shutil.copy(os.path.join(wd, "mixbconds"), os.path.join(HERE, f"{case}.mixbconds"))
    finally:
        shutil.rmtree(wd, ignore_errors=True)
    print(f"{case}: grid_out and stage.log kept", flush=True)


if __name__ == "__main__":
    for c in (sys.argv[1:] or CASES):
        make(c)
