"""EFICOOL fixtures: the reference's stage and machine performance after a
fixed number of steps.

Runs the reference built by oracle/build_ref.sh (this container only) in its
FP32 and FP64 (-fdefault-real-8) builds on decks already under
tests/golden/decks/ for the deck's own step count, and stores the last
EFICOOL block the reference printed at the end of the run (mf:11253-11548,
parsed by make_converged.eficool_records): per stage, and for the whole
machine when there is more than one stage.

  eficool_<case>.json   {"deck": <deck file>, "steps": N, "fp32": [...], "fp64": [...]}

Cases: the three-row synthetic decks with the mixing-length and
Spalart-Allmaras models (rows in stages 1, 1, 2: two stages and the whole
machine; 100 steps) and the single turbine stator row (one stage; 100 steps).

Usage: python tests/golden/make_eficool.py [case ...]
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "multall-open-18.3_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

from make_converged import eficool_records  # noqa: E402

CASES = ["tiny3_ml", "tiny3_sa", "config0"]


def make(case: str) -> None:
    import refdump
    deck = f"decks/{case}.dat"
    txt = open(os.path.join(HERE, deck)).read()
    out = {"deck": deck}
    for variant in ("fp32", "fp64"):
        _, final, stdout = refdump.run_reference(txt, variant)
        out["steps"] = int(final["NSTEP"])
        out[variant] = eficool_records(stdout)
        print(f"{case} {variant}: {len(out[variant])} records after {out['steps']} steps", flush=True)
    with open(os.path.join(HERE, f"eficool_{case}.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    for c in (sys.argv[1:] or CASES):
        make(c)
