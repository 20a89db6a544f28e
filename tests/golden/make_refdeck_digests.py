"""Digests of the reference's own set-up of its own test decks.

For each deck under tests/golden/decks/ref/ (copied verbatim from the
reference's multall-test-cases/), runs the reference built by
oracle/build_ref.sh for one time step and records, for every array the
Fortran host hands to the library, the SHA-256 of the reference's
post-SETUP COMMON values (FP32, in the host's dense layout), with a float64
sum for diagnosis, and the reference's scalars.  Output:
tests/golden/refdeck_setup_digests.json.

Running the reference needs two edits to the deck text that change no
input value: NSTEPS_MAX = 1, and a comment line that opens with "(" loses
that parenthesis (flang's list-directed read of the comment into a
CHARACTER variable stops on it).

Usage: python tests/golden/make_refdeck_digests.py [deck.dat ...]
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "multall-open-18.3_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

DECKS = os.path.join(HERE, "decks", "ref")
OUT = os.path.join(HERE, "refdeck_setup_digests.json")
# state the reference's SETUP forms with its first LOSS call; the host leaves
# that to the library
LOSS_STATE = {"XFORCE", "RFORCE", "TFORCE", "QSOURCE", "Y_PLUS", "VISC_RAT", "VISC_LAM", "TRANS_DYN_VIS",
              "TRANS_KVIS", "T_SOURCE"}


def runnable(text: str) -> str:
    lines = text.split("\n")
    for i, line in enumerate(lines):
        if "NSTEPS_MAX" in line.upper():
            parts = lines[i + 1].split()
            parts[0] = "1"
            lines[i + 1] = "  ".join(parts)
            break
    lines = [ln.replace("(", " ", 1) if ln.strip().startswith("(") else ln for ln in lines]
    return "\n".join(lines)


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def normalise(host_arr: np.ndarray, ref_arr) -> np.ndarray:
    """The reference array in the host array's dtype and dense layout."""
    b = np.asarray(ref_arr)
    if b.size != host_arr.size:
        b = b.ravel()[: host_arr.size]
    return b.astype(host_arr.dtype).reshape(host_arr.shape)


def make(name: str) -> dict:
    import orc
    import refdump
    from multall_hip import host
    path = os.path.join(DECKS, name)
    st = host.load_deck(path)          # only for the list of arrays and their layout
    variant = "fp32l" if st["scalars"]["JM"] > 600 else "fp32"     # the long build for many rows
    setup, _, _ = refdump.run_reference(runnable(open(path, errors="replace").read()), variant)
    ref = orc.from_dump(setup)
    skip = set(LOSS_STATE)
    if st["scalars"]["ILOS"] == 0:
        skip |= {"DWALLSQ", "XLENGTH", "FST_RAT"}      # formed for viscous runs only
    arrays = {}
    for k, a in st.items():
        if k == "scalars" or k in skip or k not in ref:
            continue
        b = normalise(np.asarray(a), ref[k])
        arrays[k] = {"sha256": digest(b), "shape": list(b.shape), "dtype": str(b.dtype),
                     "sum": float(np.asarray(b, np.float64).sum())}
    scal = {k: float(ref["scalars"][k]) for k in st["scalars"]
            if k in ref["scalars"] and k not in ("NSTEPS_MAX", "VISLAM", "FTCOND")}
    print(f"{name}: {len(arrays)} arrays", flush=True)
    return {"arrays": arrays, "scalars": scal}


if __name__ == "__main__":
    # the six-stage compressor is stepped on the GPU but not digested: the
    # reference's long build reads an unset local in GRID_DOWN
    # (SNEW(J1-NEXTRAP-1)), so its wake line behind the last row differs from
    # the host's (which takes the value as zero) in the last bits
    names = sys.argv[1:] or sorted(n for n in os.listdir(DECKS) if not n.startswith("Six-stage"))
    out = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for n in names:
        out[n] = make(n)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
