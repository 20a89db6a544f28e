"""Generate the golden fixtures under tests/golden/ from the reference solver.

Runs the reference built by oracle/build_ref.sh (this container only; the
reference is not available where the tests run) on the synthetic decks (the
IM = 80 deck on its pitchwise-wide build) and stores, per case:

  <case>.dat          the NEW-format input deck
  <case>_setup.npz    COMMON state after the reference's SETUP (dense arrays
                      plus a JSON list of scalars, and the printout steps
                      NOUT where they fall inside the run): the input of
                      every test
  <case>_ref<N>.npz   conserved variables, SSS histories and time steps after
                      N time steps of the reference (N = 1 and 100, or the
                      STEPS_OF pair)

Usage: python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "multall-open-18.3_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from multall_hip import deckgen  # noqa: E402

OUT_VARS = ["RO", "ROVX", "ROVR", "RORVT", "ROE", "DRO", "DROVX", "DROVR", "DRORVT", "DROE", "STEP",
            "XFORCE", "RFORCE", "TFORCE", "QSOURCE"]
# ROVT as the reference leaves it (refreshed from RORVT at the start of the
# last step, mf:3833): the angular-momentum record of its flow_out
OUT_VARS_LATER = ["ROVT"]
SA_VARS = ["TRANS_DYN_VIS", "DEL_DYNVIS", "TRANS_KVIS", "T_SOURCE", "Y_PLUS"]   # ILOS >= 200 decks

CASES = {
    "tiny1_inv": lambda n: deckgen.deck_tiny(ilos=0, nrows=1, nsteps=n),
    "tiny3_inv": lambda n: deckgen.deck_tiny(ilos=0, nrows=3, nsteps=n),
    "tiny3_ml": lambda n: deckgen.deck_tiny(ilos=10, nrows=3, nsteps=n),
    "tiny3_ml2": lambda n: deckgen.deck_tiny_ml_variant(nsteps=n),
    "tiny1_mlw": lambda n: deckgen.deck_tiny_wall_variant(nsteps=n),
    "tiny3_sa": lambda n: deckgen.deck_tiny_sa(nsteps=n),
    "tiny3_sa2": lambda n: deckgen.deck_tiny_sa_variant(nsteps=n),
    "tiny3_exit": lambda n: deckgen.deck_tiny_exit(nsteps=n),
    "tiny3_wide": lambda n: deckgen.deck_tiny_wide(nsteps=n),
    "tiny3_nl": lambda n: deckgen.deck_tiny_nl(nsteps=n),
    "tiny3_nl2": lambda n: deckgen.deck_tiny_nl_variant(nsteps=n),
    "tiny3_sss": lambda n: deckgen.deck_tiny_sss(nsteps=n),
    "tiny3_flow2": lambda n: deckgen.deck_tiny_flow(in_flow=2, nsteps=n),
    "tiny3_flow3": lambda n: deckgen.deck_tiny_flow(in_flow=3, nsteps=n),
    "tiny3_ml3": lambda n: deckgen.deck_tiny_wallfun(ilos=10, nsteps=n),
    "tiny3_nl3": lambda n: deckgen.deck_tiny_wallfun(ilos=100, nsteps=n),
    "tiny3_sa3": lambda n: deckgen.deck_tiny_wallfun(ilos=200, nsteps=n),
    "tiny3_gap": lambda n: deckgen.deck_tiny_gap(nsteps=n),
    "tiny3_bc1": lambda n: deckgen.deck_tiny_bc("bc1", nsteps=n),
    "tiny3_bc2": lambda n: deckgen.deck_tiny_bc("bc2", nsteps=n),
    "tiny2_bc3": lambda n: deckgen.deck_tiny_bc("bc3", nsteps=n),
    "tiny3_sup": lambda n: deckgen.deck_tiny_supersonic(nsteps=n),
    "tiny3_nout": lambda n: deckgen.deck_tiny_nout(nsteps=n),
    "tiny3_cusp2": lambda n: deckgen.deck_tiny_cusp2(nsteps=n),
    "tiny2_rep": lambda n: deckgen.deck_tiny_repeat(nsteps=n),
}
# decks wider than the main reference build (IM > 65) run on its wide build
VARIANT = {"tiny3_wide": "fp32w"}
STEPS = (1, 100)
# the supersonic cascade is compared after 20 steps: between steps 20 and
# 30 a one-ulp change of any single initial value grows to ~1e-3 of the
# solution (a limiter or sonic switch flips at one node), so agreement after
# 100 steps would mean nothing there
STEPS_OF = {"tiny3_sup": (1, 20)}


def save_state(path, st):
    arrays = {k: v for k, v in st.items() if k != "scalars"}
    np.savez_compressed(path, __scalars__=np.array(json.dumps(st["scalars"])), **arrays)


def load_state(path):
    z = np.load(path, allow_pickle=False)
    st = {k: z[k] for k in z.files if k != "__scalars__"}
    st["scalars"] = json.loads(str(z["__scalars__"]))
    return st


def read_global_plt(path, real=np.float32):
    """The records EFICOOL writes to unit 11 at the end of a run
    (mf:10748-10755): (1,1), then BLADE_FLOW, SUMPO, SUMTO, SUMRVT,
    SUM_ENTPY per J, after what OUTPUT wrote there.  Fortran sequential
    records with 4-byte markers; REAL is 4 or 8 bytes by build."""
    raw = open(path, "rb").read()
    recs, pos = [], 0
    while pos < len(raw):
        n = int(np.frombuffer(raw, np.int32, 1, pos)[0])
        recs.append(raw[pos + 4: pos + 4 + n])
        pos += 8 + n
    k = max(i for i, r in enumerate(recs) if len(r) == 8)   # EFICOOL's (1,1) record
    return np.stack([np.frombuffer(r, real) for r in recs[k + 1:k + 6]])


def make(case):
    import tempfile
    import shutil
    import refdump, orc
    setup_ref = None
    steps = STEPS_OF.get(case, STEPS)
    for n in steps:
        spec = CASES[case](n)
        txt = deckgen.write_deck(spec)
        wd = tempfile.mkdtemp(prefix="mhgold_")
        try:
            setup, final, _ = refdump.run_reference(txt, VARIANT.get(case, "fp32"), workdir=wd)
            sumflx = read_global_plt(os.path.join(wd, "global.plt"))
        finally:
            shutil.rmtree(wd, ignore_errors=True)
        st = orc.from_dump(setup)
        st["scalars"].pop("NSTEPS_MAX")
        # the printout steps matter only where a deck sets them inside the
        # run (the others put them just past NSTEPS_MAX, which differs by run)
        if spec.nout is None:
            st.pop("NOUT")
        if setup_ref is None:
            setup_ref = st
            with open(os.path.join(HERE, "decks", f"{case}.dat"), "w") as f:
                f.write(deckgen.write_deck(CASES[case](max(steps))))
            save_state(os.path.join(HERE, f"{case}_setup.npz"), st)
        else:
            for k, v in st.items():
                if k == "scalars":
                    assert v == setup_ref["scalars"], case
                else:
                    assert np.array_equal(v, setup_ref[k]), (case, k)
        names = OUT_VARS + OUT_VARS_LATER + (SA_VARS if st["scalars"]["ILOS"] >= 200 else [])
        if st["scalars"]["YPLUSWALL"] < -0.001:
            names = names + ["DPDS_CELL"] + (["Y_PLUS"] if st["scalars"]["ILOS"] < 200 else [])
        out = {k: final.node(k).astype(np.float32) for k in names}
        if st["scalars"]["YPLUSWALL"] < -0.001:   # the Shih wall functions' y+ state
            jm, im = st["scalars"]["JM"], st["scalars"]["IM"]
            for k in ("YPLUS_I1", "YPLUS_IM"):
                out[k] = final.jk(k).astype(np.float32)
            for k in ("YPLUS_K1", "YPLUS_KM"):
                out[k] = np.ascontiguousarray(final[k][:jm, :im]).astype(np.float32)
        out["NSTEP"] = np.array(int(final["NSTEP"]))
        out["SUMFLX"] = sumflx          # SUMFLX's per-J integrals at the end of the run
        np.savez_compressed(os.path.join(HERE, f"{case}_ref{n}.npz"), **out)
        print(f"{case}: {n} steps done")


if __name__ == "__main__":
    os.makedirs(os.path.join(HERE, "decks"), exist_ok=True)
    for c in (sys.argv[1:] or list(CASES)):
        make(c)
