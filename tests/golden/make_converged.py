"""Generate the converged-run fixtures under tests/golden/ from the reference.

Runs the reference built by oracle/build_ref.sh (this container only) in its
FP64 build (-fdefault-real-8) and its FP32 build on a deck for the deck's
whole step budget or until its own convergence test stops it (the
reference's own Rotor 37 deck, as it ships and with plain Spalart-Allmaras),
and stores what EFICOOL wrote at the end of the run (mf:10748-10755): the per-J SUMFLX integrals BLADE_FLOW, SUMPO, SUMTO,
SUMRVT, SUM_ENTPY, with the step the run stopped at.

  converged_<case>.json   {"deck": <deck file>, "fp64": {...}, "fp32": {...}}

and, per build, the last EFICOOL block the reference printed (mf:11253-11548):
per stage and for the whole machine, the stagnation and static pressures,
stagnation temperatures and mass flows at the stage's inlet and exit planes,
the pressure ratio, the total-to-total, total-to-static and polytropic
efficiencies, the reactions and the work.

Usage: python tests/golden/make_converged.py [case ...]
"""
from __future__ import annotations

import json
import os
import shutil
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "multall-open-18.3_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

from make_golden import read_global_plt  # noqa: E402

_NUM = r"[-+]?(?:NaN|Infinity|\d+\.?\d*|\.\d+)(?:[EeDd][-+]?\d+)?"


def eficool_records(stdout: str) -> list:
    """The last EFICOOL performance block the reference printed (the end of
    the run; IFCOOL = 0, mf:11253-11548): one record per stage, then the
    whole machine when there is more than one stage, each with the numbers
    the reference printed (list-directed output: a line that runs out of
    width continues on the next, which then holds numbers only)."""
    import re
    lines = stdout.split("\n")
    starts = [i for i, l in enumerate(lines) if "RESULTS FOR STAGE NUMBER" in l or
              "THE FOLLOWING RESULTS ARE FOR THE WHOLE MACHINE" in l]
    if not starts:
        return []
    first = max(i for i in starts if "RESULTS FOR STAGE NUMBER" in lines[i] and
                int(re.findall(_NUM, lines[i])[-1]) == 1)
    numeric = re.compile(r"^\s*(?:" + _NUM + r"\s*)+$")

    def vals(i, k, after="="):
        """The first k numbers after `after` on line i and its continuation lines."""
        text = lines[i].split(after, 1)[1] if after else lines[i]
        out = [float(x.replace("D", "E").replace("d", "e")) for x in re.findall(_NUM, text)]
        j = i + 1
        while len(out) < k and j < len(lines) and numeric.match(lines[j]):
            out += [float(x.replace("D", "E")) for x in re.findall(_NUM, lines[j])]
            j += 1
        return out[:k]
    recs = []
    keys = [("PRESSURE RATIO, POinlet/POexit", ("po_ratio_printed",)),
            ("PRESSURE RATIO, POexit/POinlet", ("po_ratio_printed",)),
            ("EFFICIENCIES BASED ON VALUES AT J", ("j1", "j2m1")), ("JMIX", ("jmix",)),
            ("INLET AND EXIT STAGNATION PRESSURES", ("po_in", "po_out")),
            ("INLET AND EXIT STATIC PRESSURES", ("ps_in", "ps_out")),
            ("INLET AND EXIT STAGNATION TEMPERATURES", ("to_in", "to_out")),
            ("INLET AND OUTLET MASS FLOW RATES", ("flow_in", "flow_out")),
            ("TOTAL TO TOTAL ISENTROPIC EFFICIENCY", ("eta_tt",)),
            ("TOTAL TO STATIC ISENTROPIC EFFICIENCY", ("eta_ts",)),
            ("TOTAL TO TOTAL POLYTROPIC EFFICIENCY", ("eta_poly",)),
            ("AVERAGE STAGE REACTION BASED ON MASS AVERAGED", ("reaction",)),
            ("STAGE WORK NEGLECTING", ("work_kw",)), ("OVERALL POWER OUTPUT", ("work_kw",))]
    pressure_keys = [("HUB REACTION BASED ON PRESSURE CHANGES", "reaction_hub"),
                     ("MID SPAN REACTION BASED ON PRESSURE CHANGES", "reaction_mid"),
                     ("TIP REACTION BASED ON PRESSURE CHANGES", "reaction_tip")]
    for s0 in [x for x in starts if x >= first]:
        r = {"stage": 0 if "WHOLE MACHINE" in lines[s0] else int(re.findall(_NUM, lines[s0])[-1])}
        for i in range(s0 + 1, min(s0 + 90, len(lines))):
            l = lines[i]
            if "RESULTS FOR STAGE NUMBER" in l or "FOLLOWING RESULTS ARE FOR THE WHOLE MACHINE" in l:
                break
            for label, names in keys:
                if l.strip().startswith(label) and names[0] not in r:
                    if label.startswith("PRESSURE RATIO"):
                        r["turbine"] = "POinlet/POexit" in label
                    v = vals(i, len(names))
                    for n, x in zip(names, v):
                        r[n] = int(x) if n in ("j1", "j2m1", "jmix") else x
                    break
            for label, name in pressure_keys:
                if l.strip().startswith(label):
                    r[name] = vals(i, 1, after=label)[0]
        recs.append(r)
    return recs

R37 = "/root/reference/Multall-open-18.3/multall-test-cases/r37+samods-17.5.dat"


def r37_plain_sa() -> str:
    """The reference's own Rotor 37 Spalart-Allmaras deck (multall-test-cases/
    r37+samods-17.5.dat) with its two SA modifications (FAC_VORT, FAC_PGRAD)
    and the Shih wall functions (YPLUSWALL = -11) switched off: plain SA with
    the log-law wall treatment."""
    lines = open(R37).read().split("\n")
    i = next(n for n, l in enumerate(lines) if l.strip().upper().startswith("REYNOLDS NUMBER"))
    v = lines[i + 1].split()
    assert len(v) == 6 and float(v[5]) == -11.0, v
    lines[i + 1] = "  " + "    ".join(v[:5] + ["0.0000"])
    f = lines[i + 3].split()
    assert len(f) == 8, f
    lines[i + 3] = "    " + "    ".join(f[:6] + ["0.0", "0.0"])
    return "\n".join(lines)


def r37_samods() -> str:
    """The reference's own Rotor 37 Spalart-Allmaras deck as it ships: the
    helicity and pressure-gradient source factors (FAC_VORT = FAC_PGRAD = 1)
    and the Shih wall functions with the pressure-gradient term
    (YPLUSWALL = -11)."""
    return open(R37).read()


def centrifugal() -> str:
    from multall_hip import deckgen
    return deckgen.config4_deck(open(os.path.join(os.path.dirname(R37), deckgen.C4_BASE)).read(), nsteps=9000)


def shipped(name: str):
    """One of the reference's own multi-row test decks as it ships, except
    that a comment line opening with "(" loses that parenthesis (flang's
    list-directed read of the comment into a CHARACTER variable stops on it;
    no input value changes)."""
    def text() -> str:
        lines = open(os.path.join(os.path.dirname(R37), name), errors="replace").read().split("\n")
        return "\n".join(ln.replace("(", " ", 1) if ln.strip().startswith("(") else ln for ln in lines)
    return text


CASES = {
    "r37sa": r37_plain_sa,
    "r37samods": r37_samods,
    # a one-stage axial compressor and a one-stage axial turbine (rotor and
    # stator across a mixing plane, gaps re-gridded by the reader)
    "axialcompr": shipped("simple-axial-compr-17.1.dat"),
    "axialturb": shipped("Single-stage-turbine-17.1.dat"),
    # the six-stage compressor: twelve rows, eleven mixing planes
    "sixstage": shipped("Six-stage-compr-17.1.dat"),
    # the centrifugal stage: the centrifugal compressor deck (impeller and
    # vaned diffuser) on 97 x 81 with Spalart-Allmaras
    "centrifugal": centrifugal,
}
# the reference builds (FP64, FP32) a case runs on: decks longer than the
# main builds' JD = 600 on the long builds, the centrifugal stage on its own
BUILDS = {"sixstage": ("fp64l", "fp32l"), "centrifugal": ("fp64c", "fp32c")}


def run_one(txt: str, variant: str, real) -> dict:
    """One reference run to its stop: the step it stopped at, the final
    per-J integrals and the last EFICOOL block it printed."""
    import refdump
    wd = tempfile.mkdtemp(prefix="mhconv_")
    try:
        _, final, stdout = refdump.run_reference(txt, variant, workdir=wd, timeout=14400)
        integ = read_global_plt(os.path.join(wd, "global.plt"), real)
    finally:
        shutil.rmtree(wd, ignore_errors=True)
    return {"nstep": int(final["NSTEP"]), "integrals": integ.astype(np.float64).tolist(),
            "eficool": eficool_records(stdout)}


def make(case):
    txt = CASES[case]()
    deck = f"converged_{case}.dat"
    with open(os.path.join(HERE, "decks", deck), "w") as f:
        f.write(txt)
    out = {"deck": f"decks/{deck}"}
    for (variant, real), key in zip(zip(BUILDS.get(case, ("fp64", "fp32")), (np.float64, np.float32)),
                                    ("fp64", "fp32")):
        out[key] = run_one(txt, variant, real)
        print(f"{case} {variant}: stopped at step {out[key]['nstep']}", flush=True)
    with open(os.path.join(HERE, f"converged_{case}.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    for c in (sys.argv[1:] or list(CASES)):
        make(c)
