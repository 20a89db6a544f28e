#!/usr/bin/env python3
"""Throughput of the time-marching step on one MI355X.

Workload: the 1.5-stage axial turbine (stator-rotor-stator, 3 rows of
64 x 64 x 193 nodes, 48/60/48 blades, rotor at 12000 rpm with a 0.5 mm tip
gap over 4 K-cells, p0 = 2.0e5 Pa, T0 = 1000 K, exit p = 1.1e5 Pa,
Spalart-Allmaras turbulence), built from a synthetic NEW-format deck by the
Fortran host set-up and marched by libmultall_hip.  A "step" is one time step
of the solver (the body of the reference's LOOP) over all 2,371,584 nodes.
`value` is the plain FP32 form; the mixed form (SSS histories stored in BF16,
FP32 arithmetic, FP64 reductions) is reported beside it.

Prints one JSON line: node-steps/s (whole job), a check of the same run
against the reference's 100-step summary of this deck, the roofline of the
dominant kernel (HIP-event durations on the library's stream, HBM traffic
from the committed counter table), the step-level fraction of the 8 TB/s HBM
roofline at the step's algorithmic bytes per node, the reference solver
timed on this host's CPU (one core, 100 steps differenced), the throughput
of the other machines of the benchmark set in both forms (the reference's
six-stage compressor deck; its centrifugal compressor deck, impeller and
vaned diffuser, on 97 x 81 with Spalart-Allmaras), the time to
converge Rotor 37 (the reference's own deck as it ships), and the
16-point compressor speedline run to per-point convergence: plain FP32 one
point after another, batched FP32, batched mixed, and the reference on 16
host cores at once.

  python bench.py [--steps K] [--warmup W]
  python bench.py --pmc-table FETCH.csv WRITE.csv COPY_FETCH.csv COPY_WRITE.csv
      (writes the HBM-traffic table under profiles/ from rocprofv3 --pmc runs)
  python bench.py --stats-table KERNEL_STATS.csv STEPS
      (prints the per-kernel time table of a rocprofv3 --stats run)
  python bench.py --converged-profile STEPS
      (per-kernel HIP-event times of the time-to-converged workload, Rotor 37)
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import platform
import resource
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "multall-open-18.3_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per node-step of a fused step: U and its change (5 FP32
# read + written), the SSS history of each equation (read + written), the
# face areas and volumes, the body forces and the time step; +20 B for the
# Spalart-Allmaras variable and its history.  The mixed form keeps the six
# histories in BF16: 2 B less each way per history.
B_ALG = {("plain", False): 132.0, ("plain", True): 152.0, ("mixed", False): 112.0, ("mixed", True): 128.0}

# Algorithmic bytes per node and step of each kernel family on the bench
# workload (Spalart-Allmaras, the fast mode's batched launches): every array
# a launch reads or writes, counted once (stencil neighbours are re-use, not
# traffic; (J,K)- and I-shaped arrays are negligible), summed over the
# step's launches.  Cell arrays are node-shaped, so per-node == per-cell.
#   k_massflux_eq: ROVX ROVR RO ROVT and six face areas, HO PEFF QSOURCE,
#     TRANS_KVIS T_SOURCE read; FLOWX/T/R ROWT WT and the energy and
#     transport fluxes (4 each) written: 15 + 13 arrays
#   k_flux_mom: PEFF VX VT VR ROVT FLOWX/T/R, eight face areas and VOLOR,
#     three body forces read; 3 x 4 fluxes written: 18 + 12
#   k_resid_line: per variable XFLUX TFLUX RFLUX SOURCE and the history
#     read, STORE and the history written (7; the mixed form's BF16 history
#     2 B each way), six variables a step in two launches; continuity has
#     no source array
#   k_mg_apply_lines: per variable STORE read and written, STEP and RSTEP
#     once per launch (the variables of a launch share them in L2): 3 + 3
#     variables
#   k_damp_scatter*: per variable STORE and the variable read, DSC written,
#     five variables; k_smooth_reg* / k_smooth_plane: DSC read, the variable
#     written, five variables
#   k_sa_tail1: the transport variable and its STORE read, TMP written;
#     k_sa_tail2: TMP read, the variable written
KERNEL_BYTES_PER_NODE_STEP = {
    "k_prim": 4 * (5 + 7), "k_massflux_eq": 4 * (15 + 13), "k_flux_mom": 4 * (18 + 12),
    "k_resid_line": 6 * 4 * 7 - 4, "k_resid_line_bf": 6 * (4 * 5 + 2 * 2) - 4,
    "k_mg_apply_lines": 2 * (3 * 4 * 2 + 4 * 2), "k_damp_scatter1": 5 * 4 * 3, "k_damp_scatter2": 5 * 4 * 3,
    "k_damp_scatter4": 5 * 4 * 3,
    "k_smooth_reg1": 5 * 4 * 2, "k_smooth_reg2": 5 * 4 * 2, "k_smooth_reg4": 5 * 4 * 2,
    "k_smooth_reg8": 5 * 4 * 2, "k_smooth_reg16": 5 * 4 * 2, "k_smooth_plane": 5 * 4 * 2,
    "k_sa_cellavg_kvis": 4 * (2 + 3), "k_sa_tail1": 4 * (2 + 1), "k_sa_tail2": 4 * 2, "k_peff": 4 * 2,
}

# HBM traffic per launch of the hot kernels on this workload, written by
# `bench.py --pmc-table` from rocprofv3 FETCH_SIZE and WRITE_SIZE passes, each
# scaled by what the same counter reads for a copy of known size.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "tstep_kernels_config1_sa_pmc_traffic.json")
# launches per step of each family on that workload
LAUNCHES_PER_STEP = {"k_prim": 1, "k_massflux_eq": 1, "k_flux_mom": 1, "k_resid_line": 2, "k_resid_line_bf": 2,
                     "k_mg_apply_lines": 2, "k_damp_scatter1": 2, "k_damp_scatter2": 2, "k_damp_scatter4": 2, "k_smooth_reg1": 2,
                     "k_smooth_reg2": 2, "k_smooth_reg4": 2, "k_smooth_reg8": 2, "k_smooth_reg16": 2,
                     "k_smooth_plane": 2, "k_sa_cellavg_kvis": 1, "k_sa_tail1": 1, "k_sa_tail2": 1, "k_peff": 1}


def alg_bytes_per_launch(kernel: str, nodes: int) -> float:
    fam = base_name(kernel)
    n = LAUNCHES_PER_STEP.get(fam)
    return KERNEL_BYTES_PER_NODE_STEP.get(fam, 0) * nodes / n if n else 0.0


PMC_KERNELS = ["k_massflux_eq", "k_resid_line", "k_flux_mom", "k_damp_scatter1", "k_smooth_reg8", "k_mg_apply_lines",
               "k_prim", "k_sa_tail1", "k_sa_tail2", "k_sa_cellavg_kvis"]
NODES_C1 = 3 * 64 * 64 * 193


def base_name(kernel: str) -> str:
    """A kernel's family name: tile-shape variants (k_damp_scatter1_k4,
    k_smooth_reg4w16) share their family's bytes per node."""
    import re
    return re.sub(r"w\d+$", "", re.sub(r"_k\d+$", "", kernel))


def traffic_bytes(kernel: str):
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    if not k:
        return None
    cal = t["calibration"]
    return (k["fetch_kib"] * cal["fetch_scale"] + k["write_kib"] * cal["write_scale"]) * 1024.0


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_cores() -> list:
    try:
        return sorted(os.sched_getaffinity(0))
    except AttributeError:
        return list(range(os.cpu_count() or 1))


# ---------------------------------------------------------------- workloads

def workload_deck(nsteps: int, ilos: int) -> str:
    from multall_hip import deckgen
    return deckgen.write_deck(deckgen.deck_config1(nsteps=nsteps, ilos=ilos))


def load_state(deck_text: str) -> dict:
    from multall_hip import host
    with tempfile.NamedTemporaryFile("w", suffix=".dat", delete=False) as f:
        f.write(deck_text)
        path = f.name
    try:
        return host.load_deck(path)
    finally:
        os.unlink(path)


def _copy_state(st: dict) -> dict:
    return {k: (v if k != "scalars" else dict(v)) for k, v in st.items()}


def timed_form(st: dict, mixed: bool, warmup: int, steps: int, tiles=None) -> tuple:
    """One Solver on the workload: warm-up steps, then `steps` timed steps
    between two device synchronisations.  Returns (node-steps/s, ms/step,
    the open Solver).  tiles = (TK, NW) fixes the tile shapes instead of
    timing them at the start (so a counter run sees only the step's
    launches)."""
    from multall_hip.runtime import Solver
    s = st["scalars"]
    nodes = int(s["IM"] * s["JM"] * s["KM"])
    g = Solver(_copy_state(st), exact_order=False)
    g.set_scalar("MIXED", 1 if mixed else 0)
    if tiles:
        g.set_scalar("SCAT_TK", tiles[0])
        g.set_scalar("SMOOTH_WAVES", tiles[1])
    g.loss_init()          # SETUP's initial viscous forces
    g.start()
    g.step(warmup)
    g.sync()
    t0 = time.perf_counter()
    r = g.step(steps)
    g.sync()
    dt = time.perf_counter() - t0
    assert r["nstep"] == warmup + steps, r
    return nodes * steps / dt, dt / steps * 1e3, g


def check_vs_reference(st: dict, mixed: bool) -> dict:
    """The workload marched 100 steps from the reference's set-up and
    compared with the reference's own 100-step summaries of this deck
    (tests/golden/config1_summary100_fp64.npz and config1_summary100.npz,
    its FP64 and FP32 builds): the largest difference of any pitchwise line
    mean from the FP64 build's, in each conserved variable's scale, beside
    the FP32 build's own distance from it."""
    from multall_hip.runtime import Solver
    path = os.path.join(GOLDEN, "config1_summary100_fp64.npz")
    path32 = os.path.join(GOLDEN, "config1_summary100.npz")
    if not os.path.exists(path) or not os.path.exists(path32):
        return None
    ref = np.load(path, allow_pickle=False)
    ref32 = np.load(path32, allow_pickle=False)
    st = _copy_state(st)
    st["scalars"]["NSTEPS_MAX"] = 100
    g = Solver(st, exact_order=False)
    g.set_scalar("MIXED", 1 if mixed else 0)
    g.loss_init()
    g.start()
    r = g.step(110)
    worst = worst32 = 0.0
    for v in ("RO", "ROVX", "ROVR", "RORVT", "ROE"):
        lines = g.get(v).astype(np.float64).mean(axis=2)
        want = ref[f"{v}_lines"].astype(np.float64)
        scale = np.abs(ref[f"{v}_plane"]).max() if v in ("RO", "ROE", "RORVT") else np.sqrt(
            ref["ROVX_plane"].astype(np.float64) ** 2 + ref["ROVR_plane"].astype(np.float64) ** 2).max()
        worst = max(worst, float(np.abs(lines - want).max() / scale))
        worst32 = max(worst32, float(np.abs(ref32[f"{v}_lines"].astype(np.float64) - want).max() / scale))
    g.close()
    return {"steps": int(r["nstep"]), "max_line_mean_err_vs_fp64_reference": worst,
            "fp32_reference_vs_fp64_reference": worst32}


CONVERGED_DECK = os.path.join(GOLDEN, "decks", "converged_r37samods.dat")
# The other machines of the benchmark set, as the reference can state them
# (one IM and KM for all rows): the multi-stage compressor is the reference's
# own six-stage compressor test deck (12 rows of 37 x 37 on 1282 J-points,
# mixing length, total pressure ratio 3.44 at convergence), the centrifugal
# stage its centrifugal compressor deck (impeller and vaned diffuser) on a
# uniform 97 x 81 grid with Spalart-Allmaras (deckgen.config4_deck).
MULTISTAGE_DECK = os.path.join(GOLDEN, "decks", "converged_sixstage.dat")
CENTRIFUGAL_DECK = os.path.join(GOLDEN, "decks", "config4.dat")


def other_workload(path: str, title: str, warmup: int, steps: int) -> dict:
    """Throughput of another machine: its deck read and set up by the host
    (wall distances and mixing lengths on the device), then warm-up steps
    and `steps` timed steps, plain FP32 and mixed, as for the main
    workload."""
    from multall_hip import host
    from multall_hip.runtime import Solver
    st = host.load_deck(path, xlength_on_host=False)
    sc = st["scalars"]
    sc["NSTEPS_MAX"] = warmup + steps + 1000
    nodes = int(sc["IM"] * sc["JM"] * sc["KM"])
    out = {"workload": title, "nodes": nodes, "grid": [int(sc["IM"]), int(sc["JM"]), int(sc["KM"])],
           "rows": int(sc["NROWS"]), "viscous_model": f"ILOS={int(sc['ILOS'])}"}
    sa = int(sc["ILOS"]) >= 100
    for mixed in (False, True):
        g = Solver(_copy_state(st), exact_order=False)
        g.set_scalar("MIXED", 1 if mixed else 0)
        if sa:
            g.set_xlength(host.array("THETA", (sc["KM"], sc["JM"], sc["IM"])),
                          *[host.array(n, (sc["NROWS"],)) for n in ("XLLIM_IN", "XLLIM_LE", "XLLIM_TE", "XLLIM_DN")])
        g.loss_init()
        g.start()
        g.step(warmup)
        g.sync()
        t0 = time.perf_counter()
        r = g.step(steps)
        g.sync()
        dt = time.perf_counter() - t0
        g.close()
        assert r["nstep"] == warmup + steps, r
        b = B_ALG[("mixed" if mixed else "plain", sa)]
        out["mixed" if mixed else "plain"] = {
            "value": nodes * steps / dt, "ms_per_step": dt / steps * 1e3,
            "step_roofline": {"alg_bytes_per_node_step": b, "frac": nodes * steps / dt * b / 1e9 / HBM_PEAK_GBS}}
    return out


def time_to_converged(mixed: bool) -> dict:
    """Rotor 37 (the reference's own test deck as it ships: Spalart-Allmaras
    with its helicity and pressure-gradient source factors, Shih wall
    functions; 45 x 150 x 49 nodes) marched from the reference's initial
    guess until the reference's stop test ends the run (mf:3762-3764): wall
    seconds for the steps alone and including the deck read and host
    set-up (the wall distances and mixing lengths formed on the device)."""
    from multall_hip import host
    from multall_hip.runtime import Solver
    t0 = time.perf_counter()
    st = host.load_deck(CONVERGED_DECK, xlength_on_host=False)
    sc = st["scalars"]
    g = Solver(st, exact_order=False)
    g.set_scalar("MIXED", 1 if mixed else 0)
    g.set_xlength(host.array("THETA", (sc["KM"], sc["JM"], sc["IM"])),
                  *[host.array(n, (sc["NROWS"],)) for n in ("XLLIM_IN", "XLLIM_LE", "XLLIM_TE", "XLLIM_DN")])
    g.loss_init()
    g.start()
    g.sync()
    t1 = time.perf_counter()
    r = g.step(int(sc["NSTEPS_MAX"]) + 10)
    g.sync()
    t2 = time.perf_counter()
    nodes = int(sc["IM"] * sc["JM"] * sc["KM"])
    q = machine(g.integrals(), float(sc["GA"]))
    g.close()
    return {"form": "mixed" if mixed else "plain FP32",
            "converged": bool(r["ifend"] == 1 and r["nstep"] < int(sc["NSTEPS_MAX"])),
            "steps": int(r["nstep"]), "seconds": t2 - t1, "seconds_with_setup": t2 - t0,
            "node_steps_per_s": nodes * r["nstep"] / (t2 - t1), "eta_tt": q["eta"],
            "final": {"eavg": r["eavg"], "emax": r["emax"], "econt": r["econt"]}}


def machine(integ, ga: float) -> dict:
    """Mass flow, mass-averaged total-pressure and total-temperature ratios
    and total-to-total isentropic efficiency from the per-J FP64 integrals
    (BLADE_FLOW, SUMPO, SUMTO)."""
    flow, spo, sto = (np.asarray(integ[q], np.float64) for q in range(3))
    p0r = (spo[-1] / flow[-1]) / (spo[0] / flow[0])
    t0r = (sto[-1] / flow[-1]) / (sto[0] / flow[0])
    x = (ga - 1.0) / ga
    eta = (p0r ** x - 1.0) / (t0r - 1.0) if p0r > 1.0 else (1.0 - t0r) / (1.0 - p0r ** x)
    return {"flow": float(flow[0]), "p0_ratio": float(p0r), "eta": float(eta)}


# ---------------------------------------------------------------- speedline

SPEEDLINE_POINTS = 16
FLOW_CONV = 1.0e-5


def _speedline_base() -> str:
    from multall_hip import deckgen
    with open(os.path.join(GOLDEN, "decks", "ref", deckgen.SPEEDLINE_BASE)) as f:
        return f.read()


def speedline_text(i: int, nsteps: int = 6000) -> str:
    from multall_hip import deckgen
    return deckgen.speedline_deck(_speedline_base(), i, SPEEDLINE_POINTS, nsteps=nsteps)


def _speedline_state(i: int):
    """Deck read and host set-up of speedline point i, in a worker process
    (the Fortran host holds one deck at a time); the wall distances are
    left to the device."""
    from multall_hip import host
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, f"point{i}.dat")
        with open(path, "w") as f:
            f.write(speedline_text(i))
        st = host.load_deck(path, xlength_on_host=False)
        s = st["scalars"]
        st["_xl"] = (host.array("THETA", (s["KM"], s["JM"], s["IM"])),
                     [host.array(n, (s["NROWS"],)) for n in ("XLLIM_IN", "XLLIM_LE", "XLLIM_TE", "XLLIM_DN")])
        return st


def speedline_states() -> list:
    """The speedline's set-up states, read in worker processes before this
    process touches the GPU."""
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(min(8, SPEEDLINE_POINTS)) as ex:
        return list(ex.map(_speedline_state, range(SPEEDLINE_POINTS)))


def _speedline_solvers(states: list, mixed: bool) -> list:
    from multall_hip.runtime import Solver
    out = []
    for st in states:
        g = Solver({k: v for k, v in st.items() if k != "_xl"}, exact_order=False)
        g.set_scalar("MIXED", 1 if mixed else 0)
        g.set_scalar("FLOW_CONV", FLOW_CONV)
        theta, xl = st["_xl"]
        g.set_xlength(theta, *xl)
        g.loss_init()
        g.start()
        out.append(g)
    for g in out:
        g.sync()
    return out


def speedline(states: list) -> dict:
    """The 16-point speedline run to per-point convergence, three ways from
    the same set-up: plain FP32 one point after another (mh_step), batched
    plain FP32 and batched mixed (mh_step_batch, every point's chunks of
    steps captured as HIP graphs on a stream of its own).  A point stops at
    the reference's own test or when its mass-flow change between the
    every-5-step checks stays below 1e-5 of the flow for 10 checks in a row
    with the continuity error ECONT below 0.01 (all in FP64 on the device)."""
    from multall_hip.runtime import step_batch
    nodes = [int(s["scalars"]["IM"] * s["scalars"]["JM"] * s["scalars"]["KM"]) for s in states]
    nmax = int(states[0]["scalars"]["NSTEPS_MAX"]) + 10
    ga = float(states[0]["scalars"]["GA"])

    def finish(sol, res, secs):
        pts = []
        for g, r, n in zip(sol, res, nodes):
            q = machine(g.integrals(), ga)
            pts.append({"nstep": int(r["nstep"]), "converged": bool(r["nstep"] < nmax - 10), "econt": r["econt"],
                        **q})
            g.close()
        work = sum(n * p["nstep"] for n, p in zip(nodes, pts))
        return {"seconds": secs, "node_steps": work, "node_steps_per_s": work / secs,
                "all_converged": all(p["converged"] for p in pts)}, pts

    print("speedline: plain FP32, one point after another", file=sys.stderr, flush=True)
    sol = _speedline_solvers(states, False)
    t0 = time.perf_counter()
    res = []
    for g in sol:
        res.append(g.step(nmax))
        g.sync()
    seq, pts_seq = finish(sol, res, time.perf_counter() - t0)
    out = {}
    for name, mixed in (("batched_fp32", False), ("batched_mixed", True)):
        print(f"speedline: {name}", file=sys.stderr, flush=True)
        sol = _speedline_solvers(states, mixed)
        t0 = time.perf_counter()
        res = step_batch(sol, nmax)
        for g in sol:
            g.sync()
        out[name], pts = finish(sol, res, time.perf_counter() - t0)
        out[name]["points"] = pts if name == "batched_mixed" else None
        out[name]["max_eta_diff_pts_vs_sequential_fp32"] = max(
            abs(a["eta"] - b["eta"]) * 100 for a, b in zip(pts, pts_seq))
        out[name]["max_flow_diff_rel_vs_sequential_fp32"] = max(
            abs(a["flow"] - b["flow"]) / b["flow"] for a, b in zip(pts, pts_seq))
        if name == "batched_fp32":
            out[name]["same_steps_as_sequential"] = [p["nstep"] for p in pts] == [p["nstep"] for p in pts_seq]
            out[name]["points"] = None
    return {"workload": f"{SPEEDLINE_POINTS}-point speedline: the reference's one-stage axial compressor deck "
                        f"(rotor + stator, 2 x 111 J-points, mixing length) re-gridded IM from {{37, 49}}, KM from "
                        f"{{33, 49, 65}} (seeded), exit p 0.95-1.08 x design; {sum(nodes)} nodes "
                        f"({min(nodes)}-{max(nodes)} per point)",
            "stop": "reference stop test, or mass-flow change < 1e-5 for 10 checks with ECONT < 0.01",
            "sequential_fp32": seq, **out,
            "batched_mixed_vs_sequential_fp32_time": seq["seconds"] / out["batched_mixed"]["seconds"],
            "batched_fp32_vs_sequential_fp32_time": seq["seconds"] / out["batched_fp32"]["seconds"]}


# ---------------------------------------------------------------- CPU baselines

REF_EXE = os.path.join(ROOT, "oracle", "_ref", "multall_ref_fp32")


def _ref_loop_seconds(deck: str, core: int):
    """Start the reference (FP32, one core) on a deck; returns the Popen and
    its work directory.  The loop time is read back by _ref_wait."""
    wd = tempfile.mkdtemp(prefix="mhcpu_")
    with open(os.path.join(wd, "intype"), "w") as f:
        f.write("N\n")
    env = dict(os.environ, MH_NODUMP="1", OMP_NUM_THREADS="1")

    def pin():
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
        os.sched_setaffinity(0, {core})
    p = subprocess.Popen([REF_EXE], stdin=subprocess.PIPE, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                         cwd=wd, env=env, preexec_fn=pin, text=True)
    p.stdin.write(deck)
    p.stdin.close()
    return p, wd


def _ref_wait(p, wd):
    import shutil
    try:
        p.wait(timeout=900)
        if p.returncode != 0:
            return None
        raw = open(os.path.join(wd, "final.bin"), "rb").read(24)
        return float(np.frombuffer(raw, np.float64, 1, 16)[0])
    finally:
        shutil.rmtree(wd, ignore_errors=True)


def _ref_per_step(decks_a, decks_b, steps: int, cores: list):
    """Loop seconds per step of each deck: runs of n_a and n_a + steps steps
    (the decks_a / decks_b texts), all of one length at once, one process per
    core; per-step time from the difference, which cancels the set-up and the
    end-of-run output."""
    out = []
    for decks in (decks_a, decks_b):
        procs = [_ref_loop_seconds(d, cores[i % len(cores)]) for i, d in enumerate(decks)]
        out.append([_ref_wait(p, wd) for p, wd in procs])
    if any(t is None for t in out[0] + out[1]):
        return None
    return [(b - a) / steps for a, b in zip(*out)]


def cpu_baseline(ilos: int, steps: int = 100, base: int = 5) -> dict:
    """The reference solver (oracle/_ref, FP32 -O3, single-threaded) on the
    bench workload's own deck, one core: per-step loop time over `steps`
    steps (runs of `base` and `base + steps` steps differenced)."""
    if not os.path.exists(REF_EXE):
        return None
    cores = host_cores()
    # the two runs on two cores at once (the second core when the host has one)
    t = []
    for n in (base, base + steps):
        p, wd = _ref_loop_seconds(workload_deck(n, ilos), cores[len(t) % len(cores)])
        t.append((p, wd))
    times = [_ref_wait(p, wd) for p, wd in t]
    if any(x is None for x in times):
        return None
    per = (times[1] - times[0]) / steps
    return {"value": NODES_C1 / per, "unit": "node-steps/s", "cores": 1, "kind": "reference",
            "cpu": cpu_model(),
            "sample": f"reference FP32 -O3 (amdflang), the bench workload's deck, loop time of {base + steps} "
                      f"minus {base} steps over {steps} steps (20 viscous refreshes at NLOS=5), one process "
                      f"pinned to one core; {per:.3f} s/step"}


def cpu_baseline_speedline(nodes: list, steps: int = 100, base: int = 5) -> dict:
    """The reference on the 16 speedline points at once, one process per
    host core (at most 16), each pinned: aggregate node-steps/s from each
    point's per-step loop time over `steps` steps."""
    if not os.path.exists(REF_EXE):
        return None
    cores = host_cores()[:SPEEDLINE_POINTS]
    per = _ref_per_step([speedline_text(i, base) for i in range(SPEEDLINE_POINTS)],
                        [speedline_text(i, base + steps) for i in range(SPEEDLINE_POINTS)], steps, cores)
    if per is None:
        return None
    return {"value": sum(n / s for n, s in zip(nodes, per)), "unit": "node-steps/s", "cores": len(cores),
            "kind": "reference", "cpu": cpu_model(),
            "sample": f"reference FP32 -O3, {SPEEDLINE_POINTS} speedline points at once on {len(cores)} cores "
                      f"({-(-SPEEDLINE_POINTS // len(cores))} per core), loop time of {base + steps} minus {base} "
                      f"steps each; per point {min(per):.3f}-{max(per):.3f} s/step"}


# ---------------------------------------------------------------- profiles

def _counter_rows(path: str, counter: str) -> dict:
    """Mean counter value (KiB) per kernel name from a rocprofv3
    counter_collection.csv."""
    acc = {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] != counter:
                continue
            name = row["Kernel_Name"].split("(")[0].split("<")[0].strip().split("::")[-1]
            s, n = acc.get(name, (0.0, 0))
            acc[name] = (s + float(row["Counter_Value"]), n + 1)
    return {k: s / n for k, (s, n) in acc.items()}


def pmc_table(fetch_csv: str, write_csv: str, copy_fetch_csv: str, copy_write_csv: str) -> None:
    """Write the HBM-traffic table (JSON for this script, markdown for a
    reader) from two counter passes over the bench workload and the same two
    counters on a 256 MiB device copy (bench.py --calibrate-copy), which
    reads and writes 262144 KiB per launch."""
    fetch, write = _counter_rows(fetch_csv, "FETCH_SIZE"), _counter_rows(write_csv, "WRITE_SIZE")
    cf = max(v for k, v in _counter_rows(copy_fetch_csv, "FETCH_SIZE").items())
    cw = max(v for k, v in _counter_rows(copy_write_csv, "WRITE_SIZE").items())
    cal = {"copy_kib": 262144.0, "fetch_scale": 262144.0 / cf, "write_scale": 262144.0 / cw}
    kernels = {k: {"fetch_kib": fetch[k], "write_kib": write.get(k, 0.0)} for k in fetch
               if any(base_name(k).startswith(m) for m in PMC_KERNELS)}
    with open(TRAFFIC_FILE, "w") as f:
        json.dump({"workload": "1.5-stage turbine, 3 rows x 64x64x193 nodes, Spalart-Allmaras, plain FP32",
                   "calibration": cal, "kernels": kernels}, f, indent=1)
    rows = []
    for k, v in sorted(kernels.items(), key=lambda kv: -(kv[1]["fetch_kib"] + kv[1]["write_kib"])):
        mb = (v["fetch_kib"] * cal["fetch_scale"] + v["write_kib"] * cal["write_scale"]) * 1024 / 1e6
        alg = alg_bytes_per_launch(k, NODES_C1) / 1e6
        rows.append(f"| {k} | {v['fetch_kib']:.0f} | {v['write_kib']:.0f} | {mb:.1f} | {alg:.1f} | "
                    f"{mb / alg if alg else float('nan'):.2f} |")
    md = TRAFFIC_FILE[:-5] + ".md"
    with open(md, "w") as f:
        f.write("# HBM traffic per launch of the hot kernels: the bench workload (1.5-stage turbine, 2,371,584 "
                "nodes, Spalart-Allmaras, plain FP32)\n\n")
        f.write("Made by `python3 bench.py --pmc-table` from four rocprofv3 counter passes (from the repository "
                "root on the MI355X box; `cd /tmp && export TMPDIR=/tmp` first):\n\n```\n")
        f.write("rocprofv3 --pmc FETCH_SIZE --kernel-include-regex '" + "|".join(PMC_KERNELS) + "' \\\n"
                "    --output-format csv -d <out>/fetch -o run -- python3 bench.py --steps 3 --warmup 1 --tiles 4 8 "
                "--no-cpu-baseline --no-converged --no-speedline --no-mixed --no-check --no-machines\n"
                "rocprofv3 --pmc WRITE_SIZE  (same filter and command, -d <out>/write)\n"
                "rocprofv3 --pmc FETCH_SIZE --kernel-include-regex 'copy|Copy' --output-format csv -d <out>/cfetch "
                "-o run -- python3 bench.py --calibrate-copy\n"
                "rocprofv3 --pmc WRITE_SIZE  (same, -d <out>/cwrite)\n```\n\n")
        f.write(f"Scale: the copy reads and writes 262144 KiB per launch; FETCH_SIZE reports {cf:.0f} KiB "
                f"(factor {cal['fetch_scale']:.3f}), WRITE_SIZE {cw:.0f} KiB (factor {cal['write_scale']:.3f}). "
                f"Traffic = FETCH_SIZE x {cal['fetch_scale']:.3f} + WRITE_SIZE x {cal['write_scale']:.3f}, mean "
                f"over the launches in the run.\n\n")
        f.write("| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | traffic MB | algorithmic MB | traffic / algorithmic |\n"
                "|---|---:|---:|---:|---:|---:|\n")
        f.write("\n".join(rows) + "\n")
    print(f"wrote {TRAFFIC_FILE} and {md}")


def stats_table(stats_csv: str, steps: int) -> None:
    """Per-kernel time table (markdown) of a rocprofv3 --stats
    kernel_stats.csv over a run of `steps` library steps."""
    rows = []
    with open(stats_csv, newline="") as f:
        for row in csv.DictReader(f):
            name = row["Name"].split("(")[0].split("<")[0].strip().split("::")[-1]
            rows.append((name, int(row["Calls"]), float(row["TotalDurationNs"]) / 1e3,
                         float(row["AverageNs"]) / 1e3, float(row["Percentage"])))
    tot = sum(r[2] for r in rows)
    print(f"Sum of kernel durations {tot / 1e3:.2f} ms = {tot / steps / 1e3:.3f} ms per step over {steps} steps.\n")
    print("| kernel | calls | avg us | us/step | % |\n|---|---:|---:|---:|---:|")
    for name, calls, tus, avg, pct in sorted(rows, key=lambda r: -r[2]):
        print(f"| {name} | {calls} | {avg:.2f} | {tus / steps:.1f} | {pct:.2f} |")


def converged_profile(steps: int) -> None:
    """Per-kernel durations of the time-to-converged workload (Rotor 37 as
    it ships, fast mode, plain FP32): 100 steps from the initial guess, the
    wall time of 500 steps, then `steps` steps launched one by one with
    dispatch-timestamped HIP events on the library's stream (a markdown
    table: launches, mean duration, time per step)."""
    from multall_hip import host
    from multall_hip.runtime import Solver
    st = host.load_deck(CONVERGED_DECK, xlength_on_host=False)
    sc = st["scalars"]
    g = Solver(st, exact_order=False)
    g.set_xlength(host.array("THETA", (sc["KM"], sc["JM"], sc["IM"])),
                  *[host.array(n, (sc["NROWS"],)) for n in ("XLLIM_IN", "XLLIM_LE", "XLLIM_TE", "XLLIM_DN")])
    g.loss_init()
    g.start()
    g.step(100)
    g.sync()
    t = time.perf_counter()
    g.step(500)
    g.sync()
    wall = (time.perf_counter() - t) / 500
    g.profile(True)
    g.step(steps)
    rep = g.profile_report()
    g.profile(False)
    g.close()
    busy = sum(ms for c, ms in rep.values()) / steps
    launches = sum(c for c, ms in rep.values()) / steps
    print(f"Grid {sc['IM']} x {sc['JM']} x {sc['KM']}; {wall * 1e3:.4f} ms/step wall over 500 steps as captured "
          f"graphs; {steps} steps launched one by one: {busy * 1e3:.1f} us/step of kernel time, "
          f"{launches:.1f} launches/step.\n")
    print("| kernel | calls | avg us | us/step |\n|---|---:|---:|---:|")
    for k, (c, ms) in sorted(rep.items(), key=lambda kv: -kv[1][1]):
        print(f"| {k} | {c} | {ms / c * 1e3:.2f} | {ms / steps * 1e3:.2f} |")


def calibrate_copy() -> None:
    """One device-to-device copy of 256 MiB, three times, for scaling the
    FETCH_SIZE / WRITE_SIZE counters (each launch reads and writes exactly
    268435456 bytes)."""
    import torch
    n = 1 << 26
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": 4 * n, "copies": 3}))


# ---------------------------------------------------------------- main

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--ilos", type=int, default=200,
                    help="viscous model of the deck: 200 Spalart-Allmaras (the workload's), 10 mixing length, 0 inviscid")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-converged", action="store_true", help="skip the time-to-converged runs")
    ap.add_argument("--no-speedline", action="store_true", help="skip the speedline runs")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-form run of the workload")
    ap.add_argument("--no-check", action="store_true", help="skip the 100-step checks against the reference")
    ap.add_argument("--no-machines", action="store_true",
                    help="skip the multi-stage compressor and centrifugal stage throughput runs")
    ap.add_argument("--tiles", nargs=2, type=int, metavar=("TK", "NW"),
                    help="fix the corner-distribution and SMOOTH_VAR tile shapes (no tuning launches)")
    ap.add_argument("--converged-profile", type=int, metavar="STEPS",
                    help="print the per-kernel time table of the time-to-converged workload (Rotor 37)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 256 MiB device copy (for scaling the PMC byte counters)")
    ap.add_argument("--pmc-table", nargs=4, metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="write the traffic table under profiles/ from rocprofv3 counter_collection.csv files")
    ap.add_argument("--stats-table", nargs=2, metavar=("KERNEL_STATS_CSV", "STEPS"),
                    help="print the per-kernel time table of a rocprofv3 --stats run")
    args = ap.parse_args()
    if args.pmc_table:
        pmc_table(*args.pmc_table)
        return
    if args.stats_table:
        stats_table(args.stats_table[0], int(args.stats_table[1]))
        return
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.converged_profile:
        converged_profile(args.converged_profile)
        return
    if args.gpus != 1:
        print("bench.py measures one MI355X: multi-GPU runs are not part of this build "
              "(rows and operating points would shard across GPUs)", file=sys.stderr)
        sys.exit(2)

    sl_states = None if args.no_speedline else speedline_states()
    st = load_state(workload_deck(args.warmup + args.steps + 1000, args.ilos))
    s = st["scalars"]
    nodes = int(s["IM"] * s["JM"] * s["KM"])
    assert nodes == NODES_C1
    sa = args.ilos >= 100

    print("workload: plain FP32", file=sys.stderr, flush=True)
    value, ms, g = timed_form(st, False, args.warmup, args.steps, args.tiles)
    # kernel durations from dispatch-timestamped HIP events on the library
    # stream (hipExtLaunchKernel), same workload, launched one by one
    g.profile(True)
    g.step(args.steps)
    rep = g.profile_report()
    g.profile(False)
    g.close()
    kern = dict(rep)
    dom = max(kern, key=lambda k: kern[k][1])
    calls, tot_ms = kern[dom]
    avg_s = tot_ms / calls * 1e-3
    bytes_launch = alg_bytes_per_launch(dom, nodes)
    achieved = bytes_launch / avg_s / 1e9
    gpu_ms_step = sum(ms_ for _, ms_ in kern.values()) / args.steps
    mixed = None
    if not args.no_mixed:
        print("workload: mixed form", file=sys.stderr, flush=True)
        mv, mms, gm = timed_form(st, True, args.warmup, args.steps, args.tiles)
        gm.close()
        b = B_ALG[("mixed", sa)]
        mixed = {"value": mv, "ms_per_step": mms, "dtype": "f32 arithmetic, bf16 SSS histories",
                 "vs_plain": mv / value,
                 "step_roofline": {"alg_bytes_per_node_step": b, "achieved_GBs": mv * b / 1e9,
                                   "frac": mv * b / 1e9 / HBM_PEAK_GBS}}
    print("checks against the reference's 100-step summary", file=sys.stderr, flush=True)
    check = None if args.no_check else {"plain": check_vs_reference(st, False),
                                        "mixed": None if args.no_mixed else check_vs_reference(st, True)}
    machines = None
    if not args.no_machines:
        print("other machines: multi-stage compressor, centrifugal stage", file=sys.stderr, flush=True)
        machines = [other_workload(MULTISTAGE_DECK, "multi-stage axial compressor: the reference's six-stage "
                                   "compressor deck, 12 rows", 10, 50),
                    other_workload(CENTRIFUGAL_DECK, "centrifugal stage: the reference's centrifugal compressor "
                                   "deck (impeller + vaned diffuser, mixing plane) on 97 x 81, Spalart-Allmaras",
                                   10, 50)]
    ttc = None
    if not args.no_converged:
        print("time to converged: Rotor 37", file=sys.stderr, flush=True)
        ttc = [time_to_converged(False), time_to_converged(True)]
    sl = None if args.no_speedline else speedline(sl_states)
    cpu = cpu_sl = None
    if not args.no_cpu_baseline:
        print("CPU baseline: reference on the workload", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.ilos)
        if sl is not None:
            print("CPU baseline: reference on the speedline", file=sys.stderr, flush=True)
            cpu_sl = cpu_baseline_speedline([int(x["scalars"]["IM"] * x["scalars"]["JM"] * x["scalars"]["KM"])
                                             for x in sl_states])
            if cpu_sl is not None:
                sl["cpu_baseline"] = cpu_sl

    b = B_ALG[("plain", sa)]
    name = "1.5-stage axial turbine, 3 rows x 64x64x193 nodes" + (", Spalart-Allmaras" if sa else "")
    out = {
        "metric": "node-steps/s", "value": value, "unit": "node-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic deck (seeded geometry), initial guess from the reference set-up",
        "config": {"workload": name + ", plain FP32 form", "nodes": nodes, "viscous_model": f"ILOS={args.ilos}",
                   "scheme": "scree, multigrid 3/9"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_bytes(dom),
                     "alg_bytes_per_launch": bytes_launch, "avg_launch_us": avg_s * 1e6},
        "step_roofline": {"alg_bytes_per_node_step": b, "achieved_GBs": value * b / 1e9,
                          "frac": value * b / 1e9 / HBM_PEAK_GBS, "gpu_busy_ms_per_step": gpu_ms_step},
        "mixed": mixed,
        "check": check,
        "kernels_ms_per_step": {k: round(ms_ / args.steps, 4) for k, (c, ms_) in
                                sorted(kern.items(), key=lambda kv: -kv[1][1])},
        "cpu_baseline": cpu,
        "machines": machines,
        "time_to_converged": ttc,
        "speedline": sl,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
