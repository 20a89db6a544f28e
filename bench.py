#!/usr/bin/env python3
"""Throughput of the time-marching step on one MI355X.

Workload: the 1.5-stage axial turbine (stator-rotor-stator, 3 rows of
64 x 64 x 193 nodes, 48/60/48 blades, rotor at 12000 rpm with a 0.5 mm tip
gap over 4 K-cells, p0 = 2.0e5 Pa, T0 = 1000 K, exit p = 1.1e5 Pa,
Spalart-Allmaras turbulence), built from a synthetic NEW-format deck by the
Fortran host set-up and marched by libmultall_hip in FP32.  A "step" is one time step of the solver (the body
of the reference's LOOP) over all 2,371,584 nodes.

Prints one JSON line: node-steps/s (whole job), the roofline of the
dominant kernel (dispatch-timestamped event durations on the library's
stream, HBM traffic from the committed counter runs), the step-level fraction
of the 8 TB/s HBM roofline at the 132 B/node-step algorithmic traffic of a
fused FP32 step, the reference solver timed on this host's CPU
(single-threaded, as built by oracle/build_ref.sh), the time to converge
Rotor 37 (the reference's own test deck as it ships) by the reference's stop
test, and the batched speedline: 16 operating points of a single-stage
compressor with uneven grids (11.5 M nodes in all) stepped concurrently
(mh_step_batch) against the same points stepped one after another.
"""
from __future__ import annotations

import argparse
import json
import os
import resource
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "multall-open-18.3_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0
B_ALG_STEP = 132.0          # bytes per node-step of a fused FP32 step (U, history, geometry, forces)

# Algorithmic bytes per node for one launch of each kernel: every array the
# kernel reads or writes, counted once (stencil neighbours are re-use, not
# traffic).  Cell arrays are node-shaped, so per-node == per-cell.
KERNEL_BYTES_PER_NODE = {
    "k_prim": 4 * (5 + 7), "k_massflux": 4 * (10 + 9), "k_flux_energy": 4 * (6 + 4),
    "k_flux_axial": 4 * (9 + 4), "k_flux_angular": 4 * (6 + 4), "k_flux_radial": 4 * (12 + 4),
    "k_resid_line": 4 * (5 + 2), "k_mg_part": 4, "k_mg_apply_lines": 4 * (3 + 1), "k_mg_apply": 4 * (3 + 1),
    "k_damp_scatter1": 4 * (2 + 1), "k_damp_scatter4": 4 * (2 + 1), "k_smooth_plane": 4 * 2, "k_peff": 4 * 2,
    "k_smooth_reg1": 4 * 2, "k_smooth_reg2": 4 * 2, "k_smooth_reg4": 4 * 2, "k_smooth_reg8": 4 * 2,
    "k_smooth_reg16": 4 * 2,
    "k_mg_sum_seq": 4, "k_mg_sum_wave": 4, "k_jplane_sums": 4, "k_rolim_check": 4, "k_mach_apply": 4 * 4,
    "k_stepup_temp": 4 * 6, "k_stepup_step": 4 * 5, "k_smooth_resid": 4 * 2, "k_sa_addstore": 4 * 3,
}

# HBM traffic per launch of the TSTEP kernels from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes over this workload, each scaled by what the same counter
# reads for a copy of known size (see --calibrate-copy); written by hand from
# those runs into this file, read here when present.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "tstep_kernels_config1_sa_pmc_traffic.json")


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


CONVERGED_DECK = os.path.join(ROOT, "tests", "golden", "decks", "converged_r37samods.dat")


def time_to_converged() -> dict:
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
    # the wall distances and mixing lengths on the device (SET_XLENGTH)
    g.set_xlength(host.array("THETA", (sc["KM"], sc["JM"], sc["IM"])),
                  *[host.array(n, (sc["NROWS"],)) for n in ("XLLIM_IN", "XLLIM_LE", "XLLIM_TE", "XLLIM_DN")])
    g.loss_init()
    g.start()
    g.sync()
    t1 = time.perf_counter()
    r = g.step(int(st["scalars"]["NSTEPS_MAX"]) + 10)
    g.sync()
    t2 = time.perf_counter()
    nodes = int(st["scalars"]["IM"] * st["scalars"]["JM"] * st["scalars"]["KM"])
    g.close()
    return {"workload": f"Rotor 37, reference test deck r37+samods as shipped (SA, Shih wall functions), {nodes} nodes",
            "converged": bool(r["ifend"] == 1 and r["nstep"] < int(st["scalars"]["NSTEPS_MAX"])),
            "steps": int(r["nstep"]), "seconds": t2 - t1, "seconds_with_setup": t2 - t0,
            "node_steps_per_s": nodes * r["nstep"] / (t2 - t1),
            "final": {"eavg": r["eavg"], "emax": r["emax"], "econt": r["econt"]}}


def _speedline_state(i: int):
    """Deck read and host set-up of speedline point i (in a worker process:
    the Fortran host holds one deck at a time)."""
    from multall_hip import deckgen, host
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, f"point{i}.dat")
        with open(path, "w") as f:
            f.write(deckgen.write_deck(deckgen.deck_speedline_point(i)))
        return host.load_deck(path)


def speedline_states(points: int = 16) -> list:
    """The speedline's set-up states, read in worker processes before this
    process touches the GPU."""
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(min(8, points)) as ex:
        return list(ex.map(_speedline_state, range(points)))


def speedline(states: list, steps: int = 100, warmup: int = 5) -> dict:
    """The batched speedline: all points stepped concurrently by
    mh_step_batch (one host thread and HIP stream per point) against the same
    points, freshly set up, stepped one after another.  Both runs start from
    the reference's initial guess after `warmup` steps; their results are
    bit-identical."""
    from multall_hip.runtime import Solver, step_batch
    points = len(states)
    nodes = [int(s["scalars"]["IM"] * s["scalars"]["JM"] * s["scalars"]["KM"]) for s in states]

    def fresh():
        sol = [Solver(s) for s in states]
        for g in sol:
            g.loss_init()
            g.start()
        return sol

    seq = fresh()
    for g in seq:
        g.step(warmup)
        g.sync()
    t0 = time.perf_counter()
    r_seq = [g.step(steps) for g in seq]
    for g in seq:
        g.sync()
    t_seq = time.perf_counter() - t0
    for g in seq:
        g.close()
    bat = fresh()
    step_batch(bat, warmup)
    for g in bat:
        g.sync()
    t0 = time.perf_counter()
    r_bat = step_batch(bat, steps)
    for g in bat:
        g.sync()
    t_bat = time.perf_counter() - t0
    for g in bat:
        g.close()
    total = sum(nodes)
    return {"workload": f"{points}-point speedline, single-stage axial compressor, exit p 1.2-1.6 x inlet p0, "
                        f"mixing length, {total} nodes ({min(nodes)}-{max(nodes)} per point)",
            "points": points, "steps": steps, "node_steps_per_s_batched": total * steps / t_bat,
            "node_steps_per_s_sequential": total * steps / t_seq, "speedup": t_seq / t_bat,
            "identical": r_seq == r_bat}


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


def workload_deck(nsteps: int, ilos: int):
    from multall_hip import deckgen
    spec = deckgen.deck_config1(nsteps=nsteps, ilos=ilos)
    return spec, deckgen.write_deck(spec)


def cpu_baseline(deck_text_for, nodes: int, steps_a: int = 2, steps_b: int = 6):
    """Time the reference solver (oracle/_ref, FP32, one core) on the same
    deck: two runs of steps_a and steps_b steps, per-step time from the
    difference of the loop times (cancels set-up and end-of-run output)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "multall_ref_fp32")
    if not os.path.exists(exe):
        return None
    times = []
    for n in (steps_a, steps_b):
        with tempfile.TemporaryDirectory(prefix="mhcpu_") as wd:
            with open(os.path.join(wd, "intype"), "w") as f:
                f.write("N\n")
            env = dict(os.environ, MH_NODUMP="1", OMP_NUM_THREADS="1")

            def unlimit():
                resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
            p = subprocess.run(["taskset", "-c", "0", exe], input=deck_text_for(n), cwd=wd, env=env,
                               capture_output=True, text=True, preexec_fn=unlimit, timeout=600)
            if p.returncode != 0:
                return None
            import numpy as np
            raw = open(os.path.join(wd, "final.bin"), "rb").read(24)
            times.append(float(np.frombuffer(raw, np.float64, 1, 16)[0]))
    per_step = (times[1] - times[0]) / (steps_b - steps_a)
    return {"value": nodes / per_step, "unit": "node-steps/s", "cores": 1, "kind": "reference",
            "sample": f"reference FP32 -O3 (amdflang), same deck, loop time of {steps_b} minus "
                      f"{steps_a} steps over {steps_b - steps_a} steps; {per_step:.3f} s/step"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--ilos", type=int, default=200,
                    help="viscous model of the deck: 200 Spalart-Allmaras (the workload's), 10 mixing length, 0 inviscid")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-converged", action="store_true", help="skip the time-to-converged run")
    ap.add_argument("--no-speedline", action="store_true", help="skip the batched speedline run")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 256 MiB device copy (for scaling the PMC byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print("bench.py measures one MI355X: multi-GPU runs are not part of this build "
              "(rows and operating points would shard across GPUs)", file=sys.stderr)
        sys.exit(2)

    sl_states = None if args.no_speedline else speedline_states()
    total = args.warmup + args.steps + 1000
    spec, deck = workload_deck(total, args.ilos)
    nodes = sum(r.jm for r in spec.rows) * spec.im * spec.km

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(lambda n: workload_deck(n, args.ilos)[1], nodes)

    from multall_hip import host
    from multall_hip.runtime import Solver
    with tempfile.NamedTemporaryFile("w", suffix=".dat", delete=False) as f:
        f.write(deck)
        deck_path = f.name
    st = host.load_deck(deck_path)
    os.unlink(deck_path)
    assert st["scalars"]["IM"] * st["scalars"]["JM"] * st["scalars"]["KM"] == nodes
    g = Solver(st, exact_order=False)
    g.loss_init()          # SETUP's initial viscous forces (no-op when inviscid)
    g.start()
    g.step(args.warmup)
    g.sync()
    t0 = time.perf_counter()
    r = g.step(args.steps)
    g.sync()
    t1 = time.perf_counter()
    assert r["nstep"] == args.warmup + args.steps, r
    dt = t1 - t0
    value = nodes * args.steps / dt

    # kernel durations from dispatch-timestamped HIP events on the library
    # stream (hipExtLaunchKernel), same workload
    g.profile(True)
    g.step(args.steps)
    rep = g.profile_report()
    g.profile(False)
    kern = {k: (c, ms) for k, (c, ms) in rep.items()}
    dom = max(kern, key=lambda k: kern[k][1])
    calls, tot_ms = kern[dom]
    avg_s = tot_ms / calls * 1e-3
    bytes_launch = KERNEL_BYTES_PER_NODE.get(dom, 0) * nodes
    achieved = bytes_launch / avg_s / 1e9
    gpu_ms_step = sum(ms for _, ms in kern.values()) / args.steps
    g.close()   # the other measurements run with only their own contexts on the GPU
    ttc = None if args.no_converged else time_to_converged()
    sl = None if args.no_speedline else speedline(sl_states)

    out = {
        "metric": "node-steps/s", "value": value, "unit": "node-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic deck (seeded geometry), initial guess from the reference set-up",
        "config": {"workload": "1.5-stage axial turbine, 3 rows x 64x64x193 nodes, Spalart-Allmaras, FP32 plain form"
                   if args.ilos >= 200 else "1.5-stage axial turbine, 3 rows x 64x64x193 nodes, FP32 plain form",
                   "nodes": nodes, "viscous_model": f"ILOS={args.ilos}", "scheme": "scree, multigrid 3/9"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_bytes(dom),
                     "alg_bytes_per_launch": bytes_launch, "avg_launch_us": avg_s * 1e6},
        "step_roofline": {"alg_bytes_per_node_step": B_ALG_STEP,
                          "achieved_GBs": value * B_ALG_STEP / 1e9,
                          "frac": value * B_ALG_STEP / 1e9 / HBM_PEAK_GBS,
                          "gpu_busy_ms_per_step": gpu_ms_step},
        "kernels_ms_per_step": {k: round(ms / args.steps, 4) for k, (c, ms) in
                                sorted(kern.items(), key=lambda kv: -kv[1][1])},
        "cpu_baseline": cpu,
        "time_to_converged": ttc,
        "speedline": sl,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
