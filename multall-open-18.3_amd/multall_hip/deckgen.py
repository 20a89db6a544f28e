"""Synthetic-cascade deck generator in Multall's NEW input format.

The writer emits the label-line / value-line pairs in the order the
reference's list-directed reader consumes them (``NEW_READIN``,
multall-open-18.3.f:148-2158; row block at :593-1550, section block at
:832-1000, inlet/exit block at :1838-1940, mixing-length limits at
:2100-2125).  Only the options this build supports are emitted: no cooling,
no bleed, no shroud, ``ISHIFT=0`` (the generator lays the inter-row grid
out itself, with the two J-lines of a mixing plane a fraction of a
spacing apart), no
throughflow or Q3D.

Geometry: circular-arc camber line between the inlet and exit metal
angles, NACA-4-digit thickness distribution (closed trailing edge form),
straight camber extensions upstream and downstream, constant hub and tip
radii, and a seeded uniform jitter of +/-``jitter`` (relative) applied to
the streamwise grid spacing and to the tangential camber coordinate.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


@dataclass
class RowSpec:
    name: str = "STATOR"
    nblades: int = 60
    jm: int = 97
    jle: int = 26
    jte: int = 72
    chord_x: float = 0.04          # axial chord, m
    alpha_in: float = 0.0          # camber angle at the leading edge, degrees
    alpha_out: float = 70.0        # camber angle at the trailing edge, degrees
    tmax: float = 0.20             # maximum thickness / chord
    rpm: float = 0.0
    up_len: float = 1.0            # upstream length in axial chords
    dn_len: float = 1.0            # downstream length in axial chords
    ktips: int = 0                 # tip-gap start K (0 = no gap)
    ktipe: int = 0
    fractip: float = 0.0           # tip gap as a fraction of span
    jtran: tuple = (0, 0, 0, 0)    # transition J from the row start: I=1, I=IM, hub, casing surfaces
    pguess: tuple = (1.0e5, 1.0e5, 0.8e5, 0.8e5)   # PUP, PLE, PTE, PDN
    cusp: Optional[tuple] = None   # (ICUSP, LCUSP, LCUSPUP): IF_CUSP = 1, a new trailing-edge cusp
    # (NUP_I1, NUP_IM, N_WAKE, SEP_THIK, SEP_DRAG): IF_CUSP = 2, the trailing-edge separation force
    sep_force: Optional[tuple] = None
    # (span fractions, upstream, downstream-start, downstream-end grid angles in degrees): IF_ANGLES = 1
    angles: Optional[tuple] = None


@dataclass
class DeckSpec:
    title: str = "SYNTHETIC CASCADE"
    im: int = 37
    km: int = 33
    rows: List[RowSpec] = field(default_factory=lambda: [RowSpec()])
    r_hub: float = 0.30
    r_tip: float = 0.35
    cp: float = 1004.5
    ga: float = 1.4
    itimst: int = 3
    cfl: float = 0.4
    damp: float = 10.0
    machlim: float = 2.0
    f_pdown: float = 0.0
    nsteps: int = 2000
    conlim: float = 0.005
    sfx: float = 0.005
    sft: float = 0.005
    fac_4th: float = 0.8
    nchange: int = 500
    mg: tuple = (3, 3, 3, 9, 9, 9)
    fblk: tuple = (0.4, 0.2, 0.1)
    ifmix: int = 1
    rfmix: float = 0.025
    fextrap: float = 0.8
    fsmthb: float = 1.0
    fangle: float = 0.8
    in_press: int = 0
    in_vtan: int = 0
    in_vr: int = 1
    rfin: float = 0.1
    ishift: int = 0                # 2: the reader re-grids the gaps between rows (GRID_UP/GRID_DOWN)
    nextrap: tuple = (10, 10)      # NEXTRAP_LE, NEXTRAP_TE
    in_flow: int = 0               # mass-flow forcing: 3 towards FLOWIN, 2 towards the mean over J
    repeat: Optional[tuple] = None  # (NINMOD, RFINBC): IF_REPEAT = 1, the repeating-stage inlet condition
    flowin: float = 0.0            # kg/s through the whole annulus
    rflow: float = 0.0
    ipout: int = 0
    sfexit: float = 0.0
    nsfexit: int = 0
    ilos: int = 10
    nlos: int = 5
    reyno: float = 5.0e5
    rf_vis: float = 0.5
    ftrans: float = 0.0001
    turbvis_lim: float = 3000.0
    prandtl: float = 1.0
    yplam: float = 5.0
    ypturb: float = 25.0
    facmixup: float = 2.0
    ibound: int = 0
    ypluswall: float = 0.0
    if_rough: int = 0
    rough: tuple = (0.0, 0.0, 0.0, 0.0)   # hub, casing, I=1 and I=IM surfaces, microns
    nmixup: int = 1000
    sa_factors: tuple = (0.0, 1.0, 1.0, 1.0, 1.0, 2.0, 0.0, 0.0)
    po1: float = 1.0e5
    to1: float = 300.0
    vm1: float = 100.0
    vtin: float = 0.0              # inlet tangential velocity (read with IN_VTAN = 1), m/s
    yaw: float = 0.0               # inlet yaw angle BS, degrees
    pitch: float = 0.0             # inlet meridional pitch angle BR, degrees
    pd_exit: Optional[tuple] = None   # IPOUT = 3: exit static pressure at the KIN inlet-data points
    nout: Optional[tuple] = None   # the 5 printout steps (default: past the run)
    pdown_hub: float = 0.8e5
    pdown_tip: float = 0.8e5
    fp_ratio: float = 1.0          # geometric expansion of FP towards mid-pitch
    fr_ratio: float = 1.0
    seed: int = 0
    jitter: float = 0.002
    nsecs: int = 2


def _geometric(n: int, ratio: float) -> np.ndarray:
    """Symmetric spacing weights growing by ``ratio`` towards the middle."""
    w = np.ones(n)
    for i in range(1, n):
        w[i] = w[i - 1] * ratio
    return np.minimum(w, w[::-1])


def _camber(xf: np.ndarray, a1: float, a2: float, cx: float):
    """Circular-arc camber line: rtheta offset and slope angle at axial
    fraction ``xf`` in [0, 1] of an arc turning from ``a1`` to ``a2`` (rad)."""
    if abs(a2 - a1) < 1e-9:
        return xf * cx * math.tan(a1), np.full_like(xf, a1)
    rad = cx / (math.sin(a2) - math.sin(a1))
    s = np.clip(xf * cx / rad + math.sin(a1), -1.0, 1.0)
    phi = np.arcsin(s)
    y = -rad * (np.cos(phi) - math.cos(a1))
    return y, phi


def _naca_thickness(xf: np.ndarray, tmax: float) -> np.ndarray:
    x = np.clip(xf, 0.0, 1.0)
    return 5.0 * tmax * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2
                         + 0.2843 * x ** 3 - 0.1036 * x ** 4)


def row_sections(spec: DeckSpec, row: RowSpec, x_start: float, rng):
    """Return (x, rt_upp, rt_thick) along J for one row, and the row's end x."""
    cx = row.chord_x
    nup = row.jle - 1                       # intervals upstream of the LE
    nbl = row.jte - row.jle                 # intervals on the blade
    ndn = row.jm - row.jte                  # intervals downstream of the TE
    x_le = x_start + row.up_len * cx
    x_te = x_le + cx
    x_end = x_te + row.dn_len * cx
    # upstream: spacing shrinking towards the LE; blade: cosine clustering.
    fu = _geometric(2 * nup, 1.08)[:nup][::-1] if nup > 0 else np.zeros(0)
    fu = fu[::-1]
    xu = x_start + np.concatenate([[0.0], np.cumsum(fu[::-1])]) / max(fu.sum(), 1e-30) * (x_le - x_start)
    tb = 0.5 * (1.0 - np.cos(np.linspace(0.0, math.pi, nbl + 1)))
    xb = x_le + tb * cx
    fd = _geometric(2 * ndn, 1.08)[:ndn] if ndn > 0 else np.zeros(0)
    xd = x_te + np.concatenate([[0.0], np.cumsum(fd)]) / max(fd.sum(), 1e-30) * (x_end - x_te)
    x = np.concatenate([xu[:-1], xb, xd[1:]])
    assert x.size == row.jm, (x.size, row.jm)
    # relative jitter of the interior streamwise spacing
    dx = np.diff(x)
    jit = 1.0 + spec.jitter * rng.uniform(-1.0, 1.0, size=dx.size)
    jit[row.jle - 2:row.jle] = 1.0        # keep the LE/TE points exactly on the blade
    jit[row.jte - 2:row.jte] = 1.0
    x = x[0] + np.concatenate([[0.0], np.cumsum(dx * jit)])
    x *= (x_end - x_start) / (x[-1] - x[0])
    x += x_start - x[0]
    a1 = math.radians(row.alpha_in)
    a2 = math.radians(row.alpha_out)
    xf = (x - x[row.jle - 1]) / (x[row.jte - 1] - x[row.jle - 1])
    yb, phi = _camber(np.clip(xf, 0.0, 1.0), a1, a2, cx)
    y_te = yb[row.jte - 1]
    y = np.where(xf < 0.0, (x - x[row.jle - 1]) * math.tan(a1),
                 np.where(xf > 1.0, y_te + (x - x[row.jte - 1]) * math.tan(a2), yb))
    y = y + spec.jitter * cx * 0.01 * rng.uniform(-1.0, 1.0, size=y.size) * (xf > 0.0) * (xf < 1.0)
    t = _naca_thickness(xf, row.tmax) * cx
    t = np.where((xf <= 0.0) | (xf >= 1.0), 0.0, t / np.cos(phi))
    rt_upp = y + 0.5 * t
    return x, rt_upp, t, x_end


def _fmt(vals, per_line=8, fmt="{:12.6f}"):
    out = []
    vals = list(vals)
    for i in range(0, len(vals), per_line):
        out.append("".join(fmt.format(v) for v in vals[i:i + per_line]))
    return "\n".join(out)


def _ints(vals):
    return " ".join(str(int(v)) for v in vals)


def write_deck(spec: DeckSpec) -> str:
    """Render ``spec`` as NEW-format deck text."""
    rng = np.random.default_rng(spec.seed)
    L: List[str] = []
    a = L.append
    a(spec.title[:72])
    a("    CP   and   GAMMA")
    a(f"{spec.cp:12.4f}{spec.ga:12.4f}")
    a("       ITIMST")
    a(f"{spec.itimst:10d}")
    a("     CFL,    DAMP,    MACHLIM,    F_PDOWN")
    a(f"{spec.cfl:12.6f}{spec.damp:12.6f}{spec.machlim:12.6f}{spec.f_pdown:12.6f}")
    a("  IF_RESTART")
    a("         0")
    a("  NSTEPS_MAX, CONLIM")
    a(f"{spec.nsteps:10d}  {spec.conlim:.9g}")
    a("   SFX,      SFT,      FAC_4TH,     NCHANGE")
    a(f"{spec.sfx:12.6f}{spec.sft:12.6f}{spec.fac_4th:12.6f}{spec.nchange:10d}")
    a("       NUMBER OF BLADE ROWS")
    a(f"{len(spec.rows):10d}")
    a("        IM        KM")
    a(f"{spec.im:10d}{spec.km:10d}")
    a("  FP(I),I=1,IMM1")
    a(_fmt(_geometric(spec.im - 1, spec.fp_ratio)))
    a("  FR(K),K=1,KMM1")
    a(_fmt(_geometric(spec.km - 1, spec.fr_ratio)))
    a("        IR        JR        KR        IRBB      JRBB      KRBB")
    a("".join(f"{v:10d}" for v in spec.mg))
    a("   FBLK1,     FBLK2,     FBLK3")
    a("".join(f"{v:12.6f}" for v in spec.fblk))
    a("       IFMIX")
    a(f"{spec.ifmix:10d}")
    if spec.ifmix != 0:
        a("   RFMIX,    FEXTRAP,   FSMTHB,    FANGLE")
        a(f"{spec.rfmix:12.6f}{spec.fextrap:12.6f}{spec.fsmthb:12.6f}{spec.fangle:12.6f}")
    a("      IFCOOL    IFBLEED    IF_ROUGH")
    a(f"         0         0{spec.if_rough:10d}")
    a("       NSECS_IN")
    a(f"{spec.nsecs:10d}")
    a("       IN_PRESS  IN_VTAN   IN_VR    IN_FLOW  IF_REPEAT  RFIN")
    a(f"{spec.in_press:10d}{spec.in_vtan:10d}{spec.in_vr:10d}{spec.in_flow:10d}"
      f"{0 if spec.repeat is None else 1:10d}{spec.rfin:12.5f}")
    a("  IPOUT SFEXIT  NSFEXIT")
    a(f"{spec.ipout:5d}{spec.sfexit:12.6f}{spec.nsfexit:5d}")
    a("  PLATE_LOSS THROTTLE_EXIT")
    a("  0.000000  0.000000")
    if spec.in_flow != 0:
        a("  FLOWIN  RFLOW")
        a(f"{spec.flowin:12.5f}{spec.rflow:10.4f}")
    if spec.repeat is not None:
        a("  NINMOD  RFINBC")
        a(f"{spec.repeat[0]:10d}{spec.repeat[1]:10.4f}")
    a("        ILOS      NLOS      IBOUND")
    a(f"{spec.ilos:10d}{spec.nlos:10d}{spec.ibound:10d}")
    if spec.ilos != 0:
        a("   REYNO,     RF_VIS,   FTRANS, TURBVIS_LIM, PRANDTL, YPLUSWALL")
        a(f"{spec.reyno:12.1f}{spec.rf_vis:10.3f}{spec.ftrans:10.5f}{spec.turbvis_lim:10.3f}"
          f"{spec.prandtl:10.3f}{spec.ypluswall:10.3f}")
    if spec.ilos >= 200:
        a("  FAC_STMIX, FAC_ST0, FAC_ST1, FAC_ST2, FAC_ST3, FAC_SFVIS, FAC_VORT, FAC_PGRAD")
        a("".join(f"{v:10.4f}" for v in spec.sa_factors))
    if spec.ilos != 0:
        a("   YPLAM      YPTURB")
        a(f"{spec.yplam:12.6f}{spec.ypturb:12.6f}")
    a("      ISHIFT    NEXTRAP_LE  NEXTRAP_TE")
    a(f"{spec.ishift:10d}{spec.nextrap[0]:10d}{spec.nextrap[1]:10d}")
    a("  NSTG(N),N=1,NROWS")
    a("".join(f"{1 + n // 2:10d}" for n in range(len(spec.rows))))
    a("  5  TIME STEPS FOR PRINTOUT")
    big = spec.nsteps + 10
    a("".join(f"{n:10d}" for n in (spec.nout or (big,) * 5)))
    a("  MARKER FOR VARIABLES TO BE SENT TO THE OUTPUT FILE.")
    a(" " + _ints([0] * 13))
    a("  STREAM SURFACES ON WHICH RESULTS ARE TO BE SENT TO   THE OUTPUT FILE")
    a(" " + _ints([0] * spec.km))

    x_start = 0.0
    span = spec.r_tip - spec.r_hub
    for nr, row in enumerate(spec.rows):
        x, rtu, rtt, x_end = row_sections(spec, row, x_start, rng)
        # the next row starts a small fraction of a spacing downstream, so the
        # two J-lines of the mixing plane are nearly but not exactly coincident
        x_start = x_end + 0.02 * (x[-1] - x[-2])
        a(" ******************************************************************************")
        a(" ***********STARTING THE INPUT FOR EACH BLADE ROW******************************")
        a(f"{row.name:<72s}")
        a("    NUMBER OF BLADES IN ROW")
        a(f"{row.nblades:10d}")
        a("        JM        JLE       JTE")
        a(f"{row.jm:10d}{row.jle:10d}{row.jte:10d}")
        a("      KTIPSTART  KTIPEND")
        a(f"{row.ktips:10d}{row.ktipe:10d}")
        if row.ktips > 0:
            a("  TIP GAPS")
            a(f"{row.fractip:12.6f}{row.fractip:12.6f}")
            a("  BLADE THINNING")
            a(_fmt([1.0] * spec.km, 16, "{:5.1f}"))
        a("       BOUNDARY LAYER TRANSITION POINTS")
        a("".join(f"{j:10d}" for j in row.jtran))
        a("  SET NEWGRID = 1 TO GENERATE A NEW GRID")
        a("         0")
        a("   RPMROW,    RPMHUB")
        hub_rpm = row.rpm
        a(f"{row.rpm:10.1f}{hub_rpm:10.1f}")
        a("       JROTHS    JROTHE    JROTTS    JROTTE")
        a("         0         0         0         0")
        a("   PUPROW    PLEROW   PTEROW    PDROW")
        a("".join(f"{v:10.1f}" for v in row.pguess))
        a("      NSECS_ROW   INSURF")
        a(f"{spec.nsecs:10d}         0")
        a("  IF_CUSP   IFANGLES")
        ifcusp = 2 if row.sep_force is not None else (0 if row.cusp is None else 1)
        a(f"{ifcusp:10d}{0 if row.angles is None else 1:10d}")
        if row.cusp is not None and row.sep_force is None:
            a("".join(f"{v:10d}" for v in row.cusp))
        if row.sep_force is not None:
            n1, nm, nw, thik, drag = row.sep_force
            a(f"{n1:10d}{nm:10d}{nw:10d}{thik:10.4f}{drag:10.4f}")
        for ks in range(spec.nsecs):
            fr = ks / (spec.nsecs - 1)
            r = spec.r_hub + fr * span
            a(" ******************************************************************************")
            a(f"  ROW NUMBER {nr + 1:11d}  SECTION NUMBER {ks + 1:11d}")
            a("           0           0           0  IF_DESIGN etc")
            a("   1.00000   0.00000")
            a(_fmt(x))
            a("   1.00000   0.00000")
            a(_fmt(rtu))
            a("   1.00000")
            a(_fmt(rtt))
            a("   1.00000   0.00000")
            a(_fmt(np.full(row.jm, r)))
        if row.angles is not None:
            a("  GRID ANGLES: N, SPAN FRACTIONS, UPSTREAM, DOWNSTREAM START AND END")
            a(f"{len(row.angles[0]):10d}")
            for vals in row.angles:
                a("".join(f"{v:10.4f}" for v in vals))
    kin = 2
    a("  STARTING INLET BOUNDARY CONDITION DATA .")
    a("  NUMBER OF POINTS FOR INLET BOUNDARY CONDITIONS")
    a(f"{kin:10d}")
    a("  SPACING OF INLET BOUNDARY CONDITION POINTS")
    a(_fmt([1.0] * (kin - 1)))
    a("   INLET STAGNATION PRESSURES")
    a(_fmt([spec.po1] * kin, 8, "{:12.1f}"))
    if spec.ipout == 3:
        a("   EXIT STATIC PRESSURES")
        a(_fmt(spec.pd_exit, 8, "{:12.1f}"))
    a("   INLET STAGNATION TEMPERATURES")
    a(_fmt([spec.to1] * kin, 8, "{:12.3f}"))
    a("   INLET TANGENTIAL VELOCITY")
    a(_fmt([spec.vtin] * kin, 8, "{:12.4f}"))
    a("   INLET MERIDIONAL VELOCITY")
    a(_fmt([spec.vm1] * kin, 8, "{:12.4f}"))
    a("   INLET YAW ANGLE")
    a(_fmt([spec.yaw] * kin, 8, "{:12.4f}"))
    a("   INLET PITCH ANGLE")
    a(_fmt([spec.pitch] * kin, 8, "{:12.4f}"))
    a("   PDOWN_HUB   PDOWN_TIP")
    a(f"{spec.pdown_hub:12.1f}{spec.pdown_tip:12.1f}")
    if spec.ilos != 0:
        a(" MIXING LENGTH LIMITS ON ALL BLADE ROWS")
        for _ in spec.rows:
            if spec.ilos == 10:
                a("  0.030000  0.030000  0.030000  0.030000  0.030000  0.020000")
            else:
                a("  0.020000  0.030000  0.040000  0.050000  1.000000  0.500000")
        a("  FACTOR TO INCREASE THE TURBULENT VISCOSITY OVER THE FIRST NMIXUP STEPS")
        a(f"{spec.facmixup:10.5f}{spec.nmixup:6d}")
    if spec.if_rough > 0:
        a("  SURFACE ROUGHNESS IN MICRONS, HUB, CASING, I=1 AND I=IM SURFACES, EACH ROW")
        for _ in spec.rows:
            a("".join(f"{v:10.3f}" for v in spec.rough))
    return "\n".join(L) + "\n"


# --------------------------------------------------------------------------
# Named workloads.  Sizes and operating points follow the task definitions.

def deck_tiny(ilos: int = 0, nrows: int = 1, nsteps: int = 100) -> DeckSpec:
    """A small cascade for exact-parity tests: IM=9, KM=9, 41 J-points/row."""
    rows = []
    for n in range(nrows):
        rotor = (n % 2 == 1)
        rows.append(RowSpec(name="ROTOR" if rotor else "STATOR", nblades=48 if rotor else 40,
                            jm=41, jle=11, jte=31,
                            alpha_in=-20.0 if rotor else 0.0, alpha_out=-60.0 if rotor else 65.0,
                            rpm=6000.0 if rotor else 0.0,
                            pguess=(1.0e5, 0.97e5, 0.85e5, 0.83e5) if n == 0 else (0.83e5, 0.82e5, 0.8e5, 0.8e5)))
    return DeckSpec(title=f"TINY CASCADE ILOS={ilos} ROWS={nrows}", im=9, km=9, rows=rows,
                    ilos=ilos, nsteps=nsteps, conlim=1e-9, nchange=50, nmixup=50,
                    pdown_hub=0.8e5, pdown_tip=0.8e5, seed=1)


def deck_tiny_gap(nsteps: int = 100) -> DeckSpec:
    """The three-row mixing-length cascade with the gaps re-gridded by the
    reader (ISHIFT = 2, NEXTRAP 6 and 8): a new centred cusp on the first
    stator, input grid angles varying over the span on the rotor, and a cusp
    that keeps the I = IM surface straight with input angles on the last
    stator."""
    spec = deck_tiny(ilos=10, nrows=3, nsteps=nsteps)
    spec.title = "TINY CASCADE REGRIDDED GAPS"
    spec.ishift = 2
    spec.nextrap = (6, 8)
    spec.rows[0].cusp = (0, 6, 2)
    spec.rows[1].angles = ((0.0, 0.5, 1.0), (-15.0, -18.0, -22.0), (-55.0, -58.0, -62.0), (-50.0, -52.0, -56.0))
    spec.rows[2].cusp = (-1, 4, 1)
    spec.rows[2].angles = ((0.0, 1.0), (2.0, 4.0), (60.0, 63.0), (55.0, 58.0))
    return spec


def deck_tiny_ml_variant(nsteps: int = 100) -> DeckSpec:
    """Three-row tiny cascade with the mixing-length model's other branches:
    power-law laminar viscosity (REYNO < 0), rough walls, and a rotor tip
    gap over the two outer K-cells."""
    spec = deck_tiny(ilos=10, nrows=3, nsteps=nsteps)
    spec.title = "TINY CASCADE ILOS=10 POWER-LAW VISCOSITY, ROUGH, TIP GAP"
    spec.reyno = -2.0
    spec.if_rough = 1
    spec.rough = (20.0, 30.0, 10.0, 15.0)
    r = spec.rows[1]
    r.ktips, r.ktipe, r.fractip = 7, 9, 0.02
    return spec


def deck_tiny_wall_variant(nsteps: int = 100) -> DeckSpec:
    """One-row tiny cascade with wall stresses from a fixed wall y+
    (YPLUSWALL > 5) and a shear-free hub (IBOUND = 1)."""
    spec = deck_tiny(ilos=10, nrows=1, nsteps=nsteps)
    spec.title = "TINY CASCADE ILOS=10 YPLUSWALL=12 IBOUND=1"
    spec.ypluswall = 12.0
    spec.ibound = 1
    return spec


def deck_tiny_sa(nsteps: int = 100) -> DeckSpec:
    """Three-row tiny cascade with the Spalart-Allmaras model (ILOS=200) and
    a rotor tip gap over the two outer K-cells."""
    spec = deck_tiny(ilos=200, nrows=3, nsteps=nsteps)
    spec.title = "TINY CASCADE ILOS=200 SPALART-ALLMARAS, TIP GAP"
    r = spec.rows[1]
    r.ktips, r.ktipe, r.fractip = 7, 9, 0.02
    return spec


def deck_tiny_sa_variant(nsteps: int = 100) -> DeckSpec:
    """The Spalart-Allmaras tiny cascade with power-law laminar viscosity,
    rough walls, a shear-free casing (IBOUND = 2) and the mixing-length
    relaxation source switched on (FAC_STMIX = 1)."""
    spec = deck_tiny_sa(nsteps=nsteps)
    spec.title = "TINY CASCADE ILOS=200 POWER-LAW VISCOSITY, ROUGH, IBOUND=2"
    spec.reyno = -2.0
    spec.if_rough = 1
    spec.rough = (20.0, 30.0, 10.0, 15.0)
    spec.ibound = 2
    spec.sa_factors = (1.0, 1.0, 1.0, 1.0, 1.0, 2.0, 0.0, 0.0)
    return spec


def deck_tiny_nl(nsteps: int = 100) -> DeckSpec:
    """Three-row tiny cascade with the vorticity mixing-length model of
    NEW_LOSS (ILOS=100) and a rotor tip gap over the two outer K-cells."""
    spec = deck_tiny(ilos=100, nrows=3, nsteps=nsteps)
    spec.title = "TINY CASCADE ILOS=100 NEW MIXING LENGTH, TIP GAP"
    r = spec.rows[1]
    r.ktips, r.ktipe, r.fractip = 7, 9, 0.02
    return spec


def deck_tiny_nl_variant(nsteps: int = 100) -> DeckSpec:
    """The ILOS=100 tiny cascade with power-law laminar viscosity, rough
    walls, a shear-free hub (IBOUND = 1), transition points part-way along
    every surface of each row and a transition threshold (FTRANS) that
    leaves part of the flow laminar downstream of them."""
    spec = deck_tiny_nl(nsteps=nsteps)
    spec.title = "TINY CASCADE ILOS=100 POWER-LAW VISCOSITY, ROUGH, IBOUND=1, TRANSITION"
    spec.reyno = -2.0
    spec.if_rough = 1
    spec.rough = (20.0, 30.0, 10.0, 15.0)
    spec.ibound = 1
    spec.ftrans = 4.0
    for n, r in enumerate(spec.rows):
        r.jtran = (14 + 2 * n, 18, 9, 24 - 2 * n)
    return spec


def deck_tiny_sss(nsteps: int = 100) -> DeckSpec:
    """The three-row mixing-length tiny cascade with the SSS scheme
    (ITIMST = -3: F2 = -1.65, F3 = -0.65 and one residual-smoothing pass,
    NRSMTH = 1, on every conserved variable)."""
    spec = deck_tiny(ilos=10, nrows=3, nsteps=nsteps)
    spec.title = "TINY CASCADE ILOS=10 SSS ITIMST=-3"
    spec.itimst = -3
    return spec


def deck_tiny_flow(in_flow: int = 2, nsteps: int = 100) -> DeckSpec:
    """The three-row mixing-length tiny cascade with SETFLO's mass-flow
    forcing: towards the mean flow over J (IN_FLOW = 2), or towards a fixed
    annulus flow of 12 kg/s, within the range the unforced cascade passes in
    its first 100 steps (IN_FLOW = 3)."""
    spec = deck_tiny(ilos=10, nrows=3, nsteps=nsteps)
    spec.title = f"TINY CASCADE ILOS=10 MASS-FLOW FORCING IN_FLOW={in_flow}"
    spec.in_flow = in_flow
    spec.flowin = 12.0
    spec.rflow = 0.05
    return spec


def deck_tiny_wallfun(ilos: int, nsteps: int = 100) -> DeckSpec:
    """Three-row tiny cascades with the Shih et al. wall functions: without
    the pressure-gradient term for the mixing-length model (ILOS = 10,
    YPLUSWALL = -6), with it (YPLUSWALL = -11) for NEW_LOSS (ILOS = 100)
    and for Spalart-Allmaras (ILOS = 200).  The Spalart-Allmaras deck also
    takes the Rotor 37 test deck's helicity and pressure-gradient source
    factors (FAC_VORT = FAC_PGRAD = 1) and transition points with a
    threshold that leaves part of the flow laminar."""
    spec = deck_tiny(ilos=ilos, nrows=3, nsteps=nsteps)
    spec.ypluswall = -6.0 if ilos == 10 else -11.0
    spec.title = f"TINY CASCADE ILOS={ilos} SHIH WALL FUNCTIONS YPLUSWALL={spec.ypluswall:g}"
    r = spec.rows[1]
    r.ktips, r.ktipe, r.fractip = 7, 9, 0.02
    if ilos >= 200:
        spec.sa_factors = (0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)
        spec.ftrans = 4.0
        for n, row in enumerate(spec.rows):
            row.jtran = (14 + 2 * n, 18, 9, 24 - 2 * n)
    return spec


def deck_tiny_exit(nsteps: int = 100) -> DeckSpec:
    """The three-row mixing-length tiny cascade with exit smoothing on the
    last three J planes (SFEXIT = 0.1, NSFEXIT = 3)."""
    spec = deck_tiny(ilos=10, nrows=3, nsteps=nsteps)
    spec.title = "TINY CASCADE ILOS=10 EXIT SMOOTHING"
    spec.sfexit = 0.1
    spec.nsfexit = 3
    return spec


def deck_tiny_wide(nsteps: int = 100) -> DeckSpec:
    """A three-row mixing-length cascade with IM=72 (more than 64 pitchwise
    cells per line) and short rows of 21 J-points, with exit smoothing on the
    last three J planes (SFEXIT = 0.1, NSFEXIT = 3)."""
    spec = deck_tiny(ilos=10, nrows=3, nsteps=nsteps)
    spec.title = "TINY CASCADE ILOS=10 IM=72 EXIT SMOOTHING"
    spec.im = 72
    for r in spec.rows:
        r.jm, r.jle, r.jte = 21, 6, 16
    spec.sfexit = 0.1
    spec.nsfexit = 3
    return spec


def deck_config0(nsteps: int = 2000) -> DeckSpec:
    """Single axial turbine stator row, 37x33x97, 60 blades, hub/tip 0.30/0.35 m,
    camber 0->70 deg, 20% NACA thickness, p0=1e5 Pa, T0=300 K, p_exit=0.8e5 Pa,
    mixing-length model, scree, multigrid 3^3/9^3, FP32."""
    row = RowSpec(name="STATOR", nblades=60, jm=97, jle=26, jte=72, chord_x=0.04,
                  alpha_in=0.0, alpha_out=70.0, tmax=0.20,
                  pguess=(1.0e5, 0.97e5, 0.82e5, 0.80e5))
    return DeckSpec(title="CONFIG0 TURBINE STATOR 37X33X97", im=37, km=33, rows=[row],
                    ilos=10, nsteps=nsteps, conlim=0.005, nchange=500, seed=0,
                    po1=1.0e5, to1=300.0, vm1=100.0, pdown_hub=0.8e5, pdown_tip=0.8e5)


def deck_config1(nsteps: int = 5000, ilos: int = 200) -> DeckSpec:
    """1.5-stage axial turbine, 3 rows of 64x64x193, blades 48/60/48, rotor at
    12000 rpm with a 0.5 mm tip gap over 4 K-cells, p0=2e5 Pa, T0=1000 K,
    p_exit=1.1e5 Pa, Spalart-Allmaras (ILOS=200)."""
    span = 0.05
    s1 = RowSpec(name="STATOR 1", nblades=48, jm=193, jle=49, jte=145, alpha_in=0.0, alpha_out=70.0,
                 pguess=(2.0e5, 1.95e5, 1.55e5, 1.5e5))
    r1 = RowSpec(name="ROTOR 1", nblades=60, jm=193, jle=49, jte=145, alpha_in=-20.0, alpha_out=-65.0,
                 rpm=12000.0, ktips=60, ktipe=64, fractip=0.0005 / span,
                 pguess=(1.5e5, 1.48e5, 1.25e5, 1.22e5))
    s2 = RowSpec(name="STATOR 2", nblades=48, jm=193, jle=49, jte=145, alpha_in=0.0, alpha_out=70.0,
                 pguess=(1.22e5, 1.2e5, 1.12e5, 1.1e5))
    return DeckSpec(title="CONFIG1 1.5-STAGE TURBINE 3X64X64X193", im=64, km=64, rows=[s1, r1, s2],
                    ilos=ilos, nsteps=nsteps, conlim=0.005, nchange=1000, seed=1,
                    po1=2.0e5, to1=1000.0, vm1=150.0, pdown_hub=1.1e5, pdown_tip=1.1e5)


def deck_tiny_bc(variant: str, nsteps: int = 100) -> DeckSpec:
    """Tiny mixing-length cascades for the inlet and exit condition options
    (LOOP, multall-open-18.3.f:4506-4757):

    - ``bc1``: the reference's Rotor 37 test deck's choice: exit pressure
      held at its set-up values (IPOUT = 1), inlet pressure extrapolated
      from the pitch-averaged pressures of J = 2 and 3 (IN_PRESS = -3) and
      the inlet radial velocity taken from J = 2 (IN_VR = -3), 10 degrees of
      inlet yaw;
    - ``bc2``: radial equilibrium shifted to the casing pressure
      (IPOUT = -1, with hub and casing pressures apart), the inlet pressure
      extrapolated node by node (IN_PRESS = 1), a fixed inlet tangential
      velocity (IN_VTAN = 1) and a 5 degree inlet pitch angle (IN_VR = 1);
    - ``bc3``: a rotor first (rotor and stator), a spanwise exit pressure
      profile from the deck (IPOUT = 3), the inlet pressure extrapolated from
      mid-pitch, mid-span (IN_PRESS = 4), a fixed relative inlet flow angle
      against the blade speed at J = 1 (IN_VTAN = 2) and the inlet radial
      velocity from J = 2 (IN_VR = 0)."""
    if variant == "bc3":
        spec = deck_tiny(ilos=10, nrows=2, nsteps=nsteps)
        rot, sta = spec.rows
        rot.name, rot.nblades, rot.alpha_in, rot.alpha_out, rot.rpm = "ROTOR", 48, -45.0, -20.0, 6000.0
        sta.name, sta.nblades, sta.alpha_in, sta.alpha_out, sta.rpm = "STATOR", 40, 30.0, 0.0, 0.0
        rot.pguess = (1.0e5, 0.97e5, 0.93e5, 0.92e5)
        sta.pguess = (0.92e5, 0.91e5, 0.90e5, 0.90e5)
        spec.ipout, spec.pd_exit = 3, (0.89e5, 0.91e5)
        spec.in_press, spec.in_vtan, spec.in_vr, spec.yaw = 4, 2, 0, 5.0
        spec.pdown_hub = spec.pdown_tip = 0.9e5
    else:
        spec = deck_tiny(ilos=10, nrows=3, nsteps=nsteps)
        if variant == "bc1":
            spec.ipout, spec.in_press, spec.in_vr, spec.yaw = 1, -3, -3, 10.0
        elif variant == "bc2":
            spec.ipout, spec.in_press, spec.in_vtan, spec.vtin, spec.pitch = -1, 1, 1, 20.0, 5.0
            spec.pdown_hub, spec.pdown_tip = 0.79e5, 0.81e5
        else:
            raise ValueError(variant)
    spec.title = (f"TINY CASCADE ILOS=10 IPOUT={spec.ipout} IN_PRESS={spec.in_press} "
                  f"IN_VTAN={spec.in_vtan} IN_VR={spec.in_vr}")
    return spec


def deck_tiny_supersonic(nsteps: int = 100) -> DeckSpec:
    """A three-row inviscid cascade whose first stator turns the flow to 70
    degrees at a high subsonic exit Mach number (low guessed and imposed exit
    pressures) ahead of a rotor turning against it (negative speed), so the
    flow relative to the rotor is supersonic at JMIX + 2 and NEW_MIXPLAN's
    Prandtl-Meyer angle sweeps (multall-open-18.3.f:15139-15182) run."""
    spec = deck_tiny(ilos=0, nrows=3, nsteps=nsteps)
    s1, r1, s2 = spec.rows
    s1.alpha_out = 70.0
    s1.pguess = (1.0e5, 0.95e5, 0.45e5, 0.42e5)
    r1.alpha_in, r1.alpha_out, r1.rpm = 65.0, 30.0, -6000.0
    r1.pguess = (0.42e5, 0.41e5, 0.36e5, 0.35e5)
    s2.alpha_in, s2.alpha_out = 20.0, 40.0
    s2.pguess = (0.35e5, 0.35e5, 0.34e5, 0.34e5)
    spec.pdown_hub = spec.pdown_tip = 0.34e5
    spec.title = "TINY CASCADE INVISCID SUPERSONIC MIXING PLANES"
    return spec


def deck_tiny_cusp2(nsteps: int = 100) -> DeckSpec:
    """The three-row mixing-length cascade with the trailing-edge separation
    force (IF_CUSP = 2, multall-open-18.3.f:5184-5233) on both stators: from
    4 (I = 1 side) and 3 (I = IM side) points upstream of the trailing edge
    to 3 points into the wake, a layer of 0.05 of the blade thickness,
    momentum factor 0.97."""
    spec = deck_tiny(ilos=10, nrows=3, nsteps=nsteps)
    spec.rows[0].sep_force = (4, 3, 3, 0.05, 0.97)
    spec.rows[2].sep_force = (3, 4, 2, 0.05, 0.97)
    spec.title = "TINY CASCADE ILOS=10 IF_CUSP=2 ON THE STATORS"
    return spec


def deck_tiny_repeat(nsteps: int = 100) -> DeckSpec:
    """A rotor-stator stage (the two-row cascade of ``deck_tiny_bc("bc3")``
    with the plain inlet and exit conditions) under the repeating-stage
    condition (IF_REPEAT = 1, multall-open-18.3.f:5323-5331, 12085-12115):
    every 10 steps the inlet stagnation pressure, yaw angle and tangential
    velocity are relaxed by 0.05 towards the exit profiles."""
    spec = deck_tiny(ilos=10, nrows=2, nsteps=nsteps)
    rot, sta = spec.rows
    rot.name, rot.nblades, rot.alpha_in, rot.alpha_out, rot.rpm = "ROTOR", 48, -45.0, -20.0, 6000.0
    sta.name, sta.nblades, sta.alpha_in, sta.alpha_out, sta.rpm = "STATOR", 40, 30.0, 0.0, 0.0
    rot.pguess = (1.0e5, 0.97e5, 0.93e5, 0.92e5)
    sta.pguess = (0.92e5, 0.91e5, 0.90e5, 0.90e5)
    spec.pdown_hub = spec.pdown_tip = 0.9e5
    spec.repeat = (10, 0.05)
    spec.title = "TINY STAGE ILOS=10 REPEATING STAGE NINMOD=10 RFINBC=0.05"
    return spec


def deck_tiny_nout(nsteps: int = 100) -> DeckSpec:
    """The three-row mixing-length cascade with printout steps inside the
    run: NOUT = 10, 25, 40 and 55 land on convergence checks, so the
    reference refreshes the viscous forces there
    (multall-open-18.3.f:5338-5341); NOUT = 73 does not (the printout block
    runs only on the every-5-step check)."""
    spec = deck_tiny(ilos=10, nrows=3, nsteps=nsteps)
    spec.title = "TINY CASCADE ILOS=10 PRINTOUT AT STEPS 10 25 40 55 73"
    spec.nout = (10, 25, 40, 55, 73)
    return spec


def speedline_sizes(n_points: int = 16, seed: int = 3):
    """The seeded grid-size draw of the compressor speedline: (IM, KM), IM
    from {37, 49} and KM from {33, 49, 65}.  (At IM = 65 this machine's
    pitchwise spacing leaves a steady continuity error, ECONT, of 1.4 % on
    the choke side, above the 1 % a converged point is held to.)"""
    rng = np.random.default_rng(seed)
    return [(int(rng.choice([37, 49])), int(rng.choice([33, 49, 65]))) for _ in range(n_points)]


def speedline_pexit(i: int, n_points: int = 16, lo: float = 0.95, hi: float = 1.08,
                    design: float = 109140.9) -> float:
    """Exit static pressure of point i: linearly spaced from `lo` (towards
    choke) to `hi` (towards stall) times the design exit pressure of the
    reference's one-stage axial compressor deck."""
    return design * (lo + (hi - lo) * i / max(1, n_points - 1))


def regrid_deck(text: str, im: int, km: int, pexit: Optional[float] = None, nsteps: Optional[int] = None) -> str:
    """A NEW-format deck with its pitchwise and spanwise grid counts changed
    (IM, KM and the FP / FR spacing lists, geometric towards mid-passage like
    the reference's own decks, and the KM stream-surface output flags), and
    optionally the exit static pressure (PDOWN_HUB = PDOWN_TIP) and
    NSTEPS_MAX.  The blade geometry is given per section and interpolated
    onto the grid by the reader, so any IM and KM describe the same machine."""
    lines = text.split("\n")
    out = []
    i = 0

    def spacing(n):
        w = np.ones(n)
        for q in range(1, n):
            w[q] = min(w[q - 1] * 1.25, 20.0)
        return np.minimum(w, w[::-1])

    while i < len(lines):
        ln = lines[i]
        up = ln.upper().strip()
        if up.startswith("IM") and "KM" in up and i + 1 < len(lines):
            out += [ln, f"{im:10d}{km:10d}"]
            i += 2
            continue
        if up.startswith("FP(I)") or up.startswith("FR(K)"):
            n = (im if up.startswith("FP") else km) - 1
            out.append(ln)
            out.append(_fmt(spacing(n)))
            i += 1
            while i < len(lines) and re.match(r"^\s*[-0-9.]", lines[i]):
                i += 1
            continue
        if up.startswith("STREAM SURFACES"):
            out += [ln, " " + _ints([0] * km)]
            i += 2
            continue
        if up.startswith("PDOWN_HUB") and pexit is not None:
            out += [ln, f"{pexit:10.1f}{pexit:10.1f}"]
            i += 2
            continue
        if up.startswith("NSTEPS_MAX") and nsteps is not None:
            rest = lines[i + 1].split()
            out += [ln, f"{nsteps:10d}  " + "  ".join(rest[1:])]
            i += 2
            continue
        out.append(ln)
        i += 1
    return "\n".join(out)


SPEEDLINE_BASE = "simple-axial-compr-17.1.dat"   # the reference's one-stage axial compressor (test deck)


def speedline_deck(base_text: str, i: int, n_points: int = 16, nsteps: int = 6000, seed: int = 3) -> str:
    """Point i of the 16-point compressor speedline: the reference's
    one-stage axial compressor (rotor 50 blades at 5000 rpm, stator; 2 rows
    of 111 J-points; mixing length) with the pitchwise and spanwise grid of a
    seeded draw, IM from {37, 49} and KM from {33, 49, 65}, and the exit
    static pressure linearly spaced from 0.95 to 1.08 times the deck's own
    (choke side towards stall).  A comment line opening with "(" loses that
    parenthesis, so that flang's list-directed read of the comment into a
    CHARACTER variable takes the whole line; no input value changes."""
    im, km = speedline_sizes(n_points, seed)[i]
    return _plain_comments(regrid_deck(base_text, im, km, speedline_pexit(i, n_points), nsteps))


def _plain_comments(text: str) -> str:
    """A comment line opening with "(" loses that parenthesis, so that
    flang's list-directed read of the comment into a CHARACTER variable
    takes the whole line; no input value changes."""
    return "\n".join(ln.replace("(", " ", 1) if ln.strip().startswith("(") else ln for ln in text.split("\n"))


def with_spalart_allmaras(text: str, limits=(0.03, 0.03, 0.03, 0.03, 1.0, 0.5),
                          factors=(0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.0, 0.0)) -> str:
    """A mixing-length (ILOS = 10) NEW-format deck switched to the
    Spalart-Allmaras model (ILOS = 200): the ILOS value, the source-factor
    line the reader takes after the Reynolds-number line when ILOS >= 200
    (FAC_STMIX .. FAC_PGRAD), and each row's limits line read as XLLIM_IN,
    XLLIM_LE, XLLIM_TE, XLLIM_DN, FSTURB, TURBVIS_DAMP (multall-open-18.3.f:
    2062-2075)."""
    lines = text.split("\n")
    out = []
    i = 0
    nlim = 0
    while i < len(lines):
        ln = lines[i]
        up = ln.upper().strip()
        if up.startswith("ILOS") and i + 1 < len(lines):
            rest = lines[i + 1].split()
            out += [ln, f"{200:10d}" + "".join(f"{int(v):10d}" for v in rest[1:])]
            i += 2
            continue
        if up.startswith("REYNO") and i + 1 < len(lines):
            out += [ln, lines[i + 1], "  FAC_STMIX, FAC_ST0, FAC_ST1, FAC_ST2, FAC_ST3, FAC_SFVIS, FAC_VORT, FAC_PGRAD",
                    "".join(f"{v:10.4f}" for v in factors)]
            i += 2
            continue
        if up.startswith("MIXING LENGTH LIMITS"):
            out.append(ln)
            i += 1
            while i < len(lines) and re.match(r"^\s*[-0-9.]", lines[i]):
                out.append("".join(f"{v:10.4f}" for v in limits))
                nlim += 1
                i += 1
            continue
        out.append(ln)
        i += 1
    assert nlim > 0, "no mixing-length limits block"
    return "\n".join(out)


C4_BASE = "centrif-compr-17.1.dat"   # the reference's centrifugal compressor (impeller + vaned diffuser)


def config4_deck(base_text: str, nsteps: int = 100, im: int = 97, km: int = 81) -> str:
    """The centrifugal-stage workload: the reference's centrifugal compressor
    test deck (impeller and vaned diffuser, 2 rows of 141 and 146 J-points,
    a mixing plane between them) with a uniform pitchwise and spanwise grid
    of IM = 97 by KM = 81 on both rows and the Spalart-Allmaras model:
    2,254,959 nodes.  (The reference reads IM and KM once for all rows, so
    the impeller cannot keep a finer grid than the diffuser.)  Comment
    lines lose a leading parenthesis, as in speedline_deck."""
    return _plain_comments(with_spalart_allmaras(regrid_deck(base_text, im, km, nsteps=nsteps)))
