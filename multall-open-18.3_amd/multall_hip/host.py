"""ctypes front end of libmultall_host.so (the Fortran deck reader and grid
set-up, host/mh_host.f90): turns a NEW-format input deck into the state dict
that ``runtime.Solver`` uploads -- the reference's COMMON state after SETUP.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict

import numpy as np

from .runtime import PKG_ROOT

HOST_LIB = os.path.join(PKG_ROOT, "lib", "libmultall_host.so")

# arrays the time-marching library takes (see include/multall_hip.h)
N3 = ["RO", "ROVX", "ROVR", "RORVT", "ROE", "DRO", "DROVX", "DROVR", "DRORVT", "DROE", "ROVT", "ROWT",
      "P", "HO", "PEFF", "VX", "VR", "VT", "WT", "T_STATIC", "AQX", "AQR", "ABX", "ABR", "ASX", "ASR",
      "VOL", "VOLOR", "STEP", "BSTEP", "RSTEP", "XFORCE", "RFORCE", "TFORCE", "QSOURCE", "FLOWX",
      "FLOWT", "FLOWR", "DWALLSQ", "XLENGTH", "FST_RAT", "ROAVG_CELL", "TRANS_DYN_VIS", "RTHETA"]
JK = ["R", "X", "RAVG_CELL", "ABT", "WRABT", "UBLADE", "DX", "DR", "DS", "SMERID", "RT_THICK"]
IK = ["FTL", "FBL", "FTR", "FBR"]
VI = ["FP", "FU", "FD"]
VK = ["FR", "FKU", "FKD", "PO1", "TO1", "RPO1", "BSSIN", "BSCOS", "BRSIN", "BRCOS", "VTIN", "PD", "VM1", "BS", "BR"]
VJ = ["FACUP", "FACDWN", "WRAD", "WHUB", "WTIP"]
RN = ["XLLIM_I1", "XLLIM_IM", "XLLIM_K1", "XLLIM_KM", "XLLIM_DWN", "XLLIM_UP", "ROUGH_H", "ROUGH_T", "ROUGH_L",
      "ROUGH_U", "FSTURB", "TURBVIS_DAMP", "CHORD", "SEP_THIK", "SEP_DRAG"]
RNI = ["JTRAN_I1", "JTRAN_IM", "JTRAN_K1", "JTRAN_KM", "NSTAGE", "IF_CUSP", "NUP_I1", "NUP_IM", "N_WAKE"]
MG = ["STEP1", "STEP2", "STEPSBK"]
II = ["IB1", "IB2"]
IKK = ["KB1", "KB2"]
IJ = ["NBLADE", "IND", "INDLE", "INDMIX", "INDTE", "NROW", "JSBLK", "JB1", "JB2"]
IR_ = ["JSTART", "JMIX", "JLE", "JTE", "JLED", "KTIPS", "KTIPE", "KS1", "KS2"]
SCALARS = ["IM", "JM", "KM", "NROWS", "IMID", "KMID", "NIB1", "NJB1", "NKB1", "NIB2", "NJB2", "NKB2",
           "NSBLK", "ITIMST", "NRSMTH", "IF_RESTART", "NSTEPS_MAX", "NCHANGE", "IR", "JR", "KR",
           "IRBB", "JRBB", "KRBB", "IFMIX", "IN_PRESS", "IN_VTAN", "IN_VR", "IN_FLOW", "IF_REPEAT",
           "IPOUT", "NSFEXIT", "ILOS", "NLOS", "IBOUND", "NMIXUP", "IFGAS", "CP", "GA", "F1", "F2",
           "F3", "RSMTH", "CFL", "DAMPIN", "MACHLIM", "F_PDOWN", "CONLIM", "SFXIN", "SFTIN", "FAC_4TH",
           "FBLK1", "FBLK2", "FBLK3", "RFMIX", "FEXTRAP", "FSMTHB", "FANGLE", "RFIN", "SFEXIT", "REYNO",
           "RF_VIS", "FTRANS", "TURBVIS_LIM", "PRANDTL", "YPLUSWALL", "PDOWN_HUB", "PDOWN_TIP",
           "FACMIXUP", "GA1", "FGA", "RCP", "CV", "RGAS", "VSOUND", "VORT_MAX", "YPLAM", "YPTURB", "CFWALL", "FAC_STMIX", "FAC_ST0",
           "FAC_ST1", "FAC_ST2", "FAC_ST3", "FAC_SFVIS", "FAC_VORT", "FAC_PGRAD", "FLOWIN", "RFLOW",
           "NINMOD", "RFINBC", "PO_IN_MID", "VT_IN_MID"]

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"libmultall_host.so not built: {HOST_LIB}")
        L = ctypes.CDLL(HOST_LIB)
        L.mhh_load.argtypes = [ctypes.c_char_p]
        L.mhh_load.restype = ctypes.c_int
        L.mhh_error.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.mhh_scalar.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]
        L.mhh_scalar.restype = ctypes.c_int
        L.mhh_array.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_longlong]
        L.mhh_array.restype = ctypes.c_longlong
        L.mhh_set_host_xlength.argtypes = [ctypes.c_int]
        L.mhh_write_grid_out.argtypes = [ctypes.c_char_p]
        L.mhh_write_grid_out.restype = ctypes.c_int
        _lib = L
    return _lib


def write_grid_out(path: str) -> None:
    """The plotting grid file of the deck loaded last, in the record layout
    the reference's SETUP writes (grid_out)."""
    L = _load()
    if L.mhh_write_grid_out(os.fspath(path).encode()) != 0:
        buf = ctypes.create_string_buffer(512)
        L.mhh_error(buf, 512)
        raise OSError(buf.value.decode(errors="replace"))


def array(name: str, shape, dtype=np.float32) -> np.ndarray:
    """One more array of the deck loaded last (e.g. THETA, XLLIM_IN), by its
    reference name."""
    L = _load()
    out = np.empty(int(np.prod(shape)), dtype)
    n = L.mhh_array(name.encode(), out.ctypes.data, out.size)
    if n != out.size:
        raise RuntimeError(f"{name}: host returned {n}, expected {out.size}")
    return out.reshape(shape)


def load_deck(path: str, xlength_on_host: bool = True) -> Dict:
    """Read and set up a deck; return the state dict (dense [k, j, i] arrays).
    With xlength_on_host False the wall distances and mixing lengths
    (DWALLSQ, XLENGTH, FST_RAT) are left zero for ``Solver.set_xlength``."""
    L = _load()
    L.mhh_set_host_xlength(1 if xlength_on_host else 0)
    if L.mhh_load(os.fspath(path).encode()) != 0:
        buf = ctypes.create_string_buffer(512)
        L.mhh_error(buf, 512)
        raise ValueError(f"{path}: {buf.value.decode(errors='replace')}")
    sc = {}
    for n in SCALARS:
        v = ctypes.c_double()
        if L.mhh_scalar(n.encode(), ctypes.byref(v)) != 0:
            raise KeyError(n)
        sc[n] = float(v.value)
    for n in list(sc):
        if n in ("CP", "GA", "F1", "F2", "F3", "RSMTH", "CFL", "DAMPIN", "MACHLIM", "F_PDOWN", "CONLIM",
                 "SFXIN", "SFTIN", "FAC_4TH", "FBLK1", "FBLK2", "FBLK3", "RFMIX", "FEXTRAP", "FSMTHB",
                 "FANGLE", "RFIN", "SFEXIT", "REYNO", "RF_VIS", "FTRANS", "TURBVIS_LIM", "PRANDTL",
                 "YPLUSWALL", "PDOWN_HUB", "PDOWN_TIP", "FACMIXUP", "GA1", "FGA", "RCP", "CV", "RGAS",
                 "VSOUND", "VORT_MAX", "YPLAM", "YPTURB", "CFWALL", "FAC_STMIX", "FAC_ST0", "FAC_ST1",
                 "FAC_ST2", "FAC_ST3", "FAC_SFVIS", "FAC_VORT", "FAC_PGRAD", "FLOWIN", "RFLOW",
                 "RFINBC", "PO_IN_MID", "VT_IN_MID"):
            continue
        sc[n] = int(sc[n])
    im, jm, km = sc["IM"], sc["JM"], sc["KM"]
    st: Dict = {"scalars": sc}

    def arr(name, shape, dt):
        out = np.empty(int(np.prod(shape)), dt)
        n = L.mhh_array(name.encode(), out.ctypes.data, out.size)
        if n != out.size:
            raise RuntimeError(f"{name}: host returned {n}, expected {out.size}")
        return out.reshape(shape)

    for n in N3:
        st[n] = arr(n, (km, jm, im), np.float32)
    for n in JK:
        st[n] = arr(n, (km, jm), np.float32)
    for n in IK:
        st[n] = arr(n, (km, im), np.float32)
    for n in VI:
        st[n] = arr(n, (im,), np.float32)
    for n in VK:
        st[n] = arr(n, (km,), np.float32)
    for n in VJ:
        st[n] = arr(n, (jm,), np.float32)
    st["STEP1"] = arr("STEP1", (sc["NKB1"], sc["NJB1"] + 1, sc["NIB1"]), np.float32)
    st["STEP2"] = arr("STEP2", (sc["NKB2"], sc["NJB2"] + 1, sc["NIB2"]), np.float32)
    st["STEPSBK"] = arr("STEPSBK", (sc["NSBLK"],), np.float32)
    for n in II:
        st[n] = arr(n, (im,), np.int32)
    for n in IKK:
        st[n] = arr(n, (km,), np.int32)
    for n in IJ:
        st[n] = arr(n, (jm,), np.int32)
    for n in IR_:
        st[n] = arr(n, (sc["NROWS"] + 1,), np.int32)
    for n in RN:
        st[n] = arr(n, (sc["NROWS"],), np.float32)
    for n in RNI:
        st[n] = arr(n, (sc["NROWS"],), np.int32)
    st["NOUT"] = arr("NOUT", (5,), np.int32)     # the printout steps

    # the turbulence state SETUP starts from (mf:6903-6909); the viscous
    # models' first call (mh_loss_init) fills in the rest
    st["Y_PLUS"] = np.full((km, jm, im), 25.0, np.float32)
    st["VISC_RAT"] = np.ones((km, jm, im), np.float32)
    # (TRANS_DYN_VIS is the host's: zero, or the restart file's)
    for n in ("DEL_DYNVIS", "TRANS_KVIS", "T_SOURCE", "VISC_LAM"):
        st[n] = np.zeros((km, jm, im), np.float32)
    return st
