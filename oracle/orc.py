"""TEST INFRASTRUCTURE ONLY (the oracle).  Never imported by the product.

ctypes front end of the C restatement (mh_oracle.c).  A case is described by
a dict of named numpy arrays in the dense [k, j, i] layout plus a dict of
scalars -- the same "state dict" the fixtures and the HIP path use.  The
``from_dump`` helper builds that dict from a reference COMMON dump.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libmh_oracle.so")

# Names bound by pointer, grouped by shape kind (see mh_oracle.c).
NODE3 = ["RO", "ROVX", "ROVR", "RORVT", "ROE", "ROVT", "ROWT", "P", "HO", "PEFF", "VX", "VR", "VT",
         "WT", "T_STATIC", "DRO", "DROVX", "DROVR", "DRORVT", "DROE", "AQX", "AQR", "ABX", "ABR",
         "ASX", "ASR", "VOL", "VOLOR", "STEP", "BSTEP", "RSTEP", "XFORCE", "RFORCE", "TFORCE",
         "QSOURCE", "FLOWX", "FLOWT", "FLOWR", "XFLUX", "TFLUX", "RFLUX", "SOURCE", "STORE", "TEMP1",
         "SGEN", "DWALLSQ", "Y_PLUS", "TRANS_DYN_VIS", "DEL_DYNVIS", "TRANS_KVIS", "T_SOURCE", "ROAVG_CELL",
         "VISC_LAM", "VISC_RAT", "VISC_TURB", "XLENGTH", "FST_RAT", "DYN_VIS_ML", "ABS_VORT", "TRANS_VORT",
         "GRAD_TVIS", "TVISGRADX", "TVISGRADR", "TVISGRADT", "DPDS_CELL", "RTHETA"]
JK2 = ["R", "X", "RAVG_CELL", "ABT", "WRABT", "UBLADE", "DX", "DR", "DS", "SMERID", "YPLUS_I1", "YPLUS_IM",
       "RT_THICK"]
IJ2 = ["YPLUS_K1", "YPLUS_KM"]
# SETUP's initial values of the wall-function state (mf:6904-6910), for
# fixtures written before these arrays were part of the state
WALL_STATE_DEFAULTS = {"DPDS_CELL": ("N", 0.0), "YPLUS_I1": ("JK", 25.0), "YPLUS_IM": ("JK", 25.0),
                       "YPLUS_K1": ("IJ", 25.0), "YPLUS_KM": ("IJ", 25.0)}
IK2 = ["FTL", "FBL", "FTR", "FBR"]
VEC_I = ["FP", "FU", "FD"]
VEC_K = ["FR", "FKU", "FKD", "PO1", "TO1", "RPO1", "BSSIN", "BSCOS", "BRSIN", "BRCOS", "VTIN", "PD",
         "PLIM", "VM1", "BS", "BR"]
VEC_J = ["FACUP", "FACDWN", "WRAD", "WHUB", "WTIP"]
VEC_NR = ["VLIMIT", "ROLIMIT", "AVG_CHG", "AVG_BLK", "XLLIM_I1", "XLLIM_IM", "XLLIM_K1", "XLLIM_KM",
          "XLLIM_DWN", "XLLIM_UP", "CHORD", "ROUGH_H", "ROUGH_T", "ROUGH_L", "ROUGH_U", "FSTURB", "TURBVIS_DAMP", "VISCOSY", "REYNOLDS",
          "ROVEXIT", "SEP_THIK", "SEP_DRAG"]
NRTAB = ["FILAM", "FITURB", "FIWAKE", "FIUP", "FKLAM", "FKTURB"]     # (NRS, MAXKI) tables
INT_I = ["IB1", "IB2"]
INT_K = ["KB1", "KB2"]
INT_J = ["NBLADE", "IND", "INDLE", "INDMIX", "INDTE", "NROW", "JSBLK", "JB1", "JB2"]
INT_NR = ["JSTART", "JMIX", "JLE", "JTE", "JLED", "KTIPS", "KTIPE", "KS1", "KS2", "JTRAN_I1",
          "JTRAN_IM", "JTRAN_K1", "JTRAN_KM"]
INT_RN = ["IF_CUSP", "NUP_I1", "NUP_IM", "N_WAKE"]     # one per row
SCALARS_F = ["CFL", "DAMP", "DAMPIN", "MACHLIM", "F_PDOWN", "CONLIM", "SFXIN", "SFTIN", "SFX", "SFT",
             "FAC_4TH", "F1", "F2", "F3", "RSMTH", "SFEXIT", "FSMTHB", "RFIN", "PDOWN_HUB",
             "PDOWN_TIP", "FEXTRAP", "RFMIX", "FANGLE", "CP", "GA", "RGAS", "CV", "FGA", "GA1", "RCP",
             "VSOUND", "FACMIXUP", "REYNO", "RF_VIS", "FTRANS", "PRANDTL", "YPLUSWALL",
             "TURBVIS_LIM", "VISLAM", "FTCOND", "YPLAM", "YPTURB", "CFWALL", "FAC_STMIX", "FAC_ST0",
             "FAC_ST1", "FAC_ST2", "FAC_ST3", "FAC_SFVIS", "FAC_VORT", "FAC_PGRAD", "VORT_MAX", "FLOWIN", "RFLOW",
             "RFINBC", "PO_IN_MID", "VT_IN_MID"]
SCALARS_I = ["IM", "JM", "KM", "NROWS", "IMID", "KMID", "NSTEPS_MAX", "NCHANGE", "NRSMTH", "NSFEXIT",
             "IN_PRESS", "IN_VTAN", "IN_VR", "IPOUT", "IFMIX", "IR", "JR", "KR", "ILOS", "NLOS",
             "NMIXUP", "NIB1", "NJB1", "NKB1", "NIB2", "NJB2", "NKB2", "NSBLK", "ITIMST",
             "IF_RESTART", "IFGAS", "IN_FLOW", "IF_REPEAT", "IBOUND", "IRBB", "JRBB", "KRBB", "NINMOD"]



# arrays a host state may hold that LOOP never reads (EFICOOL's stage
# numbering of the rows, mf:11271-11295): not bound
NOT_READ_BY_LOOP = {"NSTAGE"}

def with_wall_defaults(state: Dict) -> Dict:
    """The state with SETUP's initial wall-function arrays added where absent."""
    sc = state["scalars"]
    im, jm, km = int(sc["IM"]), int(sc["JM"]), int(sc["KM"])
    shape = {"N": (km, jm, im), "JK": (km, jm), "IJ": (jm, im)}
    out = dict(state)
    for n, (kind, v) in WALL_STATE_DEFAULTS.items():
        if n not in out:
            out[n] = np.full(shape[kind], v, np.float32)
    return out


def with_workspace(state: Dict) -> Dict:
    """A state set up by the Fortran host (which hands over only what the
    library takes) completed with zeros, as in a fresh COMMON block, for the
    oracle's work arrays and the per-row LOSS outputs its viscous calls
    recompute before reading."""
    sc = state["scalars"]
    im, jm, km, nr = int(sc["IM"]), int(sc["JM"]), int(sc["KM"]), int(sc["NROWS"])
    out = with_wall_defaults(state)
    for names, shape in ((NODE3, (km, jm, im)), (VEC_K, (km,)), (VEC_NR, (nr,)),
                         (NRTAB, (max(im, km, 65), 21))):
        for n in names:
            if n not in out:
                out[n] = np.zeros(shape, np.float32)
    return out


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def from_dump(d) -> Dict:
    """State dict (dense arrays + scalars) from a reference ``refdump.Dump``."""
    im, jm, km, nr = int(d["IM"]), int(d["JM"]), int(d["KM"]), int(d["NROWS"])
    rdt = np.float32
    st: Dict = {"scalars": {}}
    for n in NODE3:
        st[n] = d.node(n).astype(rdt)
    for n in JK2:
        st[n] = d.jk(n).astype(rdt)
    for n in IK2:
        st[n] = np.ascontiguousarray(d[n][:km, :im]).astype(rdt)
    for n in IJ2:
        st[n] = np.ascontiguousarray(d[n][:jm, :im]).astype(rdt)
    for n in VEC_I:
        st[n] = np.array(d[n][:im], rdt)
    for n in VEC_K:
        st[n] = np.array(d[n][:km], rdt)
    for n in VEC_J:
        st[n] = np.array(d[n][:jm], rdt)
    for n in VEC_NR:
        st[n] = np.array(d[n][:nr], rdt)
    for n in NRTAB:
        st[n] = np.ascontiguousarray(d[n]).astype(rdt)   # [MAXKI][NRS]
    for n in INT_I:
        st[n] = np.array(d[n][:im], np.int32)
    for n in INT_K:
        st[n] = np.array(d[n][:km], np.int32)
    for n in INT_J:
        st[n] = np.array(d[n][:jm], np.int32)
    for n in INT_NR:
        st[n] = np.array(d[n][:nr + 1], np.int32)
    for n in INT_RN:
        st[n] = np.array(d[n][:nr], np.int32)
    st["NOUT"] = np.array(d["NOUT"][:5], np.int32)
    nib1, njb1, nkb1 = int(d["NIB1"]), int(d["NJB1"]), int(d["NKB1"])
    nib2, njb2, nkb2 = int(d["NIB2"]), int(d["NJB2"]), int(d["NKB2"])
    st["STEP1"] = np.ascontiguousarray(d["STEP1"][:nkb1, :njb1 + 1, :nib1]).astype(rdt)
    st["STEP2"] = np.ascontiguousarray(d["STEP2"][:nkb2, :njb2 + 1, :nib2]).astype(rdt)
    st["STEPSBK"] = np.array(d["STEPSBK"][:int(d["NSBLK"])], rdt)
    for n in SCALARS_F:
        st["scalars"][n] = float(d[n])
    for n in SCALARS_I:
        st["scalars"][n] = int(d[n])
    return st


class Oracle:
    """One oracle case bound to (copies of) the arrays of a state dict."""

    def __init__(self, state: Dict):
        if not os.path.exists(LIB):
            build()
        self.lib = ctypes.CDLL(LIB)
        L = self.lib
        L.orc_create.restype = ctypes.c_void_p
        L.orc_bind.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
        L.orc_scalar.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double]
        L.orc_get_scalar.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.orc_get_scalar.restype = ctypes.c_double
        L.orc_start.argtypes = [ctypes.c_void_p]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for fn in ("orc_primitives", "orc_new_mixplan", "orc_loss", "orc_spal_loss", "orc_new_loss"):
            getattr(L, fn).argtypes = [ctypes.c_void_p]
        L.orc_stepup.argtypes = [ctypes.c_void_p, ctypes.c_float]
        L.orc_tstep.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.orc_smooth_var.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.h = L.orc_create()
        self.a: Dict[str, np.ndarray] = {}
        state = with_wall_defaults(state)
        for k, v in state.items():
            if k == "scalars" or k in NOT_READ_BY_LOOP:
                continue
            arr = np.ascontiguousarray(v).copy()
            self.a[k] = arr
            if L.orc_bind(self.h, k.encode(), arr.ctypes.data) != 0:
                raise KeyError(k)
        for k, v in state["scalars"].items():
            L.orc_scalar(self.h, k.encode(), float(v))   # names the oracle does not use are ignored

    def start(self):
        rc = self.lib.orc_start(self.h)
        if rc != 0:
            raise RuntimeError(f"orc_start rc={rc}")

    def step(self, n: int) -> int:
        return self.lib.orc_step(self.h, n)

    def scalar(self, name: str) -> float:
        return self.lib.orc_get_scalar(self.h, name.encode())

    def __getitem__(self, k):
        return self.a[k]

    def __del__(self):
        try:
            self.lib.orc_destroy(self.h)
        except Exception:
            pass
