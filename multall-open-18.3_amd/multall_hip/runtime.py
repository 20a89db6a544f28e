"""ctypes binding of libmultall_hip.so (include/multall_hip.h).

``Solver`` owns one ``mh_ctx``: it uploads a state dict (named numpy arrays
in the dense [k, j, i] layout plus a ``scalars`` dict, the COMMON-block
contents after the reference's SETUP), runs time steps on the GPU and
reads arrays back.  The library is required: there is no CPU fallback, and
a missing or unloadable library raises immediately.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libmultall_hip.so")

EXPORTS = ["mh_create", "mh_destroy", "mh_last_error", "mh_array_id", "mh_scalar_id", "mh_array_len",
           "mh_array_is_int", "mh_set_array", "mh_get_array", "mh_set_array_strided",
           "mh_get_array_strided", "mh_set_scalar", "mh_get_scalar", "mh_loss_init", "mh_set_xlength", "mh_start", "mh_step",
           "mh_integrals", "mh_sync", "mh_profile_enable", "mh_profile_report", "mh_step_batch",
           "mh_request_stop", "mh_set_xlength_noalloc", "mh_eficool", "mh_mix_bconds"]


class MhDims(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("im", "jm", "km", "nrows", "nib1", "njb1", "nkb1", "nib2", "njb2", "nkb2", "nsblk")]


class MhResid(ctypes.Structure):
    _fields_ = [("nstep", ctypes.c_int), ("ifend", ctypes.c_int), ("emax", ctypes.c_float),
                ("eavg", ctypes.c_float), ("econt", ctypes.c_float), ("vref", ctypes.c_float),
                ("vmax", ctypes.c_float), ("flow_in", ctypes.c_float), ("flow_out", ctypes.c_float),
                ("imax", ctypes.c_int), ("jmax", ctypes.c_int), ("kmax", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class MhPerf(ctypes.Structure):
    """One stage's (stage >= 1) or the whole machine's (stage 0) performance
    record of mh_eficool."""
    _fields_ = [(n, ctypes.c_int) for n in ("stage", "j1", "j2m1", "jmix", "compressor", "pad_")] + \
               [(n, ctypes.c_double) for n in ("po_in", "po_out", "ps_in", "ps_out", "to_in", "to_out", "flow_in",
                                               "flow_out", "po_ratio", "eta_tt", "eta_ts", "eta_poly", "reaction",
                                               "reaction_hub", "reaction_mid", "reaction_tip", "work_kw")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad_"}


class MhError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH):
    """Load the library (once) and declare the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MhError(f"libmultall_hip.so not built: {path} (run __graft_entry__.build())")
    L = ctypes.CDLL(path)
    vp, i, ll, cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_char_p
    L.mh_create.argtypes = [ctypes.POINTER(MhDims), ctypes.POINTER(vp)]
    L.mh_destroy.argtypes = [vp]
    L.mh_last_error.argtypes = [vp]
    L.mh_last_error.restype = cp
    L.mh_array_id.argtypes = [cp]
    L.mh_scalar_id.argtypes = [cp]
    L.mh_array_len.argtypes = [vp, i]
    L.mh_array_len.restype = ll
    L.mh_array_is_int.argtypes = [i]
    L.mh_set_array.argtypes = [vp, i, vp, ll]
    L.mh_get_array.argtypes = [vp, i, vp, ll]
    L.mh_set_array_strided.argtypes = [vp, i, vp, i, i]
    L.mh_get_array_strided.argtypes = [vp, i, vp, i, i]
    L.mh_set_scalar.argtypes = [vp, i, ctypes.c_double]
    L.mh_get_scalar.argtypes = [vp, i]
    L.mh_get_scalar.restype = ctypes.c_double
    L.mh_start.argtypes = [vp]
    L.mh_loss_init.argtypes = [vp]
    L.mh_set_xlength.argtypes = [vp] + [ctypes.c_void_p] * 5
    L.mh_set_xlength_noalloc.argtypes = [vp] + [ctypes.c_void_p] * 5
    L.mh_step.argtypes = [vp, i, ctypes.POINTER(MhResid)]
    L.mh_request_stop.argtypes = [vp]
    L.mh_step_batch.argtypes = [ctypes.POINTER(vp), i, i, ctypes.POINTER(MhResid), ctypes.POINTER(i)]
    L.mh_integrals.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ll]
    L.mh_eficool.argtypes = [vp, ctypes.POINTER(MhPerf), i, ctypes.POINTER(i)]
    L.mh_mix_bconds.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ll]
    L.mh_sync.argtypes = [vp]
    L.mh_profile_enable.argtypes = [vp, i]
    L.mh_profile_report.argtypes = [vp, ctypes.c_char_p, i]
    _lib = L
    return L


def dims_from_state(st: Dict) -> MhDims:
    s = st["scalars"]
    return MhDims(im=int(s["IM"]), jm=int(s["JM"]), km=int(s["KM"]), nrows=int(s["NROWS"]),
                  nib1=int(s["NIB1"]), njb1=int(s["NJB1"]), nkb1=int(s["NKB1"]),
                  nib2=int(s["NIB2"]), njb2=int(s["NJB2"]), nkb2=int(s["NKB2"]), nsblk=int(s["NSBLK"]))


class Solver:
    """One machine's time-marching state on the GPU."""

    def __init__(self, state: Dict, exact_order: bool = False):
        self.L = load()
        self.dims = dims_from_state(state)
        h = ctypes.c_void_p()
        rc = self.L.mh_create(ctypes.byref(self.dims), ctypes.byref(h))
        if rc != 0:
            raise MhError(f"mh_create failed rc={rc}")
        self.h = h
        self._keep = []
        for name, arr in state.items():
            if name == "scalars":
                continue
            aid = self.L.mh_array_id(name.encode())
            if aid < 0:
                continue
            n = self.L.mh_array_len(self.h, aid)
            want = np.int32 if self.L.mh_array_is_int(aid) == 1 else np.float32
            a = np.ascontiguousarray(np.asarray(arr).ravel()[:n], dtype=want)
            if a.size != n:
                raise MhError(f"{name}: have {a.size} elements, need {n}")
            self._check(self.L.mh_set_array(self.h, aid, a.ctypes.data, n), f"set {name}")
        for name, v in state["scalars"].items():
            sid = self.L.mh_scalar_id(name.encode())
            if sid >= 0:
                self._check(self.L.mh_set_scalar(self.h, sid, float(v)), f"scalar {name}")
        self._check(self.L.mh_set_scalar(self.h, self.L.mh_scalar_id(b"EXACT_ORDER"),
                                         1.0 if exact_order else 0.0), "exact")

    def _check(self, rc, what):
        if rc != 0:
            msg = self.L.mh_last_error(self.h).decode(errors="replace") if self.h else ""
            raise MhError(f"{what}: rc={rc} {msg}")

    def loss_init(self):
        """The SETUP-time viscous-force initialisation (one LOSS call)."""
        self._check(self.L.mh_loss_init(self.h), "mh_loss_init")

    def set_xlength(self, theta, xllim_in, xllim_le, xllim_te, xllim_dn):
        """SET_XLENGTH on the device (wall distances, XLENGTH, FST_RAT) from
        the nodal THETA and the per-row mixing-length limits."""
        arrs = [np.ascontiguousarray(a, np.float32) for a in (theta, xllim_in, xllim_le, xllim_te, xllim_dn)]
        self._check(self.L.mh_set_xlength(self.h, *[a.ctypes.data for a in arrs]), "mh_set_xlength")

    def set_xlength_noalloc(self, theta, xllim_in, xllim_le, xllim_te, xllim_dn):
        """set_xlength on the library's per-device scratch array (no device
        allocation; grids up to MH_XLENGTH_SCRATCH_NODES nodes)."""
        arrs = [np.ascontiguousarray(a, np.float32) for a in (theta, xllim_in, xllim_le, xllim_te, xllim_dn)]
        self._check(self.L.mh_set_xlength_noalloc(self.h, *[a.ctypes.data for a in arrs]), "mh_set_xlength_noalloc")

    def start(self):
        self._check(self.L.mh_start(self.h), "mh_start")

    def set_scalar(self, name: str, value: float):
        self._check(self.L.mh_set_scalar(self.h, self.L.mh_scalar_id(name.encode()), float(value)), f"scalar {name}")

    def scalar(self, name: str) -> float:
        return float(self.L.mh_get_scalar(self.h, self.L.mh_scalar_id(name.encode())))

    def step(self, n: int) -> Dict:
        r = MhResid()
        self._check(self.L.mh_step(self.h, int(n), ctypes.byref(r)), "mh_step")
        return r.as_dict()

    def request_stop(self):
        """The `stopit` request: the next step divisible by 10 ends the run."""
        self._check(self.L.mh_request_stop(self.h), "mh_request_stop")

    def get(self, name: str, shape=None) -> np.ndarray:
        aid = self.L.mh_array_id(name.encode())
        if aid < 0:
            raise KeyError(name)
        n = self.L.mh_array_len(self.h, aid)
        dt = np.int32 if self.L.mh_array_is_int(aid) == 1 else np.float32
        out = np.empty(n, dt)
        self._check(self.L.mh_get_array(self.h, aid, out.ctypes.data, n), f"get {name}")
        d = self.dims
        if shape is None and n == d.im * d.jm * d.km:
            shape = (d.km, d.jm, d.im)
        return out.reshape(shape) if shape else out

    def integrals(self) -> np.ndarray:
        jm = self.dims.jm
        out = np.empty(7 * jm, np.float64)
        self._check(self.L.mh_integrals(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 7 * jm),
                    "mh_integrals")
        return out.reshape(7, jm)

    def eficool(self) -> list:
        """Stage and machine performance of the current state (mh_eficool):
        a list of dicts, stages first, then the whole machine (stage 0)
        when there is more than one stage."""
        cap = int(self.dims.nrows) + 1
        out = (MhPerf * cap)()
        n = ctypes.c_int(0)
        self._check(self.L.mh_eficool(self.h, out, cap, ctypes.byref(n)), "mh_eficool")
        return [out[i].as_dict() for i in range(n.value)]

    MIXB_NAMES = ["PO", "P", "TO", "TS", "VT", "VM", "YAW", "PITCH"]

    def mix_bconds(self) -> np.ndarray:
        """MIX_BCONDS profiles (mh_mix_bconds): [plane, quantity, K] for
        K = 1..KM, then the plane's mass average at index KM; planes: inlet,
        then each row's JMIX; quantities in MIXB_NAMES order."""
        np_, km = int(self.dims.nrows) + 1, int(self.dims.km)
        out = np.empty(np_ * 8 * (km + 1), np.float64)
        self._check(self.L.mh_mix_bconds(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size),
                    "mh_mix_bconds")
        return out.reshape(np_, 8, km + 1)

    def profile(self, on: bool):
        self._check(self.L.mh_profile_enable(self.h, 1 if on else 0), "mh_profile_enable")

    def profile_report(self) -> Dict[str, tuple]:
        """{kernel: (launches, total_ms)} since profiling was enabled."""
        buf = ctypes.create_string_buffer(1 << 16)
        n = self.L.mh_profile_report(self.h, buf, len(buf))
        if n < 0:
            self._check(n, "mh_profile_report")
        out = {}
        for line in buf.value.decode().splitlines():
            name, calls, ms = line.split()
            out[name] = (int(calls), float(ms))
        return out

    def sync(self):
        self._check(self.L.mh_sync(self.h), "mh_sync")

    def close(self):
        if getattr(self, "h", None):
            self.L.mh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def step_batch(solvers, n: int):
    """Advance several solvers (distinct machines) by up to n steps each,
    concurrently on the GPU (mh_step_batch); returns their residual dicts."""
    L = load()
    k = len(solvers)
    hs = (ctypes.c_void_p * k)(*[s.h.value for s in solvers])
    out = (MhResid * k)()
    failed = ctypes.c_int(-1)
    rc = L.mh_step_batch(hs, k, int(n), out, ctypes.byref(failed))
    if rc != 0:
        i = failed.value
        msg = L.mh_last_error(solvers[i].h).decode(errors="replace") if 0 <= i < k else ""
        raise MhError(f"mh_step_batch: context {i} rc={rc} {msg}")
    return [out[i].as_dict() for i in range(k)]
