"""TEST INFRASTRUCTURE ONLY (the oracle).  Never imported by the product.

Runs the reference solver built by ``build_ref.sh`` on a deck and decodes
the raw COMMON-block dumps written by ``ref_harness.f``/``ref_dump.c``.

The member layout of each block is recovered by parsing the reference's
COMMON include file (commall-open-18.3:13-203) with FORTRAN's implicit
typing rule (names starting I-N are INTEGER*4, others default REAL, plus
the explicit ``REAL MACHLIM, MACHREL, MACHCOOL`` at :205).  COMMON storage
is a contiguous sequence, so offsets are running sums of member sizes.
Arrays are returned in C order ``[k, j, i]`` (I fastest, as in the file)
and trimmed to the logical extents where the caller asks for it.
"""
from __future__ import annotations

import os
import resource
import re
import shutil
import subprocess
import tempfile
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
COMMALL = "/root/reference/Multall-open-18.3/multall-18.3/commall-open-18.3"

# The dimensions build_ref.sh compiles with (commall-open-18.3:4-9 defaults
# except the first five).
BUILD_PARAMS = dict(ID=65, JD=600, KD=65, MAXKI=65, NRS=21, IG1=30, JG1=1000, KG1=30,
                    IG2=25, JG2=500, KG2=20, JG3=100, NCWL=100, NBLED=25)
# per executable: fp32w is the pitchwise-wide build (DIMS_WIDE in build_ref.sh)
VARIANT_PARAMS = {"fp32": BUILD_PARAMS, "fp64": BUILD_PARAMS,
                  "fp32w": dict(BUILD_PARAMS, ID=96, JD=200, KD=33, MAXKI=96),
                  "fp32l": dict(BUILD_PARAMS, ID=38, JD=1300, KD=38, MAXKI=38),
                  "fp64l": dict(BUILD_PARAMS, ID=38, JD=1300, KD=38, MAXKI=38),
                  "fp32c": dict(BUILD_PARAMS, ID=97, JD=290, KD=81, MAXKI=97, IG1=40),
                  "fp64c": dict(BUILD_PARAMS, ID=97, JD=290, KD=81, MAXKI=97, IG1=40)}

_EXPLICIT_REAL = {"MACHLIM", "MACHREL", "MACHCOOL"}


def parse_commall(path: str = COMMALL, params: Dict[str, int] = BUILD_PARAMS):
    """Return {block: [(name, is_int, dims_tuple), ...]} from the include file."""
    text = open(path).read().splitlines()
    # join fixed-form continuation lines (non-blank in column 6)
    stmts = []
    for line in text:
        if not line.strip() or line[0] in "Cc*!":
            continue
        if len(line) > 5 and line[5] not in " 0" and stmts:
            stmts[-1] += line[6:72]
        else:
            stmts.append(line[6:72])
    blocks = {}
    for st in stmts:
        s = st.strip()
        m = re.match(r"COMMON\s*/\s*(\w+)\s*/(.*)", s, re.I)
        if not m:
            continue
        name, body = m.group(1).upper(), m.group(2)
        members = []
        depth, cur = 0, ""
        for ch in body + ",":
            if ch == "(":
                depth += 1
            elif ch == ")":
                depth -= 1
            if ch == "," and depth == 0:
                if cur.strip():
                    members.append(cur.strip())
                cur = ""
            else:
                cur += ch
        out = []
        for mem in members:
            mm = re.match(r"(\w+)\s*(?:\((.*)\))?", mem)
            nm = mm.group(1).upper()
            dims = ()
            if mm.group(2):
                dims = tuple(int(eval(d.strip().upper(), {}, dict(params)))
                             for d in mm.group(2).split(","))
            is_int = nm[0] in "IJKLMN" and nm not in _EXPLICIT_REAL
            out.append((nm, is_int, dims))
        blocks[name] = out
    return blocks


class Dump:
    """Decoded contents of one dump file: ``dump['RO']`` -> ndarray [k,j,i]."""

    def __init__(self, path: str, layout=None):
        layout = layout or parse_commall()
        raw = open(path, "rb").read()
        assert raw[:7] == b"MHDUMP1", path
        self.real_size = int(np.frombuffer(raw, np.int32, 1, 8)[0])
        nblk = int(np.frombuffer(raw, np.int32, 1, 12)[0])
        self.loop_seconds = float(np.frombuffer(raw, np.float64, 1, 16)[0])
        off = 24
        rdt = np.float32 if self.real_size == 4 else np.float64
        self.vars: Dict[str, np.ndarray] = {}
        for _ in range(nblk):
            bname = raw[off:off + 16].split(b"\0")[0].decode()
            nbytes = int(np.frombuffer(raw, np.int64, 1, off + 16)[0])
            data = raw[off + 24: off + 24 + nbytes]
            off += 24 + nbytes
            pos = 0
            for (nm, is_int, dims) in layout[bname]:
                n = int(np.prod(dims)) if dims else 1
                dt = np.int32 if is_int else rdt
                isz = np.dtype(dt).itemsize
                if pos % isz:  # the compiler aligns an 8-byte REAL after a 4-byte INTEGER
                    pos += isz - pos % isz
                sz = isz * n
                arr = np.frombuffer(data, dt, n, pos)
                pos += sz
                self.vars[nm] = arr.reshape(dims[::-1]) if dims else arr[0]
            assert pos == nbytes, (bname, pos, nbytes)

    def __getitem__(self, k):
        return self.vars[k.upper()]

    def __contains__(self, k):
        return k.upper() in self.vars

    def node(self, name):
        """3D (ID,JD,KD) array trimmed to (KM, JM, IM)."""
        im, jm, km = int(self["IM"]), int(self["JM"]), int(self["KM"])
        return np.ascontiguousarray(self[name][:km, :jm, :im])

    def jk(self, name):
        """2D (JD,KD) array trimmed to (KM, JM)."""
        jm, km = int(self["JM"]), int(self["KM"])
        return np.ascontiguousarray(self[name][:km, :jm])


def run_reference(deck_text: str, variant: str = "fp32", workdir: Optional[str] = None,
                  timeout: float = 3600.0):
    """Run the reference on ``deck_text``; return (setup Dump, final Dump, stdout)."""
    exe = os.path.join(REF_DIR, f"multall_ref_{variant}")
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    own = workdir is None
    wd = workdir or tempfile.mkdtemp(prefix="mhref_")
    try:
        with open(os.path.join(wd, "intype"), "w") as f:
            f.write("N\n")
        def unlimit_stack():  # SPAL_LOSS keeps four (ID,JD,KD) work arrays on the stack
            resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
        p = subprocess.run([exe], input=deck_text, cwd=wd, capture_output=True, text=True,
                           timeout=timeout, preexec_fn=unlimit_stack)
        if p.returncode != 0 or not os.path.exists(os.path.join(wd, "final.bin")):
            raise RuntimeError(f"reference failed rc={p.returncode}\n{p.stdout[-3000:]}\n{p.stderr[-2000:]}")
        layout = parse_commall(params=VARIANT_PARAMS[variant])
        return Dump(os.path.join(wd, "setup.bin"), layout), Dump(os.path.join(wd, "final.bin"), layout), p.stdout
    finally:
        if own:
            shutil.rmtree(wd, ignore_errors=True)
