"""The C-ABI library loads and exports every entry point its header declares
(no GPU needed: nothing is called)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^[\w \*]*?\b(mh_\w+)\s*\([^;{]*\);", txt, flags=re.M)))


def test_library_exports_header():
    from multall_hip import runtime
    lib = ctypes.CDLL(runtime.LIB_PATH)
    names = _declared("multall_hip.h")
    assert len(names) >= 20, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(runtime.EXPORTS) == names


def test_array_and_scalar_ids_cover_header():
    from multall_hip import runtime
    lib = runtime.load()
    txt = open(os.path.join(ROOT, "include", "multall_hip.h")).read()
    arrays = re.findall(r"\bMH_([A-Z0-9_]+)\s*(?:=\s*0)?\s*[,}]", txt.split("enum mh_array_id")[1].split("};")[0])
    arrays = [a for a in arrays if a != "NARRAYS"]
    for i, a in enumerate(arrays):
        assert lib.mh_array_id(a.encode()) == i, a
    scal = re.findall(r"\bMH_S_([A-Z0-9_]+)", txt.split("enum mh_scalar_id")[1].split("};")[0])
    for i, n in enumerate(scal):
        assert lib.mh_scalar_id(n.encode()) == i, n
    assert lib.mh_array_id(b"NOT_AN_ARRAY") == -1
