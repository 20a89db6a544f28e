"""mh_set_xlength_noalloc, the allocation-free form of mh_set_xlength (its
work buffers in the library's per-device scratch array): bit-identical
wall distances and mixing lengths to mh_set_xlength on two calls in a row,
the second on a smaller grid (so it runs on scratch memory the first call
left behind), and a clear error for a grid beyond the array's size."""
import ctypes
import os

import numpy as np
import pytest

from conftest import GOLDEN

DECKS = ["ref/simple-axial-compr-17.1.dat", "tiny3_ml"]      # 303,918 nodes, then 9,963


def _inputs(deck):
    from multall_hip import host
    path = os.path.join(GOLDEN, "decks", deck if deck.endswith(".dat") else deck + ".dat")
    st = host.load_deck(path)
    sc = st["scalars"]
    km, jm, im, nr = sc["KM"], sc["JM"], sc["IM"], sc["NROWS"]
    theta = host.array("THETA", (km, jm, im))
    xl = [host.array(n, (nr,)) for n in ("XLLIM_IN", "XLLIM_LE", "XLLIM_TE", "XLLIM_DN")]
    for n in ("DWALLSQ", "XLENGTH", "FST_RAT"):
        st[n] = np.full_like(st[n], -1.0)       # both forms must write every value
    return st, theta, xl


@pytest.mark.gpu
def test_noalloc_equals_allocating_form_twice():
    from multall_hip.runtime import Solver
    for deck in DECKS:
        st, theta, xl = _inputs(deck)
        print(f"{deck}: IM={st['scalars']['IM']} JM={st['scalars']['JM']} KM={st['scalars']['KM']}", flush=True)
        a = Solver(st, exact_order=True)
        a.set_xlength(theta, *xl)
        b = Solver(st, exact_order=True)
        b.set_xlength_noalloc(theta, *xl)
        for n in ("DWALLSQ", "XLENGTH", "FST_RAT"):
            x, y = a.get(n), b.get(n)
            assert not np.any(y == -1.0), (deck, n)
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (deck, n)
        a.close()
        b.close()


@pytest.mark.gpu
def test_noalloc_rejects_a_grid_beyond_the_scratch():
    from multall_hip.runtime import MhDims, load
    L = load()
    dims = MhDims(im=65, jm=1000, km=65, nrows=1, nib1=22, njb1=333, nkb1=22, nib2=8, njb2=111, nkb2=8, nsblk=4)
    n = dims.im * dims.jm * dims.km
    assert n > 4194304
    h = ctypes.c_void_p()
    assert L.mh_create(ctypes.byref(dims), ctypes.byref(h)) == 0
    try:
        theta = np.zeros(n, np.float32)
        row = np.zeros(1, np.float32)
        rc = L.mh_set_xlength_noalloc(h, theta.ctypes.data, row.ctypes.data, row.ctypes.data,
                                      row.ctypes.data, row.ctypes.data)
        msg = L.mh_last_error(h).decode()
        print(msg)
        assert rc == -1 and "4194304" in msg, (rc, msg)
    finally:
        L.mh_destroy(h)
