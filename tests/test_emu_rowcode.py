"""CPU check of the device row code (rowterms.hpp) through the host emulator.

The emulator runs the exact functions the HIP kernels run; every index they
form is bounds-checked, every row / primal entry must be produced exactly
once, and the results must match the oracle to 1e-12 (FP64 summation order
is the only difference).  This validates the kernels' index arithmetic on the
CPU before any GPU launch.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle_py as orc

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu", "build", "libv2g_emu.so")

LAYOUTS = [
    ([4], 2, 2, 1, "DQG", True),
    ([5], 3, 2, 2, "DQG", True),
    ([6], 4, 2, 3, "DQG", True),
    ([2, 1, 1, 0], 2, 2, 1, "DQG", True),
    ([2, 1, 1, 1, 0, 1, 1, 1], 3, 3, 1, "DQG", True),
    ([4], 2, 2, 1, "DQG", False),
    ([5], 2, 2, 1, "DQ", True),
    ([5], 2, 2, 1, "DG", True),
    ([5], 2, 2, 1, "D", True),
    ([5], 3, 2, 2, "DQGT2", True),
    ([2, 1, 1, 1], 2, 2, 1, "DQGT2", True),
    ([3, 0, 1, 1, 0, 2, 1, 1], 4, 3, 2, "DQGT2", True),
    ([9], 4, 4, 1, "DQG", True),
    ([5, 0, 0, 0, 0, 5, 0, 0], 5, 5, 1, "DQG", True),
    ([5], 3, 2, 2, "DQGT1", True),
    ([6], 3, 3, 1, "DQGT1", True),
    ([3, 0, 1, 1, 0, 2, 1, 1], 4, 3, 2, "DQGT1", True),
    ([2, 1, 1, 1], 3, 2, 2, "DQGT1T2", True),
    ([3, 2], 3, 3, 1, "DQGT", False),
]


@pytest.fixture(scope="module")
def emu():
    if not os.path.exists(EMU):
        subprocess.check_call(["make", "-C", os.path.join(HERE, "emu")])
    lib = ctypes.CDLL(EMU)
    lib.emu_create.restype = ctypes.c_void_p
    lib.emu_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)] + [ctypes.c_int] * 9
    lib.emu_destroy.argtypes = [ctypes.c_void_p]
    lib.emu_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong),
                             ctypes.POINTER(ctypes.c_longlong)]
    lib.emu_violations.restype = ctypes.c_longlong
    lib.emu_violations.argtypes = [ctypes.c_void_p]
    for f in ("emu_Au", "emu_ATu"):
        getattr(lib, f).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.emu_row_norm2.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    return lib


def _make(lib, amopi, na, nb, mult, pos, sz):
    a = (ctypes.c_int * 8)(*(amopi + [0] * (8 - len(amopi))))
    q2, g2, t1, t2 = orc.positivity_flags(pos)
    h = lib.emu_create(len(amopi), a, na, nb, mult, q2, g2, t2, 1, int(sz), t1)
    assert h
    dx, nc = ctypes.c_longlong(), ctypes.c_longlong()
    lib.emu_dims(h, ctypes.byref(dx), ctypes.byref(nc))
    return h, dx.value, nc.value


@pytest.mark.parametrize("amopi,na,nb,mult,pos,sz", LAYOUTS)
def test_device_rowcode_matches_oracle(emu, amopi, na, nb, mult, pos, sz):
    h, dimx, ncon = _make(emu, amopi, na, nb, mult, pos, sz)
    L = orc.Layout(amopi, na, nb, mult, pos, constrain_sz=sz)
    assert (dimx, ncon) == (L.dimx, L.ncon)
    rng = np.random.default_rng(11)
    x = rng.normal(size=dimx)
    ax = np.full(ncon, np.nan)
    emu.emu_Au(h, x.ctypes.data, ax.ctypes.data)
    guarded = np.full(3 * ncon, np.nan)
    y = rng.normal(size=ncon)
    guarded[ncon:2 * ncon] = y
    aty = np.full(dimx, np.nan)
    emu.emu_ATu(h, guarded[ncon:].ctypes.data, aty.ctypes.data)
    d = np.full(ncon, np.nan)
    emu.emu_row_norm2(h, d.ctypes.data)
    assert emu.emu_violations(h) == 0
    ref = L.Au(x)
    assert np.linalg.norm(ax - ref) <= 1e-12 * np.linalg.norm(ref)
    ref = L.ATu(y)
    assert np.all(np.isfinite(aty))
    assert np.linalg.norm(aty - ref) <= 1e-12 * np.linalg.norm(ref)
    if dimx <= 6000:
        A = np.stack([L.Au(e) for e in np.eye(dimx)], axis=1)
        assert np.allclose(d, (A * A).sum(axis=1), rtol=1e-13, atol=1e-13)
    emu.emu_destroy(h)
