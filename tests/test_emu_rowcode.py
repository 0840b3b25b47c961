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
    lib.emu_create.argtypes = ([ctypes.c_int, ctypes.POINTER(ctypes.c_int)] + [ctypes.c_int] * 11
                               + [ctypes.POINTER(ctypes.c_int)])
    lib.emu_destroy.argtypes = [ctypes.c_void_p]
    lib.emu_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong),
                             ctypes.POINTER(ctypes.c_longlong)]
    lib.emu_violations.restype = ctypes.c_longlong
    lib.emu_violations.argtypes = [ctypes.c_void_p]
    for f in ("emu_Au", "emu_ATu"):
        getattr(lib, f).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.emu_row_norm2.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    return lib


# spin-adapted Q2 / G2 (SPIN_ADAPT_Q2 / _G2; singlets only, G2 with
# CONSTRAIN_SPIN off): amopi, na, nb, pos, sz, spin, sa_q2, sa_g2, energy order
SA_LAYOUTS = [
    ([5], 2, 2, "DQG", True, True, True, False, None),
    ([5], 2, 2, "DQG", True, True, True, False, [3, 0, 4, 2, 1]),
    ([5], 2, 2, "DQG", True, False, True, True, [1, 4, 0, 2, 3]),
    ([2, 1, 1, 1, 0, 1, 1, 1], 3, 3, "DQG", True, False, True, True, [0, 4, 1, 5, 2, 6, 3, 7]),
    ([3, 1, 2, 0], 2, 2, "DG", False, False, False, True, None),
    ([3, 0, 1, 1, 0, 2, 1, 1], 4, 4, "DQ", True, True, True, False, [8, 0, 3, 1, 2, 7, 4, 5, 6]),
]


def _make(lib, amopi, na, nb, mult, pos, sz, spin=True, sa_q2=False, sa_g2=False, eorder=None):
    a = (ctypes.c_int * 8)(*(amopi + [0] * (8 - len(amopi))))
    q2, g2, t1, t2 = orc.positivity_flags(pos)
    eo = (ctypes.c_int * len(eorder))(*eorder) if eorder else None
    h = lib.emu_create(len(amopi), a, na, nb, mult, q2, g2, t2, int(spin), int(sz), t1,
                       int(sa_q2), int(sa_g2), eo)
    assert h
    dx, nc = ctypes.c_longlong(), ctypes.c_longlong()
    lib.emu_dims(h, ctypes.byref(dx), ctypes.byref(nc))
    return h, dx.value, nc.value


def _check(emu, L, h, dimx, ncon):
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


@pytest.mark.parametrize("amopi,na,nb,mult,pos,sz", LAYOUTS)
def test_device_rowcode_matches_oracle(emu, amopi, na, nb, mult, pos, sz):
    h, dimx, ncon = _make(emu, amopi, na, nb, mult, pos, sz)
    _check(emu, orc.Layout(amopi, na, nb, mult, pos, constrain_sz=sz), h, dimx, ncon)


@pytest.mark.parametrize("amopi,na,nb,pos,sz,spin,saq,sag,eo", SA_LAYOUTS)
def test_device_rowcode_matches_oracle_spin_adapted(emu, amopi, na, nb, pos, sz, spin, saq, sag,
                                                    eo):
    h, dimx, ncon = _make(emu, amopi, na, nb, 1, pos, sz, spin, saq, sag, eo)
    L = orc.Layout(amopi, na, nb, 1, pos, constrain_spin=spin, constrain_sz=sz,
                   spin_adapt_q2=saq, spin_adapt_g2=sag, energy_order=eo)
    _check(emu, L, h, dimx, ncon)


def test_spin_adaptation_refusals(emu):
    """The layouts the reference refuses (v2rdm_solver.cc:538-561), and
    spin-adapted G2 beside the maximal-spin rows, which address G2ba."""
    a = (ctypes.c_int * 8)(5, 0, 0, 0, 0, 0, 0, 0)
    # na, nb, mult, q2, g2, t2, spin, sz, t1, sa_q2, sa_g2
    for args in [(2, 2, 1, 1, 1, 1, 1, 1, 0, 1, 0),      # T2
                 (2, 2, 1, 1, 1, 0, 1, 1, 1, 1, 0),      # T1
                 (3, 1, 3, 1, 1, 0, 1, 1, 0, 1, 0),      # triplet
                 (2, 2, 1, 1, 1, 0, 1, 1, 0, 0, 1)]:     # G2 with spin rows
        assert not emu.emu_create(1, a, *args, None)
    with pytest.raises(ValueError):
        orc.Layout([5], 2, 2, 1, "DQGT2", spin_adapt_q2=True)
    with pytest.raises(ValueError):
        orc.Layout([5], 2, 2, 1, "DQG", spin_adapt_g2=True)
