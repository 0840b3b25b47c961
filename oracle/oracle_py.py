"""ctypes loader for the CPU oracle (test infrastructure only).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
The product package never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle_v2rdm.so")

OFF = dict(D2AB=0, D2AA=1, D2BB=2, D200=3, D1A=4, D1B=5, Q1A=6, Q1B=7, Q2AB=8, Q2AA=9,
           Q2BB=10, G2AB=11, G2BA=12, G2AA=13, T2AAA=14, T2BBB=15, T2AAB=16, T2BBA=17,
           T1AAA=18, T1BBB=19, T1AAB=20, T1BBA=21, Q2S=22, Q2T=23, Q2TP1=24, Q2TM1=25,
           G2S=26, G2T=27, G2TP1=28, G2TM1=29)


class _Options(ctypes.Structure):
    _fields_ = [("nirrep", ctypes.c_int), ("amopi", ctypes.c_int * 8),
                ("nalpha", ctypes.c_int), ("nbeta", ctypes.c_int),
                ("multiplicity", ctypes.c_int), ("constrain_q2", ctypes.c_int),
                ("constrain_g2", ctypes.c_int), ("constrain_t2", ctypes.c_int),
                ("constrain_spin", ctypes.c_int), ("constrain_sz", ctypes.c_int),
                ("constrain_t1", ctypes.c_int), ("spin_adapt_q2", ctypes.c_int),
                ("spin_adapt_g2", ctypes.c_int), ("energy_order", ctypes.POINTER(ctypes.c_int))]


def positivity_flags(positivity):
    """(q2, g2, t1, t2) of a POSITIVITY string, as plugin.cc / v2rdm_solver.cc:488-523."""
    p = positivity.upper()
    if p == "DQGT":
        return 1, 1, 1, 1
    return int("Q" in p), int("G" in p), int("T1" in p), int("T2" in p)


class _Params(ctypes.Structure):
    _fields_ = [("tau", ctypes.c_double), ("r_convergence", ctypes.c_double),
                ("e_convergence", ctypes.c_double), ("cg_convergence", ctypes.c_double),
                ("maxiter", ctypes.c_int), ("cg_maxiter", ctypes.c_int),
                ("mu_update_frequency", ctypes.c_int), ("verbose", ctypes.c_int),
                ("mu0", ctypes.c_double), ("oiter0", ctypes.c_int), ("ep0", ctypes.c_double),
                ("ed0", ctypes.c_double)]


class _Stats(ctypes.Structure):
    _fields_ = [("oiter", ctypes.c_int), ("cg_iters", ctypes.c_long),
                ("converged", ctypes.c_int), ("energy_primal", ctypes.c_double),
                ("energy_dual", ctypes.c_double), ("ep", ctypes.c_double),
                ("ed", ctypes.c_double), ("mu", ctypes.c_double), ("t_cg", ctypes.c_double),
                ("t_xz", ctypes.c_double), ("t_total", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, ip, fp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), \
            ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)
        L.orc_layout_create.restype = vp
        L.orc_layout_create.argtypes = [ctypes.POINTER(_Options)]
        L.orc_layout_destroy.argtypes = [vp]
        for f in ("orc_dimx", "orc_ncon"):
            getattr(L, f).restype = ctypes.c_long
            getattr(L, f).argtypes = [vp]
        L.orc_nblocks.restype = ctypes.c_int
        L.orc_nblocks.argtypes = [vp]
        L.orc_block_dims.argtypes = [vp, ip]
        L.orc_offset.restype = ctypes.c_long
        L.orc_offset.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        for f in ("orc_gems_ab", "orc_gems_aa", "orc_trip_aab", "orc_trip_aba", "orc_trip_aaa",
                  "orc_orbital_irrep", "orc_gems_00"):
            getattr(L, f).restype = ctypes.c_int
            getattr(L, f).argtypes = [vp, ctypes.c_int]
        for f in ("orc_ibas_ab", "orc_ibas_aa"):
            getattr(L, f).restype = ctypes.c_int
            getattr(L, f).argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.orc_bas_00.argtypes = [vp, ctypes.c_int, ip]
        for f in ("orc_spatial_d1_dim", "orc_spatial_d2_dim"):
            getattr(L, f).restype = ctypes.c_long
            getattr(L, f).argtypes = [vp]
        L.orc_pack_spatial_density.argtypes = [vp, dp, dp, dp]
        L.orc_row_families.restype = ctypes.c_int
        L.orc_row_families.argtypes = [vp, ctypes.c_int, ctypes.c_char_p,
                                       ctypes.POINTER(ctypes.c_long),
                                       ctypes.POINTER(ctypes.c_long)]
        L.orc_Au.argtypes = [vp, dp, dp]
        L.orc_ATu.argtypes = [vp, dp, dp]
        L.orc_cg_Ax.argtypes = [vp, dp, dp, dp]
        L.orc_cg_solve.restype = ctypes.c_int
        L.orc_cg_solve.argtypes = [vp, dp, dp, ctypes.c_double, ctypes.c_int, dp]
        L.orc_update_xz.argtypes = [vp, ctypes.c_double, dp, dp, dp, dp]
        L.orc_psd_split.argtypes = [ctypes.c_int, dp, ctypes.c_double, dp, dp]
        L.orc_build_b.argtypes = [vp, dp]
        L.orc_build_c.argtypes = [vp, dp, dp, dp]
        L.orc_guess_random.argtypes = [vp, dp, dp, dp]
        L.orc_bpsdp.restype = ctypes.c_int
        L.orc_bpsdp.argtypes = [vp, ctypes.POINTER(_Params), dp, dp, dp, dp, dp,
                                ctypes.c_double, ctypes.POINTER(_Stats)]
        L.orc_lapack_backend.restype = ctypes.c_int
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_mlp_forward.argtypes = [ctypes.c_int] * 3 + [fp] * 6 + [ctypes.c_float, fp]
        L.orc_mlp_loss_grad.restype = ctypes.c_double
        L.orc_mlp_loss_grad.argtypes = [ctypes.c_int] * 3 + [fp] * 8 + [ctypes.c_float] + [dp] * 6
        _lib = L
    return _lib


def _d(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Layout:
    """Oracle layout: geminal tables, x offsets, constraint rows."""

    def __init__(self, amopi, nalpha, nbeta, multiplicity=None, positivity="DQG",
                 constrain_spin=True, constrain_sz=True, spin_adapt_q2=False,
                 spin_adapt_g2=False, energy_order=None):
        L = lib()
        amopi = list(amopi)
        nirrep = len(amopi)
        o = _Options()
        o.nirrep = nirrep
        for h in range(8):
            o.amopi[h] = amopi[h] if h < nirrep else 0
        o.nalpha, o.nbeta = nalpha, nbeta
        o.multiplicity = multiplicity if multiplicity else (nalpha - nbeta + 1)
        p = positivity.upper()
        o.constrain_q2, o.constrain_g2, o.constrain_t1, o.constrain_t2 = positivity_flags(p)
        o.constrain_spin = int(constrain_spin)
        o.constrain_sz = int(constrain_sz)
        o.spin_adapt_q2 = int(bool(spin_adapt_q2))
        o.spin_adapt_g2 = int(bool(spin_adapt_g2))
        amo = sum(amopi)
        self._eorder = None
        if energy_order is not None:
            eo = [int(v) for v in energy_order]
            if sorted(eo) != list(range(amo)):
                raise ValueError("energy_order must be a permutation of the active orbitals")
            self._eorder = (ctypes.c_int * amo)(*eo)
            o.energy_order = ctypes.cast(self._eorder, ctypes.POINTER(ctypes.c_int))
        if (spin_adapt_q2 or spin_adapt_g2) and (o.constrain_t1 or o.constrain_t2):
            raise ValueError("If constraining T1/T2, G2/Q2 cannot currently be spin adapted.")
        if (spin_adapt_q2 or spin_adapt_g2) and o.multiplicity > 1:
            raise ValueError("Q2 / G2 not spin adapted for S = M != 0")
        if spin_adapt_g2 and o.constrain_g2 and constrain_spin:
            raise ValueError("SPIN_ADAPT_G2 needs CONSTRAIN_SPIN false: the maximal-spin rows "
                             "address the G2ba block, which spin adaptation replaces")
        self.opts = o
        self.amopi = amopi
        self.nirrep = nirrep
        self.amo = sum(amopi)
        self.h = L.orc_layout_create(ctypes.byref(o))
        if not self.h:
            raise ValueError("layout creation failed")
        self.dimx = L.orc_dimx(self.h)
        self.ncon = L.orc_ncon(self.h)
        nb = L.orc_nblocks(self.h)
        dims = (ctypes.c_int * nb)()
        L.orc_block_dims(self.h, dims)
        self.block_dims = list(dims)

    def __del__(self):
        try:
            if self.h:
                lib().orc_layout_destroy(self.h)
        except Exception:
            pass

    def offset(self, name, h):
        return lib().orc_offset(self.h, OFF[name], h)

    def gems_ab(self, h):
        return lib().orc_gems_ab(self.h, h)

    def gems_aa(self, h):
        return lib().orc_gems_aa(self.h, h)

    def gems_00(self, h):
        return lib().orc_gems_00(self.h, h)

    def bas_00(self, h):
        """singlet geminals (i, j) of irrep h in their order"""
        n = lib().orc_gems_00(self.h, h)
        buf = (ctypes.c_int * max(1, 2 * n))()
        lib().orc_bas_00(self.h, h, buf)
        return [(buf[2 * a], buf[2 * a + 1]) for a in range(n)]

    def pack_spatial_density(self, x):
        """(d1, d2) packed as PackSpatialDensity (v2rdm_solver.cc:3664-3773)"""
        d1 = np.zeros(lib().orc_spatial_d1_dim(self.h))
        d2 = np.zeros(lib().orc_spatial_d2_dim(self.h))
        lib().orc_pack_spatial_density(self.h, _d(np.ascontiguousarray(x, dtype=np.float64)),
                                       _d(d1), _d(d2))
        return d1, d2

    def trip_aab(self, h):
        return lib().orc_trip_aab(self.h, h)

    def trip_aba(self, h):
        return lib().orc_trip_aba(self.h, h)

    def trip_aaa(self, h):
        return lib().orc_trip_aaa(self.h, h)

    def ibas_ab(self, p, q):
        return lib().orc_ibas_ab(self.h, p, q)

    def ibas_aa(self, p, q):
        return lib().orc_ibas_aa(self.h, p, q)

    def irrep(self, p):
        return lib().orc_orbital_irrep(self.h, p)

    def row_families(self):
        names = ctypes.create_string_buffer(32 * 24)
        start = (ctypes.c_long * 32)()
        count = (ctypes.c_long * 32)()
        n = lib().orc_row_families(self.h, 32, names, start, count)
        out = []
        for f in range(n):
            nm = names.raw[24 * f:24 * (f + 1)].split(b"\0")[0].decode()
            out.append((nm, start[f], count[f]))
        return out

    def Au(self, x):
        out = np.empty(self.ncon)
        lib().orc_Au(self.h, _d(np.ascontiguousarray(x, dtype=np.float64)), _d(out))
        return out

    def ATu(self, y):
        out = np.empty(self.dimx)
        lib().orc_ATu(self.h, _d(np.ascontiguousarray(y, dtype=np.float64)), _d(out))
        return out

    def cg_Ax(self, u):
        out = np.empty(self.ncon)
        scr = np.empty(self.dimx)
        lib().orc_cg_Ax(self.h, _d(np.ascontiguousarray(u, dtype=np.float64)), _d(out), _d(scr))
        return out

    def cg_solve(self, rhs, y0, tol, maxiter):
        y = np.array(y0, dtype=np.float64, copy=True)
        it = lib().orc_cg_solve(self.h, _d(np.ascontiguousarray(rhs)), _d(y), tol, maxiter, None)
        return y, it

    def update_xz(self, mu, y, c, x):
        x = np.array(x, dtype=np.float64, copy=True)
        z = np.zeros(self.dimx)
        lib().orc_update_xz(self.h, mu, _d(np.ascontiguousarray(y)), _d(np.ascontiguousarray(c)),
                            _d(x), _d(z))
        return x, z

    def build_b(self):
        b = np.empty(self.ncon)
        lib().orc_build_b(self.h, _d(b))
        return b

    def build_c(self, h1, eri):
        c = np.empty(self.dimx)
        lib().orc_build_c(self.h, _d(np.ascontiguousarray(h1, dtype=np.float64)),
                          _d(np.ascontiguousarray(eri, dtype=np.float64)), _d(c))
        return c

    def guess_random(self):
        x = np.empty(self.dimx)
        z = np.empty(self.dimx)
        y = np.empty(self.ncon)
        lib().orc_guess_random(self.h, _d(x), _d(z), _d(y))
        return x, z, y

    def bpsdp(self, c, b, x, z, y, r_conv=1e-6, e_conv=1e-6, cg_conv=1e-9, maxiter=10000,
              cg_maxiter=10000, tau=1.0, mu_update_frequency=500, verbose=False,
              max_seconds=0.0, resume=None):
        """The reference loop; resume = dict(mu, oiter, ep, ed) of an earlier
        run continues it (x, z, y passed in hold that run's state), and maxiter
        counts from the start of the whole solve, as in the reference."""
        p = _Params(tau, r_conv, e_conv, cg_conv, maxiter, cg_maxiter, mu_update_frequency,
                    int(verbose))
        if resume:
            p.mu0, p.oiter0 = float(resume["mu"]), int(resume["oiter"])
            p.ep0, p.ed0 = float(resume["ep"]), float(resume["ed"])
        st = _Stats()
        lib().orc_bpsdp(self.h, ctypes.byref(p), _d(c), _d(b), _d(x), _d(z), _d(y),
                        float(max_seconds), ctypes.byref(st))
        return {k: getattr(st, k) for k, _ in _Stats._fields_}


def psd_split(M, mu):
    n = M.shape[0]
    M = np.ascontiguousarray(M, dtype=np.float64)
    xb = np.empty_like(M)
    zb = np.empty_like(M)
    lib().orc_psd_split(n, _d(M), mu, _d(xb), _d(zb))
    return xb, zb


def lapack_backend():
    return lib().orc_lapack_backend()


def num_threads():
    return lib().orc_num_threads()


def mlp_forward(feat, W1, b1, W2, b2, w3, b3):
    n, nin = feat.shape
    nh = W1.shape[1]
    out = np.empty(n, dtype=np.float32)
    lib().orc_mlp_forward(n, nin, nh, _f(feat), _f(W1), _f(b1), _f(W2), _f(b2), _f(w3),
                          float(b3), _f(out))
    return out


def mlp_loss_grad(feat, target, wt, W1, b1, W2, b2, w3, b3):
    n, nin = feat.shape
    nh = W1.shape[1]
    g = [np.zeros(nin * nh), np.zeros(nh), np.zeros(nh * nh), np.zeros(nh), np.zeros(nh),
         np.zeros(1)]
    tp = _f(target) if target is not None else ctypes.POINTER(ctypes.c_float)()
    loss = lib().orc_mlp_loss_grad(n, nin, nh, _f(feat), tp, _f(wt), _f(W1), _f(b1),
                                   _f(W2), _f(b2), _f(w3), float(b3), *[_d(a) for a in g])
    return loss, g
