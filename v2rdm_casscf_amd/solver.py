"""Host-side mirror of the reference's v2RDMSolver hot path over the v2g C ABI.

Method names and argument meaning follow the reference (v2rdm_solver.h:85,
366-417; cg_solver.h:49-53):

  bpsdp_Au(A, u)     -> A = A(u)                  v2rdm_solver.cc:3020-3065
  bpsdp_ATu(A, u)    -> A = A^T(u)                v2rdm_solver.cc:3114-3158
  cg_Ax(N, A, u)     -> A = A(A^T u)              v2rdm_solver.cc:3207-3213
  cg_solve(...)      -> CGSolver::solve           cg_solver.cc:125-177
  Update_xz()        -> per-block eigen split     v2rdm_solver.cc:3216-3289
  BuildConstraints() -> b                         v2rdm_solver.cc:2518-3017
  GetIntegrals(h,eri)-> c (RepackIntegrals)       tei.cc:144-313
  Guess()            -> RANDOM guess              v2rdm_solver.cc:2386-2490
  compute_energy()   -> BPSDP outer loop          v2rdm_solver.cc:1685-1912

Vectors are torch float64 tensors on the HIP device (torch is used only to
hold device memory).  Errors raise V2GError with the library's message, the
analogue of the reference's PsiException.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check

# POSITIVITY -> (q2, g2, t1, t2), v2rdm_solver.cc:495-523
POSITIVITY = {"D": (0, 0, 0, 0), "DQ": (1, 0, 0, 0), "DG": (0, 1, 0, 0), "DQG": (1, 1, 0, 0),
              "DQGT1": (1, 1, 1, 0), "DQGT2": (1, 1, 0, 1), "DQGT1T2": (1, 1, 1, 1),
              "DQGT": (1, 1, 1, 1)}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _np_ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def fill_desc(d, amopi, nalpha, nbeta, multiplicity=None, positivity="DQG", constrain_spin=True,
              constrain_sz=True, spin_adapt_q2=False, spin_adapt_g2=False, energy_order=None):
    """Fills a v2g_layout_desc; returns the energy-order array the
    descriptor points into (keep it alive as long as the descriptor).

    spin_adapt_q2 / spin_adapt_g2 are the reference's SPIN_ADAPT_Q2 /
    SPIN_ADAPT_G2 (plugin.cc:108-110); energy_order lists the active orbitals
    (Pitzer index) from lowest to highest orbital energy, the order of the
    singlet geminals of spin-adapted Q2 (basis.cc:626-651), None = Pitzer."""
    pos = positivity.upper()
    if pos not in POSITIVITY:
        raise ValueError(f"POSITIVITY {positivity} is not supported ({', '.join(POSITIVITY)})")
    amopi = list(amopi)
    if len(amopi) not in (1, 2, 4, 8):
        raise ValueError("amopi must have 1, 2, 4 or 8 irreps")
    d.nirrep = len(amopi)
    for h in range(8):
        d.amopi[h] = amopi[h] if h < len(amopi) else 0
    d.nalpha, d.nbeta = int(nalpha), int(nbeta)
    d.multiplicity = int(multiplicity) if multiplicity else int(nalpha - nbeta + 1)
    d.constrain_q2, d.constrain_g2, d.constrain_t1, d.constrain_t2 = POSITIVITY[pos]
    d.constrain_spin, d.constrain_sz = int(constrain_spin), int(constrain_sz)
    d.spin_adapt_q2, d.spin_adapt_g2 = int(bool(spin_adapt_q2)), int(bool(spin_adapt_g2))
    keep = None
    if energy_order is not None:
        eo = [int(v) for v in energy_order]
        if sorted(eo) != list(range(sum(amopi))):
            raise ValueError("energy_order must be a permutation of the active orbitals")
        keep = (ctypes.c_int * len(eo))(*eo)
        d.energy_order = ctypes.cast(keep, ctypes.POINTER(ctypes.c_int))
    return keep


class V2RDMSolver:
    """One active space / constraint set on one GPU (`device`)."""

    def __init__(self, amopi, nalpha, nbeta, multiplicity=None, positivity="DQG",
                 constrain_spin=True, constrain_sz=True, device="cuda:0", options=None,
                 spin_adapt_q2=False, spin_adapt_g2=False, energy_order=None):
        pos = positivity.upper()
        amopi = list(amopi)
        self.lib = _lib.load()
        self.device = torch.device(device)
        d = _lib.LayoutDesc()
        self._eorder = fill_desc(d, amopi, nalpha, nbeta, multiplicity, pos, constrain_spin,
                                 constrain_sz, spin_adapt_q2, spin_adapt_g2, energy_order)
        self.desc = d
        self.amopi, self.nalpha, self.nbeta = amopi, nalpha, nbeta
        self.positivity = pos
        with torch.cuda.device(self.device):
            h = ctypes.c_void_p()
            check(self.lib.v2g_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        dx, nc, nb = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_int()
        check(self.lib.v2g_dims(self.h, ctypes.byref(dx), ctypes.byref(nc), ctypes.byref(nb)))
        self.dimx, self.nconstraints = dx.value, nc.value
        dims = (ctypes.c_int * nb.value)()
        check(self.lib.v2g_block_dims(self.h, dims))
        self.dimensions = list(dims)
        # reference option defaults (plugin.cc:97-124), E/R_CONVERGENCE as the
        # configurations here use them
        self.options = dict(TAU_PARAMETER=1.0, R_CONVERGENCE=1e-6, E_CONVERGENCE=1e-6,
                            CG_CONVERGENCE=1e-9, MAXITER=10000, CG_MAXITER=10000,
                            MU_UPDATE_FREQUENCY=500, PRECONDITIONER="NONE", FLEXIBLE_CG=False,
                            MLP_TRAIN_ITERS=30, MLP_TRAIN_BATCH=16384, MLP_TRAIN_STEPS=10,
                            MLP_TRAIN_LR=2e-3, EIG_WARM_SWEEPS=1, EIG_INNER_SWEEPS=1,
                            CG_EXTRAPOLATE=True, MLP_BASE="IDENTITY", EIG_PAIR_ORDER=1,
                            MLP_TRAIN_PENALTY=0.0)
        if options:
            self.options.update({k.upper(): v for k, v in options.items()})
        self.x = self.z = self.y = self.c = self.b = None
        self.mu = 1.0
        self.last_stats = None
        self._cb = None

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.v2g_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # ---- allocation helpers --------------------------------------------
    def _vec(self, n):
        return torch.empty(n, dtype=torch.float64, device=self.device)

    def _check_vec(self, t, n, name):
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() == n):
            raise ValueError(f"{name}: expected a contiguous float64 device tensor of {n} entries")

    def _sync(self):
        torch.cuda.synchronize(self.device)

    # ---- maps ----------------------------------------------------------
    def bpsdp_Au(self, A, u):
        self._check_vec(u, self.dimx, "u")
        self._check_vec(A, self.nconstraints, "A")
        self._sync()
        check(self.lib.v2g_Au(self.h, _ptr(u), _ptr(A)))
        return A

    def bpsdp_ATu(self, A, u):
        self._check_vec(u, self.nconstraints, "u")
        self._check_vec(A, self.dimx, "A")
        self._sync()
        check(self.lib.v2g_ATu(self.h, _ptr(u), _ptr(A)))
        return A

    def cg_Ax(self, N, A, u):
        if N != self.nconstraints:
            raise ValueError("dimension does not match dimension from initialization")
        self._check_vec(u, N, "u")
        self._check_vec(A, N, "A")
        self._sync()
        check(self.lib.v2g_cg_Ax(self.h, _ptr(u), _ptr(A)))
        return A

    def row_norm2(self):
        d = self._vec(self.nconstraints)
        self._sync()
        check(self.lib.v2g_row_norm2(self.h, _ptr(d)))
        return d

    # ---- CG ------------------------------------------------------------
    def cg_solve(self, rhs, y, tol, maxiter, precond=0, flexible=False):
        """CGSolver::solve on A A^T; y updated in place; returns (iters, ||r||)."""
        self._check_vec(rhs, self.nconstraints, "rhs")
        self._check_vec(y, self.nconstraints, "y")
        self._sync()
        it = ctypes.c_int()
        res = ctypes.c_double()
        check(self.lib.v2g_cg_solve(self.h, _ptr(rhs), _ptr(y), float(tol), int(maxiter),
                                    int(precond), int(bool(flexible)), ctypes.byref(it),
                                    ctypes.byref(res)))
        return it.value, res.value

    def set_precond_diag(self, minv):
        self._check_vec(minv, self.nconstraints, "minv")
        self._sync()
        check(self.lib.v2g_set_precond_diag(self.h, _ptr(minv)))

    # ---- learned preconditioner (no reference counterpart) ---------------
    def mlp_num_params(self):
        return self.lib.v2g_mlp_num_params()

    def mlp_init(self, seed=12345):
        check(self.lib.v2g_mlp_init(self.h, int(seed)))

    def mlp_weights(self):
        w = np.empty(self.mlp_num_params(), dtype=np.float32)
        check(self.lib.v2g_mlp_get_weights(self.h, _np_ptr(w)))
        return w

    def mlp_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        check(self.lib.v2g_mlp_set_weights(self.h, _np_ptr(w)))

    def mlp_loss_grad(self, feat, tgt, wt):
        """feat: (nb, 14) float32, tgt/wt: (nb,) float32 device tensors; with
        tgt None, wt is the upstream gradient d(loss)/d(out) per sample."""
        nb = feat.shape[0]
        g = torch.empty(self.mlp_num_params(), dtype=torch.float64, device=self.device)
        loss = ctypes.c_double()
        feat, wt = feat.contiguous(), wt.contiguous()
        tp = _ptr(tgt.contiguous()) if tgt is not None else None
        self._sync()
        check(self.lib.v2g_mlp_loss_grad(self.h, int(nb), _ptr(feat), tp, _ptr(wt), _ptr(g),
                                         ctypes.byref(loss)))
        return loss.value, g

    def mlp_spectral_grad(self, rhs, y, r_hi, r_lo):
        """Per-row d L / d g of the spectral training objective and L."""
        out = torch.empty(self.nconstraints, dtype=torch.float32, device=self.device)
        obj = ctypes.c_double()
        self._sync()
        check(self.lib.v2g_mlp_spectral_grad(self.h, _ptr(rhs), _ptr(y), _ptr(r_hi), _ptr(r_lo),
                                             _ptr(out), ctypes.byref(obj)))
        return out, obj.value

    def mlp_adam_step(self, grad, lr):
        check(self.lib.v2g_mlp_adam_step(self.h, _ptr(grad), float(lr)))

    def mlp_minv(self, rhs, y):
        out = self._vec(self.nconstraints)
        self._sync()
        check(self.lib.v2g_mlp_minv(self.h, _ptr(rhs), _ptr(y), _ptr(out)))
        return out

    # ---- Update_xz ------------------------------------------------------
    def update_xz(self, mu, y, c, x, z):
        for t, n, nm in ((y, self.nconstraints, "y"), (c, self.dimx, "c"), (x, self.dimx, "x"),
                         (z, self.dimx, "z")):
            self._check_vec(t, n, nm)
        self._sync()
        check(self.lib.v2g_update_xz(self.h, float(mu), _ptr(y), _ptr(c), _ptr(x), _ptr(z)))
        return x, z

    def Update_xz(self):
        return self.update_xz(self.mu, self.y, self.c, self.x, self.z)

    def reset_eig(self):
        check(self.lib.v2g_reset_eig(self.h))

    # ---- launch shape of the eigen split (no reference counterpart) ------
    def set_eig_shape(self, lanes=0, pair_threads=0):
        """lanes 1/2, pair_threads 256/1024 (0 keeps the current value)."""
        check(self.lib.v2g_set_eig_shape(self.h, int(lanes), int(pair_threads)))

    def eig_shape(self):
        lanes, pt = ctypes.c_int(), ctypes.c_int()
        check(self.lib.v2g_get_eig_shape(self.h, ctypes.byref(lanes), ctypes.byref(pt)))
        return lanes.value, pt.value

    def autotune(self, on=True):
        """Time every launch shape on the next warm splits and keep the fastest."""
        check(self.lib.v2g_set_autotune(self.h, int(bool(on))))

    # ---- path inputs ----------------------------------------------------
    # Built on the device by default (bit-identical to the host builders,
    # which stay available with on_device=False)
    def BuildConstraints(self, on_device=True):
        if on_device:
            self.b = self._vec(self.nconstraints)
            self._sync()
            check(self.lib.v2g_build_b_device(self.h, _ptr(self.b)))
            return self.b
        b = np.empty(self.nconstraints)
        check(self.lib.v2g_build_b(self.h, _np_ptr(b)))
        self.b = torch.from_numpy(b).to(self.device)
        return self.b

    def GetIntegrals(self, h1, eri, on_device=True):
        """c from active-space integrals (h1: amo x amo, eri: amo^4 chemists')."""
        amo = sum(self.amopi)
        h1 = np.ascontiguousarray(h1, dtype=np.float64).reshape(amo, amo)
        eri = np.ascontiguousarray(eri, dtype=np.float64).reshape(amo, amo, amo, amo)
        if on_device:
            dh1 = torch.from_numpy(h1).to(self.device)
            deri = torch.from_numpy(eri).to(self.device)
            self.c = self._vec(self.dimx)
            self._sync()
            check(self.lib.v2g_build_c_device(self.h, _ptr(dh1), _ptr(deri), _ptr(self.c)))
            return self.c
        c = np.empty(self.dimx)
        check(self.lib.v2g_build_c(self.h, _np_ptr(h1), _np_ptr(eri), _np_ptr(c)))
        self.c = torch.from_numpy(c).to(self.device)
        return self.c

    def Guess(self, on_device=True):
        if on_device:
            self.x, self.z = self._vec(self.dimx), self._vec(self.dimx)
            self.y = self._vec(self.nconstraints)
            self._sync()
            check(self.lib.v2g_guess_random_device(self.h, _ptr(self.x), _ptr(self.z),
                                                   _ptr(self.y)))
            return self.x, self.z, self.y
        x = np.empty(self.dimx)
        z = np.empty(self.dimx)
        y = np.empty(self.nconstraints)
        check(self.lib.v2g_guess_random(self.h, _np_ptr(x), _np_ptr(z), _np_ptr(y)))
        self.x = torch.from_numpy(x).to(self.device)
        self.z = torch.from_numpy(z).to(self.device)
        self.y = torch.from_numpy(y).to(self.device)
        return self.x, self.z, self.y

    # ---- RDM outputs and checkpoint ---------------------------------------
    def get_opdm(self):
        """v2RDMSolver::get_opdm (python_api.cc:376-396): list of per-irrep
        amopi[h] x amopi[h] numpy arrays, D1a + D1b."""
        out = np.empty(sum(n * n for n in self.amopi))
        self._sync()
        check(self.lib.v2g_get_opdm(self.h, _ptr(self.x), _np_ptr(out)))
        blocks, o = [], 0
        for n in self.amopi:
            blocks.append(out[o:o + n * n].reshape(n, n))
            o += n * n
        return blocks

    def get_tpdm(self):
        """v2RDMSolver::get_tpdm (python_api.cc:398-458): amo^2 x amo^2 numpy
        array, element [i amo + k, j amo + l]; C1 only."""
        amo = sum(self.amopi)
        out = np.empty((amo * amo, amo * amo))
        self._sync()
        check(self.lib.v2g_get_tpdm(self.h, _ptr(self.x), _np_ptr(out)))
        return out

    # ---- orbital-optimization hand-off ------------------------------------
    def spatial_density_dims(self):
        d1, d2 = ctypes.c_longlong(), ctypes.c_longlong()
        check(self.lib.v2g_spatial_density_dims(self.h, ctypes.byref(d1), ctypes.byref(d2)))
        return d1.value, d2.value

    def singlet_geminals(self):
        """per irrep, the singlet geminals (i, j) the packed D2 runs over"""
        nirrep = len(self.amopi)
        counts = (ctypes.c_int * nirrep)()
        check(self.lib.v2g_singlet_geminals(self.h, counts, None))
        tot = sum(counts)
        pairs = (ctypes.c_int * max(1, 2 * tot))()
        check(self.lib.v2g_singlet_geminals(self.h, None, pairs))
        out, o = [], 0
        for h in range(nirrep):
            out.append([(pairs[2 * (o + n)], pairs[2 * (o + n) + 1]) for n in range(counts[h])])
            o += counts[h]
        return out

    def pack_spatial_density(self, x=None):
        """v2RDMSolver::PackSpatialDensity (v2rdm_solver.cc:3664-3773): the
        packed spin-free active D1 and D2 (numpy) of x (default: the solver's)"""
        x = self.x if x is None else x
        self._check_vec(x, self.dimx, "x")
        d1dim, d2dim = self.spatial_density_dims()
        d1, d2 = np.empty(max(1, d1dim)), np.empty(max(1, d2dim))
        self._sync()
        check(self.lib.v2g_pack_spatial_density(self.h, _ptr(x), _np_ptr(d1), _np_ptr(d2)))
        return d1[:d1dim], d2[:d2dim]

    def write_fcidump_rdm(self, path):
        """The .rdm file of v2RDMSolver::FCIDUMP (fcidump.cc:176-268)."""
        self._sync()
        check(self.lib.v2g_write_fcidump_rdm(self.h, _ptr(self.x), str(path).encode()))

    def write_checkpoint(self, path):
        """WriteCheckpointFile (checkpoint.cc:56-87): entries MU, PRIMAL (x),
        DUAL 1 (y), DUAL 2 (z), plus the iteration count and the last errors
        so that compute_energy(resume=True) continues the solve; a numpy .npz
        file (no pickled objects) instead of the psio file 269."""
        if self.x is None:
            raise ValueError("nothing to write: no solve has run")
        st = self.last_stats or dict(mu=self.mu, oiter=0, ep=0.0, ed=0.0)
        with open(path, "wb") as f:
            np.savez(f, **{"MU": np.float64(st["mu"]), "PRIMAL": self.x.cpu().numpy(),
                           "DUAL 1": self.y.cpu().numpy(), "DUAL 2": self.z.cpu().numpy(),
                           "OITER": np.int64(st["oiter"]), "EP": np.float64(st["ep"]),
                           "ED": np.float64(st["ed"]),
                           "DIMS": np.array([self.dimx, self.nconstraints], dtype=np.int64)})

    def read_checkpoint(self, path):
        """ReadFromCheckpointFile (checkpoint.cc:89-121): x, y, z and mu; the
        next compute_energy(resume=True) continues from them (with a fresh
        eigen split, as the reference starts one)."""
        with np.load(path, allow_pickle=False) as d:
            if tuple(d["DIMS"]) != (self.dimx, self.nconstraints):
                raise ValueError("checkpoint dimensions do not match this layout")
            self.x = torch.from_numpy(d["PRIMAL"].copy()).to(self.device)
            self.y = torch.from_numpy(d["DUAL 1"].copy()).to(self.device)
            self.z = torch.from_numpy(d["DUAL 2"].copy()).to(self.device)
            self.mu = float(d["MU"])
            self.last_stats = dict(mu=self.mu, oiter=int(d["OITER"]), ep=float(d["EP"]),
                                   ed=float(d["ED"]))
        self.reset_eig()

    # ---- outer loop ---------------------------------------------------
    def set_iter_callback(self, fn):
        """fn(oiter, stats_dict) -> bool (True stops); None clears."""
        if fn is None:
            self._cb = None
            check(self.lib.v2g_set_iter_callback(self.h, _lib.ITER_CALLBACK(0), None))
            return

        def tramp(user, oiter, st):
            s = st.contents
            return int(bool(fn(oiter, {k: getattr(s, k) for k, _ in _lib.Stats._fields_})))

        self._cb = _lib.ITER_CALLBACK(tramp)
        check(self.lib.v2g_set_iter_callback(self.h, self._cb, None))

    def compute_energy(self, h1=None, eri=None, enuc=0.0, verbose=False, max_seconds=0.0,
                       guess=True, resume=False):
        """BPSDP loop; returns a stats dict with energies (enuc added).

        resume=True continues the previous call (or a loaded checkpoint) from
        its x, z, y, mu, iteration count and errors instead of restarting the
        iteration at mu = 1 (implies guess=False)."""
        if h1 is not None:
            self.GetIntegrals(h1, eri)
        if self.c is None:
            raise ValueError("no integrals: pass h1/eri or call GetIntegrals first")
        if self.b is None:
            self.BuildConstraints()
        if resume and self.x is None:
            raise ValueError("resume needs x, z, y from a previous call or a checkpoint")
        if (guess and not resume) or self.x is None:
            self.Guess()
            self.last_stats = None
        o = self.options
        pc = {"NONE": 0, "JACOBI": 1, "LEARNED": 3}[str(o["PRECONDITIONER"]).upper()]
        p = _lib.Params(o["TAU_PARAMETER"], o["R_CONVERGENCE"], o["E_CONVERGENCE"],
                        o["CG_CONVERGENCE"], int(o["MAXITER"]), int(o["CG_MAXITER"]),
                        int(o["MU_UPDATE_FREQUENCY"]), pc, int(bool(o["FLEXIBLE_CG"])),
                        int(bool(verbose)), float(max_seconds), int(o["MLP_TRAIN_ITERS"]),
                        int(o["MLP_TRAIN_BATCH"]), int(o["MLP_TRAIN_STEPS"]),
                        float(o["MLP_TRAIN_LR"]), int(o["EIG_WARM_SWEEPS"]),
                        int(o["EIG_INNER_SWEEPS"]), int(bool(o["CG_EXTRAPOLATE"])))
        p.mlp_base = {"IDENTITY": 0, "JACOBI": 1}[str(o["MLP_BASE"]).upper()]
        p.eig_pair_order = int(o["EIG_PAIR_ORDER"])
        p.train_penalty = float(o["MLP_TRAIN_PENALTY"])
        prev = self.last_stats if resume else None
        if prev:
            p.mu0, p.oiter0, p.ep0, p.ed0 = prev["mu"], prev["oiter"], prev["ep"], prev["ed"]
        st = _lib.Stats()
        self._sync()
        check(self.lib.v2g_bpsdp(self.h, ctypes.byref(p), _ptr(self.c), _ptr(self.b),
                                 _ptr(self.x), _ptr(self.z), _ptr(self.y), ctypes.byref(st)))
        out = {k: getattr(st, k) for k, _ in _lib.Stats._fields_}
        out["energy"] = out["energy_primal"] + enuc
        self.mu = out["mu"]
        self.last_stats = out
        return out
