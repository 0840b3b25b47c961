"""Batches of independent BPSDP problems (bond scans) solved together on one
GPU through the v2g batch C ABI (include/v2g.h, v2g_batch_*).

All problems' vectors are concatenated on the device; every kernel launch of
the iteration covers the whole batch, each problem keeps its own mu, errors,
CG scalars and stopping rule (v2rdm_solver.cc:1778-1912) and leaves the batch
when it converges.  The reference has no counterpart: it solves one problem
per Psi4 process.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check
from .solver import _ptr, fill_desc


class BatchSolver:
    """problems: list of dict(amopi, nalpha, nbeta[, multiplicity][, positivity]
    [, constrain_spin][, constrain_sz][, spin_adapt_q2][, spin_adapt_g2]
    [, energy_order]), the keywords of V2RDMSolver."""

    def __init__(self, problems, device="cuda:0", options=None):
        self.lib = _lib.load()
        self.device = torch.device(device)
        n = len(problems)
        if n == 0:
            raise ValueError("empty batch")
        descs = (_lib.LayoutDesc * n)()
        self._eorders = []
        for d, pb in zip(descs, problems):
            self._eorders.append(fill_desc(
                d, pb["amopi"], pb["nalpha"], pb["nbeta"], pb.get("multiplicity"),
                pb.get("positivity", "DQG"), pb.get("constrain_spin", True),
                pb.get("constrain_sz", True), pb.get("spin_adapt_q2", False),
                pb.get("spin_adapt_g2", False), pb.get("energy_order")))
        self.problems = problems
        self.nprob = n
        with torch.cuda.device(self.device):
            h = ctypes.c_void_p()
            check(self.lib.v2g_batch_create(descs, n, ctypes.byref(h)))
        self.h = h
        xo = (ctypes.c_longlong * (n + 1))()
        yo = (ctypes.c_longlong * (n + 1))()
        check(self.lib.v2g_batch_offsets(self.h, xo, yo))
        self.xoff, self.yoff = list(xo), list(yo)
        self.dimx, self.ncon = self.xoff[-1], self.yoff[-1]
        self.options = dict(TAU_PARAMETER=1.0, R_CONVERGENCE=1e-6, E_CONVERGENCE=1e-6,
                            CG_CONVERGENCE=1e-9, MAXITER=100000, CG_MAXITER=10000,
                            MU_UPDATE_FREQUENCY=500, EIG_WARM_SWEEPS=1, EIG_INNER_SWEEPS=1,
                            CG_EXTRAPOLATE=True, EIG_PAIR_ORDER=1)
        if options:
            self.options.update({k.upper(): v for k, v in options.items()})
        self.c = self.b = self.x = self.z = self.y = None

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.v2g_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def _vec(self, n):
        return torch.empty(n, dtype=torch.float64, device=self.device)

    def x_of(self, q, v=None):
        v = self.x if v is None else v
        return v[self.xoff[q]:self.xoff[q + 1]]

    def y_of(self, q, v=None):
        v = self.y if v is None else v
        return v[self.yoff[q]:self.yoff[q + 1]]

    def Au(self, x):
        out = self._vec(self.ncon)
        torch.cuda.synchronize(self.device)
        check(self.lib.v2g_batch_Au(self.h, _ptr(x), _ptr(out)))
        return out

    def ATu(self, y):
        out = self._vec(self.dimx)
        torch.cuda.synchronize(self.device)
        check(self.lib.v2g_batch_ATu(self.h, _ptr(y), _ptr(out)))
        return out

    def set_integrals(self, h1s, eris):
        """Active-space integrals per problem (numpy); builds b, c and the
        RANDOM guess of every problem on the device."""
        dh = [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(self.device)
              for a in h1s]
        de = [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(self.device)
              for a in eris]
        ph = (ctypes.c_void_p * self.nprob)(*[t.data_ptr() for t in dh])
        pe = (ctypes.c_void_p * self.nprob)(*[t.data_ptr() for t in de])
        self.c, self.b = self._vec(self.dimx), self._vec(self.ncon)
        self.x, self.z, self.y = self._vec(self.dimx), self._vec(self.dimx), self._vec(self.ncon)
        torch.cuda.synchronize(self.device)
        check(self.lib.v2g_batch_inputs(self.h, ph, pe, _ptr(self.c), _ptr(self.b), _ptr(self.x),
                                        _ptr(self.z), _ptr(self.y)))

    def compute_energy(self, enucs=None, max_seconds=0.0, progress=False):
        """The BPSDP loop of every problem; returns a list of stats dicts."""
        if self.c is None:
            raise ValueError("no integrals: call set_integrals first")
        o = self.options
        p = _lib.Params()
        p.tau, p.r_convergence, p.e_convergence = o["TAU_PARAMETER"], o["R_CONVERGENCE"], o["E_CONVERGENCE"]
        p.cg_convergence, p.maxiter, p.cg_maxiter = o["CG_CONVERGENCE"], int(o["MAXITER"]), int(o["CG_MAXITER"])
        p.mu_update_frequency, p.precond = int(o["MU_UPDATE_FREQUENCY"]), 0
        p.max_seconds = float(max_seconds)
        p.verbose = 2 if progress else 0     # a progress line on stderr every 20 s
        p.eig_warm_sweeps, p.eig_inner_sweeps = int(o["EIG_WARM_SWEEPS"]), int(o["EIG_INNER_SWEEPS"])
        p.cg_extrapolate, p.eig_pair_order = int(bool(o["CG_EXTRAPOLATE"])), int(o["EIG_PAIR_ORDER"])
        st = (_lib.Stats * self.nprob)()
        torch.cuda.synchronize(self.device)
        check(self.lib.v2g_batch_bpsdp(self.h, ctypes.byref(p), _ptr(self.c), _ptr(self.b),
                                       _ptr(self.x), _ptr(self.z), _ptr(self.y), st))
        out = []
        for q in range(self.nprob):
            d = {k: getattr(st[q], k) for k, _ in _lib.Stats._fields_}
            d["energy"] = d["energy_primal"] + (enucs[q] if enucs is not None else 0.0)
            out.append(d)
        return out
