"""ctypes binding of libv2g.so (the HIP engine; include/v2g.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C v2rdm_casscf_amd/csrc``).  There is no CPU fallback: loading fails
loudly when the library is missing.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libv2g.so")


class LayoutDesc(ctypes.Structure):
    _fields_ = [("nirrep", ctypes.c_int), ("amopi", ctypes.c_int * 8),
                ("nalpha", ctypes.c_int), ("nbeta", ctypes.c_int),
                ("multiplicity", ctypes.c_int), ("constrain_q2", ctypes.c_int),
                ("constrain_g2", ctypes.c_int), ("constrain_t2", ctypes.c_int),
                ("constrain_spin", ctypes.c_int), ("constrain_sz", ctypes.c_int),
                ("constrain_t1", ctypes.c_int), ("spin_adapt_q2", ctypes.c_int),
                ("spin_adapt_g2", ctypes.c_int), ("energy_order", ctypes.POINTER(ctypes.c_int))]


class Params(ctypes.Structure):
    _fields_ = [("tau", ctypes.c_double), ("r_convergence", ctypes.c_double),
                ("e_convergence", ctypes.c_double), ("cg_convergence", ctypes.c_double),
                ("maxiter", ctypes.c_int), ("cg_maxiter", ctypes.c_int),
                ("mu_update_frequency", ctypes.c_int), ("precond", ctypes.c_int),
                ("flexible", ctypes.c_int), ("verbose", ctypes.c_int),
                ("max_seconds", ctypes.c_double), ("train_iters", ctypes.c_int),
                ("train_batch", ctypes.c_int), ("train_steps", ctypes.c_int),
                ("train_lr", ctypes.c_double), ("eig_warm_sweeps", ctypes.c_int),
                ("eig_inner_sweeps", ctypes.c_int), ("cg_extrapolate", ctypes.c_int),
                ("mu0", ctypes.c_double), ("oiter0", ctypes.c_int), ("ep0", ctypes.c_double),
                ("ed0", ctypes.c_double), ("mlp_base", ctypes.c_int),
                ("train_penalty", ctypes.c_double), ("eig_pair_order", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("oiter", ctypes.c_int), ("cg_iters", ctypes.c_longlong),
                ("converged", ctypes.c_int), ("energy_primal", ctypes.c_double),
                ("energy_dual", ctypes.c_double), ("ep", ctypes.c_double),
                ("ed", ctypes.c_double), ("mu", ctypes.c_double), ("t_cg", ctypes.c_double),
                ("t_xz", ctypes.c_double), ("t_total", ctypes.c_double),
                ("eig_sweeps", ctypes.c_int), ("precond_active", ctypes.c_int),
                ("cg_iters_train", ctypes.c_longlong), ("train_outer", ctypes.c_int),
                ("mlp_loss", ctypes.c_double), ("fallback", ctypes.c_int),
                ("exact_splits", ctypes.c_int)]


ITER_CALLBACK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                 ctypes.POINTER(Stats))

# name -> (restype, argtypes); every exported symbol of include/v2g.h
_VP, _DP, _IP = ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)
SIGNATURES = {
    "v2g_create": (ctypes.c_int, [ctypes.POINTER(LayoutDesc), ctypes.POINTER(ctypes.c_void_p)]),
    "v2g_destroy": (ctypes.c_int, [_VP]),
    "v2g_dims": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_longlong),
                                ctypes.POINTER(ctypes.c_longlong), _IP]),
    "v2g_block_dims": (ctypes.c_int, [_VP, _IP]),
    "v2g_set_stream": (ctypes.c_int, [_VP, _VP]),
    "v2g_Au": (ctypes.c_int, [_VP, _DP, _DP]),
    "v2g_ATu": (ctypes.c_int, [_VP, _DP, _DP]),
    "v2g_cg_Ax": (ctypes.c_int, [_VP, _DP, _DP]),
    "v2g_row_norm2": (ctypes.c_int, [_VP, _DP]),
    "v2g_cg_solve": (ctypes.c_int, [_VP, _DP, _DP, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, _IP, ctypes.POINTER(ctypes.c_double)]),
    "v2g_cg_solve_noalloc": (ctypes.c_int, [_VP, _DP, _DP, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, _IP,
                                            ctypes.POINTER(ctypes.c_double)]),
    "v2g_set_precond_diag": (ctypes.c_int, [_VP, _DP]),
    "v2g_update_xz": (ctypes.c_int, [_VP, ctypes.c_double, _DP, _DP, _DP, _DP]),
    "v2g_reset_eig": (ctypes.c_int, [_VP]),
    "v2g_set_eig_shape": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int]),
    "v2g_get_eig_shape": (ctypes.c_int, [_VP, _IP, _IP]),
    "v2g_set_autotune": (ctypes.c_int, [_VP, ctypes.c_int]),
    "v2g_bpsdp": (ctypes.c_int, [_VP, ctypes.POINTER(Params), _DP, _DP, _DP, _DP, _DP,
                                 ctypes.POINTER(Stats)]),
    "v2g_set_iter_callback": (ctypes.c_int, [_VP, ITER_CALLBACK, _VP]),
    "v2g_build_b": (ctypes.c_int, [_VP, _DP]),
    "v2g_build_c": (ctypes.c_int, [_VP, _DP, _DP, _DP]),
    "v2g_guess_random": (ctypes.c_int, [_VP, _DP, _DP, _DP]),
    "v2g_build_b_device": (ctypes.c_int, [_VP, _DP]),
    "v2g_build_c_device": (ctypes.c_int, [_VP, _DP, _DP, _DP]),
    "v2g_guess_random_device": (ctypes.c_int, [_VP, _DP, _DP, _DP]),
    "v2g_mlp_num_params": (ctypes.c_int, []),
    "v2g_mlp_init": (ctypes.c_int, [_VP, ctypes.c_ulonglong]),
    "v2g_mlp_get_weights": (ctypes.c_int, [_VP, _DP]),
    "v2g_mlp_set_weights": (ctypes.c_int, [_VP, _DP]),
    "v2g_mlp_loss_grad": (ctypes.c_int, [_VP, ctypes.c_int, _DP, _DP, _DP, _DP,
                                         ctypes.POINTER(ctypes.c_double)]),
    "v2g_mlp_adam_step": (ctypes.c_int, [_VP, _DP, ctypes.c_double]),
    "v2g_mlp_spectral_grad": (ctypes.c_int, [_VP, _DP, _DP, _DP, _DP, _DP,
                                             ctypes.POINTER(ctypes.c_double)]),
    "v2g_mlp_minv": (ctypes.c_int, [_VP, _DP, _DP, _DP]),
    "v2g_get_opdm": (ctypes.c_int, [_VP, _DP, _DP]),
    "v2g_get_tpdm": (ctypes.c_int, [_VP, _DP, _DP]),
    "v2g_spatial_density_dims": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_longlong),
                                                ctypes.POINTER(ctypes.c_longlong)]),
    "v2g_pack_spatial_density": (ctypes.c_int, [_VP, _DP, _DP, _DP]),
    "v2g_singlet_geminals": (ctypes.c_int, [_VP, _IP, _IP]),
    "v2g_write_fcidump_rdm": (ctypes.c_int, [_VP, _DP, ctypes.c_char_p]),
    "v2g_batch_create": (ctypes.c_int, [ctypes.POINTER(LayoutDesc), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "v2g_batch_destroy": (ctypes.c_int, [_VP]),
    "v2g_batch_offsets": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_longlong),
                                         ctypes.POINTER(ctypes.c_longlong)]),
    "v2g_batch_Au": (ctypes.c_int, [_VP, _DP, _DP]),
    "v2g_batch_ATu": (ctypes.c_int, [_VP, _DP, _DP]),
    "v2g_batch_inputs": (ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_void_p),
                                        ctypes.POINTER(ctypes.c_void_p), _DP, _DP, _DP, _DP, _DP]),
    "v2g_batch_bpsdp": (ctypes.c_int, [_VP, ctypes.POINTER(Params), _DP, _DP, _DP, _DP, _DP,
                                       ctypes.POINTER(Stats)]),
    "v2g_last_error": (ctypes.c_char_p, []),
    "v2g_version": (ctypes.c_char_p, []),
}

_lib = None


def load():
    """Load libv2g.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "or `make -C v2rdm_casscf_amd/csrc` (no CPU fallback exists)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class V2GError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        msg = load().v2g_last_error().decode()
        raise V2GError(f"v2g error {rc}: {msg}")
