"""MI355X-native boundary-point SDP engine for variational 2-RDM (v2RDM-CASSCF).

The hot path of edeprince3/v2rdm_casscf (constraint maps A.x / A^T.y, the CG
solve of A A^T y = rhs, the per-block eigen split of Update_xz and the outer
BPSDP loop) runs as hand-written HIP kernels in lib/libv2g.so behind the C ABI
of include/v2g.h.  ``V2RDMSolver`` mirrors the reference's v2RDMSolver method
names over that ABI.
"""
from .solver import V2RDMSolver  # noqa: F401

__all__ = ["V2RDMSolver"]
