"""RDM outputs and the checkpoint through the C ABI: get_opdm / get_tpdm
(python_api.cc:376-458), the RDM file of FCIDUMP (fcidump.cc:176-268) and
WriteCheckpointFile / ReadFromCheckpointFile (checkpoint.cc:56-121).

The output conventions are pinned with the exact RDMs of a brute-force FCI
state uploaded as x: each output must give back the FCI energy."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import fci_tools as ft  # noqa: E402
import oracle_py as orc  # noqa: E402

pytestmark = pytest.mark.gpu


def _exact_solver(amopi, na, nb, seed):
    from v2rdm_casscf_amd import V2RDMSolver
    L = orc.Layout(amopi, na, nb, None, "DQG")
    n = L.amo
    irreps = [L.irrep(p) for p in range(n)]
    h1, eri = ft.random_hamiltonian(n, seed, irreps if len(amopi) > 1 else None)
    F = ft.Fock(n)
    e, psi = F.ground_state(h1, eri, na, nb, s_target=(na - nb) / 2)
    x = ft.exact_x(F, psi, L)
    s = V2RDMSolver(amopi, na, nb, None, "DQG")
    s.x = torch.from_numpy(x).cuda()
    return s, L, h1, eri, e, x


@pytest.mark.parametrize("amopi,na,nb", [([4], 2, 2), ([5], 3, 2), ([2, 1, 1, 0], 2, 2)])
def test_outputs_of_the_exact_state_give_its_energy(tmp_path, amopi, na, nb):
    from v2rdm_casscf_amd import fcidump
    s, L, h1, eri, e, x = _exact_solver(amopi, na, nb, 3)
    n = L.amo
    # get_opdm: per-irrep blocks of D1a + D1b
    blocks = s.get_opdm()
    d1 = np.zeros((n, n))
    off = 0
    for hh, m in enumerate(amopi):
        d1[off:off + m, off:off + m] = blocks[hh]
        off += m
    # the RDM file: its 1-RDM equals get_opdm, and the energy comes back
    path = tmp_path / "x.rdm"
    s.write_fcidump_rdm(path)
    f1, f2 = fcidump.read_rdm(path, n)
    assert np.allclose(f1, d1, atol=1e-14)
    assert abs(fcidump.rdm_energy(f1, f2, h1, eri) - e) < 1e-10
    assert abs(np.einsum("ij,ij->", h1, d1) + 0.5 * np.einsum("ijkl,ikjl->", f2, eri) - e) < 1e-10
    if len(amopi) == 1:
        tp = s.get_tpdm().reshape(n, n, n, n)           # [i, k, j, l]
        e_tp = np.einsum("ij,ij->", h1, d1) + 0.5 * np.einsum("ikjl,ikjl->", tp, eri)
        assert abs(e_tp - e) < 1e-10
    else:
        from v2rdm_casscf_amd._lib import V2GError
        with pytest.raises(V2GError, match="CIWavefunction"):
            s.get_tpdm()


def test_checkpoint_resume_reaches_the_same_energy(tmp_path):
    """Write a checkpoint part way through, read it into a fresh solver and
    resume: the same converged energy as one uninterrupted solve."""
    from v2rdm_casscf_amd import V2RDMSolver
    from v2rdm_casscf_amd.hamiltonians import hydrogen_chain
    d = hydrogen_chain(6, 1.8)
    args = (d["amopi"], 3, 3, 1, "DQG")
    one = V2RDMSolver(*args, options=dict(MAXITER=100000))
    ref = one.compute_energy(d["h1"], d["eri"], enuc=d["enuc"])
    a = V2RDMSolver(*args, options=dict(MAXITER=400))
    part = a.compute_energy(d["h1"], d["eri"], enuc=d["enuc"])
    assert part["oiter"] == 400 and not part["converged"]
    path = tmp_path / "v2rdm.chk.npz"
    a.write_checkpoint(path)
    b = V2RDMSolver(*args, options=dict(MAXITER=100000))
    b.GetIntegrals(d["h1"], d["eri"])
    b.read_checkpoint(path)
    assert b.mu == part["mu"]
    out = b.compute_energy(enuc=d["enuc"], resume=True)
    assert out["converged"] and ref["converged"]
    assert abs(out["energy"] - ref["energy"]) < 1e-6
    assert abs(out["oiter"] - ref["oiter"]) <= 0.05 * ref["oiter"] + 10
