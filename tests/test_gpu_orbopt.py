"""The orbital-optimization hand-off on the GPU (v2g_pack_spatial_density,
v2g_singlet_geminals) and the v2RDM-CASSCF cycle of orbopt.v2rdm_casscf.

* the packed densities of a device x equal the oracle's packing of the same x
  (PackSpatialDensity, v2rdm_solver.cc:3664-3773), and the singlet geminals
  the oracle's;
* with two active electrons the DQG conditions are exact, so v2RDM-CASSCF
  (BPSDP on the GPU, orbital steps on the host, c rebuilt on the device after
  each step, the solve resumed) must reach the CASSCF energy that FCI
  densities in the active space give with the same orbital step, 1e-6 Eh;
* a (4e,4o) active space: converged, orbital gradient of the final densities
  below the threshold, energy below CAS-CI of the starting orbitals.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle_py as orc  # noqa: E402
import test_orbopt as to  # noqa: E402
from v2rdm_casscf_amd import orbopt as oo  # noqa: E402
from v2rdm_casscf_amd.hamiltonians import hydrogen_chain  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("amopi,na,nb,eo", [([4], 2, 2, [2, 0, 3, 1]),
                                            ([3, 1, 2, 0, 1, 2, 0, 1], 4, 4,
                                             [9, 0, 5, 1, 8, 2, 7, 3, 6, 4])])
def test_packed_densities_match_oracle(amopi, na, nb, eo):
    from v2rdm_casscf_amd import V2RDMSolver
    s = V2RDMSolver(amopi, na, nb, 1, "DQG", energy_order=eo)
    L = orc.Layout(amopi, na, nb, 1, "DQG", energy_order=eo)
    assert s.singlet_geminals() == [L.bas_00(h) for h in range(L.nirrep)]
    x = np.random.default_rng(4).normal(size=L.dimx)
    d1, d2 = s.pack_spatial_density(torch.from_numpy(x).cuda())
    r1, r2 = L.pack_spatial_density(x)
    assert np.array_equal(d1, r1)
    assert np.allclose(d2, r2, rtol=0, atol=1e-14)


@pytest.mark.parametrize("n,spacing,docc,active", [(6, 1.9, [1, 1], [1, 1]),
                                                   (8, 2.2, [2, 1], [1, 1])])
def test_v2rdm_casscf_two_active_electrons_equals_casscf(n, spacing, docc, active):
    d = hydrogen_chain(n, spacing)
    sp = oo.Spaces([n // 2, n // 2], docc, active)
    print(f"[casscf H{n} docc {docc} active {active}]", flush=True)
    ref, _, _, ref_it = to.casscf_fci(d["h1"], d["eri"], sp, 1, 1)
    out = oo.v2rdm_casscf(d["h1"], d["eri"], d["enuc"], sp, 1, 1, options=dict(MAXITER=100000))
    print(f"  E {out['energy']:.10f} CASSCF(FCI) {ref + d['enuc']:.10f}, {out['macro']} macro "
          f"iterations (FCI {ref_it})", flush=True)
    assert out["converged"]
    assert abs(out["energy"] - (ref + d["enuc"])) < 1e-6


def test_v2rdm_casscf_four_active_electrons():
    d = hydrogen_chain(8, 2.0)
    sp = oo.Spaces([4, 4], [1, 1], [2, 2])       # (4e,4o), 2 core, 2 virtual
    e_ci, _, _ = to._cas_fci_densities(d["h1"], d["eri"], sp, 2, 2)
    out = oo.v2rdm_casscf(d["h1"], d["eri"], d["enuc"], sp, 2, 2, options=dict(MAXITER=100000))
    print(f"[casscf H8 (4e,4o)] E {out['energy']:.10f}, CAS-CI {e_ci + d['enuc']:.10f}, "
          f"{out['macro']} macro iterations", flush=True)
    assert out["converged"]
    assert out["energy"] < e_ci + d["enuc"]
    s = out["solver"]
    D1, G = oo.unpack_spatial_density(*s.pack_spatial_density(), sp.active, s.singlet_geminals())
    grad, _ = oo.gradient_hessian(out["h"], out["eri"], sp, D1, G)
    assert np.linalg.norm([grad[m, q] for m, q in sp.pairs()]) < 1e-5
