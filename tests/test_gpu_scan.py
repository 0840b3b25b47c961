"""Bond scans: independent H-chain problems solved together -- as one batch of
the engine, and with one handle and stream each on host threads -- give the
same energies as solving them one by one, and known answers hold (H2:
v2RDM(DQG) = FCI; H4: lower bound to FCI)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("engine", ["batch", "threads"])
def test_scan_matches_individual_solves_and_fci(engine):
    import fci_tools as ft
    from v2rdm_casscf_amd import V2RDMSolver
    from v2rdm_casscf_amd.bondscan import solve_scan, solve_scan_threads
    from v2rdm_casscf_amd.hamiltonians import ANGSTROM_TO_BOHR, hydrogen_chain
    pts = [(2, 0.74), (4, 1.0), (2, 1.5), (4, 2.0), (6, 1.2)]
    if engine == "batch":
        opts = dict(R_CONVERGENCE=1e-6, E_CONVERGENCE=1e-6)
        res, wall = solve_scan(pts, options=opts)
    else:
        opts = dict(PRECONDITIONER="JACOBI", R_CONVERGENCE=1e-6, E_CONVERGENCE=1e-6)
        res, wall = solve_scan_threads(pts, workers=4, options=opts)
    for (n, r), out in zip(pts, res):
        assert out["converged"]
        d = hydrogen_chain(n, r * ANGSTROM_TO_BOHR)
        s = V2RDMSolver(d["amopi"], n // 2, n // 2, 1, "DQG", options=opts)
        single = s.compute_energy(d["h1"], d["eri"], enuc=d["enuc"])
        assert abs(single["energy"] - out["energy"]) < 1e-6
        if n <= 4:
            F = ft.Fock(n)
            efci, _ = F.ground_state(d["h1"], d["eri"], n // 2, n // 2, s_target=0)
            efci += d["enuc"]
            if n == 2:
                assert abs(out["energy"] - efci) < 2e-5
            else:
                assert out["energy"] <= efci + 1e-5
