"""Device-side path inputs (v2g_build_b_device, v2g_build_c_device,
v2g_guess_random_device) equal the host builders bit for bit: b and c of
BuildConstraints / RepackIntegrals, and the RANDOM guess from glibc's rand()
after srand(0), on small layouts and on H50."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("amopi,na,nb,pos", [
    ([4], 2, 2, "DQG"), ([5], 3, 2, "DQG"), ([2, 1, 1, 1, 0, 1, 1, 1], 3, 3, "DQG"),
    ([5], 3, 2, "DQGT2"), ([2, 0, 1, 1, 0, 2, 1, 1], 5, 5, "DQG"),
    ([25, 0, 0, 0, 0, 25, 0, 0], 25, 25, "DQG"),
])
def test_device_inputs_equal_host_inputs(amopi, na, nb, pos):
    import fci_tools as ft
    from v2rdm_casscf_amd import V2RDMSolver
    from v2rdm_casscf_amd._lib import check
    print(f"[inputs {amopi} {na}/{nb} {pos}]", flush=True)
    s = V2RDMSolver(amopi, na, nb, None, pos)
    n = sum(amopi)
    irreps = [h for h, m in enumerate(amopi) for _ in range(m)]
    h1, eri = ft.random_hamiltonian(n, 2, irreps if len(amopi) > 1 else None) if n <= 12 else \
        (np.random.default_rng(2).normal(size=(n, n)), np.random.default_rng(3).normal(size=(n,) * 4))
    b_host = s.BuildConstraints(on_device=False).clone()
    c_host = s.GetIntegrals(h1, eri, on_device=False).clone()
    x_h, z_h, y_h = (t.clone() for t in s.Guess(on_device=False))
    b = torch.empty_like(b_host)
    c = torch.empty_like(c_host)
    x, z, y = torch.empty_like(x_h), torch.empty_like(z_h), torch.empty_like(y_h)
    dh1 = torch.from_numpy(np.ascontiguousarray(h1)).cuda()
    deri = torch.from_numpy(np.ascontiguousarray(eri)).cuda()
    torch.cuda.synchronize()
    check(s.lib.v2g_build_b_device(s.h, _p(b)))
    check(s.lib.v2g_build_c_device(s.h, _p(dh1), _p(deri), _p(c)))
    check(s.lib.v2g_guess_random_device(s.h, _p(x), _p(z), _p(y)))
    assert torch.equal(b, b_host)
    assert torch.equal(c, c_host)
    assert torch.equal(x, x_h) and torch.equal(z, z_h) and torch.equal(y, y_h)
