"""The C++ binding a maintainer adds to the reference plugin, compiled against
include/v2g.h and linked with libv2g.so (tests/abi/plugin_binding.cc, built by
__graft_entry__.build()), runs the whole integration sequence -- create, dims,
build_b / build_c, guess, bpsdp, copy x back -- and the call-by-call seams
bpsdp_Au / bpsdp_ATu / cg_Ax; its results must equal the oracle's."""
import json
import os
import subprocess

import numpy as np
import pytest


pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "abi", "build", "plugin_binding")


@pytest.mark.parametrize("key", ["d2h_2110_s0", "c1_4_s0"])
def test_cpp_binding_matches_oracle(tmp_path, key):
    import oracle_loops as ol
    assert os.path.exists(EXE), "build it with __graft_entry__.build() (make -C tests/abi)"
    amopi, na, nb, _, _, _ = ol.CASES[key]
    L, h1, eri = ol.inputs(key)
    ints = tmp_path / "ints.bin"
    out = tmp_path / "out.bin"
    np.concatenate([h1.ravel(), eri.ravel()]).astype(np.float64).tofile(ints)
    args = [EXE, str(len(amopi)), *map(str, amopi), str(na), str(nb), str(ints), str(out)]
    print("[plugin_binding]", " ".join(args[1:]), flush=True)
    res = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    rec = json.loads(res.stdout.strip().splitlines()[-1])
    assert (rec["dimx"], rec["ncon"]) == (L.dimx, L.ncon)
    c = L.build_c(h1, eri)
    ref = ol.converged(key)     # oracle solve at R/E_CONVERGENCE 1e-6
    assert rec["converged"] == 1 and ref["converged"]
    assert abs(rec["energy"] - ref["energy_primal"]) < 1e-6
    data = np.fromfile(out, dtype=np.float64)
    x, au, atw = np.split(data, [L.dimx, L.dimx + L.ncon])
    assert abs(c @ x - rec["energy"]) < 1e-10
    u = np.sin(np.arange(1, L.dimx + 1, dtype=np.float64))
    w = np.cos(np.arange(1, L.ncon + 1, dtype=np.float64))
    assert np.linalg.norm(au - L.Au(u)) < 1e-12 * np.linalg.norm(L.Au(u))
    assert np.linalg.norm(atw - L.ATu(w)) < 1e-12 * np.linalg.norm(L.ATu(w))
    assert rec["cg_Ax_consistency"] < 1e-12
