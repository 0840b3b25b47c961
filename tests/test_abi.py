"""The C ABI library loads without a GPU and exports every symbol that
include/v2g.h declares (no compute call is made here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "v2g.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*[\s*]+(v2g_[A-Za-z0-9_]+)\s*\(",
                       src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from v2rdm_casscf_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert set(names) <= set(_lib.SIGNATURES), sorted(set(names) - set(_lib.SIGNATURES))


def test_binding_loads_and_reports_version():
    from v2rdm_casscf_amd import _lib
    lib = _lib.load()
    assert lib.v2g_version().startswith(b"v2g")
    assert lib.v2g_mlp_num_params() == 14 * 64 + 64 + 64 * 64 + 64 + 64 + 1
