"""The ctypes mirror of the v2g.h structures (v2rdm_casscf_amd/_lib.py) has the
sizes and field offsets a C compiler gives the header: tests/abi/struct_layout.c
is compiled with the host C compiler and its offsetof() dump compared field by
field (a mismatch would pass garbage across the boundary silently)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _c_layout(tmp_path):
    exe = str(tmp_path / "struct_layout")
    subprocess.check_call(["gcc", "-std=c99", "-O0", "-o", exe,
                           os.path.join(HERE, "abi", "struct_layout.c")])
    out = subprocess.check_output([exe], text=True)
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


@pytest.mark.parametrize("cname,pyname", [("v2g_layout_desc", "LayoutDesc"),
                                          ("v2g_params", "Params"), ("v2g_stats", "Stats")])
def test_ctypes_structs_match_header(tmp_path, cname, pyname):
    import ctypes
    from v2rdm_casscf_amd import _lib
    c = _c_layout(tmp_path)
    cls = getattr(_lib, pyname)
    assert ctypes.sizeof(cls) == c[f"{cname}.sizeof"]
    fields = [k.split(".", 1)[1] for k in c if k.startswith(cname + ".") and not k.endswith("sizeof")]
    assert [f for f, _ in cls._fields_] == fields
    for f in fields:
        assert getattr(cls, f).offset == c[f"{cname}.{f}"], f
