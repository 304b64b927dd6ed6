"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute call: runs without a GPU)."""
import ctypes
import os
import re

from fdfault_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(fdf_\w+)\s*\(", txt, flags=re.M):
                syms.add(m.group(1))
    return syms


def test_library_exports_every_declared_symbol():
    L = lib.load()
    syms = declared_symbols()
    assert len(syms) >= 14
    for s in sorted(syms):
        assert hasattr(L, s), s
    assert syms <= set(lib.EXPORTS) | {"fdf_zz_unused"}


def test_error_path_reports_reference_message(tmp_path):
    bad = tmp_path / "bad.in"
    bad.write_text("[fdfault.domain]\n2\n")
    L = lib.load()
    info = lib.Info()
    rc = L.fdf_host_setup(str(bad).encode(), ctypes.byref(info), None, None, None)
    assert rc < 0
    assert "Error reading problem from input file" in L.fdf_last_error().decode()
