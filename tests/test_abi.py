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


def test_null_handles_and_buffers_are_errors():
    """A null handle or a null required pointer returns a negative code with
    a message; nothing here reaches the GPU."""
    L = lib.load()
    calls = [
        ("fdf_step", lambda: L.fdf_step(None, 1)),
        ("fdf_stage", lambda: L.fdf_stage(None, 0, 0)),
        ("fdf_read_field", lambda: L.fdf_read_field(None, 0, None, None, None, None)),
        ("fdf_read_array", lambda: L.fdf_read_array(None, 0, None, 0)),
        ("fdf_write_array", lambda: L.fdf_write_array(None, 0, None, 0)),
        ("fdf_save_state", lambda: L.fdf_save_state(None, b"/tmp/x")),
        ("fdf_load_state", lambda: L.fdf_load_state(None, b"/tmp/x")),
        ("fdf_read_iface", lambda: L.fdf_read_iface(None, 0, 0, None, 0)),
        ("fdf_read_front", lambda: L.fdf_read_front(None, 0, None, 0)),
        ("fdf_finish", lambda: L.fdf_finish(None)),
        ("fdf_batch_step", lambda: L.fdf_batch_step(None, 1)),
        ("fdf_batch_read_iface", lambda: L.fdf_batch_read_iface(None, 0, 0, 0, None, 0)),
        ("fdf_batch_read_front", lambda: L.fdf_batch_read_front(None, 0, 0, None, 0)),
        ("fdf_batch_read_array", lambda: L.fdf_batch_read_array(None, 0, 0, None, 0)),
        ("fdf_batch_finish", lambda: L.fdf_batch_finish(None)),
        ("fdf_stz", lambda: L.fdf_stz_log(None, 10)),
        ("fdf_stz", lambda: L.fdf_stz_stats(None, (ctypes.c_ulonglong * 4)())),
    ]
    for name, call in calls:
        assert call() < 0, name
        msg = L.fdf_last_error().decode()
        assert msg.startswith(name) and ("null handle" in msg or "null argument" in msg), (name, msg)
    assert L.fdf_bulk_kernel_ms(None, None) < 0
    assert L.fdf_get_info(None, None) < 0 and L.fdf_batch_size(None) < 0
