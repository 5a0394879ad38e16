"""The C ABI: libqdfn_hip.so loads and exports every function include/qdfn.h
declares, the Python bindings declare the same set, and error reporting works
without a device.  No compute call is made (CPU-only)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qdfn.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(qdfn_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_parses():
    names = declared_functions()
    assert "qdfn_okada_stiffness_rect" in names and "qdfn_sim_run" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from qdfn import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    from qdfn import _lib
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_error_path_without_device():
    import qdfn
    L = qdfn.lib()
    # argument validation fails before any device call
    rc = L.qdfn_okada_stiffness_rect(None, 1, 1, None, 1, 1, 1.0, 0.25, None, None, 1)
    assert rc == qdfn._lib.QDFN_ERR_ARG
    assert "qdfn_okada_stiffness_rect" in qdfn._lib.last_error()
    assert L.qdfn_dense_create(None, None, 4, 4, 0) is None
    assert "null" in qdfn._lib.last_error()
