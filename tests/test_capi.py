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


def test_sampled_build_argument_checks_without_device():
    import numpy as np
    import qdfn
    from qdfn import _lib
    L = qdfn.lib()
    PD, PI64 = _lib.PD, _lib.PI64
    assert L.qdfn_hmat_build_tri_sampled(4, None, None, None, None, None, None, None, 1.0, 1.0, 1, None, None,
                                         1e-6, 0.0, _lib.PREC_FP64) is None
    assert "null input" in _lib.last_error()
    assert L.qdfn_hmat_build_rect_sampled(None, 4, 4, 1.0, 0.25, 1, None, None, 1e-6, 0.0,
                                          _lib.PREC_FP64) is None
    assert "qdfn_hmat_build_rect_sampled" in _lib.last_error()
    # a block range outside [1, n] is rejected before the device is touched
    seg = np.zeros((4, 16), order="F")
    rng = np.array([[1, 4, 1, 5]], dtype=np.int64)
    adm = np.array([0], dtype=np.int64)
    assert L.qdfn_hmat_build_rect_sampled(seg.ctypes.data_as(PD), 4, 4, 1.0, 0.25, 1, rng.ctypes.data_as(PI64),
                                          adm.ctypes.data_as(PI64), 1e-6, 0.0, _lib.PREC_FP64) is None
    assert "out of bounds" in _lib.last_error()
    assert L.qdfn_hmat_build_info(None, None) == _lib.QDFN_ERR_ARG


def test_precision_checks_without_device():
    import numpy as np
    import qdfn
    from qdfn import _lib
    L = qdfn.lib()
    PD, PI32, PI64 = _lib.PD, _lib.PI32, _lib.PI64
    # BF16 low-rank factors are an H-operator storage: dense operators refuse it
    K = np.eye(4, order="F")
    assert L.qdfn_dense_create(K.ctypes.data_as(PD), K.ctypes.data_as(PD), 4, 4, _lib.PREC_BF16_FAR) is None
    assert "prec" in _lib.last_error()
    # an unknown storage precision is refused before the device is touched
    rng = np.array([[1, 4, 1, 4]], dtype=np.int64)
    adm = np.array([0], dtype=np.int64)
    kind = np.zeros(1, dtype=np.int32)
    rank = np.zeros(1, dtype=np.int64)
    A = (PD * 1)(K.ctypes.data_as(PD))
    B = (PD * 1)()
    import ctypes as C
    args = [kind.ctypes.data_as(PI32), rank.ctypes.data_as(PI64), C.cast(A, C.POINTER(PD)), C.cast(B, C.POINTER(PD))]
    assert L.qdfn_hmat_create(4, 1, rng.ctypes.data_as(PI64), adm.ctypes.data_as(PI64), *args, *args, 5) is None
    assert "precision" in _lib.last_error()
