"""GPU: qdfn_matvec_device_noalloc (the coupling product on device vectors,
split partial sums in the library's static device scratch, no allocation)
against qdfn_matvec on the same operator: bitwise equal, in both forms, at
FP64 and FP32 storage, on two calls in a row with the second on a smaller
operator (the shared scratch keeps the first call's partials); an operator
whose partials exceed the scratch is refused with QDFN_ERR_ARG."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """The HIP runtime the library links (libamdhip64.so.7, already loaded by
    it): device buffers and a stream for the test's inputs and outputs."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.h.hipStreamDestroy.argtypes = [C.c_void_p]

    def ok(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def malloc(self, nbytes):
        p = self.C.c_void_p()
        self.ok(self.h.hipMalloc(self.C.byref(p), nbytes))
        return p.value

    def to_device(self, p, a):
        self.ok(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1))  # host to device

    def to_host(self, a, p):
        self.ok(self.h.hipMemcpy(a.ctypes.data, p, a.nbytes, 2))  # device to host


def _device_product(op, x, form):
    hip = _Hip()
    x = np.ascontiguousarray(x, dtype=np.float64)
    nb = 8 * op.n
    dx, dS, dN = hip.malloc(nb), hip.malloc(nb), hip.malloc(nb)
    nan = np.full(op.n, np.nan)
    hip.to_device(dx, x)
    hip.to_device(dS, nan)
    hip.to_device(dN, nan)
    s = hip.C.c_void_p()
    hip.ok(hip.h.hipStreamCreate(hip.C.byref(s)))
    try:
        op.matvec_device_noalloc(dx, dS, dN, form, s.value)
        hip.ok(hip.h.hipStreamSynchronize(s))
        yS, yN = np.empty(op.n), np.empty(op.n)
        hip.to_host(yS, dS)
        hip.to_host(yN, dN)
    finally:
        hip.h.hipStreamSynchronize(s)
        hip.h.hipStreamDestroy(s)
        for p in (dx, dS, dN):
            hip.h.hipFree(p)
    return yS, yN


@pytest.mark.parametrize("prec_name", ["PREC_FP64", "PREC_FP32"])
def test_noalloc_matches_matvec_on_two_calls(qd, prec_name):
    from qdfn import configs
    prec = getattr(qd, prec_name)
    rng = np.random.default_rng(11)
    for n_faults, nx, nz in ((3, 16, 16), (2, 12, 8)):  # 768 elements, then 192
        cfg = configs.fault_network(n_faults=n_faults, nx=nx, nz=nz)
        op = qd.DenseOperator.from_rect_segments(cfg["seg"], cfg["mu"], cfg["nu"], prec)
        assert not op.normal_zero()
        x = rng.standard_normal(op.n)
        for form in (True, False):
            ref = op.matvec(x, loading_form=form)
            got = _device_product(op, x, form)
            for r, g in zip(ref, got):
                assert np.isfinite(g).all()
                np.testing.assert_array_equal(g, r)
        op.close()


def test_noalloc_refuses_partials_beyond_scratch(qd, monkeypatch):
    from qdfn import configs
    monkeypatch.setenv("QDFN_COUPLING_WGS", "100000")  # as many column splits as the size allows
    cfg = configs.single_fault(nx=65, nz=64)  # 4,160 elements: 130 splits x 4,160 rows x 2 > 2^20 doubles
    op = qd.DenseOperator.from_rect_segments(cfg["seg"], cfg["mu"], cfg["nu"], qd.PREC_FP64)
    with pytest.raises(qd.QdfnError) as e:
        _device_product(op, np.ones(op.n), True)
    assert e.value.code == -1 and "maximum" in str(e.value)
    op.close()
