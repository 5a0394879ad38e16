"""GPU parity: dense coupling product (scripts/Functions_RSFDFN3DMain_D.jl:54-63,
:199-202) vs the oracle's FP64 product and numpy/OpenBLAS dgemv.

FP64 storage: <= 1e-12 relative to sum_j |E_ij x_j| (summation order differs).
FP32 storage: <= 2e-7 relative to the same scale (one FP32 rounding per entry,
FP64 accumulation).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mats(n, seed):
    rng = np.random.default_rng(seed)
    KS = rng.standard_normal((n, n))
    KS[np.diag_indices(n)] = -np.abs(KS[np.diag_indices(n)]) - n  # dominant negative self term
    KN = rng.standard_normal((n, n))
    return np.asfortranarray(KS), np.asfortranarray(KN)


@pytest.mark.parametrize("n", [1, 7, 64, 513, 2048, 3001])
def test_fp64_matches_oracle_and_blas(qd, orc, n):
    KS, KN = _mats(n, n)
    x = np.random.default_rng(n + 1).standard_normal(n)
    op = qd.DenseOperator.from_matrices(KS, KN, qd.PREC_FP64)
    tl, sp = op.matvec(x, loading_form=True)
    E, ks = orc.dense_prepare(KS)
    tlo, spo = orc.dense_coupling(E, KN, x)
    scaleS = np.abs(E) @ np.abs(x) + 1e-300
    scaleN = np.abs(KN) @ np.abs(x) + 1e-300
    assert (np.abs(tl - tlo) / scaleS).max() <= 1e-12
    assert (np.abs(sp - spo) / scaleN).max() <= 1e-12
    assert (np.abs(sp - KN @ x) / scaleN).max() <= 1e-12  # numpy/OpenBLAS dgemv
    yS, _ = op.matvec(x, loading_form=False)
    assert (np.abs(yS - KS @ x) / (np.abs(KS) @ np.abs(x))).max() <= 1e-12
    op.close()


@pytest.mark.parametrize("n", [64, 1000, 4096])
def test_fp32_storage_error_bound(qd, orc, n):
    KS, KN = _mats(n, 100 + n)
    x = np.random.default_rng(n).standard_normal(n)
    op = qd.DenseOperator.from_matrices(KS, KN, qd.PREC_FP32)
    tl, sp = op.matvec(x)
    E, _ = orc.dense_prepare(KS)
    tlo, spo = orc.dense_coupling(E, KN, x)
    # exact product of the FP32-rounded matrices, accumulated in FP64
    E32 = E.astype(np.float32).astype(np.float64)
    K32 = KN.astype(np.float32).astype(np.float64)
    scaleS = np.abs(E) @ np.abs(x) + 1e-300
    scaleN = np.abs(KN) @ np.abs(x) + 1e-300
    assert (np.abs(tl - (-(E32 @ x))) / scaleS).max() <= 1e-12
    assert (np.abs(sp - K32 @ x) / scaleN).max() <= 1e-12
    assert (np.abs(tl - tlo) / scaleS).max() <= 2e-7
    assert (np.abs(sp - spo) / scaleN).max() <= 2e-7
    op.close()


def test_normal_zero_flag(qd):
    n = 100
    KS, _ = _mats(n, 5)
    op = qd.DenseOperator.from_matrices(KS, np.zeros((n, n)))
    assert op.normal_zero()
    _, sp = op.matvec(np.ones(n))
    assert np.all(sp == 0.0)
    KN = np.zeros((n, n))
    KN[3, 7] = 1e-300
    op2 = qd.DenseOperator.from_matrices(KS, KN)
    assert not op2.normal_zero()


def test_bad_args_raise(qd):
    with pytest.raises(qd.QdfnError):
        qd.DenseOperator.from_matrices(np.zeros((0, 0)), np.zeros((0, 0)))


def test_tiled_layout_bitwise_equal(qd, monkeypatch):
    """The split-tiled dense layout (default from 6,144 elements) gives the
    column-major layout's products bit for bit (same per-row column order),
    in both storage precisions and both forms, and the adjoint's transposed
    product reads it through the same index."""
    from qdfn import configs
    cfg = configs.fault_network(n_faults=3, nx=16, nz=16)
    x = np.random.default_rng(21).standard_normal(cfg["seg"].shape[0])
    for prec in (qd.PREC_FP64, qd.PREC_FP32):
        out = []
        for tiled in ("0", "1"):
            monkeypatch.setenv("QDFN_DENSE_TILED", tiled)
            op = qd.DenseOperator.from_rect_segments(cfg["seg"], cfg["mu"], cfg["nu"], prec)
            out.append([op.matvec(x, form) for form in (True, False)])
            op.close()
        for a, b in zip(out[0], out[1]):
            np.testing.assert_array_equal(a[0], b[0])
            np.testing.assert_array_equal(a[1], b[1])


@pytest.mark.parametrize("prec", ["PREC_FP32", "PREC_FP64"])
def test_split_count_keeps_the_product(qd, orc, prec):
    """qdfn_dense_set_splits: any split count gives the oracle's product to the
    storage's bar; the default count (0) restores the default product bit for
    bit; the split-tiled FP64 layout (6,144 or more elements) refuses."""
    n = 2048
    KS, KN = _mats(n, 17)
    x = np.random.default_rng(18).standard_normal(n)
    E, _ = orc.dense_prepare(KS)
    tlo, spo = orc.dense_coupling(E, KN, x)
    scaleS = np.abs(E) @ np.abs(x) + 1e-300
    bar = 1e-12 if prec == "PREC_FP64" else 2e-7
    op = qd.DenseOperator.from_matrices(KS, KN, getattr(qd, prec))
    tl0, sp0 = op.matvec(x, loading_form=True)
    for sp in (1, 3, 64, n):
        taken = op.set_splits(sp)
        assert 1 <= taken <= sp
        tl, s = op.matvec(x, loading_form=True)
        assert (np.abs(tl - tlo) / scaleS).max() <= bar
    op.set_splits(0)
    tl, s = op.matvec(x, loading_form=True)
    np.testing.assert_array_equal(tl, tl0)
    np.testing.assert_array_equal(s, sp0)
    op.close()
    big = 6144
    rng = np.random.default_rng(19)
    KS2 = np.asfortranarray(rng.standard_normal((big, big)))
    op2 = qd.DenseOperator.from_matrices(KS2, KS2, qd.PREC_FP64)
    with pytest.raises(qd.QdfnError):
        op2.set_splits(4)
    op2.close()
