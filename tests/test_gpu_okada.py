"""GPU parity: Okada DC3D and rectangle stiffness (HIP) vs the CPU oracle.

Bar: <= 1e-10 relative, row-scaled (|gpu - oracle| / max|oracle row|), NaN
where the oracle has NaN.  Reference: scripts/Functions_OKADA3D.jl:464-654,
scripts/Functions_BuildInputFile.jl:386-512.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NU, MU = 0.25, 32e9
LAM = 2 * MU * NU / (1 - 2 * NU)
ALPHA = (LAM + MU) / (LAM + 2 * MU)


def _points(rng, n):
    X = rng.uniform(-3000, 3000, n)
    Y = rng.uniform(-3000, 3000, n)
    Z = -rng.uniform(0, 4000, n)
    # edge cases: surface, on-plane, near edges, corner extensions
    X[:8] = [0, 500, -500, 500.0000001, 0, 123, -500, 0]
    Y[:8] = [0, 0, 0, 0, 1e-7, 0, 0, 0]
    Z[:8] = [0, -1000, -1000, -1300, -1000, 0, -700, -700]
    return X, Y, Z


@pytest.mark.parametrize("dip", [90.0, 60.0, 30.0, 89.99999999, 0.0])
@pytest.mark.parametrize("disl", [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.3, -0.7, 0.2)])
def test_dc3d_matches_oracle(qd, orc, dip, disl):
    rng = np.random.default_rng(7)
    X, Y, Z = _points(rng, 4096)
    args = (ALPHA, X, Y, Z, 1000.0, dip, 500.0, 500.0, 300.0, 300.0, *disl)
    Uo, io = orc.dc3d(*args)
    out = qd.Okada_DC3D_Vector(*args)
    Ug = np.stack(out[:12], axis=1)
    assert np.array_equal(np.isnan(Ug), np.isnan(Uo))
    fin = np.isfinite(Uo)
    scale = np.nanmax(np.where(np.isfinite(Uo), np.abs(Uo), 0), axis=1, keepdims=True)
    err = np.where(fin, np.abs(Ug - Uo), 0) / np.maximum(scale, 1e-300)
    assert err.max() <= 1e-10, err.max()
    np.testing.assert_array_equal(out[12], io)


def _row_scaled_err(A, B):
    scale = np.abs(B).max(axis=1, keepdims=True)
    return (np.abs(A - B) / np.maximum(scale, 1e-300)).max()


def test_stiffness_rect_single_fault(qd, orc):
    from qdfn import configs
    c = configs.single_fault()
    seg = c["seg"]
    KSo, KNo = orc.stiffness_rect(seg, seg, c["mu"], c["nu"])
    KSg, KNg = qd.StiffnessMatrix_ByParts_Calculation_Rec(seg, seg, c["mu"], c["nu"])
    assert _row_scaled_err(KSg, KSo) <= 1e-10
    # K_N of a vertical planar strike-slip fault is analytically zero: compare
    # against the shear scale instead of its own (round-off) scale
    assert np.abs(KNg - KNo).max() <= 1e-10 * np.abs(KSo).max()


def test_stiffness_rect_network_columns(qd, orc):
    from qdfn import configs
    c = configs.fault_network()
    seg = c["seg"]
    rng = np.random.default_rng(3)
    cols = np.sort(rng.choice(seg.shape[0], 192, replace=False))
    src = seg[cols]
    KSo, KNo = orc.stiffness_rect(src, seg, c["mu"], c["nu"])
    KSg, KNg = qd.StiffnessMatrix_ByParts_Calculation_Rec(src, seg, c["mu"], c["nu"])
    assert np.isfinite(KSg).all() and np.isfinite(KNg).all()
    assert _row_scaled_err(KSg, KSo) <= 1e-10
    assert _row_scaled_err(KNg, KNo) <= 1e-10


def test_dense_operator_assembly_matches_host_path(qd):
    """On-device assembly + operator build == host assembly uploaded."""
    from qdfn import configs
    c = configs.single_fault()
    seg = c["seg"]
    KS, KN = qd.StiffnessMatrix_ByParts_Calculation_Rec(seg, seg, c["mu"], c["nu"])
    a = qd.DenseOperator.from_rect_segments(seg, c["mu"], c["nu"])
    b = qd.DenseOperator.from_matrices(KS, KN)
    ka, da = a.kself()
    kb, db = b.kself()
    np.testing.assert_array_equal(da, db)
    np.testing.assert_array_equal(da, np.diag(KS))
    np.testing.assert_array_equal(ka, np.abs(np.diag(KS)))
    x = np.random.default_rng(1).standard_normal(seg.shape[0])
    for lf in (True, False):
        ya, na = a.matvec(x, lf)
        yb, nb = b.matvec(x, lf)
        np.testing.assert_array_equal(ya, yb)
        np.testing.assert_array_equal(na, nb)
