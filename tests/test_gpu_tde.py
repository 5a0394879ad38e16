"""GPU parity: triangular dislocations (HIP) vs the CPU oracle.

TDstressHS (scripts/Functions_TDstressHS.jl:2-130) at random points, on the
free surface and near the triangle, for several triangles and slips: <= 1e-10
relative to the field's scale per point (row-scaled), identical NaN pattern
(on-edge points).  Triangle stiffness (Functions_BuildInputFile.jl:551-594) of
a small rough mesh and RotVerts_UnitVectors (:338-377) against the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MU, NU = 32e9, 0.25
LAM = 2 * MU * NU / (1 - 2 * NU)

TRIS = [
    (np.array([-400.0, 100, -300]), np.array([350.0, -150, -900]), np.array([0.0, 300, -1200])),
    (np.array([-400.0, 0, 0.0]), np.array([400.0, 0, 0.0]), np.array([0.0, 0, -700])),      # surface edge
    (np.array([-300.0, -200, -500]), np.array([300.0, -200, -500]), np.array([0.0, 300, -500])),  # horizontal
    (np.array([0.0, 0, -100]), np.array([10.0, 400, -1500]), np.array([0.0, 0, -1500])),
]


def _points(rng, P1, P2, P3, n=3000):
    X = rng.uniform(-2500, 2500, n)
    Y = rng.uniform(-2500, 2500, n)
    Z = -rng.uniform(0, 3000, n)
    Z[:200] = 0.0
    c = (P1 + P2 + P3) / 3
    nrm = np.cross(P2 - P1, P3 - P1)
    nrm /= np.linalg.norm(nrm)
    off = rng.uniform(-5, 5, 200)
    near = c + off[:, None] * nrm + rng.uniform(-50, 50, (200, 3))
    X[200:400], Y[200:400], Z[200:400] = near.T
    Z[200:400] = np.minimum(Z[200:400], 0.0)
    mid = (P1 + P2) / 2
    X[400], Y[400], Z[400] = mid
    return X, Y, Z


@pytest.mark.parametrize("k", range(len(TRIS)))
@pytest.mark.parametrize("slip", [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (-0.6, 0.5, 0.3)])
def test_tdstress_hs_matches_oracle(qd, orc, k, slip):
    P1, P2, P3 = TRIS[k]
    X, Y, Z = _points(np.random.default_rng(10 + k), P1, P2, P3)
    So, Eo = orc.tdstress_hs(X, Y, Z, P1, P2, P3, *slip, MU, LAM)
    Sg, Eg = qd.TDstressHS(X, Y, Z, P1, P2, P3, *slip, MU, LAM)
    np.testing.assert_array_equal(np.isnan(Sg), np.isnan(So))
    for G, O in ((Sg, So), (Eg, Eo)):
        fin = np.isfinite(O).all(axis=1)
        scale = np.abs(O[fin]).max(axis=1, keepdims=True)
        err = np.abs(G[fin] - O[fin]) / np.maximum(scale, 1e-300)
        assert err.max() <= 1e-10, err.max()


def _mesh(nx=16, nz=10):
    from qdfn import configs
    return configs.rough_fault(nx=nx, nz=nz, h=60.0, ratio=2e-2)


def test_rotverts_matches_oracle(qd, orc):
    m = _mesh()
    g = qd.RotVerts_UnitVectors(np.hstack([m["P1"], m["P2"], m["P3"]]), m["P1"].shape[0], m["rake"])
    o = orc.rotverts_unitvectors(m["P1"], m["P2"], m["P3"], m["rake"])
    for a, b in zip(g, o):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-13 * max(1.0, np.abs(b).max()))


def test_stiffness_tri_matches_oracle(qd, orc):
    m = _mesh()
    P1, P2, P3, nrm, st, dp, slip = orc.rotverts_unitvectors(m["P1"], m["P2"], m["P3"], m["rake"])
    ctr = (P1 + P2 + P3) / 3
    KSo, KNo = orc.stiffness_tri(P1, P2, P3, m["rake"], ctr, nrm, slip, MU, LAM)
    KSg, KNg = qd.StiffnessMatrix_ByParts_Calculation_Tri(P1, P2, P3, m["rake"], ctr, MU, LAM, nrm, slip)
    assert np.isfinite(KSg).all() and np.isfinite(KNg).all()
    sc = np.abs(KSo).max(axis=1, keepdims=True)
    assert (np.abs(KSg - KSo) / sc).max() <= 1e-10
    assert (np.abs(KNg - KNo) / sc).max() <= 1e-10
    assert np.all(np.diag(KSo) < 0)
    op = qd.DenseOperator.from_triangles(P1, P2, P3, m["rake"], ctr, nrm, slip, MU, LAM)
    ks, dg = op.kself()
    np.testing.assert_allclose(dg, np.diag(KSg), rtol=1e-14)
    op.close()
