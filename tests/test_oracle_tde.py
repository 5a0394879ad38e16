"""Pins for the oracle's triangular-dislocation restatement (oracle/tde.c,
Nikkhoo & Walter 2015 as scripts/Functions_TDstressHS.jl states it):

* a rectangle is the sum of its two triangles: TDstressHS of the two halves
  equals the stress from the independently pinned Okada DC3D oracle (to
  1e-11 of the field's scale) for strike, dip and tensile slip at several
  dips - this exercises the main, image and free-surface-correction terms;
* the free surface z = 0 is traction-free;
* equilibrium div(sigma) = 0 by finite differences;
* points on a triangle edge give NaN (trimodefinder's configuration 0).
"""
import numpy as np
import pytest

MU, NU = 32e9, 0.25
LAM = 2 * MU * NU / (1 - 2 * NU)
ALPHA = (LAM + MU) / (LAM + 2 * MU)


def okada_stress(orc, X, Y, Z, depth, dip, L, W, disl):
    U, _ = orc.dc3d(ALPHA, X, Y, Z, depth, dip, L, L, W, W, *disl)
    uxx, uyx, uzx, uxy, uyy, uzy, uxz, uyz, uzz = U[:, 3:12].T
    tr = uxx + uyy + uzz
    return np.stack([LAM * tr + 2 * MU * uxx, LAM * tr + 2 * MU * uyy, LAM * tr + 2 * MU * uzz,
                     MU * (uxy + uyx), MU * (uxz + uzx), MU * (uyz + uzy)], 1)


def rect_as_triangles(depth, dip, L, W):
    d = np.deg2rad(dip)
    sdip = np.array([0, np.cos(d), np.sin(d)])
    c = np.array([0, 0, -depth])
    A = c - [L, 0, 0] - W * sdip
    B = c + [L, 0, 0] - W * sdip
    C = c + [L, 0, 0] + W * sdip
    D = c - [L, 0, 0] + W * sdip
    return sdip, np.array([0, np.sin(d), -np.cos(d)]), [(A, B, C), (A, C, D)]


def tde_sum(orc, X, Y, Z, tris, b, tensile_sign):
    tot = np.zeros((X.size, 6))
    for P1, P2, P3 in tris:
        Vn = np.cross(P2 - P1, P3 - P1)
        Vn /= np.linalg.norm(Vn)
        Vs = np.cross([0, 0, 1.0], Vn)
        Vs /= np.linalg.norm(Vs)
        Vd = np.cross(Vn, Vs)
        S, _ = orc.tdstress_hs(X, Y, Z, P1, P2, P3, b @ Vs, b @ Vd, tensile_sign * (b @ Vn), MU, LAM)
        tot += S
    return tot


@pytest.mark.parametrize("dip", [90.0, 60.0, 35.0, 15.0])
@pytest.mark.parametrize("disl", [(1, 0, 0), (0, 1, 0), (0, 0, 1)])
def test_rectangle_equals_two_triangles(orc, dip, disl):
    rng = np.random.default_rng(0)
    n = 60
    X, Y, Z = rng.uniform(-3000, 3000, n), rng.uniform(-3000, 3000, n), -rng.uniform(0, 4000, n)
    depth, L, W = 1500.0, 500.0, 300.0
    sdip, nrm, tris = rect_as_triangles(depth, dip, L, W)
    b = disl[0] * np.array([1.0, 0, 0]) + disl[1] * sdip + disl[2] * nrm
    # The triangles' normal (B-A)x(C-A) is -nrm: Okada's shear slips are the
    # hanging wall's motion and its opening is along +nrm, so in the
    # triangles' frame the opening component changes sign.
    S = tde_sum(orc, X, Y, Z, tris, b, tensile_sign=-1.0)
    S_ok = okada_stress(orc, X, Y, Z, depth, dip, L, W, disl)
    assert np.abs(S - S_ok).max() <= 1e-11 * np.abs(S_ok).max()


@pytest.mark.parametrize("slip", [(1, 0, 0), (0, 1, 0), (0, 0, 1), (0.3, -0.8, 0.2)])
def test_free_surface_traction_free(orc, slip):
    P1, P2, P3 = np.array([-400.0, 100, -300]), np.array([350.0, -150, -900]), np.array([0.0, 300, -1200])
    rng = np.random.default_rng(1)
    X, Y = rng.uniform(-3000, 3000, 200), rng.uniform(-3000, 3000, 200)
    S, _ = orc.tdstress_hs(X, Y, np.zeros(200), P1, P2, P3, *slip, MU, LAM)
    scale = np.abs(S[:, :4]).max()
    assert np.abs(S[:, 2]).max() <= 1e-9 * scale
    assert np.abs(S[:, 4]).max() <= 1e-9 * scale
    assert np.abs(S[:, 5]).max() <= 1e-9 * scale


def test_equilibrium(orc):
    P1, P2, P3 = np.array([-400.0, 100, -300]), np.array([350.0, -150, -900]), np.array([0.0, 300, -1200])
    rng = np.random.default_rng(2)
    P = np.stack([rng.uniform(-1500, 1500, 40), rng.uniform(-1500, 1500, 40), -rng.uniform(100, 2500, 40)], 1)
    Vn = np.cross(P2 - P1, P3 - P1)
    Vn /= np.linalg.norm(Vn)
    P = P[np.abs((P - P1) @ Vn) > 100]
    h = 0.05
    comps = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (1, 0): 3, (0, 2): 4, (2, 0): 4, (1, 2): 5, (2, 1): 5}
    div = np.zeros((P.shape[0], 3))
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        Sp, _ = orc.tdstress_hs(*(P + e).T, P1, P2, P3, 0.4, 0.7, -0.2, MU, LAM)
        Sm, _ = orc.tdstress_hs(*(P - e).T, P1, P2, P3, 0.4, 0.7, -0.2, MU, LAM)
        for i in range(3):
            div[:, i] += (Sp[:, comps[(i, j)]] - Sm[:, comps[(i, j)]]) / (2 * h)
    S0, _ = orc.tdstress_hs(*P.T, P1, P2, P3, 0.4, 0.7, -0.2, MU, LAM)
    assert np.abs(div).max() <= 1e-5 * np.abs(S0).max() / 100


def test_edge_points_are_nan(orc):
    P1, P2, P3 = np.array([-400.0, 0, -300]), np.array([400.0, 0, -300]), np.array([0.0, 0, -900])
    mid = (P1 + P2) / 2
    S, _ = orc.tdstress_hs([mid[0]], [mid[1]], [mid[2]], P1, P2, P3, 1.0, 0.0, 0.0, MU, LAM)
    assert np.all(np.isnan(S))
    S, _ = orc.tdstress_hs([0.0], [50.0], [-500.0], P1, P2, P3, 1.0, 0.0, 0.0, MU, LAM)
    assert np.all(np.isfinite(S))


def test_rotverts_frames(orc):
    """RotVerts_UnitVectors (Functions_BuildInputFile.jl:338-377): vertices
    reordered so the normal has n_y > 0; unit, orthogonal normal/strike/dip;
    horizontal strike; slip = strike cos(rake) + dip sin(rake)."""
    from qdfn import configs
    m = configs.rough_fault(nx=6, nz=5, h=60.0, ratio=2e-2)
    rake = np.linspace(-170, 170, m["P1"].shape[0])
    P1, P2, P3, nrm, st, dp, slip = orc.rotverts_unitvectors(m["P1"], m["P2"], m["P3"], rake)
    c = np.cross(P2 - P1, P3 - P1)
    np.testing.assert_allclose(c / np.linalg.norm(c, axis=1)[:, None], nrm, atol=1e-14)
    assert (nrm[:, 1] > 0).all()
    for v in (nrm, st, dp, slip):
        np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-14)
    assert np.abs((nrm * st).sum(1)).max() < 1e-14 and np.abs((nrm * dp).sum(1)).max() < 1e-14
    assert np.abs(st[:, 2]).max() < 1e-14
    r = np.deg2rad(rake)[:, None]
    np.testing.assert_allclose(slip, st * np.cos(r) + dp * np.sin(r), atol=1e-14)


def test_triangle_stiffness_is_traction_projection(orc):
    """StiffnessMatrix_ByParts_Calculation_Tri (:551-594): column s equals
    TDstressHS of source s (Ss = -cos rake, Ds = sin rake) at the receiver
    centroids, traction t = sigma n, projected on n (K_N) and on the slip
    vector (K_S); self-stiffness negative."""
    from qdfn import configs
    m = configs.rough_fault(nx=4, nz=3, h=60.0, ratio=2e-2)
    rake = np.full(m["P1"].shape[0], 30.0)
    P1, P2, P3, nrm, st, dp, slip = orc.rotverts_unitvectors(m["P1"], m["P2"], m["P3"], rake)
    ctr = (P1 + P2 + P3) / 3
    KS, KN = orc.stiffness_tri(P1, P2, P3, rake, ctr, nrm, slip, MU, LAM)
    assert (np.diag(KS) < 0).all()
    for s in (0, 5, 17):
        S, _ = orc.tdstress_hs(*ctr.T, P1[s], P2[s], P3[s], -np.cos(np.deg2rad(30.0)), np.sin(np.deg2rad(30.0)),
                               0.0, MU, LAM)
        sig = np.zeros((ctr.shape[0], 3, 3))
        for (i, j), k in {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (0, 2): 4, (1, 2): 5}.items():
            sig[:, i, j] = sig[:, j, i] = S[:, k]
        t = np.einsum("rij,rj->ri", sig, nrm)
        np.testing.assert_allclose(KN[:, s], (t * nrm).sum(1), rtol=1e-12, atol=1e-12 * np.abs(KN).max())
        np.testing.assert_allclose(KS[:, s], (t * slip).sum(1), rtol=1e-12, atol=1e-12 * np.abs(KS).max())
