"""Pins for the oracle's Okada DC3D restatement (oracle/okada.c), from what the
elasticity of a dislocation in a half-space fixes independently of the
formula's algebra:

* the displacement jump across the fault plane equals the slip vector
  (Volterra dislocation), component by component, for all three slip types;
* the returned displacement gradients equal finite differences of the returned
  displacements (the two are separate expressions in Okada's solution);
* the free surface z = 0 is traction-free;
* the stress field is in equilibrium (div sigma = 0 by finite differences);
* symmetry of a vertical strike-slip source (u_x odd in y, even in x);
* stiffness: reciprocity for identical coplanar elements, invariance under
  horizontal translation and rotation of the whole system.
A dropped term, a sign error or a transposed gradient index fails at least one.
"""
import numpy as np
import pytest

NU, MU = 0.25, 32e9
LAM = 2 * MU * NU / (1 - 2 * NU)
ALPHA = (LAM + MU) / (LAM + 2 * MU)
SRC = dict(depth=1000.0, al1=500.0, al2=500.0, aw1=300.0, aw2=300.0)


def dc3d(orc, X, Y, Z, dip, disl):
    U, _ = orc.dc3d(ALPHA, X, Y, Z, SRC["depth"], dip, SRC["al1"], SRC["al2"], SRC["aw1"], SRC["aw2"], *disl)
    return U


@pytest.mark.parametrize("dip", [90.0, 70.0, 45.0, 20.0])
@pytest.mark.parametrize("disl", [(1, 0, 0), (0, 1, 0), (0, 0, 1)])
def test_displacement_jump_equals_slip(orc, dip, disl):
    d = np.deg2rad(dip)
    n = np.array([0.0, np.sin(d), -np.cos(d)])       # fault normal
    s_dip = np.array([0.0, np.cos(d), np.sin(d)])    # up-dip unit vector
    rng = np.random.default_rng(1)
    pts = []
    for _ in range(20):
        x = rng.uniform(-450, 450)
        w = rng.uniform(-250, 250)
        pts.append(np.array([x, 0.0, -SRC["depth"]]) + w * s_dip)
    pts = np.array(pts)
    h = 1e-4
    Up = dc3d(orc, *(pts + h * n).T, dip, disl)
    Um = dc3d(orc, *(pts - h * n).T, dip, disl)
    jump = Up[:, :3] - Um[:, :3]
    # n points into the footwall, so the jump is footwall minus hanging wall:
    # Okada's positive DISL1 is left-lateral (hanging wall moves +x), positive
    # DISL2 is reverse (hanging wall moves up-dip), positive DISL3 is opening.
    b = -disl[0] * np.array([1.0, 0, 0]) - disl[1] * s_dip + disl[2] * n
    np.testing.assert_allclose(jump, np.broadcast_to(b, jump.shape), atol=2e-6)


@pytest.mark.parametrize("dip", [90.0, 60.0, 30.0])
@pytest.mark.parametrize("disl", [(1, 0, 0), (0, 1, 0), (0, 0, 1)])
def test_gradients_are_derivatives_of_displacement(orc, dip, disl):
    rng = np.random.default_rng(2)
    P = np.stack([rng.uniform(-2000, 2000, 50), rng.uniform(-2000, 2000, 50), -rng.uniform(50, 3000, 50)], 1)
    # keep away from the fault plane
    d = np.deg2rad(dip)
    n = np.array([0.0, np.sin(d), -np.cos(d)])
    P = P[np.abs((P - [0, 0, -SRC["depth"]]) @ n) > 50]
    U = dc3d(orc, *P.T, dip, disl)
    h = 1e-2
    for ax, cols in ((0, (3, 4, 5)), (1, (6, 7, 8)), (2, (9, 10, 11))):
        e = np.zeros(3)
        e[ax] = h
        Up = dc3d(orc, *(P + e).T, dip, disl)
        Um = dc3d(orc, *(P - e).T, dip, disl)
        fd = (Up[:, :3] - Um[:, :3]) / (2 * h)
        scale = np.abs(U[:, 3:]).max()
        np.testing.assert_allclose(U[:, cols], fd, atol=1e-6 * scale + 1e-14)


def _stress(U):
    g = U[:, 3:12]
    uxx, uyx, uzx, uxy, uyy, uzy, uxz, uyz, uzz = g.T
    tr = uxx + uyy + uzz
    return dict(xx=LAM * tr + 2 * MU * uxx, yy=LAM * tr + 2 * MU * uyy, zz=LAM * tr + 2 * MU * uzz,
                xy=MU * (uxy + uyx), xz=MU * (uxz + uzx), yz=MU * (uyz + uzy))


@pytest.mark.parametrize("dip", [90.0, 50.0])
@pytest.mark.parametrize("disl", [(1, 0, 0), (0, 1, 0), (0, 0, 1)])
def test_free_surface_is_traction_free(orc, dip, disl):
    rng = np.random.default_rng(3)
    X, Y = rng.uniform(-3000, 3000, 200), rng.uniform(-3000, 3000, 200)
    S = _stress(dc3d(orc, X, Y, np.zeros_like(X), dip, disl))
    scale = max(np.abs(S["xx"]).max(), np.abs(S["xy"]).max())
    for k in ("xz", "yz", "zz"):
        assert np.abs(S[k]).max() <= 1e-9 * scale, (k, np.abs(S[k]).max(), scale)


@pytest.mark.parametrize("dip", [90.0, 40.0])
@pytest.mark.parametrize("disl", [(1, 0, 0), (0, 1, 0)])
def test_stress_equilibrium(orc, dip, disl):
    rng = np.random.default_rng(4)
    P = np.stack([rng.uniform(-1500, 1500, 30), rng.uniform(-1500, 1500, 30), -rng.uniform(200, 2500, 30)], 1)
    d = np.deg2rad(dip)
    n = np.array([0.0, np.sin(d), -np.cos(d)])
    P = P[np.abs((P - [0, 0, -SRC["depth"]]) @ n) > 100]
    h = 0.05
    div = np.zeros((P.shape[0], 3))
    comps = {(0, 0): "xx", (1, 1): "yy", (2, 2): "zz", (0, 1): "xy", (1, 0): "xy", (0, 2): "xz", (2, 0): "xz",
             (1, 2): "yz", (2, 1): "yz"}
    for j in range(3):
        e = np.zeros(3)
        e[j] = h
        Sp = _stress(dc3d(orc, *(P + e).T, dip, disl))
        Sm = _stress(dc3d(orc, *(P - e).T, dip, disl))
        for i in range(3):
            div[:, i] += (Sp[comps[(i, j)]] - Sm[comps[(i, j)]]) / (2 * h)
    S0 = _stress(dc3d(orc, *P.T, dip, disl))
    # compare with the size of one stress-gradient term
    scale = np.max([np.abs(S0[k]).max() for k in S0]) / 100.0
    assert np.abs(div).max() <= 1e-5 * scale


def test_vertical_strike_slip_symmetry(orc):
    rng = np.random.default_rng(5)
    X, Y, Z = rng.uniform(10, 2000, 100), rng.uniform(10, 2000, 100), -rng.uniform(0, 2500, 100)
    U = dc3d(orc, X, Y, Z, 90.0, (1, 0, 0))
    Uy = dc3d(orc, X, -Y, Z, 90.0, (1, 0, 0))
    Ux = dc3d(orc, -X, Y, Z, 90.0, (1, 0, 0))
    s = np.abs(U[:, 0]).max()
    np.testing.assert_allclose(Uy[:, 0], -U[:, 0], atol=1e-12 * s)   # u_x odd in y
    np.testing.assert_allclose(Ux[:, 0], U[:, 0], atol=1e-12 * s)    # u_x even in x
    np.testing.assert_allclose(Ux[:, 1], -U[:, 1], atol=1e-12 * s)   # u_y odd in x


def _plane(n1, n2, ds, strike=30.0, dip=70.0, center=(100.0, -50.0, 1500.0), rake=180.0):
    from qdfn import configs
    bulk = np.zeros((1, 18))
    bulk[0, :8] = [*center, n1 * ds, n2 * ds, strike, dip, rake]
    bulk[0, 17] = ds
    return configs.BulkToSegment(bulk)


@pytest.mark.parametrize("rake", [180.0, 90.0])
def test_stiffness_reciprocity_along_strike(orc, rake):
    """Identical elements in one row along strike: translation along strike and
    the reflection x -> -x are symmetries of the half-space, so collocated
    K_ij = K_ji exactly (up to rounding)."""
    seg = _plane(10, 1, 100.0, rake=rake)
    KS, KN = orc.stiffness_rect(seg, seg, MU, NU)
    assert np.all(np.diag(KS) < 0)
    np.testing.assert_allclose(KS, KS.T, rtol=1e-9, atol=1e-10 * np.abs(KS).max())
    # normal traction from strike slip is odd under the reflection, from dip slip even
    sign = -1.0 if rake == 180.0 else 1.0
    np.testing.assert_allclose(KN, sign * KN.T, rtol=1e-9, atol=1e-10 * np.abs(KS).max())


def test_stiffness_translation_and_rotation_invariance(orc):
    seg = np.vstack([_plane(4, 3, 120.0), _plane(3, 3, 150.0, strike=75.0, dip=55.0, center=(900.0, 700.0, 2000.0))])
    KS, KN = orc.stiffness_rect(seg, seg, MU, NU)
    moved = seg.copy()
    moved[:, 0] += 12345.0
    moved[:, 1] -= 777.0
    KS2, KN2 = orc.stiffness_rect(moved, moved, MU, NU)
    sc = np.abs(KS).max()
    np.testing.assert_allclose(KS2, KS, atol=1e-9 * sc)
    np.testing.assert_allclose(KN2, KN, atol=1e-9 * sc)
    th = 37.0
    c, s = np.cos(np.deg2rad(th)), np.sin(np.deg2rad(th))
    rot = seg.copy()
    # strike is measured clockwise from the y axis in the reference frame
    # convention used here (x' = x cos + y sin, y' = -x sin + y cos)
    rot[:, 0] = c * seg[:, 0] - s * seg[:, 1]
    rot[:, 1] = s * seg[:, 0] + c * seg[:, 1]
    rot[:, 5] = seg[:, 5] + th
    KS3, KN3 = orc.stiffness_rect(rot, rot, MU, NU)
    np.testing.assert_allclose(KS3, KS, atol=1e-8 * sc)
    np.testing.assert_allclose(KN3, KN, atol=1e-8 * sc)
