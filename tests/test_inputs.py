"""Input format layer (scripts/Functions_BuildInputFile.jl:82-333) on the
reference's own example geometry files (tests/golden/, copied data files):
header values, element counts (ceil(length / MaxLeng) per direction), loading
faults last, subdivision geometry (segment centres tile the bulk rectangle in
its plane), rake from the LR flag, depth-dependent normal stress, exact degree
trigonometry, and ReadSegmentInput's column mapping for rectangles and
triangles."""
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _bulk(path):
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    return np.array([float(v) for v in lines[1]]), np.array([[float(v) for v in ln] for ln in lines[3:]])


def test_degree_trig_exact():
    from qdfn import inputs
    x = np.array([0.0, 90, 180, 270, 360, -90, 450, 30, 45])
    s, c = inputs.sind(x), inputs.cosd(x)
    assert list(s[:7]) == [0.0, 1.0, 0.0, -1.0, 0.0, -1.0, 1.0]
    assert list(c[:7]) == [1.0, 0.0, -1.0, 0.0, 1.0, 0.0, 0.0]
    np.testing.assert_allclose(s[7:], np.sin(np.deg2rad(x[7:])), rtol=1e-15)


def test_read_bulk_input_rectangles():
    from qdfn import inputs
    path = os.path.join(GOLD, "Input_BulkFaultGeometry.txt")
    params, B = _bulk(path)
    (seg, L, mu, nu, rho, sw, drop, dropm, minns, RorT, n) = inputs.ReadBulkInput(path)
    assert RorT == "R"
    assert (mu, nu, rho, sw, drop, dropm, minns) == tuple(params[[1, 2, 3, 0, 4, 5, 6]])
    counts = np.ceil(B[:, 3] / B[:, 17]) * np.ceil(B[:, 4] / B[:, 17])
    assert n == int(counts.sum()) and seg.shape == (n, 19)
    assert L == int(counts[B[:, 16] > 0].sum()) == 5
    assert (seg[n - L:, 15] > 0).all() and (seg[:n - L, 15] == 0).all()  # loading faults last
    # LR flag 1 with switch 1 -> rake 180
    assert set(np.unique(seg[:, 7])) <= {0.0, 180.0}
    # every bulk fault: segment lengths, centres average to the bulk centre, in its plane
    for k in np.unique(seg[:, 16]).astype(int):
        rows = seg[seg[:, 16] == k]
        b = B[np.argsort(B[:, 16], kind="stable")][k - 1]
        nl, nw = np.ceil(b[3] / b[17]), np.ceil(b[4] / b[17])
        assert rows.shape[0] == nl * nw
        np.testing.assert_allclose(rows[:, 3], b[3] / nl)
        np.testing.assert_allclose(rows[:, 4], b[4] / nw)
        np.testing.assert_allclose(rows[:, 0:3].mean(0), b[0:3], atol=1e-6 * max(b[3], b[4]))
        st, dp = np.deg2rad(b[5]), np.deg2rad(b[6])
        nrm = np.array([np.sin(st) * np.sin(dp), -np.cos(st) * np.sin(dp), np.cos(dp)])
        rel = rows[:, 0:3] - b[0:3]
        rel[:, 2] = -rel[:, 2]
        assert np.abs(rel @ nrm).max() <= 1e-9 * max(b[3], b[4])
        np.testing.assert_allclose(rows[:, 14], b[14] + b[15] * rows[:, 2])


def test_read_bulk_input_ridgecrest():
    from qdfn import inputs
    path = os.path.join(GOLD, "Input_BulkFaultGeometry_Ridgecrest.txt")
    params, B = _bulk(path)
    seg, L, *_, RorT, n = inputs.ReadBulkInput(path)
    assert RorT == "R" and params[0] == 0.0
    assert n == int((np.ceil(B[:, 3] / B[:, 17]) * np.ceil(B[:, 4] / B[:, 17])).sum())
    # switch 0: the rake column is used as given
    assert set(np.round(np.unique(seg[:, 7]), 9)) <= set(np.round(np.unique(B[:, 7]), 9))


def test_read_segment_input():
    from qdfn import inputs
    seg, L, *_, RorT, n = inputs.ReadBulkInput(os.path.join(GOLD, "Input_BulkFaultGeometry.txt"))
    out = inputs.ReadSegmentInput(seg, n, RorT)
    ctr, a, b, dc, th, vi, fr, sig, vc, bulk, Lst, Ldp, st, dp, rake, Lb, Wb, nz = out
    np.testing.assert_array_equal(ctr, seg[:, 0:3])
    for arr, col in ((a, 8), (b, 9), (dc, 10), (th, 11), (vi, 12), (fr, 13), (sig, 14), (vc, 15), (bulk, 16),
                     (Lst, 3), (Ldp, 4), (st, 5), (dp, 6), (rake, 7), (Lb, 17), (Wb, 18)):
        np.testing.assert_array_equal(arr, seg[:, col])
    assert nz == 0
    # triangles: 20-column rows P1 P2 P3 rake a b Dc th vi fr sig0 grad vconst
    T = np.zeros((2, 20))
    T[0, 0:9] = [0, 0, -100, 100, 0, -100, 0, 0, -200]
    T[1, 0:9] = [0, 0, -100, 50, 30, -300, 100, 0, -100]
    T[:, 9] = [0, 90]
    T[:, 10:16] = [[0.01, 0.015, 1e-3, 1e6, 1e-9, 0.6]] * 2
    T[:, 16], T[:, 17], T[:, 18] = 5e7, 1e4, 0.0
    ctr, a, b, dc, th, vi, fr, sig, vc, bulk, Lst, Ldp, st, dp, rake, Lb, Wb, nz = inputs.ReadSegmentInput(T, 2, "T")
    np.testing.assert_allclose(ctr, (T[:, 0:3] + T[:, 3:6] + T[:, 6:9]) / 3)
    np.testing.assert_allclose(sig, 5e7 - 1e4 * ctr[:, 2])
    np.testing.assert_allclose(Lst, [100.0, 200.0])  # largest |component| over the three edge vectors
    np.testing.assert_array_equal(rake, [0, 90])
