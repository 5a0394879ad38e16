"""Input format layer: the reference's bulk-geometry reader and its subdivision.

Mirrors scripts/Functions_BuildInputFile.jl:
  BulkToSegment      :82-178   bulk rectangles -> Input_Segment (19 columns)
  LRtoRake           :181-213  LR / RN flag -> rake angle
  ReadBulkInput      :216-280  Input_BulkFaultGeometry.txt -> Input_Segment, header values
  ReadSegmentInput   :283-333  Input_Segment -> per-element arrays

Host code (it runs once, before assembly).  Degree trigonometry is exact at
multiples of 90 degrees, like Julia's sind / cosd.
"""
import numpy as np


def sind(x):
    """sin of degrees with exact zeros and ones at multiples of 90."""
    x = np.asarray(x, dtype=np.float64)
    r = np.fmod(x, 360.0)
    out = np.sin(np.deg2rad(r))
    q = r / 90.0
    exact = q == np.round(q)
    k = np.mod(np.round(q), 4)
    return np.where(exact, np.choose(k.astype(int) % 4, [0.0, 1.0, 0.0, -1.0]) * (1.0 + 0 * r), out)


def cosd(x):
    """cos of degrees with exact zeros and ones at multiples of 90."""
    x = np.asarray(x, dtype=np.float64)
    r = np.fmod(x, 360.0)
    out = np.cos(np.deg2rad(r))
    q = r / 90.0
    exact = q == np.round(q)
    k = np.mod(np.round(q), 4)
    return np.where(exact, np.choose(k.astype(int) % 4, [1.0, 0.0, -1.0, 0.0]) * (1.0 + 0 * r), out)


def LRtoRake(Switch_StrikeSlip_or_ReverseNormal, Input_Bulk):
    """Column 8 (LR/RN flag, -1 or +1) -> rake angle in degrees (:181-213)."""
    B = np.array(Input_Bulk, dtype=np.float64, copy=True)
    if Switch_StrikeSlip_or_ReverseNormal == 1:
        B[:, 7] = np.where(B[:, 7] == -1.0, 0.0, 180.0)
    elif Switch_StrikeSlip_or_ReverseNormal == 2:
        shallow = B[:, 6] < 90.0
        ll = B[:, 7] == -1.0
        B[:, 7] = np.where(shallow, np.where(ll, 90.0, 270.0), np.where(ll, 270.0, 90.0))
    return B


def BulkToSegment(Input_Bulk, counts=None):
    """Subdivide bulk rectangles (18 columns) into Input_Segment rows (19 columns):
    1-3 centre (z = depth), 4 strike length, 5 dip length, 6 strike, 7 dip, 8 rake,
    9 a, 10 b, 11 Dc, 12 Theta_i, 13 V_i, 14 Friction_i, 15 normal stress
    (Sig0 + SigGrad * depth), 16 V_Const, 17 bulk index, 18-19 bulk lengths.
    Element counts along strike and dip are ceil(length / column 18) as in
    the reference, unless `counts` gives them per bulk fault."""
    B = np.asarray(Input_Bulk, dtype=np.float64)
    if counts is None:
        nl = np.ceil(B[:, 3] / B[:, 17]).astype(int)
        nw = np.ceil(B[:, 4] / B[:, 17]).astype(int)
    else:
        nl = np.array([c[0] for c in counts], dtype=int)
        nw = np.array([c[1] for c in counts], dtype=int)
    rows = []
    for k in range(B.shape[0]):
        L, W = B[k, 3], B[k, 4]
        st, dp = B[k, 5], B[k, 6]
        cs, ss, cd, sd = float(cosd(st)), float(sind(st)), float(cosd(dp)), float(sind(dp))
        Rs = np.array([[cs, -ss, 0.0], [ss, cs, 0.0], [0.0, 0.0, 1.0]])
        Rd = np.array([[1.0, 0.0, 0.0], [0.0, cd, -sd], [0.0, sd, cd]])
        R = Rs @ Rd
        for li in range(1, nl[k] + 1):
            for wi in range(1, nw[k] + 1):
                if nl[k] * nw[k] == 1:
                    c = B[k, 0:3].copy()
                else:
                    loc = np.array([-L / 2 + L * (2 * li - 1) / nl[k] / 2, -W / 2 + W * (2 * wi - 1) / nw[k] / 2, 0.0])
                    c = R @ loc + np.array([B[k, 0], B[k, 1], -B[k, 2]])
                    c[2] = -c[2]
                row = np.zeros(19)
                row[0:3] = c
                row[3] = L / nl[k]
                row[4] = W / nw[k]
                row[5:14] = B[k, 5:14]
                row[14] = B[k, 14] + B[k, 15] * c[2]
                row[15] = B[k, 16]
                row[16] = k + 1
                row[17] = L
                row[18] = W
                rows.append(row)
    return np.array(rows)


def _read_table(path):
    """readdlm of the bulk file: header line, parameter line, header line, rows."""
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    params = np.array([float(v) for v in lines[1]])
    body = np.array([[float(v) for v in ln] for ln in lines[3:]])
    return params, body


def ReadBulkInput(InputBulkFileName):
    """Returns (Input_Segment, LoadingFaultCount, ShearModulus, PoissonRatio,
    RockDensity, Switch_StrikeSlip_or_ReverseNormal, DropCrit,
    DropCritNormalStressMultiplier, MinimumNS, RorT, FaultCount) as :216-280:
    18 columns = rectangles (subdivided by BulkToSegment), 20 = triangles;
    rows sorted by V_Const so the loading faults come last."""
    params, B = _read_table(InputBulkFileName)
    if B.shape[1] == 18:
        RorT = "R"
    elif B.shape[1] == 20:
        RorT = "T"
    else:
        raise ValueError("Input Bulk Fault Geometry file should have 18 or 20 columns")
    Switch, ShearModulus, PoissonRatio, RockDensity, DropCrit, DropMult, MinimumNS = params[:7]
    if RorT == "R":
        if np.all(B[:, 2] < 0):
            B[:, 2] = -B[:, 2]
        B = B[np.argsort(B[:, 16], kind="stable")]
        B = LRtoRake(Switch, B)
        seg = BulkToSegment(B)
        seg = seg[np.argsort(seg[:, 15], kind="stable")]
        L = int(np.sum(seg[:, 15] > 0))
    else:
        B = B[np.argsort(B[:, 18], kind="stable")]
        seg = B[np.argsort(B[:, 18], kind="stable")]
        L = int(np.sum(seg[:, 18] > 0))
    return (seg, L, ShearModulus, PoissonRatio, RockDensity, Switch, DropCrit, DropMult, MinimumNS, RorT,
            seg.shape[0])


def ReadSegmentInput(Input_Segment, FaultCount, RorT):
    """Per-element arrays of an Input_Segment table (:283-333): returns
    (FaultCenter, Fault_a, Fault_b, Fault_Dc, Fault_Theta_i, Fault_V_i,
    Fault_Friction_i, Fault_NormalStress, Fault_V_Const, Fault_BulkIndex,
    FaultLengthStrike, FaultLengthDip, FaultStrikeAngle, FaultDipAngle,
    FaultRakeAngle, FaultLengthStrike_Bulk, FaultLengthDip_Bulk,
    NormalStiffnessZero)."""
    S = np.asarray(Input_Segment, dtype=np.float64)
    if RorT == "R":
        FaultCenter = S[:, 0:3].copy()
        vals = [S[:, k].copy() for k in range(8, 17)]  # a b Dc Theta_i V_i Friction_i NormalStress V_Const Bulk
        Lst, Ldp, st, dp, rake = S[:, 3].copy(), S[:, 4].copy(), S[:, 5].copy(), S[:, 6].copy(), S[:, 7].copy()
        Lst_b, Ldp_b = S[:, 17].copy(), S[:, 18].copy()
        NormalStiffnessZero = 0
        return (FaultCenter, *vals, Lst, Ldp, st, dp, rake, Lst_b, Ldp_b, NormalStiffnessZero)
    P1, P2, P3 = S[:, 0:3], S[:, 3:6], S[:, 6:9]
    rake = S[:, 9].copy()
    FaultCenter = (P1 + P2 + P3) / 3
    a, b, dc, th, vi, fr = (S[:, k].copy() for k in range(10, 16))
    sig = S[:, 16] - S[:, 17] * FaultCenter[:, 2]
    vc = S[:, 18].copy()
    bulk = rake * 0
    Lst = np.max(np.abs(np.hstack([P1 - P2, P2 - P3, P1 - P3])), axis=1)
    Ldp = Lst.copy()
    zeros = np.zeros(FaultCount)
    return (FaultCenter, a, b, dc, th, vi, fr, sig, vc, bulk, Lst, Ldp, zeros, zeros.copy(), rake, Lst.copy(),
            Ldp.copy(), 0)
