"""Stiffness assembly on the GPU with the reference's function names.

Okada_DC3D_Vector                     scripts/Functions_OKADA3D.jl:464
StiffnessMatrix_ByParts_Calculation_Rec  scripts/Functions_BuildInputFile.jl:386
BuildMatrixByParts_Even_Rec           scripts/Functions_BuildInputFile.jl:516
TDstressHS                            scripts/Functions_TDstressHS.jl:2
RotVerts_UnitVectors                  scripts/Functions_BuildInputFile.jl:338
StiffnessMatrix_ByParts_Calculation_Tri  scripts/Functions_BuildInputFile.jl:551
BuildMatrixByParts_Even_Tri           scripts/Functions_BuildInputFile.jl:599

Arrays follow the reference's shapes (Julia column-major semantics): segment
tables are (N, 19) with the Input_Segment columns, stiffness matrices are
(N_receiver, N_source).  All arithmetic runs in the HIP kernels of
libqdfn_hip.so; nothing here computes a kernel value on the CPU.
"""
import numpy as np

from ._lib import check, lib, ptr


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _fortran(a):
    return np.asfortranarray(a, dtype=np.float64)


def Okada_DC3D_Vector(ALPHA, X, Y, Z, DEPTH, DIP_Angle, AL1, AL2, AW1, AW2, DISL1, DISL2, DISL3):
    """Returns UX, UY, UZ, UXX, UYX, UZX, UXY, UYY, UZY, UXZ, UYZ, UZZ, IRET."""
    X, Y, Z = _f64(X).ravel(), _f64(Y).ravel(), _f64(Z).ravel()
    n = X.size
    if Y.size != n or Z.size != n:
        raise ValueError("X, Y, Z must have the same length")
    U = np.zeros(12 * n)
    iret = np.zeros(n)
    check(lib().qdfn_okada_dc3d(float(ALPHA), n, ptr(X), ptr(Y), ptr(Z), float(DEPTH), float(DIP_Angle),
                                float(AL1), float(AL2), float(AW1), float(AW2), float(DISL1), float(DISL2),
                                float(DISL3), ptr(U), ptr(iret)))
    U = U.reshape(12, n)
    return tuple(U[k].copy() for k in range(12)) + (iret,)


def StiffnessMatrix_ByParts_Calculation_Rec(Input_SegmentSource, Input_SegmentReceiver, ShearModulus,
                                            PoissonRatio, CurrentPart=1, TotalParts=1):
    """Shear and normal stiffness (N_r x N_s) of rectangular source segments on
    receiver segments; returns (StiffnessMatrixShear, StiffnessMatrixNormal)."""
    src = _fortran(Input_SegmentSource)
    rcv = _fortran(Input_SegmentReceiver)
    if src.ndim != 2 or rcv.ndim != 2 or src.shape[1] < 8 or rcv.shape[1] < 8:
        raise ValueError("segment tables need at least 8 columns")
    ns, nr = src.shape[0], rcv.shape[0]
    KS = np.zeros((nr, ns), order="F")
    KN = np.zeros((nr, ns), order="F")
    check(lib().qdfn_okada_stiffness_rect(ptr(src), ns, ns, ptr(rcv), nr, nr, float(ShearModulus),
                                          float(PoissonRatio), ptr(KS), ptr(KN), nr))
    return KS, KN


def BuildMatrixByParts_Even_Rec(FaultCount, Input_Segment, ShearModulus, PoissonRatio):
    """Whole N x N rectangle stiffness.  The reference tiles the matrix into
    ~2000-element parts to bound host memory; the entries do not depend on the
    tiling, so one device launch covers the whole matrix here."""
    seg = _fortran(Input_Segment)
    if seg.shape[0] != FaultCount:
        raise ValueError("FaultCount does not match Input_Segment")
    return StiffnessMatrix_ByParts_Calculation_Rec(seg, seg, ShearModulus, PoissonRatio)


def TDstressHS(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, mu, lam):
    """Returns (Stress, Strain), each (N, 6): xx, yy, zz, xy, xz, yz.  Unlike the
    reference, P1..P3 are not modified in place."""
    X, Y, Z = _f64(X).ravel(), _f64(Y).ravel(), _f64(Z).ravel()
    n = X.size
    if np.max(Z, initial=-np.inf) > 0 or P1[2] > 0 or P2[2] > 0 or P3[2] > 0:
        raise ValueError("Half-space solution: Z coordinates must be negative!")
    S = np.zeros(6 * n)
    E = np.zeros(6 * n)
    check(lib().qdfn_tde_stress_hs(n, ptr(X), ptr(Y), ptr(Z), ptr(_f64(P1)), ptr(_f64(P2)), ptr(_f64(P3)),
                                   float(Ss), float(Ds), float(Ts), float(mu), float(lam), ptr(S), ptr(E)))
    return S.reshape(6, n).T.copy(), E.reshape(6, n).T.copy()


def RotVerts_UnitVectors(Input_Segment, FaultCount, Rake):
    """Returns P1, P2, P3 (vertex order fixed so that normal_y > 0),
    UnitVector_Normal, UnitVector_StrikeSlip, UnitVector_DipSlip, UnitVector_Slip."""
    seg = np.asarray(Input_Segment, dtype=np.float64)
    P = [_fortran(seg[:, 3 * k:3 * k + 3]).copy(order="F") for k in range(3)]
    outs = [np.zeros((FaultCount, 3), order="F") for _ in range(4)]
    check(lib().qdfn_rotverts_unitvectors(FaultCount, ptr(P[0]), ptr(P[1]), ptr(P[2]), ptr(_f64(Rake)),
                                          *[ptr(o) for o in outs]))
    return (P[0], P[1], P[2], *outs)


def StiffnessMatrix_ByParts_Calculation_Tri(P1, P2, P3, Rake, FaultCenter, ShearModulus, lam,
                                            UnitVector_Normal, UnitVector_Slip):
    P1, P2, P3 = _fortran(P1), _fortran(P2), _fortran(P3)
    ctr, nrm, slip = _fortran(FaultCenter), _fortran(UnitVector_Normal), _fortran(UnitVector_Slip)
    ns, nr = P1.shape[0], ctr.shape[0]
    KS = np.zeros((nr, ns), order="F")
    KN = np.zeros((nr, ns), order="F")
    check(lib().qdfn_tde_stiffness_tri(ns, ptr(P1), ptr(P2), ptr(P3), ptr(_f64(Rake)), nr, ptr(ctr), ptr(nrm),
                                       ptr(slip), float(ShearModulus), float(lam), ptr(KS), ptr(KN), nr))
    return KS, KN


def BuildMatrixByParts_Even_Tri(P1, P2, P3, Rake, FaultCenter, UnitVector_Normal, UnitVector_Slip, FaultCount,
                                ShearModulus, PoissonRatio):
    lam = 2 * ShearModulus * PoissonRatio / (1 - 2 * PoissonRatio)
    return StiffnessMatrix_ByParts_Calculation_Tri(P1, P2, P3, Rake, FaultCenter, ShearModulus, lam,
                                                   UnitVector_Normal, UnitVector_Slip)


def ReduceTooStrongInteraction(StrongInteractionCriteriaMultiple, FaultCount, StiffnessMatrixShear,
                               StiffnessMatrixNormal):
    """ReduceTooStrongInteraction (scripts/Functions_Solvers.jl:586-607) on the
    dense matrices before they go to the device: every off-diagonal shear entry
    larger in magnitude than StrongInteractionCriteriaMultiple |K_S,ii| of its
    receiver row scales its shear and normal entries by that bound over its
    magnitude (the same arithmetic, elementwise).  Modifies and returns the
    matrices, as the reference does."""
    KS, KN = StiffnessMatrixShear, StiffnessMatrixNormal
    n = int(FaultCount)
    c = float(StrongInteractionCriteriaMultiple)
    lim = np.abs(np.diagonal(KS)[:n]).reshape(-1, 1) * c  # |K_ii| * multiple, per receiver row
    sub = KS[:n, :n]
    hit = lim < np.abs(sub)
    np.fill_diagonal(hit, False)
    ratio = np.ones_like(sub)
    ratio[hit] = (np.broadcast_to(lim, sub.shape)[hit]) / np.abs(sub[hit])
    KS[:n, :n] = np.where(hit, sub * ratio, sub)
    KN[:n, :n] = np.where(hit, KN[:n, :n] * ratio, KN[:n, :n])
    print(f"{int(hit.sum())}/{n ** 2} interaction is smoothened")
    return KS, KN
