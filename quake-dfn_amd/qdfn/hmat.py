"""H-matrix stiffness on the device (the reference's Hmatrix = true path).

Mirrors, over the C ABI of libqdfn_hip.so:
  GroupAndSort_AllLevel + BuildHierarchy   scripts/Functions_Hmatrix.jl:1-212
                                           (RUN1_HMatrix_Structure.jl)  -> hmat_structure
  HmatBuild_R / HmatBuild_T                scripts/Functions_BuildInputFile.jl:643-803
                                           (RUN2_BUILD_INPUT.jl:235-242) -> HMatrix.build_rect / build_tri
  StiffnessTransitionToLoading + HmatSolver_Pararllel
                                           scripts/Functions_Hmatrix.jl:459-477, 631-646
                                           -> HMatrix.kself / HMatrix.matvec
  main_H                                   scripts/Functions_RSFDFN3DMain_H.jl:2 -> main_H
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import HmatParams, check, check_handle, lib, ptr

PI64 = _lib.PI64
PI32 = _lib.PI32
PD = _lib.PD


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def hmat_structure(ctr, dpts, min_elements_to_cut=10, max_levels=8, dist_diam_ratio=1.0,
                   arrange_point=(10000.0, 6000.0, 10000.0), n_loading=0, kmeans_max_iter=200):
    """Cluster tree and block partition (defaults: RUN1_HMatrix_Structure.jl:29-32).
    ctr: element centres (n x 3); dpts: the rows' Input_Segment columns 1:3
    (centre for rectangles, first vertex for triangles).  Returns a dict:
    perm (sorted position -> input index), node_info (begin, end, level, parent;
    1-based ranges), node_geo (centre, diameter = Block_Ctr_Diam), range_SR
    (ElementRange_SR, 1-based), admissible (Admissible)."""
    ctr = np.asfortranarray(ctr, dtype=np.float64)
    dpts = np.asfortranarray(dpts, dtype=np.float64)
    n = ctr.shape[0]
    p = HmatParams()
    p.min_elements_to_cut = int(min_elements_to_cut)
    p.max_levels = int(max_levels)
    p.dist_diam_ratio = float(dist_diam_ratio)
    for k in range(3):
        p.arrange_point[k] = float(arrange_point[k])
    p.n_loading = int(n_loading)
    p.kmeans_max_iter = int(kmeans_max_iter)
    perm = np.zeros(n, dtype=np.int64)
    nn, nb = _lib.I64(), _lib.I64()
    rc = lib().qdfn_hmat_structure(p, n, ptr(ctr), ptr(dpts), perm.ctypes.data_as(PI64), 0, None, None,
                                   C.byref(nn), 0, None, None, C.byref(nb))
    if rc != _lib.QDFN_ERR_ARG and rc < 0:
        check(rc)
    ni = np.zeros((nn.value, 4), dtype=np.int64)
    ng = np.zeros((nn.value, 4))
    rs = np.zeros((nb.value, 4), dtype=np.int64)
    ad = np.zeros(nb.value, dtype=np.int64)
    check(lib().qdfn_hmat_structure(p, n, ptr(ctr), ptr(dpts), perm.ctypes.data_as(PI64), nn.value,
                                    ni.ctypes.data_as(PI64), ptr(ng), C.byref(nn), nb.value,
                                    rs.ctypes.data_as(PI64), ad.ctypes.data_as(PI64), C.byref(nb)))
    return {"perm": perm, "node_info": ni, "node_geo": ng, "range_SR": rs, "admissible": ad}


class HMatrix:
    """Shear/normal H-matrix operator in device memory."""

    SIM_CREATE = "qdfn_sim_create_hmat"

    def __init__(self, handle, n, prec):
        self._h = handle
        self.n = n
        self.prec = prec

    @classmethod
    def build_rect(cls, seg, mu, nu, range_SR, admissible, aca_rtol=1e-6, aca_atol=0.0, prec=_lib.PREC_FP64,
                   sampled=False):
        """HmatBuild_R on a segment table already in sorted order (n x >= 8).
        sampled=True: admissible blocks from sampled rows and columns only
        (qdfn_hmat_build_rect_sampled), never assembled in full."""
        seg = np.asfortranarray(seg, dtype=np.float64)
        n = seg.shape[0]
        rs, ad = _i64(range_SR), _i64(admissible)
        name = "qdfn_hmat_build_rect_sampled" if sampled else "qdfn_hmat_build_rect"
        h = check_handle(getattr(lib(), name)(ptr(seg), n, n, float(mu), float(nu), ad.size,
                                              rs.ctypes.data_as(PI64), ad.ctypes.data_as(PI64),
                                              float(aca_rtol), float(aca_atol), prec), name)
        return cls(h, n, prec)

    @classmethod
    def build_tri(cls, P1, P2, P3, Rake, FaultCenter, UnitVector_Normal, UnitVector_Slip, ShearModulus, lam,
                  range_SR, admissible, aca_rtol=1e-6, aca_atol=0.0, prec=_lib.PREC_FP64, sampled=False):
        """HmatBuild_T on triangles already in sorted order.  sampled=True:
        admissible blocks from sampled rows and columns only
        (qdfn_hmat_build_tri_sampled), never assembled in full."""
        f = lambda a: np.asfortranarray(a, dtype=np.float64)
        P1, P2, P3, ctr, nrm, slip = map(f, (P1, P2, P3, FaultCenter, UnitVector_Normal, UnitVector_Slip))
        n = P1.shape[0]
        rs, ad = _i64(range_SR), _i64(admissible)
        name = "qdfn_hmat_build_tri_sampled" if sampled else "qdfn_hmat_build_tri"
        h = check_handle(getattr(lib(), name)(n, ptr(P1), ptr(P2), ptr(P3), ptr(_f64(Rake)), ptr(ctr),
                                              ptr(nrm), ptr(slip), float(ShearModulus), float(lam), ad.size,
                                              rs.ctypes.data_as(PI64), ad.ctypes.data_as(PI64),
                                              float(aca_rtol), float(aca_atol), prec), name)
        return cls(h, n, prec)

    @classmethod
    def from_blocks(cls, n, range_SR, admissible, S, N, prec=_lib.PREC_FP64):
        """From ShearStiffness_H / NormalStiffness_H: S and N are lists with, per
        block, a dense (m x k) array or a (U, V) pair."""
        rs, ad = _i64(range_SR), _i64(admissible)
        nb = ad.size
        keep = []
        args = []
        for parts in (S, N):
            kind = np.zeros(nb, dtype=np.int32)
            rank = np.zeros(nb, dtype=np.int64)
            A = (PD * nb)()
            B = (PD * nb)()
            for q, blk in enumerate(parts):
                if isinstance(blk, tuple):
                    U = np.asfortranarray(blk[0], dtype=np.float64)
                    V = np.asfortranarray(blk[1], dtype=np.float64)
                    kind[q], rank[q] = 1, U.shape[1]
                    keep += [U, V]
                    A[q], B[q] = U.ctypes.data_as(PD), V.ctypes.data_as(PD)
                else:
                    M = np.asfortranarray(blk, dtype=np.float64)
                    keep.append(M)
                    A[q] = M.ctypes.data_as(PD)
            keep += [kind, rank, A, B]
            args += [kind.ctypes.data_as(PI32), rank.ctypes.data_as(PI64), C.cast(A, C.POINTER(PD)),
                     C.cast(B, C.POINTER(PD))]
        h = check_handle(lib().qdfn_hmat_create(int(n), nb, rs.ctypes.data_as(PI64), ad.ctypes.data_as(PI64),
                                                *args, prec), "qdfn_hmat_create")
        return cls(h, int(n), prec)

    @property
    def handle(self):
        if not self._h:
            raise _lib.QdfnError(_lib.QDFN_ERR_STATE, "H-matrix destroyed")
        return self._h

    def info(self):
        n, nb, db, se = _lib.I64(), _lib.I64(), _lib.I64(), _lib.I64()
        pb = (_lib.I64 * 2)()
        check(lib().qdfn_hmat_info(self.handle, C.byref(n), C.byref(nb), C.byref(db), C.byref(se), pb))
        return {"n": n.value, "blocks": nb.value, "bytes_per_product": db.value, "stored_entries": se.value,
                "phase1_bytes": pb[0], "phase2_bytes": pb[1]}

    def bytes_by_part(self):
        """Algorithmic bytes of one product by part (qdfn_hmat_bytes)."""
        out = np.zeros(6, dtype=np.int64)
        check(lib().qdfn_hmat_bytes(self.handle, out.ctypes.data_as(PI64)))
        keys = ("phase1_entries", "phase1_x_t", "phase2_entries", "phase2_index", "phase2_gather", "outputs")
        return dict(zip(keys, (int(v) for v in out)))

    def build_info(self):
        """Entries evaluated, sampled-approximation steps, parts finished by
        sampling / evaluated in full, build wall time (ms)."""
        st = _lib.HmatBuildStats()
        check(lib().qdfn_hmat_build_info(self.handle, C.byref(st)))
        out = {f: getattr(st, f) for f, _ in st._fields_ if f != "ms_phase"}
        names = ("ms_evaluation", "ms_sampling", "ms_full_pivot", "ms_recompression", "ms_layout")
        out.update({k: st.ms_phase[i] for i, k in enumerate(names)})
        return out

    def time_phases(self, reps=20):
        """(phase-1 ms, phase-2 kernel ms, whole product ms), HIP events."""
        a, b, c = _lib.D(), _lib.D(), _lib.D()
        check(lib().qdfn_hmat_time_phases(self.handle, reps, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def ranks(self):
        nb = self.info()["blocks"]
        kS, kN = np.zeros(nb, dtype=np.int32), np.zeros(nb, dtype=np.int32)
        rS, rN = np.zeros(nb, dtype=np.int64), np.zeros(nb, dtype=np.int64)
        check(lib().qdfn_hmat_get_ranks(self.handle, kS.ctypes.data_as(PI32), rS.ctypes.data_as(PI64),
                                        kN.ctypes.data_as(PI32), rN.ctypes.data_as(PI64)))
        return {"kind_S": kS, "rank_S": rS, "kind_N": kN, "rank_N": rN}

    def split(self):
        """Per block, the first low-rank component in the 16-bit pool (shear, normal)."""
        nb = self.info()["blocks"]
        s, nn = np.zeros(nb, dtype=np.int64), np.zeros(nb, dtype=np.int64)
        check(lib().qdfn_hmat_get_split(self.handle, s.ctypes.data_as(PI64), nn.ctypes.data_as(PI64)))
        return s, nn

    def split8(self):
        """Per block, the first low-rank component in the 8-bit pool (shear, normal;
        the rank where there is none)."""
        nb = self.info()["blocks"]
        s, nn = np.zeros(nb, dtype=np.int64), np.zeros(nb, dtype=np.int64)
        check(lib().qdfn_hmat_get_split8(self.handle, s.ctypes.data_as(PI64), nn.ctypes.data_as(PI64)))
        return s, nn

    def blocks(self, range_SR, stored=False):
        """The blocks in the oracle's dict format, one dict per matrix
        (n, range, kind, rank, a, b): FP64 values as built, or (stored=True)
        as the product applies them, rounded to the operator's storage."""
        get = lib().qdfn_hmat_get_block_stored if stored else lib().qdfn_hmat_get_block
        rs = _i64(range_SR)
        rk = self.ranks()
        out = []
        for w, (kk, rr) in enumerate(((rk["kind_S"], rk["rank_S"]), (rk["kind_N"], rk["rank_N"]))):
            A, B = [], []
            for q in range(rs.shape[0]):
                m = rs[q, 3] - rs[q, 2] + 1
                k = rs[q, 1] - rs[q, 0] + 1
                if kk[q] == 0:
                    a = np.zeros((m, k), order="F")
                    check(get(self.handle, q, w, a.ctypes.data_as(PD), None))
                    A.append(a)
                    B.append(None)
                else:
                    r = int(rr[q])
                    a = np.zeros((m, r), order="F")
                    b = np.zeros((r, k), order="F")
                    check(get(self.handle, q, w, a.ctypes.data_as(PD), b.ctypes.data_as(PD)))
                    A.append(a)
                    B.append(b)
            out.append({"n": self.n, "range": rs, "kind": kk.astype(np.int64), "rank": rr, "a": A, "b": B})
        return out

    def kself(self):
        ks, dg = np.zeros(self.n), np.zeros(self.n)
        check(lib().qdfn_hmat_kself(self.handle, ptr(ks), ptr(dg)))
        return ks, dg

    def reduce_too_strong_interaction(self, multiple):
        """ReduceTooStrongInteraction_Hmat (Functions_Solvers.jl:546-579) on this
        operator's inadmissible blocks (qdfn_hmat_reduce_strong_interaction);
        returns the number of entries scaled.  Before runs are created on it."""
        cnt = _lib.I64()
        check(lib().qdfn_hmat_reduce_strong_interaction(self.handle, float(multiple), C.byref(cnt)))
        return cnt.value

    def normal_zero(self):
        return bool(check(lib().qdfn_hmat_normal_zero(self.handle)))

    def matvec(self, x, loading_form=True):
        """(yS, yN): loading form yS = (H_load (0 - x)) ./ -K_Self; otherwise
        yS = H_S x.  yN = H_N x."""
        x = _f64(x).ravel()
        if x.size != self.n:
            raise ValueError(f"x has {x.size} entries; the operator has {self.n}")
        yS, yN = np.zeros(self.n), np.zeros(self.n)
        check(lib().qdfn_hmat_matvec(self.handle, ptr(x), ptr(yS), ptr(yN), 1 if loading_form else 0))
        return yS, yN

    def time_matvec(self, reps=20):
        ms = _lib.D()
        check(lib().qdfn_hmat_time_matvec(self.handle, reps, C.byref(ms)))
        return ms.value

    def bytes_per_product(self):
        return self.info()["bytes_per_product"]

    def close(self):
        if getattr(self, "_h", None):
            lib().qdfn_hmat_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def main_H(ShearModulus, FaultCount, LoadingFaultCount, Mass, NormalStiffnessZero, a, b, Dc, ThetaI, Vini,
           FrictionI, InitialNormalStress, LoadingRate, TotalStep, RecordStep, SwitchV, DtCut, RuptureDt,
           MaximumDt, SaveResultFileName, RockDensity, FaultCenter, FaultLengthStrike, FaultLengthDip,
           FaultStrikeAngle, FaultDipAngle, FaultRakeAngle, SaveStep, MinimumNormalStress, Alpha_Evo,
           Ranks_Shear, Ranks_Normal, ElementRange_SR, NormalStiffness_H, ShearStiffness_H, ThreadCount,
           EvolutionDR, Admissible, external=None, prec=_lib.PREC_FP64):
    """main_H of scripts/Functions_RSFDFN3DMain_H.jl:2 on the GPU.  The blocks
    are given as ShearStiffness_H / NormalStiffness_H lists (dense arrays or
    (U, V) pairs); Ranks_* and ThreadCount are accepted for the signature only.
    Returns the History_* arrays (saved to SaveResultFileName when it is a str:
    JLD2 with the reference's Result keys for a .jld2 name, else .npz)."""
    from .sim import Simulation
    H = HMatrix.from_blocks(FaultCount, ElementRange_SR, Admissible, ShearStiffness_H, NormalStiffness_H, prec)
    prob = dict(n=FaultCount, n_loading=LoadingFaultCount, mass=Mass, a=a, b=b, dc=Dc, theta_i=ThetaI, v_i=Vini,
                friction_i=FrictionI, sigma0=InitialNormalStress, loading_rate=LoadingRate, mu=ShearModulus,
                rho=RockDensity, switch_v=SwitchV, dt_cut=DtCut, rupture_dt=RuptureDt, maximum_dt=MaximumDt,
                min_normal_stress=MinimumNormalStress, alpha_evo=Alpha_Evo, evolution_dr=EvolutionDR,
                planar=int(NormalStiffnessZero == 1), ext=external, ext_at_told=1)
    n_rec = int(round(TotalStep / RecordStep))
    sim = Simulation(H, prob, record_step=RecordStep, record_capacity=n_rec, trace_capacity=0)
    sim.run(TotalStep)
    hist = sim.history()
    if isinstance(SaveResultFileName, str) and SaveResultFileName:
        if SaveResultFileName.endswith(".jld2"):  # the reference's Result.jld2 keys (qdfn.jld2)
            from . import jld2
            jld2.save_result(SaveResultFileName, hist)
        else:
            np.savez(SaveResultFileName, **hist)
    sim.close()
    H.close()
    return hist


def ReduceTooStrongInteraction_Hmat(StrongInteractionCriteriaMultiple, H):
    """ReduceTooStrongInteraction_Hmat (scripts/Functions_Solvers.jl:546-579, called
    at QUAKEDFN_RunARG.jl:126-131 when StrongInteractionCriteriaMultiple > 0) on
    an HMatrix built on the device: the reference takes and returns the
    ShearStiffness_H / NormalStiffness_H block arrays, here the operator holds
    them.  Prints the reference's count line and returns H."""
    cnt = H.reduce_too_strong_interaction(StrongInteractionCriteriaMultiple)
    print(f"{cnt}/{H.n ** 2} interaction is smoothened")
    return H
