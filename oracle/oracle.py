"""ORACLE — test infrastructure only.

ctypes front end of liboracle.so, the CPU FP64 restatement of the reference
(limkjae/Quake-DFN).  Only tests/, ``__graft_entry__.smoke()`` and the CPU
baseline leg of bench.py import this module, and only as the checker; the
product package never does.

Function names follow the reference functions they restate:
  dc3d                  Okada_DC3D_Vector          scripts/Functions_OKADA3D.jl:464-654
  stiffness_rect        StiffnessMatrix_ByParts_Calculation_Rec  Functions_BuildInputFile.jl:386-512
  tdstress_hs           TDstressHS                 scripts/Functions_TDstressHS.jl:2-130
  rotverts_unitvectors  RotVerts_UnitVectors       Functions_BuildInputFile.jl:338-377
  stiffness_tri         StiffnessMatrix_ByParts_Calculation_Tri  Functions_BuildInputFile.jl:551-594
  evolution_law / solver_highv / solver_lowv / solve_one_timestep
                        Functions_Solvers.jl:2-380
  interp_external       InterpolateFromStressChange Functions_Solvers.jl:465-540
  sim_dense / sim_hmat  main / main_H              Functions_RSFDFN3DMain_{D,H}.jl
  event_fragments / merge_fragments
                        get_event_fragments / merge_fragments  scripts/Functions_Plot.jl:430-579
  reduce_too_strong_interaction(_hmat)
                        ReduceTooStrongInteraction(_Hmat)      Functions_Solvers.jl:546-607
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D = C.c_double
I64 = C.c_int64
I32 = C.c_int32
PD = C.POINTER(C.c_double)
PI32 = C.POINTER(C.c_int32)
PI64 = C.POINTER(C.c_int64)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops (dense and H products,
    local solves); outputs do not depend on it.  Returns the previous count."""
    return lib().orc_set_threads(int(n))


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(PD)


def _pi32(a):
    return a.ctypes.data_as(PI32)


class ElemIn(C.Structure):
    _fields_ = [(n, D) for n in ("conv_crit disp_old friction_old theta_old v_old friction0 a b dc v0 "
                                 "k_self omega mass mu sigma tl rho dtau switch_v alpha_evo sigma_old "
                                 "tau0").split()] + [("evolution_dr", C.c_int)]


class ElemOut(C.Structure):
    _fields_ = [(n, D) for n in "v friction disp theta sigma dt".split()] + \
               [(n, C.c_int) for n in "instab flag iters stall".split()]


class StepParams(C.Structure):
    _fields_ = [(n, D) for n in "conv_crit v0 mu rho switch_v alpha_evo".split()] + \
               [("evolution_dr", C.c_int)]


class DenseOp(C.Structure):
    _fields_ = [("n", I64), ("E", PD), ("KN", PD)]


class SimIn(C.Structure):
    _fields_ = ([(n, I64) for n in "n n_loading total_step record_step n_record".split()] +
                [(n, PD) for n in "mass a b dc theta_i v_i friction_i sigma0 loading_rate k_self".split()] +
                [(n, D) for n in ("mu rho switch_v dt_cut rupture_dt maximum_dt min_normal_stress "
                                  "alpha_evo").split()] +
                [(n, C.c_int) for n in "evolution_dr planar ext_at_told".split()] +
                [("ext_nt", I64)] +
                [(n, PD) for n in "ext_t ext_n ext_s ext_p".split()] +
                [(n, PD) for n in "hist_time hist_dt hist_v hist_disp hist_theta hist_sigma".split()] +
                [(n, PD) for n in "trace_told trace_dt trace_vmax".split()])


class SimOut(C.Structure):
    _fields_ = ([(n, I64) for n in "rhs_evals retry_passes stalls steps_done n_recorded".split()] +
                [(n, D) for n in "t_end dt_last".split()] + [("final_state", PD)])


class HmatOp(C.Structure):
    _fields_ = [("n", I64), ("nblk", I64), ("range", PI64), ("kind", PI64), ("rank", PI64),
                ("a", C.POINTER(PD)), ("b", C.POINTER(PD))]


def _declare(L):
    L.orc_dc3d.argtypes = [D, I64, PD, PD, PD, D, D, D, D, D, D, D, D, D, PD, PD]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_set_threads.restype = C.c_int
    L.orc_stiffness_rect.argtypes = [PD, I64, I64, PD, I64, I64, D, D, PD, PD, I64]
    L.orc_evolution_law.argtypes = [D, D, D, D, D, D, D, D, C.c_int, C.c_int]
    L.orc_evolution_law.restype = D
    L.orc_solver_highv.argtypes = [C.POINTER(ElemIn), D, C.POINTER(ElemOut)]
    L.orc_solver_lowv.argtypes = [C.POINTER(ElemIn), D, C.POINTER(ElemOut)]
    L.orc_solve_one_timestep.argtypes = ([C.POINTER(StepParams), I64, I64] + [PD] * 9 + [D] +
                                         [PD] * 4 + [PI32] + [PD] * 4 + [PD] * 6 + [PI32] * 3)
    L.orc_solve_one_timestep.restype = I64
    L.orc_interp_external.argtypes = [D, I64, I64, PD, PD, PD, PD, PD, PD, PD]
    L.orc_dense_prepare.argtypes = [I64, PD, PD, PD]
    L.orc_dense_coupling.argtypes = [C.POINTER(DenseOp), PD, PD, PD, C.c_int]
    L.orc_sim_dense.argtypes = [C.POINTER(DenseOp), C.POINTER(SimIn), C.POINTER(SimOut)]
    L.orc_sim_dense.restype = C.c_int
    if hasattr(L, "orc_tdstress_hs"):
        L.orc_tdstress_hs.argtypes = [I64, PD, PD, PD, PD, PD, PD, D, D, D, D, D, PD, PD]
        L.orc_rotverts_unitvectors.argtypes = [I64, PD, PD, PD, PD, PD, PD, PD, PD]
        L.orc_stiffness_tri.argtypes = [I64, PD, PD, PD, PD, I64, PD, PD, PD, D, D, PD, PD, I64]
    if hasattr(L, "orc_hmat_matvec"):
        L.orc_hmat_matvec.argtypes = [C.POINTER(HmatOp), PD, PD]
        L.orc_sim_hmat.argtypes = [C.POINTER(HmatOp), C.POINTER(HmatOp), C.POINTER(SimIn),
                                   C.POINTER(SimOut)]
        L.orc_sim_hmat.restype = C.c_int
        L.orc_build_hierarchy.argtypes = [I64, PI64, PI64, PI64, PD, PD, I64, D, I64, I64, I64, PI64, PI64]
        L.orc_build_hierarchy.restype = I64
    if hasattr(L, "orc_event_fragments"):
        L.orc_event_fragments.argtypes = [I64, I64, PD, PD, I64, PD, PD, D, D, I64, PD, PD, PD, PI32]
        L.orc_event_fragments.restype = I64
        L.orc_merge_fragments.argtypes = [I64, PD, PD, PD, D, D, PD, PD, PD, PD, PI32]
        L.orc_merge_fragments.restype = I64


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def dc3d(alpha, X, Y, Z, depth, dip, al1, al2, aw1, aw2, disl1, disl2, disl3):
    """Okada_DC3D_Vector: returns (U[n,12], IRET[n]); U columns are
    UX,UY,UZ,UXX,UYX,UZX,UXY,UYY,UZY,UXZ,UYZ,UZZ."""
    X, Y, Z = _f64(X).ravel(), _f64(Y).ravel(), _f64(Z).ravel()
    n = X.size
    U = np.zeros((12, n))
    iret = np.zeros(n)
    lib().orc_dc3d(alpha, n, _p(X), _p(Y), _p(Z), depth, dip, al1, al2, aw1, aw2,
                   disl1, disl2, disl3, _p(U), _p(iret))
    return U.T.copy(), iret


def stiffness_rect(src_seg, rcv_seg, mu, nu):
    """K_S, K_N (nr x ns) for rectangular segments (Input_Segment rows)."""
    src = np.asfortranarray(src_seg, dtype=np.float64)
    rcv = np.asfortranarray(rcv_seg, dtype=np.float64)
    ns, nr = src.shape[0], rcv.shape[0]
    KS = np.zeros((nr, ns), order="F")
    KN = np.zeros((nr, ns), order="F")
    lib().orc_stiffness_rect(_p(src), ns, ns, _p(rcv), nr, nr, mu, nu,
                             KS.ctypes.data_as(PD), KN.ctypes.data_as(PD), nr)
    return KS, KN


def tdstress_hs(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, mu, lam):
    X, Y, Z = _f64(X).ravel(), _f64(Y).ravel(), _f64(Z).ravel()
    n = X.size
    S = np.zeros((6, n))
    E = np.zeros((6, n))
    lib().orc_tdstress_hs(n, _p(X), _p(Y), _p(Z), _p(_f64(P1)), _p(_f64(P2)), _p(_f64(P3)),
                          Ss, Ds, Ts, mu, lam, _p(S), _p(E))
    return S.T.copy(), E.T.copy()


def rotverts_unitvectors(P1, P2, P3, rake):
    """Returns P1,P2,P3 (possibly swapped), normal, strike, dip, slip (all n x 3)."""
    n = P1.shape[0]
    q1 = np.asfortranarray(P1, dtype=np.float64).copy(order="F")
    q2 = np.asfortranarray(P2, dtype=np.float64).copy(order="F")
    q3 = np.asfortranarray(P3, dtype=np.float64).copy(order="F")
    outs = [np.zeros((n, 3), order="F") for _ in range(4)]
    lib().orc_rotverts_unitvectors(n, _p(q1), _p(q2), _p(q3), _p(_f64(rake)),
                                   *[o.ctypes.data_as(PD) for o in outs])
    return (q1, q2, q3, *outs)


def stiffness_tri(P1, P2, P3, rake, ctr, nrm, slip, mu, lam):
    ns, nr = P1.shape[0], ctr.shape[0]
    f = lambda a: np.asfortranarray(a, dtype=np.float64)
    P1, P2, P3, ctr, nrm, slip = map(f, (P1, P2, P3, ctr, nrm, slip))
    KS = np.zeros((nr, ns), order="F")
    KN = np.zeros((nr, ns), order="F")
    lib().orc_stiffness_tri(ns, _p(P1), _p(P2), _p(P3), _p(_f64(rake)), nr, _p(ctr), _p(nrm),
                            _p(slip), mu, lam, KS.ctypes.data_as(PD), KN.ctypes.data_as(PD), nr)
    return KS, KN


def evolution_law(theta_old, dt, v, dc, alpha, sig, sig_old, b, fast, dr=1):
    return lib().orc_evolution_law(theta_old, dt, v, dc, alpha, sig, sig_old, b, fast, dr)


def _elem_in(**kw):
    e = ElemIn()
    for k, v in kw.items():
        setattr(e, k, v)
    return e


def solver_highv(dt, **kw):
    o = ElemOut()
    lib().orc_solver_highv(C.byref(_elem_in(**kw)), dt, C.byref(o))
    return o


def solver_lowv(dt, **kw):
    o = ElemOut()
    lib().orc_solver_lowv(C.byref(_elem_in(**kw)), dt, C.byref(o))
    return o


def solve_one_timestep(params, n_loading, disp_old, friction_old, theta_old, v_old, friction0, a, b,
                       dc, k_self, dt, omega, mass, sig_i, tl, switch, loading_rate, dtau, sig_old, tau0):
    """SolveOneTimeStep; returns dict of the 8 outputs plus per-element iteration counts."""
    arrs = [_f64(x) for x in (disp_old, friction_old, theta_old, v_old, friction0, a, b, dc, k_self)]
    n = arrs[0].size
    omega, mass, sig_i, tl = (_f64(x) for x in (omega, mass, sig_i, tl))
    sw = np.ascontiguousarray(switch, dtype=np.int32)
    lr, dtau, sig_old, tau0 = (_f64(x) for x in (loading_rate, dtau, sig_old, tau0))
    outs = [np.zeros(n) for _ in range(6)]
    ints = [np.zeros(n, dtype=np.int32) for _ in range(3)]
    sp = StepParams(**{k: params[k] for k in ("conv_crit", "v0", "mu", "rho", "switch_v", "alpha_evo",
                                              "evolution_dr")})
    stalls = lib().orc_solve_one_timestep(
        C.byref(sp), n, n_loading, *[_p(x) for x in arrs[:9]], dt, _p(omega), _p(mass), _p(sig_i),
        _p(tl), _pi32(sw), _p(lr), _p(dtau), _p(sig_old), _p(tau0), *[_p(o) for o in outs],
        *[_pi32(i) for i in ints])
    keys = ("V", "Friction", "Disp", "Theta", "EffNormalStress", "Dt_All")
    res = dict(zip(keys, outs))
    res.update(InstabilityThistime=ints[0], FLAG_GoodToGo=ints[1], iterations=ints[2], stalls=stalls)
    return res


def interp_external(time, ext_t, ext_n, ext_s, ext_p):
    ext_t = _f64(ext_t)
    nt = ext_t.size
    ext_n, ext_s, ext_p = (np.asfortranarray(x, dtype=np.float64) for x in (ext_n, ext_s, ext_p))
    n = ext_n.shape[1]
    o = [np.zeros(n) for _ in range(3)]
    lib().orc_interp_external(time, n, nt, _p(ext_t), ext_n.ctypes.data_as(PD),
                              ext_s.ctypes.data_as(PD), ext_p.ctypes.data_as(PD), *[_p(x) for x in o])
    return tuple(o)


def dense_prepare(KS):
    KS = np.asfortranarray(KS, dtype=np.float64)
    n = KS.shape[0]
    E = np.zeros((n, n), order="F")
    ks = np.zeros(n)
    lib().orc_dense_prepare(n, KS.ctypes.data_as(PD), E.ctypes.data_as(PD), _p(ks))
    return E, ks


def _sim_in(prob, total_step, record_step, trace=True):
    """prob: dict from the problem generator (element arrays + run parameters)."""
    n = prob["n"]
    keep = []

    def arr(x):
        a = _f64(x)
        keep.append(a)
        return _p(a)

    s = SimIn()
    s.n, s.n_loading, s.total_step, s.record_step = n, prob.get("n_loading", 0), total_step, record_step
    n_rec = total_step // record_step if record_step > 0 else 0
    s.n_record = n_rec
    for k in ("mass a b dc theta_i v_i friction_i sigma0 loading_rate k_self".split()):
        setattr(s, k, arr(prob[k]))
    for k in "mu rho switch_v dt_cut rupture_dt maximum_dt min_normal_stress".split():
        setattr(s, k, float(prob[k]))
    s.alpha_evo = float(prob.get("alpha_evo", 0.0))
    s.evolution_dr = int(prob.get("evolution_dr", 1))
    s.planar = int(prob.get("planar", 0))
    s.ext_at_told = int(prob.get("ext_at_told", 0))
    ext = prob.get("ext")
    if ext is not None:
        s.ext_nt = len(ext["t"])
        s.ext_t = arr(ext["t"])
        for k, name in (("n", "ext_n"), ("s", "ext_s"), ("p", "ext_p")):
            a = np.asfortranarray(ext[k] if k in ext else np.zeros_like(ext["n"]), dtype=np.float64)
            keep.append(a)
            setattr(s, name, a.ctypes.data_as(PD))
    else:
        s.ext_nt = 0
    res = {"hist_time": np.zeros(n_rec), "hist_dt": np.zeros(n_rec)}
    for k in "hist_v hist_disp hist_theta hist_sigma".split():
        res[k] = np.zeros((n_rec, n), order="F")
    if trace:
        for k in "trace_told trace_dt trace_vmax".split():
            res[k] = np.zeros(total_step)
    for k, v in res.items():
        setattr(s, k, v.ctypes.data_as(PD))
    return s, res, keep


def sim_dense(prob, KS, KN, total_step, record_step=1, trace=True):
    """main() with dense stiffness; returns (status, dict of histories/trace/counters)."""
    E, ks = dense_prepare(KS)
    prob = dict(prob)
    prob["k_self"] = ks
    KN = np.asfortranarray(KN, dtype=np.float64)
    op = DenseOp(KS.shape[0], E.ctypes.data_as(PD), KN.ctypes.data_as(PD))
    s, res, keep = _sim_in(prob, total_step, record_step, trace)
    n = prob["n"]
    final = np.zeros(6 * n)
    o = SimOut()
    o.final_state = _p(final)
    st = lib().orc_sim_dense(C.byref(op), C.byref(s), C.byref(o))
    res.update(rhs_evals=o.rhs_evals, retry_passes=o.retry_passes, stalls=o.stalls,
               steps_done=o.steps_done, n_recorded=o.n_recorded, t_end=o.t_end, dt_last=o.dt_last)
    res["final"] = dict(zip(("V", "Disp", "Theta", "Friction", "Sigma", "Accel"), final.reshape(6, n)))
    return st, res


def dense_coupling(E, KN, x, planar=0):
    E = np.asfortranarray(E, dtype=np.float64)
    KN = np.asfortranarray(KN, dtype=np.float64)
    n = E.shape[0]
    op = DenseOp(n, E.ctypes.data_as(PD), KN.ctypes.data_as(PD))
    x = _f64(x)
    tl, sp = np.zeros(n), np.zeros(n)
    lib().orc_dense_coupling(C.byref(op), _p(x), _p(tl), _p(sp), planar)
    return tl, sp


def _hmat_struct(H):
    """H: dict n, range (nblk x 4, 1-based src_i src_f rcv_i rcv_f), kind, rank,
    a (list: dense m x k or U m x r), b (list: None or V r x k).  Returns the
    ctypes struct and the arrays it points into (keep them alive)."""
    keep = []
    rng = np.ascontiguousarray(H["range"], dtype=np.int64)
    kind = np.ascontiguousarray(H["kind"], dtype=np.int64)
    rank = np.ascontiguousarray(H["rank"], dtype=np.int64)
    nb = rng.shape[0]
    A = (PD * nb)()
    B = (PD * nb)()
    for q in range(nb):
        a = np.asfortranarray(H["a"][q], dtype=np.float64)
        keep.append(a)
        A[q] = a.ctypes.data_as(PD)
        if H["b"][q] is not None:
            b = np.asfortranarray(H["b"][q], dtype=np.float64)
            keep.append(b)
            B[q] = b.ctypes.data_as(PD)
    keep += [rng, kind, rank, A, B]
    s = HmatOp(int(H["n"]), nb, rng.ctypes.data_as(PI64), kind.ctypes.data_as(PI64), rank.ctypes.data_as(PI64),
               C.cast(A, C.POINTER(PD)), C.cast(B, C.POINTER(PD)))
    return s, keep


def hmat_matvec(H, x):
    """HmatSolver_Pararllel (Functions_Hmatrix.jl:631-646): each row's sum in
    block order (the same for any thread count)."""
    s, keep = _hmat_struct(H)
    x = _f64(x)
    y = np.zeros(int(H["n"]))
    lib().orc_hmat_matvec(C.byref(s), _p(x), _p(y))
    return y


def reduce_too_strong_interaction_hmat(multiple, admissible, HS, HN):
    """ReduceTooStrongInteraction_Hmat (Functions_Solvers.jl:546-579), loop for
    loop on copies of the oracle's block dicts: returns (HS', HN', pairs) with
    pairs the (source, receiver) 1-based list in the reference's order."""
    load, ks = hmat_transition_to_loading(HS)
    S = dict(HS, a=[None if a is None else np.array(a, dtype=np.float64, copy=True) for a in HS["a"]])
    N = dict(HN, a=[None if a is None else np.array(a, dtype=np.float64, copy=True) for a in HN["a"]])
    pairs = []
    for blk in range(len(admissible)):
        if admissible[blk] != 0:
            continue
        s0, s1, r0, r1 = (int(v) for v in np.asarray(HS["range"])[blk])
        for js, src in enumerate(range(s0, s1 + 1)):
            for ir, rcv in enumerate(range(r0, r1 + 1)):
                L = load["a"][blk][ir, js]
                if abs(ks[rcv - 1]) * multiple < abs(L):
                    ratio = abs(ks[rcv - 1]) * multiple / abs(L)
                    S["a"][blk][ir, js] = S["a"][blk][ir, js] * ratio
                    N["a"][blk][ir, js] = N["a"][blk][ir, js] * ratio
                    pairs.append((src, rcv))
    return S, N, pairs


def reduce_too_strong_interaction(multiple, n, KS, KN):
    """ReduceTooStrongInteraction (Functions_Solvers.jl:586-607), loop for loop
    on copies of the dense matrices: returns (KS', KN', pair count)."""
    KS = np.array(KS, dtype=np.float64, copy=True)
    KN = np.array(KN, dtype=np.float64, copy=True)
    count = 0
    for src in range(n):
        for rcv in range(n):
            if src != rcv and abs(KS[rcv, rcv]) * multiple < abs(KS[rcv, src]):
                ratio = abs(KS[rcv, rcv]) * multiple / abs(KS[rcv, src])
                KS[rcv, src] = KS[rcv, src] * ratio
                KN[rcv, src] = KN[rcv, src] * ratio
                count += 1
    return KS, KN, count


def hmat_transition_to_loading(H):
    """StiffnessTransitionToLoading (Functions_Hmatrix.jl:459-477): copy of the
    shear blocks with the diagonal of every block whose source and receiver
    ranges coincide set to zero, and K_Self = -diag."""
    n = int(H["n"])
    ks = np.zeros(n)
    a = []
    for q, r in enumerate(np.asarray(H["range"])):
        blk = np.array(H["a"][q], dtype=np.float64, copy=True)
        if r[0] == r[2] and r[1] == r[3]:
            for j in range(blk.shape[0]):
                ks[r[0] - 1 + j] = -blk[j, j]
                blk[j, j] = 0.0
        a.append(blk)
    out = dict(H)
    out["a"] = a
    return out, ks


def sim_hmat(prob, HS, HN, total_step, record_step=1, trace=True):
    """main_H with H-matrix stiffness (shear blocks as built; the loading form
    and K_Self are derived here); returns (status, dict)."""
    HL, ks = hmat_transition_to_loading(HS)
    prob = dict(prob)
    prob["k_self"] = ks
    prob.setdefault("ext_at_told", 1)
    sS, keepS = _hmat_struct(HL)
    sN, keepN = _hmat_struct(HN)
    s, res, keep = _sim_in(prob, total_step, record_step, trace)
    n = prob["n"]
    final = np.zeros(6 * n)
    o = SimOut()
    o.final_state = _p(final)
    st = lib().orc_sim_hmat(C.byref(sS), C.byref(sN), C.byref(s), C.byref(o))
    res.update(rhs_evals=o.rhs_evals, retry_passes=o.retry_passes, stalls=o.stalls,
               steps_done=o.steps_done, n_recorded=o.n_recorded, t_end=o.t_end, dt_last=o.dt_last)
    res["final"] = dict(zip(("V", "Disp", "Theta", "Friction", "Sigma", "Accel"), final.reshape(6, n)))
    res["k_self"] = ks
    return st, res


def build_hierarchy(node_info, node_geo, total_levels, crit, loading_range=None, cap=None):
    """BuildHierarchy (Functions_Hmatrix.jl:180-212) on a given cluster list:
    node_info rows (begin, end, level, ...) 1-based, node_geo rows (cx, cy, cz, diam).
    Returns (range_SR nblk x 4, admissible)."""
    ni = np.asarray(node_info, dtype=np.int64)
    ng = np.asarray(node_geo, dtype=np.float64)
    nn = ni.shape[0]
    beg, end, lev = (np.ascontiguousarray(ni[:, k]) for k in range(3))
    ctr = np.ascontiguousarray(ng[:, :3])
    diam = np.ascontiguousarray(ng[:, 3])
    cap = cap or (4 * nn * nn + 16)
    rng = np.zeros((cap, 4), dtype=np.int64)
    adm = np.zeros(cap, dtype=np.int64)
    lb, le = loading_range if loading_range else (0, 0)
    nb = lib().orc_build_hierarchy(nn, beg.ctypes.data_as(PI64), end.ctypes.data_as(PI64),
                                   lev.ctypes.data_as(PI64), _p(ctr), _p(diam), total_levels, float(crit),
                                   lb, le, cap, rng.ctypes.data_as(PI64), adm.ctypes.data_as(PI64))
    if nb < 0:
        raise RuntimeError("build_hierarchy: capacity exceeded")
    return rng[:nb].copy(), adm[:nb].copy()


def event_fragments(time, disp, center, area, mu, crit):
    """get_event_fragments on arrays: time[n_rec], disp n_rec x n, center n x 3,
    area[n].  Returns (time, location count x 3, moment, element 0-based)."""
    time = _f64(time)
    disp = np.asfortranarray(disp, dtype=np.float64)
    center = np.asfortranarray(center, dtype=np.float64)
    area = _f64(area)
    n_rec, n = disp.shape
    cap = max(1, n * ((n_rec + 1) // 2))
    ft, fl, fm = np.zeros(cap), np.zeros(3 * cap), np.zeros(cap)
    fe = np.zeros(cap, dtype=np.int32)
    F = lib().orc_event_fragments(n_rec, n, _p(time), _p(disp), n_rec, _p(center), _p(area), mu, crit, cap,
                                  _p(ft), _p(fl), _p(fm), _pi32(fe))
    assert F >= 0
    return ft[:F].copy(), fl[:3 * F].reshape(3, F).T.copy(), fm[:F].copy(), fe[:F].copy()


def merge_fragments(ftime, floc, fmom, mt, md):
    """merge_fragments: returns (time, Mw, location E x 3, moment, fragment_event 0-based)."""
    ftime = _f64(ftime)
    F = ftime.shape[0]
    floc = np.asfortranarray(np.asarray(floc, dtype=np.float64).reshape(F, 3))
    fmom = _f64(fmom)
    cap = max(F, 1)
    mt_, ml, mm, mw = np.zeros(cap), np.zeros(3 * cap), np.zeros(cap), np.zeros(cap)
    fid = np.zeros(cap, dtype=np.int32)
    E = lib().orc_merge_fragments(F, _p(ftime), _p(floc), _p(fmom), mt, md, _p(mt_), _p(ml), _p(mm), _p(mw),
                                  _pi32(fid))
    return mt_[:E].copy(), mw[:E].copy(), ml.reshape(3, cap)[:, :E].T.copy(), mm[:E].copy(), fid[:F].copy()
