"""ORACLE — test infrastructure only.

Forward-mode derivative of the oracle's restated main()
(scripts/Functions_RSFDFN3DMain_D.jl:2-324, Functions_Solvers.jl:2-380):
liboracle_dual.so is sim.c / solvers.c / hmat.c built once more with every
double a dual number (dual_oracle.hpp), so one run gives the derivative of
every output along one direction in (a, b, Dc).  This stands in for a
ForwardDiff pass through the reference's Julia code, which cannot run here;
tests use it to check the device adjoint (qdfn_sim_adjoint) by the
dot-product identity <grad L, delta> = dL/ds along delta.  Only tests/ import
this module.
"""
import ctypes as C
import os

import numpy as np

from . import oracle as _orc

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

I64 = C.c_int64


class Dual(C.Structure):
    _fields_ = [("v", C.c_double), ("d", C.c_double)]


PDu = C.POINTER(Dual)


def _dualize(base):
    """A ctypes Structure like `base` with double -> Dual and double* -> Dual*."""
    fields = []
    for name, t in base._fields_:
        if t is C.c_double:
            t = Dual
        elif t is _orc.PD:
            t = PDu
        fields.append((name, t))
    return type(base.__name__ + "Dual", (C.Structure,), {"_fields_": fields})


DenseOpD = _dualize(_orc.DenseOp)
SimInD = _dualize(_orc.SimIn)
SimOutD = _dualize(_orc.SimOut)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_dual.so")
        if not os.path.exists(path):
            _orc.build()
        L = C.CDLL(path)
        L.orc_sim_dense.argtypes = [C.POINTER(DenseOpD), C.POINTER(SimInD), C.POINTER(SimOutD)]
        L.orc_sim_dense.restype = C.c_int
        L.orc_dense_prepare.argtypes = [I64, PDu, PDu, PDu]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _dual(v, d=None):
    """(n, 2) array of (value, derivative) pairs, C-contiguous (Fortran order
    is kept for matrices by the caller)."""
    v = np.asarray(v, dtype=np.float64)
    out = np.zeros(v.shape + (2,))
    out[..., 0] = v
    if d is not None:
        out[..., 1] = np.asarray(d, dtype=np.float64)
    return out


def _p(a):
    return a.ctypes.data_as(PDu)


def sim_dense(prob, KS, KN, total_step, record_step=1, seed=None):
    """main() with dense stiffness, every quantity carrying its derivative along
    seed = {"a": da, "b": db, "dc": ddc} (per-element directions; missing keys
    are zero).  Returns (status, dict of value arrays and "d_" derivative
    arrays: hist_* (n_record x n), trace_* (total_step))."""
    seed = seed or {}
    L = lib()
    n = prob["n"]
    # E and K_Self from K_S as the oracle does (zero derivative: the geometry is fixed)
    KSd = np.ascontiguousarray(np.stack([np.asfortranarray(KS).ravel(order="F"), np.zeros(n * n)], -1))
    Ed = np.zeros((n * n, 2))
    ksd = np.zeros((n, 2))
    L.orc_dense_prepare(n, _p(KSd), _p(Ed), _p(ksd))
    KNd = np.ascontiguousarray(np.stack([np.asfortranarray(KN).ravel(order="F"), np.zeros(n * n)], -1))
    op = DenseOpD(n, _p(Ed), _p(KNd))
    keep = [KSd, Ed, KNd, ksd]
    s = SimInD()
    n_rec = total_step // record_step if record_step > 0 else 0
    s.n, s.n_loading, s.total_step, s.record_step, s.n_record = n, prob.get("n_loading", 0), total_step, \
        record_step, n_rec

    def arr(name, key, d=None):
        a = _dual(prob[key], d)
        keep.append(a)
        setattr(s, name, _p(a))

    for k in "mass theta_i v_i friction_i sigma0 loading_rate".split():
        arr(k, k)
    for k in ("a", "b", "dc"):
        arr(k, k, seed.get(k))
    s.k_self = _p(ksd)
    for k in "mu rho switch_v dt_cut rupture_dt maximum_dt min_normal_stress".split():
        setattr(s, k, Dual(float(prob[k]), 0.0))
    s.alpha_evo = Dual(float(prob.get("alpha_evo", 0.0)), 0.0)
    s.evolution_dr = int(prob.get("evolution_dr", 1))
    s.planar = int(prob.get("planar", 0))
    s.ext_at_told = int(prob.get("ext_at_told", 0))
    ext = prob.get("ext")
    if ext is not None:
        s.ext_nt = len(ext["t"])
        t = _dual(ext["t"])
        keep.append(t)
        s.ext_t = _p(t)
        for k, name in (("n", "ext_n"), ("s", "ext_s"), ("p", "ext_p")):
            a = np.asfortranarray(ext[k] if k in ext else np.zeros_like(ext["n"]), dtype=np.float64)
            ad = np.ascontiguousarray(np.stack([a.ravel(order="F"), np.zeros(a.size)], -1))
            keep.append(ad)
            setattr(s, name, _p(ad))
    else:
        s.ext_nt = 0
    res = {"hist_time": np.zeros((n_rec, 2)), "hist_dt": np.zeros((n_rec, 2))}
    for k in "hist_v hist_disp hist_theta hist_sigma".split():
        res[k] = np.zeros((n_rec * n, 2))
    for k in "trace_told trace_dt trace_vmax".split():
        res[k] = np.zeros((total_step, 2))
    for k, v in res.items():
        setattr(s, k, _p(v))
    o = SimOutD()
    final = np.zeros((6 * n, 2))
    o.final_state = _p(final)
    st = L.orc_sim_dense(C.byref(op), C.byref(s), C.byref(o))
    out = {"steps_done": o.steps_done, "retry_passes": o.retry_passes, "stalls": o.stalls}
    for k, v in res.items():
        # matrices are record-major like History_*: entry (r, i) at r + i n_record
        shape = (n_rec, n) if v.shape[0] == n_rec * n and k not in ("hist_time", "hist_dt") else (v.shape[0],)
        out[k] = v[:, 0].reshape(shape, order="F")
        out["d_" + k] = v[:, 1].reshape(shape, order="F")
    out["final"] = final.reshape(6, n, 2)
    return st, out
