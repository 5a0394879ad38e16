"""Device-resident earthquake-cycle run with the reference's driver interface.

DenseOperator   the dense coupling of scripts/Functions_RSFDFN3DMain_D.jl:54-63
                (E = K_ij/K_ii, K_Self = |K_ii|) kept in HBM at FP64 or FP32.
Simulation      the TotalStep loop of main() (Functions_RSFDFN3DMain_D.jl:146-322)
                running entirely on the GPU; the host only reads a state block
                per batch of attempts.
main(...)       main() with the reference's positional arguments.
"""
import numpy as np

from . import _lib
from ._lib import SimParams, SimStats, check, check_handle, lib, ptr


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


class DenseOperator:
    """Dense shear/normal coupling operator in device memory."""

    def __init__(self, handle, n, prec):
        self._h = handle
        self.n = n
        self.prec = prec

    @classmethod
    def from_matrices(cls, KS, KN, prec=_lib.PREC_FP64):
        KS = np.asfortranarray(KS, dtype=np.float64)
        KN = np.asfortranarray(KN, dtype=np.float64)
        n = KS.shape[0]
        if KS.shape != (n, n) or KN.shape != (n, n):
            raise ValueError("K_S and K_N must be square and of equal size")
        h = check_handle(lib().qdfn_dense_create(ptr(KS), ptr(KN), n, n, prec), "qdfn_dense_create")
        return cls(h, n, prec)

    @classmethod
    def from_rect_segments(cls, Input_Segment, ShearModulus, PoissonRatio, prec=_lib.PREC_FP64):
        seg = np.asfortranarray(Input_Segment, dtype=np.float64)
        n = seg.shape[0]
        h = check_handle(lib().qdfn_dense_create_rect(ptr(seg), n, n, float(ShearModulus), float(PoissonRatio),
                                                      prec), "qdfn_dense_create_rect")
        return cls(h, n, prec)

    @classmethod
    def from_triangles(cls, P1, P2, P3, Rake, FaultCenter, UnitVector_Normal, UnitVector_Slip, ShearModulus,
                       lam, prec=_lib.PREC_FP64):
        f = lambda a: np.asfortranarray(a, dtype=np.float64)
        P1, P2, P3, ctr, nrm, slip = map(f, (P1, P2, P3, FaultCenter, UnitVector_Normal, UnitVector_Slip))
        n = P1.shape[0]
        h = check_handle(lib().qdfn_dense_create_tri(n, ptr(P1), ptr(P2), ptr(P3), ptr(_f64(Rake)), ptr(ctr),
                                                     ptr(nrm), ptr(slip), float(ShearModulus), float(lam), prec),
                         "qdfn_dense_create_tri")
        return cls(h, n, prec)

    @property
    def handle(self):
        if not self._h:
            raise _lib.QdfnError(_lib.QDFN_ERR_STATE, "operator destroyed")
        return self._h

    def kself(self):
        ks, dg = np.zeros(self.n), np.zeros(self.n)
        check(lib().qdfn_dense_kself(self.handle, ptr(ks), ptr(dg)))
        return ks, dg

    def set_splits(self, splits):
        """Column splits of the coupling product (qdfn_dense_set_splits; 0 = the
        default that fills the GPU for a run alone); returns the count taken."""
        rc = lib().qdfn_dense_set_splits(self.handle, int(splits))
        if rc < 0:
            raise _lib.QdfnError(rc, _lib.last_error())
        return rc

    def normal_zero(self):
        return bool(check(lib().qdfn_dense_normal_zero(self.handle)))

    def matvec(self, x, loading_form=True):
        """(yS, yN): loading form yS = E*(0-x); otherwise yS = K_S*x.  yN = K_N*x."""
        x = _f64(x)
        if x.size != self.n:
            raise ValueError(f"x has {x.size} entries; the operator has {self.n}")
        yS, yN = np.zeros(self.n), np.zeros(self.n)
        check(lib().qdfn_matvec(self.handle, ptr(x), ptr(yS), ptr(yN), 1 if loading_form else 0))
        return yS, yN

    def matvec_device_noalloc(self, d_x, d_yS, d_yN, loading_form=True, stream=0):
        """The coupling product on device vectors (addresses as ints, n doubles
        each), enqueued on `stream` (a hipStream_t as an int) without waiting
        and without allocating; split partials in the library's static device
        scratch (calls overlapping on one GPU race)."""
        check(lib().qdfn_matvec_device_noalloc(self.handle, _lib.C.c_void_p(d_x), _lib.C.c_void_p(d_yS),
                                               _lib.C.c_void_p(d_yN), 1 if loading_form else 0,
                                               _lib.C.c_void_p(stream)))

    def time_coupling(self, reps=20):
        ms = _lib.D()
        check(lib().qdfn_dense_time_coupling(self.handle, reps, _lib.C.byref(ms)))
        return ms.value

    def bytes_per_product(self):
        """Algorithmic HBM bytes of one fused coupling product: both matrices
        once plus the input vector and the two outputs."""
        p = 8 if self.prec == _lib.PREC_FP64 else 4
        return 2 * p * self.n * self.n + 8 * self.n + 16 * self.n

    def close(self):
        if self._h:
            lib().qdfn_dense_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Simulation:
    """One device-resident run: main() of the reference with a DenseOperator,
    main_H() with an HMatrix (set prob["ext_at_told"] = 1 for main_H's timing
    of the external tables)."""

    def __init__(self, op, prob, record_step=10, record_capacity=0, trace_capacity=0):
        self.op = op
        n = op.n
        self.n = n
        p = SimParams()
        p.n, p.n_loading = n, int(prob.get("n_loading", 0))
        p.mu, p.rho = float(prob["mu"]), float(prob["rho"])
        p.switch_v, p.dt_cut = float(prob["switch_v"]), float(prob["dt_cut"])
        p.rupture_dt, p.maximum_dt = float(prob["rupture_dt"]), float(prob["maximum_dt"])
        p.min_normal_stress = float(prob["min_normal_stress"])
        p.alpha_evo = float(prob.get("alpha_evo", 0.0))
        p.evolution_dr = int(prob.get("evolution_dr", 1))
        p.planar = int(prob.get("planar", 0))
        p.ext_at_told = int(prob.get("ext_at_told", 0))
        p.record_step = int(record_step)
        p.record_capacity = int(record_capacity)
        p.trace_capacity = int(trace_capacity)
        self.params = p
        arrs = [_f64(prob[k]) for k in ("mass", "a", "b", "dc", "theta_i", "v_i", "friction_i", "sigma0",
                                        "loading_rate")]
        for a in arrs:
            if a.size != n:
                raise ValueError("element arrays must match the operator size")
        create = getattr(op, "SIM_CREATE", "qdfn_sim_create_dense")
        self._h = check_handle(getattr(lib(), create)(op.handle, p, *[ptr(a) for a in arrs]), create)
        ext = prob.get("ext")
        if ext is not None:
            t = _f64(ext["t"])
            en = np.asfortranarray(ext["n"], dtype=np.float64)
            es = np.asfortranarray(ext["s"], dtype=np.float64)
            if en.shape != (t.size, n) or es.shape != (t.size, n):
                raise ValueError("external stress tables must be (len(t), n)")
            check(lib().qdfn_sim_set_external(self._h, t.size, ptr(t), ptr(en), ptr(es)))
        if prob.get("conv_crit") is not None:  # ConvergenceCrit of the local solves (1e-9 by default)
            check(lib().qdfn_sim_set_convergence(self._h, float(prob["conv_crit"])))

    def run(self, n_steps, timed=False, graph_attempts=32):
        """Advance n_steps accepted steps.  timed=False: replay a captured graph
        of attempts; timed=True: direct launches with HIP events around every
        coupling kernel (fills stats()["coupling_ms_avg"])."""
        mode = 1 if timed else 0
        rc = lib().qdfn_sim_run(self._h, int(n_steps), mode, int(graph_attempts))
        if rc < 0:
            raise _lib.QdfnError(rc, _lib.last_error())
        return self.stats()

    def run_checkpointed(self, n_steps, every=50, graph_attempts=32):
        """Advance a fresh dense run n_steps accepted steps, keeping its state
        every `every` steps for the adjoint's checkpointed recomputation
        (qdfn_sim_run_checkpointed; trace_capacity >= n_steps)."""
        rc = lib().qdfn_sim_run_checkpointed(self._h, int(n_steps), int(every), int(graph_attempts))
        if rc < 0:
            raise _lib.QdfnError(rc, _lib.last_error())
        return self.stats()

    def run_until(self, t_end, max_steps, on_records=None, graph_attempts=32):
        """Advance until TOld reaches t_end seconds or max_steps accepted steps
        (qdfn_sim_run_until).  on_records(first, hist) receives every batch of
        records as a dict of History_Time / History_Dt / History_V /
        History_Disp / History_Theta / History_NormalStress arrays (records x n)
        and may return True to stop the run."""
        cb = None
        if on_records is not None:
            n = self.n

            def _cb(user, first, count, t, dt, V, D, TH, SIG):
                try:
                    get = lambda p, k: np.ctypeslib.as_array(p, shape=(k,)).copy()
                    mat = lambda p: get(p, count * n).reshape((count, n), order="F")
                    hist = {"History_Time": get(t, count), "History_Dt": get(dt, count), "History_V": mat(V),
                            "History_Disp": mat(D), "History_Theta": mat(TH), "History_NormalStress": mat(SIG)}
                    return 1 if on_records(int(first), hist) else 0
                except Exception:  # a Python error stops the run instead of crossing the C boundary
                    self._cb_error = True
                    return 1
            cb = _lib.RECORD_CB(_cb)
        self._cb_error = False
        rc = lib().qdfn_sim_run_until(self._h, float(t_end), int(max_steps), int(graph_attempts), cb, None)
        if rc < 0:
            raise _lib.QdfnError(rc, _lib.last_error())
        if self._cb_error:
            raise RuntimeError("the record callback raised; the run was stopped")
        return self.stats()

    def stats(self):
        s = SimStats()
        check(lib().qdfn_sim_get_stats(self._h, s))
        return {k: getattr(s, k) for k, _ in SimStats._fields_ if k != "pad_"}

    def state(self):
        out = [np.zeros(self.n) for _ in range(6)]
        check(lib().qdfn_sim_get_state(self._h, *[ptr(o) for o in out]))
        return dict(zip(("V", "Disp", "Theta", "Friction", "Sigma", "Accel"), out))

    def history(self):
        cnt = self.stats()["n_recorded"]
        t, dt = np.zeros(cnt), np.zeros(cnt)
        mats = [np.zeros((cnt, self.n), order="F") for _ in range(4)]
        check(lib().qdfn_sim_get_history(self._h, cnt, ptr(t), ptr(dt), *[m.ctypes.data_as(_lib.PD) for m in mats]))
        return {"History_Time": t, "History_Dt": dt, "History_V": mats[0], "History_Disp": mats[1],
                "History_Theta": mats[2], "History_NormalStress": mats[3]}

    def adjoint(self, v_i, theta_i, obs, d_target, w_time=0.0, t_target=0.0, v_onset=1e-3):
        """Loss and gradients (a, b, Dc) of
        L = sum_o (Disp[obs_o] - d_target_o)^2 + w_time (t_first - t_target)^2
        over every step this run has taken (see qdfn_sim_adjoint); the run must
        have been created with record_step=1 and record_capacity >= its steps,
        or advanced by run_checkpointed.
        Returns (loss, {"a": ..., "b": ..., "dc": ...}, info dict)."""
        obs = np.ascontiguousarray(obs, dtype=np.int64).ravel()
        tgt = _f64(d_target)
        if tgt.size != obs.size:
            raise ValueError("one target per observation element")
        vi, th = _f64(v_i), _f64(theta_i)
        if vi.size != self.n or th.size != self.n:
            raise ValueError("v_i and theta_i must have one entry per element")
        if obs.size and (obs.min() < 0 or obs.max() >= self.n):
            raise ValueError("observation indices must lie in [0, n)")
        ga, gb, gd = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        loss = _lib.D()
        info = _lib.AdjointInfo()
        check(lib().qdfn_sim_adjoint(self._h, ptr(vi), ptr(th), obs.size, obs.ctypes.data_as(_lib.C.POINTER(_lib.I64)),
                                     ptr(tgt), float(w_time), float(t_target), float(v_onset), _lib.C.byref(loss),
                                     ptr(ga), ptr(gb), ptr(gd), _lib.C.byref(info)))
        return loss.value, {"a": ga, "b": gb, "dc": gd}, {k: getattr(info, k) for k, _ in info._fields_}

    def event_catalog(self, FaultCenter, FaultArea, ShearModulus, DispRateCrits=1e-2):
        """Fragments of get_event_fragments (scripts/Functions_Plot.jl:430-510)
        from this run's device-resident histories; see qdfn.catalog."""
        from .catalog import EventCatalog
        return EventCatalog.from_sim(self, FaultCenter, FaultArea, ShearModulus, DispRateCrits)

    def solver_history(self):
        """Solver of each element in each recorded step (1 HighV, 0 LowV), records x n."""
        cnt = self.stats()["n_recorded"]
        out = np.zeros((cnt, self.n), dtype=np.int8, order="F")
        check(lib().qdfn_sim_get_solver_history(self._h, cnt, out.ctypes.data_as(_lib.C.POINTER(_lib.C.c_int8))))
        return out

    def trace(self):
        cnt = self.stats()["n_traced"]
        a, b, c = np.zeros(cnt), np.zeros(cnt), np.zeros(cnt)
        check(lib().qdfn_sim_get_trace(self._h, cnt, ptr(a), ptr(b), ptr(c)))
        return {"told": a, "dt": b, "vmax": c}

    def close(self):
        if getattr(self, "_h", None):
            lib().qdfn_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def main(StiffnessMatrixShear, StiffnessMatrixNormal, NormalStiffnessZero, ShearModulus, FaultCount,
         LoadingFaultCount, Mass, a, b, Dc, ThetaI, Vini, FrictionI, InitialNormalStress, LoadingRate, TotalStep,
         RecordStep, SwitchV, DtCut, RuptureDt, MaximumDt, SaveResultFileName, RockDensity, FaultCenter,
         FaultLengthStrike, FaultLengthDip, FaultStrikeAngle, FaultDipAngle, FaultRakeAngle, SaveStep,
         MinimumNormalStress, Alpha_Evo, EvolutionDR, external=None, prec=_lib.PREC_FP64):
    """main() of scripts/Functions_RSFDFN3DMain_D.jl:2 on the GPU.  The
    reference reads Input_ExternalStressChange.jld2 from the working directory;
    here the tables are passed as external={"t","n","s"}.  Returns the History_*
    arrays (and writes them to SaveResultFileName when it is a str: the reference's
    Result.jld2 keys for a .jld2 name, else .npz)."""
    op = DenseOperator.from_matrices(StiffnessMatrixShear, StiffnessMatrixNormal, prec)
    prob = dict(n=FaultCount, n_loading=LoadingFaultCount, mass=Mass, a=a, b=b, dc=Dc, theta_i=ThetaI, v_i=Vini,
                friction_i=FrictionI, sigma0=InitialNormalStress, loading_rate=LoadingRate, mu=ShearModulus,
                rho=RockDensity, switch_v=SwitchV, dt_cut=DtCut, rupture_dt=RuptureDt, maximum_dt=MaximumDt,
                min_normal_stress=MinimumNormalStress, alpha_evo=Alpha_Evo, evolution_dr=EvolutionDR,
                planar=int(NormalStiffnessZero == 1), ext=external)
    n_rec = int(round(TotalStep / RecordStep))
    sim = Simulation(op, prob, record_step=RecordStep, record_capacity=n_rec, trace_capacity=0)
    sim.run(TotalStep)
    hist = sim.history()
    if isinstance(SaveResultFileName, str) and SaveResultFileName:
        if SaveResultFileName.endswith(".jld2"):  # the reference's Result.jld2 keys (qdfn.jld2)
            from . import jld2
            jld2.save_result(SaveResultFileName, hist)
        else:
            np.savez(SaveResultFileName, **hist)
    sim.close()
    op.close()
    return hist
