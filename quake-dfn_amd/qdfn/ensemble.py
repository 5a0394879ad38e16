"""Batched ensemble of independent runs (qdfn_ensemble_*, include/qdfn.h).

Each member is a run of main() (scripts/Functions_RSFDFN3DMain_D.jl:2-324)
with its own dense FP32 operator or one shared with members of the same
geometry, its own friction, Dt control and retries.  All members advance
together: one attempt of every member per batched launch sequence, replayed
from one captured HIP graph; members sharing an operator are read as one
multi-right-hand-side product.  A member's trajectory is bitwise the one it
has when run alone.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, check_handle, lib
from .sim import DenseOperator, Simulation


# Matrix megabytes per coupling workgroup of an ensemble member's operator.
# Alone, an operator of 500-4,000 elements is split into up to 256
# workgroups to fill the GPU, and its split partial sums (written by the
# coupling, read by the finish) cost about a third of the matrix bytes; in an
# ensemble the members fill the GPU together.  The 256-member ensemble with
# four right-hand sides per pass: 263k RHS evals/s with the single-run splits,
# 283k / 301k / 310k / 300k / 235k with 0.5 / 1 / 1.5 / 2 / 3.5 MB.
SPLIT_MB = 1.5


def ensemble_splits(n, split_mb=SPLIT_MB):
    """Column splits giving about split_mb MB of FP32 E and K_N per coupling
    workgroup (1,024 rows each) for an n-element operator."""
    ld = (n + 63) // 64 * 64
    rb = (ld + 1023) // 1024
    return max(1, int(round(2 * 4 * ld * n / (rb * split_mb * 1e6))))


class Ensemble:
    def __init__(self, sims, max_rhs=0):
        """sims: Simulation objects on FP32 DenseOperators (borrowed; kept alive
        by this object).  max_rhs: members per operator pass (1-4), 0 = autotune."""
        self.sims = list(sims)
        arr = (_lib.VP * len(self.sims))(*[s._h for s in self.sims])
        self._h = check_handle(lib().qdfn_ensemble_create(len(self.sims), arr, int(max_rhs)),
                               "qdfn_ensemble_create")
        self.ops = []

    @classmethod
    def from_members(cls, members, problem_fn, max_rhs=0, record_step=10, record_capacity=0, trace_capacity=0,
                     split_mb=SPLIT_MB):
        """Build one FP32 operator per distinct geometry (member["grid"]), one run
        per member with problem_fn(member, op) as its problem, and the ensemble.
        Each operator's coupling is split into workgroups of about split_mb MB
        of matrix (ensemble_splits; None keeps the single-run default)."""
        ops, sims = {}, []
        for mem in members:
            key = mem["grid"]
            if key not in ops:
                ops[key] = DenseOperator.from_rect_segments(mem["seg"], mem["mu"], mem["nu"], _lib.PREC_FP32)
                if split_mb:
                    ops[key].set_splits(ensemble_splits(ops[key].n, split_mb))
            op = ops[key]
            sims.append(Simulation(op, problem_fn(mem, op), record_step=record_step,
                                   record_capacity=record_capacity, trace_capacity=trace_capacity))
        e = cls(sims, max_rhs)
        e.ops = list(ops.values())
        return e

    def run(self, n_steps, t_end=0.0, graph_attempts=16):
        """Advance every member by up to n_steps accepted steps, a member stopping
        early once its simulated time reaches t_end (0: no time limit)."""
        check(lib().qdfn_ensemble_run(self._h, int(n_steps), float(t_end), int(graph_attempts)))

    def info(self):
        g, r, b = _lib.I64(), _lib.I32(), _lib.I64()
        check(lib().qdfn_ensemble_info(self._h, C.byref(g), C.byref(r), C.byref(b)))
        return {"groups": g.value, "max_rhs": r.value, "bytes_per_sweep": b.value}

    def stats(self):
        return [s.stats() for s in self.sims]

    def close(self):
        if getattr(self, "_h", None):
            lib().qdfn_ensemble_destroy(self._h)
            self._h = None
        for s in self.sims:
            s.close()
        for o in self.ops:
            o.close()
        self.sims, self.ops = [], []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def problem_for_member(mem, op):
    """The member's problem on its operator: RUN3 defaults, back-slip loading from
    the operator's own row sums (UniformBackslip.jl)."""
    from . import configs
    ks, _ = op.kself()
    ksum, knsum = op.matvec(np.ones(op.n), loading_form=False)
    return configs.problem_from_segments(mem, ks, ksum, knsum)
