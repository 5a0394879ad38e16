"""Loader for libqdfn_hip.so (the HIP/gfx950 product library) and its C-ABI.

The library is built in-tree by ``quake-dfn_amd/csrc/Makefile`` (or
``__graft_entry__.build()``).  There is no fallback: if the shared object is
missing or cannot be loaded, every entry point raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqdfn_hip.so")
_LIB = None

D = C.c_double
I32 = C.c_int32
I64 = C.c_int64
PD = C.POINTER(C.c_double)
PI32 = C.POINTER(C.c_int32)
VP = C.c_void_p

QDFN_OK = 0
QDFN_ERR_ARG = -1
QDFN_ERR_HIP = -2
QDFN_ERR_NOMEM = -3
QDFN_ERR_STATE = -4
QDFN_ERR_RETRIES = -10
QDFN_ERR_EMPTY_DTMIN = -11

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16_FAR = 2  # H operators: BF16 low-rank factors, FP32 dense blocks
PREC_MIXED = 3  # H operators: FP32, trailing low-rank components FP16 within each block's error budget
PREC_MIXED8 = 4  # H operators: PREC_MIXED with the smallest trailing components in FP8 (E4M3)


class QdfnError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qdfn error {code}: {msg}")
        self.code = code


class StepParams(C.Structure):
    _fields_ = [("conv_crit", D), ("v0", D), ("mu", D), ("rho", D), ("switch_v", D), ("alpha_evo", D),
                ("evolution_dr", I32), ("pad_", I32)]


class SimParams(C.Structure):
    _fields_ = [("n", I64), ("n_loading", I64), ("mu", D), ("rho", D), ("switch_v", D), ("dt_cut", D),
                ("rupture_dt", D), ("maximum_dt", D), ("min_normal_stress", D), ("alpha_evo", D),
                ("evolution_dr", I32), ("planar", I32), ("ext_at_told", I32), ("pad_", I32),
                ("record_step", I64), ("record_capacity", I64), ("trace_capacity", I64)]


class SimStats(C.Structure):
    _fields_ = [("steps", I64), ("rhs_evals", I64), ("retry_passes", I64), ("stalls", I64),
                ("n_recorded", I64), ("n_traced", I64), ("status", I32), ("pad_", I32), ("t", D), ("dt", D),
                ("coupling_ms_avg", D), ("coupling_timed", I64)]


class AdjointInfo(C.Structure):
    _fields_ = [("steps", I64), ("output_mismatches", I64), ("select_mismatches", I64), ("onset_step", I64),
                ("retried_steps", I64), ("t_first", D), ("ms_backward", D), ("lambda_growth", D),
                ("growth_step", I64)]


class HmatBuildStats(C.Structure):
    _fields_ = [("entries_evaluated", I64), ("steps", I64), ("parts_sampled", I64), ("blocks_fallback", I64),
                ("ms_build", D), ("ms_phase", D * 5)]


class HmatParams(C.Structure):
    _fields_ = [("min_elements_to_cut", I64), ("max_levels", I64), ("dist_diam_ratio", D),
                ("arrange_point", D * 3), ("n_loading", I64), ("kmeans_max_iter", I64)]


PI64 = C.POINTER(I64)
PPD = C.POINTER(PD)
# int cb(void* user, int64_t first, int64_t count, const double* time, const double* dt, const double* V,
#        const double* Disp, const double* Theta, const double* Sigma)
RECORD_CB = C.CFUNCTYPE(C.c_int, VP, I64, I64, PD, PD, PD, PD, PD, PD)

# name -> (restype, argtypes); mirrors include/qdfn.h
SIGNATURES = {
    "qdfn_last_error": (C.c_char_p, []),
    "qdfn_device_info": (C.c_int, [C.c_char_p, I64, PI32, C.POINTER(I64)]),
    "qdfn_okada_dc3d": (C.c_int, [D, I64, PD, PD, PD, D, D, D, D, D, D, D, D, D, PD, PD]),
    "qdfn_okada_stiffness_rect": (C.c_int, [PD, I64, I64, PD, I64, I64, D, D, PD, PD, I64]),
    "qdfn_tde_stress_hs": (C.c_int, [I64, PD, PD, PD, PD, PD, PD, D, D, D, D, D, PD, PD]),
    "qdfn_tde_stiffness_tri": (C.c_int, [I64, PD, PD, PD, PD, I64, PD, PD, PD, D, D, PD, PD, I64]),
    "qdfn_rotverts_unitvectors": (C.c_int, [I64, PD, PD, PD, PD, PD, PD, PD, PD]),
    "qdfn_dense_create": (VP, [PD, PD, I64, I64, I32]),
    "qdfn_dense_create_rect": (VP, [PD, I64, I64, D, D, I32]),
    "qdfn_dense_create_tri": (VP, [I64, PD, PD, PD, PD, PD, PD, PD, D, D, I32]),
    "qdfn_dense_destroy": (None, [VP]),
    "qdfn_dense_kself": (C.c_int, [VP, PD, PD]),
    "qdfn_dense_normal_zero": (C.c_int, [VP]),
    "qdfn_dense_set_splits": (C.c_int, [VP, I32]),
    "qdfn_matvec": (C.c_int, [VP, PD, PD, PD, I32]),
    "qdfn_matvec_device_noalloc": (C.c_int, [VP, VP, VP, VP, I32, VP]),
    "qdfn_dense_time_coupling": (C.c_int, [VP, I32, PD]),
    "qdfn_solve_one_timestep": (I64, [C.POINTER(StepParams), I64, I64] + [PD] * 9 + [D] + [PD] * 4 + [PI32] +
                                [PD] * 4 + [PD] * 6 + [PI32] * 3),
    "qdfn_hmat_structure": (C.c_int, [C.POINTER(HmatParams), I64, PD, PD, PI64, I64, PI64, PD, PI64, I64, PI64,
                                      PI64, PI64]),
    "qdfn_hmat_build_rect": (VP, [PD, I64, I64, D, D, I64, PI64, PI64, D, D, I32]),
    "qdfn_hmat_build_tri": (VP, [I64, PD, PD, PD, PD, PD, PD, PD, D, D, I64, PI64, PI64, D, D, I32]),
    "qdfn_hmat_build_rect_sampled": (VP, [PD, I64, I64, D, D, I64, PI64, PI64, D, D, I32]),
    "qdfn_hmat_build_tri_sampled": (VP, [I64, PD, PD, PD, PD, PD, PD, PD, D, D, I64, PI64, PI64, D, D, I32]),
    "qdfn_hmat_build_info": (C.c_int, [VP, C.c_void_p]),
    "qdfn_hmat_create": (VP, [I64, I64, PI64, PI64, PI32, PI64, PPD, PPD, PI32, PI64, PPD, PPD, I32]),
    "qdfn_hmat_destroy": (None, [VP]),
    "qdfn_hmat_info": (C.c_int, [VP, PI64, PI64, PI64, PI64, PI64]),
    "qdfn_hmat_time_phases": (C.c_int, [VP, I32, PD, PD, PD]),
    "qdfn_hmat_get_ranks": (C.c_int, [VP, PI32, PI64, PI32, PI64]),
    "qdfn_hmat_get_block": (C.c_int, [VP, I64, I32, PD, PD]),
    "qdfn_hmat_get_block_stored": (C.c_int, [VP, I64, I32, PD, PD]),
    "qdfn_hmat_get_split": (C.c_int, [VP, PI64, PI64]),
    "qdfn_hmat_get_split8": (C.c_int, [VP, PI64, PI64]),
    "qdfn_fp8_round": (C.c_int, [PD, PD, I64]),
    "qdfn_hmat_bytes": (C.c_int, [VP, PI64]),
    "qdfn_hmat_kself": (C.c_int, [VP, PD, PD]),
    "qdfn_hmat_normal_zero": (C.c_int, [VP]),
    "qdfn_hmat_reduce_strong_interaction": (C.c_int, [VP, D, C.POINTER(I64)]),
    "qdfn_hmat_matvec": (C.c_int, [VP, PD, PD, PD, I32]),
    "qdfn_hmat_time_matvec": (C.c_int, [VP, I32, PD]),
    "qdfn_sim_create_dense": (VP, [VP, C.POINTER(SimParams)] + [PD] * 9),
    "qdfn_sim_create_hmat": (VP, [VP, C.POINTER(SimParams)] + [PD] * 9),
    "qdfn_sim_set_external": (C.c_int, [VP, I64, PD, PD, PD]),
    "qdfn_sim_run": (C.c_int, [VP, I64, I32, I32]),
    "qdfn_sim_run_until": (C.c_int, [VP, D, I64, I32, VP, VP]),
    "qdfn_sim_set_convergence": (C.c_int, [VP, D]),
    "qdfn_sim_run_checkpointed": (C.c_int, [VP, I64, I64, I32]),
    "qdfn_sim_get_stats": (C.c_int, [VP, C.POINTER(SimStats)]),
    "qdfn_sim_get_state": (C.c_int, [VP, PD, PD, PD, PD, PD, PD]),
    "qdfn_sim_get_history": (C.c_int, [VP, I64, PD, PD, PD, PD, PD, PD]),
    "qdfn_sim_get_trace": (C.c_int, [VP, I64, PD, PD, PD]),
    "qdfn_sim_get_solver_history": (C.c_int, [VP, I64, C.POINTER(C.c_int8)]),
    "qdfn_sim_destroy": (None, [VP]),
    "qdfn_sim_adjoint": (C.c_int, [VP, PD, PD, I64, C.POINTER(I64), PD, D, D, D, PD, PD, PD, PD,
                                   C.POINTER(AdjointInfo)]),
    "qdfn_ensemble_create": (VP, [I64, C.POINTER(VP), I32]),
    "qdfn_ensemble_run": (C.c_int, [VP, I64, D, I32]),
    "qdfn_ensemble_info": (C.c_int, [VP, C.POINTER(I64), C.POINTER(I32), C.POINTER(I64)]),
    "qdfn_ensemble_destroy": (None, [VP]),
    "qdfn_catalog_from_sim": (VP, [VP, PD, PD, D, D]),
    "qdfn_catalog_from_history": (VP, [I64, I64, PD, PD, I64, PD, PD, D, D]),
    "qdfn_catalog_from_fragments": (VP, [I64, PD, PD, PD]),
    "qdfn_catalog_fragments": (C.c_int, [VP, C.POINTER(I64), PD, PD, PD, PI32]),
    "qdfn_catalog_merge": (C.c_int, [VP, D, D, C.POINTER(I64)]),
    "qdfn_catalog_events": (C.c_int, [VP, C.POINTER(I64), PD, PD, PD, PD, PI32]),
    "qdfn_catalog_destroy": (None, [VP]),
}


def lib():
    """The loaded C-ABI library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C quake-dfn_amd/csrc` "
                              "or __graft_entry__.build(); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def last_error():
    return lib().qdfn_last_error().decode(errors="replace")


def check(rc):
    if rc < 0:
        raise QdfnError(rc, last_error())
    return rc


def check_handle(h, what):
    if not h:
        raise QdfnError(QDFN_ERR_STATE, f"{what}: {last_error()}")
    return h


def ptr(a):
    """Pointer to a numpy float64 array (None passes NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(PD)


def ptr_i32(a):
    return a.ctypes.data_as(PI32)


def device_info():
    name = C.create_string_buffer(128)
    cu = I32()
    hbm = I64()
    check(lib().qdfn_device_info(name, 128, C.byref(cu), C.byref(hbm)))
    return {"arch": name.value.decode(), "cu_count": cu.value, "hbm_bytes": hbm.value}
