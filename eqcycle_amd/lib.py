"""ctypes binding of libeqcycle_hip.so (the C ABI declared in include/eqcycle_hip.h).

The library is loaded from the package directory.  There is no fallback: if
the shared object is missing, or no GPU is visible when a context is
created, the call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# EQCYCLE_HIP_LIB: another build of the library (e.g. one with other
# build-time options), as the Julia binding allows
LIB_PATH = os.environ.get("EQCYCLE_HIP_LIB", os.path.join(HERE, "libeqcycle_hip.so"))

EQ_OK, EQ_EINVAL, EQ_HIP_ERR, EQ_NR_NOCONV, EQ_NR_NAN, EQ_PCG_NOCONV, EQ_ENOTREADY, EQ_NOMEM = range(8)
STATUS_NAMES = {0: "EQ_OK", 1: "EQ_EINVAL", 2: "EQ_HIP_ERR", 3: "EQ_NR_NOCONV", 4: "EQ_NR_NAN",
                5: "EQ_PCG_NOCONV", 6: "EQ_ENOTREADY", 7: "EQ_NOMEM"}
EQ_MASK_ALL, EQ_MASK_INTERIOR, EQ_MASK_NEG, EQ_MASK_FAULT_ROWS = 0, 1, 2, 3
EQ_PREC_FP64, EQ_PREC_FP32, EQ_PREC_MIXED, EQ_PREC_F32PC = 0, 1, 2, 3
EQ_PC_NONE, EQ_PC_JACOBI, EQ_PC_AMG = 0, 1, 2
EQ_SMOOTH_JACOBI, EQ_SMOOTH_CHEBYSHEV, EQ_SMOOTH_CHEBYSHEV_TILES = 0, 1, 2
EQ_GUESS_PREVIOUS, EQ_GUESS_LEARNED, EQ_GUESS_GALERKIN, EQ_GUESS_LINEAR = 0, 1, 2, 3

# every symbol the header declares (checked by the CPU test-suite)
EXPORTS = [
    "eq_create", "eq_destroy", "eq_last_error", "eq_sizes", "eq_set_material", "eq_set_params",
    "eq_ku_apply", "eq_kni_apply", "eq_element_computation", "eq_amg_setup", "eq_amg_info", "eq_amg_l0_info", "eq_amg_coarse_mode",
    "eq_amg_ldiv", "eq_cg", "eq_pcg_jacobi", "eq_ids", "eq_ids2", "eq_nrsearch", "eq_fbc",
    "eq_slr", "eq_dtevol", "eq_set_solver", "eq_state_init", "eq_state_get", "eq_state_set",
    "eq_step", "eq_run", "eq_time_kernel", "eq_learn_init", "eq_learn_record", "eq_learn_info",
    "eq_learn_gram", "eq_learn_set_basis", "eq_learn_train", "eq_learn_loss_grad", "eq_learn_get",
    "eq_learn_predict", "eq_learn_norms", "eq_learn_galerkin", "eq_newmark_step", "eq_amg_upload",
    "eq_launch_timing", "eq_launch_timing_get", "eq_newmark_step_noalloc", "eq_learn_set_linear",
    "eq_learn_predict_linear", "eq_learn_snapshots", "eq_rs_build", "eq_rs_info", "eq_rs_level", "eq_rs_split", "eq_rs_free",
]
EQ_TIME_KU, EQ_TIME_CG_DIR, EQ_TIME_VCYCLE, EQ_TIME_CG_ITER, EQ_TIME_NEWMARK, EQ_TIME_CHEB, EQ_TIME_COPY, \
    EQ_TIME_CHEB2, EQ_TIME_SWEEP, EQ_TIME_SWEEP_POST, EQ_TIME_RESTRICT = range(11)
TK_NAMES = ["newmark_kernel", "cheb_pair_kernel<float>", "cheb_pair_kernel<double>",
            "tile_kernel<float, OpCheb>", "tile_kernel<double, OpCheb>", "tile_kernel<T, OpCgDir>",
            "cg_update_kernel<T>", "cheb_sweep_kernel<float>", "cheb_sweep_kernel<double>"]
EQ_TK_COUNT = len(TK_NAMES)

dp = C.POINTER(C.c_double)
i64p = C.POINTER(C.c_int64)
i32p = C.POINTER(C.c_int32)
fp = C.POINTER(C.c_float)


class MeshDesc(C.Structure):
    _fields_ = [("ngll", C.c_int64), ("nelx", C.c_int64), ("nely", C.c_int64), ("dxe", C.c_double),
                ("dye", C.c_double), ("iglob", i64p), ("xgll", dp), ("wgll", dp), ("H", dp)]


class ParamsDesc(C.Structure):
    _fields_ = [("FltNglob", C.c_int64), ("nglob", C.c_int64), ("IDstate", C.c_int64),
                ("Vpl", C.c_double), ("Vthres", C.c_double), ("Vevne", C.c_double), ("dt0", C.c_double),
                ("fo", dp), ("Vo", dp), ("xLf", dp), ("M", dp), ("BcLC", dp), ("BcTC", dp),
                ("FltB", dp), ("FltZ", dp), ("FltX", dp), ("cca", dp), ("ccb", dp), ("Seff", dp),
                ("tauo", dp), ("XiLf", dp), ("iFlt", i64p), ("iBcL", i64p), ("n_iBcL", C.c_int64),
                ("iBcT", i64p), ("n_iBcT", C.c_int64), ("FltIglobBC", i64p), ("n_FltIglobBC", C.c_int64),
                ("out_seis", i64p), ("n_out_seis", C.c_int64)]


class FaultParams(C.Structure):
    """eq_fault_params: the params_farray fields FBC! and slrFunc! read."""
    _fields_ = [(k, dp) for k in ("fo", "Vo", "xLf", "cca", "ccb", "Seff", "tauo", "FltZ")]


class SolverOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxiter", C.c_int64), ("pc", C.c_int), ("prec", C.c_int),
                ("guess", C.c_int), ("record_snapshots", C.c_int), ("graphs", C.c_int),
                ("cheb_single", C.c_int), ("sync_pass2", C.c_int),
                ("galerkin_full", C.c_int), ("basis_fp16", C.c_int), ("extrapolate", C.c_int),
                ("galerkin_split", C.c_int)]


class CsrDesc(C.Structure):
    _fields_ = [("n", C.c_int64), ("m", C.c_int64), ("rowptr", C.POINTER(C.c_int64)),
                ("col", C.POINTER(C.c_int64)), ("val", C.POINTER(C.c_double))]


class StepInfo(C.Structure):
    _fields_ = [("t", C.c_double), ("dt_used", C.c_double), ("dt_next", C.c_double), ("Vfmax", C.c_double),
                ("it", C.c_int64), ("isolver_used", C.c_int64), ("isolver_next", C.c_int64),
                ("pcg_iters", C.c_int64 * 2), ("nr_noconv", C.c_int64)]


class Results(C.Structure):
    _fields_ = [("cap_steps", C.c_int64), ("cap_events", C.c_int64), ("cap_is", C.c_int64),
                ("cap_seis", C.c_int64), ("cap_rec", C.c_int64),
                ("time_", dp), ("Vfmax", dp), ("isolver", i64p), ("pcg_iters", i64p),
                ("tStart", dp), ("tEnd", dp), ("hypo", dp),
                ("taubefore", dp), ("tauafter", dp), ("delfafter", dp),
                ("is_slip", dp), ("is_slipvel", dp), ("is_stress", dp),
                ("seismic_slip", dp), ("seismic_slipvel", dp), ("seismic_stress", dp),
                ("dSeis", dp), ("vSeis", dp), ("aSeis", dp),
                ("n_steps", C.c_int64), ("n_start", C.c_int64), ("n_end", C.c_int64), ("n_is", C.c_int64),
                ("n_seis", C.c_int64), ("n_rec", C.c_int64), ("n_qs", C.c_int64), ("n_dyn", C.c_int64),
                ("total_pcg_iters", C.c_int64), ("nr_noconv", C.c_int64), ("wall_qs", C.c_double),
                ("wall_dyn", C.c_double)]


_lib = None


def load():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m eqcycle_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "eq_create": ([C.POINTER(MeshDesc), C.POINTER(vp)], C.c_int),
        "eq_destroy": ([vp], None),
        "eq_last_error": ([vp], C.c_char_p),
        "eq_sizes": ([vp, i64p, i64p, i64p], C.c_int),
        "eq_set_material": ([vp, dp], C.c_int),
        "eq_set_params": ([vp, C.POINTER(ParamsDesc)], C.c_int),
        "eq_ku_apply": ([vp, dp, dp, C.c_int, C.c_int], C.c_int),
        "eq_kni_apply": ([vp, dp, dp, C.c_int], C.c_int),
        "eq_element_computation": ([vp, dp, dp], C.c_int),
        "eq_amg_setup": ([vp, C.c_double, C.c_int, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
        "eq_amg_info": ([vp, C.POINTER(C.c_int), i64p, i64p, C.c_int], C.c_int),
        "eq_amg_l0_info": ([vp, C.POINTER(C.c_int), C.POINTER(C.c_int), dp, dp], C.c_int),
        "eq_amg_coarse_mode": ([vp, C.c_int], C.c_int),
        "eq_amg_ldiv": ([vp, dp, dp, C.c_int], C.c_int),
        "eq_cg": ([vp, dp, dp, C.c_double, C.c_int64, C.c_int, C.c_int, i64p, dp], C.c_int),
        "eq_pcg_jacobi": ([vp, dp, dp, C.c_double, C.c_int64, i64p], C.c_int),
        "eq_ids": ([vp, C.c_int64, dp, dp, dp, C.c_double, dp, C.c_double, C.c_int64, dp], C.c_int),
        "eq_ids2": ([vp, C.c_int64, dp, dp, dp, dp, C.c_double, dp, dp, C.c_int64, dp], C.c_int),
        "eq_nrsearch": ([vp, C.c_int64] + [dp] * 12 + [i32p], C.c_int),
        "eq_fbc": ([vp, C.c_int64, C.POINTER(FaultParams), C.c_int64, C.c_int64] + [dp] * 9 + [C.c_double, i64p],
                   C.c_int),
        "eq_slr": ([vp, C.POINTER(FaultParams), C.c_int64, C.c_int64, dp, dp, dp, dp, C.c_int64, dp, C.c_double],
                   C.c_int),
        "eq_dtevol": ([vp, C.c_double, C.c_double, dp, C.c_int64, C.c_int64, dp, C.c_int64, dp], C.c_int),
        "eq_set_solver": ([vp, C.POINTER(SolverOpts)], C.c_int),
        "eq_newmark_step": ([vp, dp, dp, dp, dp, dp, dp, dp, C.c_double], C.c_int),
        "eq_newmark_step_noalloc": ([vp, dp, dp, dp, dp, dp, dp, dp, C.c_double], C.c_int),
        "eq_amg_upload": ([vp, C.c_int, C.POINTER(CsrDesc), C.POINTER(CsrDesc), dp, C.c_int, C.c_int], C.c_int),
        "eq_state_init": ([vp, C.c_double, dp], C.c_int),
        "eq_state_get": ([vp, dp, dp, dp, dp, dp, dp, dp, dp, i64p, i64p], C.c_int),
        "eq_state_set": ([vp, dp, dp, dp, dp, dp, C.c_double, C.c_double, C.c_int64], C.c_int),
        "eq_step": ([vp, C.POINTER(StepInfo)], C.c_int),
        "eq_run": ([vp, C.c_double, C.c_int64, C.POINTER(Results)], C.c_int),
        "eq_time_kernel": ([vp, C.c_int, C.c_int, C.c_int, dp], C.c_int),
        "eq_learn_init": ([vp, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_uint64], C.c_int),
        "eq_learn_record": ([vp, C.c_int], C.c_int),
        "eq_learn_info": ([vp, i64p, i64p, C.POINTER(C.c_int), i64p], C.c_int),
        "eq_learn_gram": ([vp, dp], C.c_int),
        "eq_learn_set_basis": ([vp, dp, dp, i32p], C.c_int),
        "eq_learn_norms": ([vp, dp], C.c_int),
        "eq_learn_train": ([vp, C.c_int, C.c_int, C.c_double, C.c_uint64, dp], C.c_int),
        "eq_learn_loss_grad": ([vp, fp, C.c_int, i32p, C.c_int, dp, fp], C.c_int),
        "eq_learn_get": ([vp, fp, fp, fp, fp], C.c_int),
        "eq_learn_snapshots": ([vp, fp], C.c_int),
        "eq_learn_predict": ([vp, dp, dp], C.c_int),
        "eq_learn_set_linear": ([vp, fp, fp], C.c_int),
        "eq_learn_predict_linear": ([vp, dp, dp], C.c_int),
        "eq_learn_galerkin": ([vp, dp, dp, dp], C.c_int),
        "eq_launch_timing": ([vp, C.c_int], C.c_int),
        "eq_launch_timing_get": ([vp, i64p, dp], C.c_int),
        "eq_rs_build": ([C.POINTER(CsrDesc), C.c_double, C.c_int, C.c_int64, C.c_int, C.POINTER(vp)], C.c_int),
        "eq_rs_info": ([vp, C.POINTER(C.c_int), i64p, i64p, i64p, C.c_int], C.c_int),
        "eq_rs_level": ([vp, C.c_int, C.c_int, i64p, i64p, dp], C.c_int),
        "eq_rs_split": ([vp, C.c_int, i32p], C.c_int),
        "eq_rs_free": ([vp], None),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def ptr(a: np.ndarray | None, ctype=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def i64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int64)
