"""ctypes binding of libaar_hip.so (the C ABI declared in include/aar_hip.h).

torch is imported before the library is loaded so that both share one HIP
runtime (libamdhip64.so.7) in the process.  There is no fallback: if the
library is not built, or no GPU is visible, loading fails loudly.
"""
from __future__ import annotations

import ctypes
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libaar_hip.so")

_lib = None


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("fdn", ctypes.c_int), ("a", ctypes.c_double), ("sigma", ctypes.c_double)]


class Instance(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("a", ctypes.c_double), ("sigma", ctypes.c_double),
                ("x", ctypes.c_void_p), ("b", ctypes.c_void_p)]


class Stats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("relres", ctypes.c_double),
        ("rhs_norm", ctypes.c_double),
        ("converged", ctypes.c_int),
        ("printed_converged", ctypes.c_int),
        ("printed_warning", ctypes.c_int),
        ("status", ctypes.c_int),
        ("hist_len", ctypes.c_int),
        ("solve_ms", ctypes.c_double),
        ("bytes", ctypes.c_double),
        ("outer_steps", ctypes.c_int),
        ("inner_iterations", ctypes.c_int),
        ("reason", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# every symbol include/*.h declares (tests check the library exports them)
REFERENCE_SYMBOLS = [
    "AAR", "AndersonExtrapolation", "PGR", "Galerkin_Richardson", "PL2R", "L2_Richardson", "CG",
    "CheckInputs", "Processor_Domain", "Comm_topologies", "Max_layer", "Initialize",
    "Deallocate_memory", "Vec_copy", "Lap_coefficient", "Lap_Vec_mult", "Precondition",
    "Lap_Vec_mult_vc", "Precondition_vc",
    "Find_size_dir", "Get_block_origin_global_coords",
    "pythag", "SingularValueDecomp", "PseudoInverseTimesVec", "Vector2Norm", "VectorDotProduct",
]
EXTENSION_SYMBOLS = [
    "aar_hip_matvec", "aar_hip_matvec_f32", "aar_hip_aar", "aar_hip_aar_f32", "aar_hip_aar_mixed",
    "aar_hip_pgr",
    "aar_hip_cg", "aar_hip_aar_refined", "aar_hip_pinv", "aar_hip_history", "aar_hip_last_stats",
    "aar_hip_set_stream", "aar_hip_set_tuning", "aar_hip_set_option", "aar_hip_profile", "aar_hip_profile_reset",
    "aar_hip_profile_query", "aar_hip_release", "aar_hip_last_error",
    "aar_hip_helm_matvec", "aar_hip_helm_aar", "aar_hip_helm_grad_c", "aar_hip_aar_batch",
    "aar_hip_autotune", "aar_hip_last_system", "aar_hip_matvec_vc", "aar_hip_aar_vc", "aar_hip_set_coefficient",
    "aar_hip_pgr_vc", "aar_hip_cg_vc", "aar_hip_gmres", "aar_hip_bicg", "aar_hip_ilu0_apply", "aar_hip_aar_petsc",
    "aar_hip_pgr_petsc", "aar_hip_helm_gmres", "aar_hip_helm_lgmres", "aar_hip_helm_bicg", "aar_hip_helm_ilu0_apply",
    "aar_hip_stream_bandwidth", "aar_hip_pinv_noalloc", "aar_hip_set_stencil", "aar_hip_kspcg",
]


def lib_path():
    return LIB_PATH


def load(require_gpu=True):
    """Load libaar_hip.so (after torch).  Raises if it is not built or, with
    require_gpu, if no GPU is visible."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (shares the HIP runtime with the library)
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python linear-solver_amd/aar_hip/build.py` "
                           "(libaar_hip has no CPU fallback)")
    if require_gpu and not torch.cuda.is_available():
        raise RuntimeError("libaar_hip needs a GPU; none is visible (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    I, D, P = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    G = ctypes.POINTER(Grid)
    S = ctypes.POINTER(Stats)
    sig = {
        "aar_hip_matvec": [G, P, P],
        "aar_hip_matvec_f32": [G, P, P],
        "aar_hip_aar": [G, P, P, D, D, I, I, I, D, I, S],
        "aar_hip_aar_f32": [G, P, P, D, D, I, I, I, D, I, S],
        "aar_hip_aar_mixed": [G, P, P, D, D, I, I, I, D, I, S],
        "aar_hip_pgr": [G, P, P, D, I, I, I, D, I, I, S],
        "aar_hip_cg": [G, P, P, I, D, I, S],
        "aar_hip_aar_refined": [G, P, P, D, D, I, I, I, I, D, D, I, S],
        "aar_hip_pinv": [P, P, P, I, P, P],
        "aar_hip_history": [P, I],
        "aar_hip_last_stats": [S],
        "aar_hip_set_stream": [P],
        "aar_hip_set_tuning": [I, I, I],
        "aar_hip_set_option": [ctypes.c_char_p, ctypes.c_long],
        "aar_hip_profile": [I],
        "aar_hip_profile_reset": [],
        "aar_hip_profile_query": [ctypes.c_char_p, ctypes.POINTER(I), ctypes.POINTER(D),
                                  ctypes.POINTER(D)],
        "aar_hip_release": [],
        "aar_hip_helm_matvec": [P, P, P],
        "aar_hip_helm_aar": [P, P, P, D, D, I, I, I, D, I, S],
        "aar_hip_helm_grad_c": [ctypes.c_long, P, P, P],
        "aar_hip_helm_gmres": [P, P, P, I, I, D, I, S],
        "aar_hip_helm_lgmres": [P, P, P, I, I, I, D, I, S],
        "aar_hip_helm_bicg": [P, P, P, I, D, I, S],
        "aar_hip_helm_ilu0_apply": [P, P, P, I],
        "aar_hip_aar_batch": [ctypes.POINTER(Instance), I, D, D, I, I, I, D, I, I, S],
        "aar_hip_autotune": [I, I, I, I, I, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I),
                             ctypes.POINTER(D)],
        "aar_hip_last_system": [I, P, P],
        "aar_hip_matvec_vc": [G, P, P, P],
        "aar_hip_aar_vc": [G, P, P, P, D, D, I, I, I, D, I, I, S],
        "aar_hip_set_coefficient": [P, ctypes.c_long],
        "aar_hip_pgr_vc": [G, P, P, P, D, I, I, I, D, I, I, S],
        "aar_hip_cg_vc": [G, P, P, P, I, D, I, S],
        "aar_hip_gmres": [G, P, P, P, I, I, D, I, S],
        "aar_hip_bicg": [G, P, P, P, I, D, I, S],
        "aar_hip_ilu0_apply": [G, P, P, P, I],
        "aar_hip_aar_petsc": [G, P, P, P, D, D, I, I, I, D, I, S],
        "aar_hip_pgr_petsc": [G, P, P, P, D, I, I, I, D, I, I, S],
        "aar_hip_stream_bandwidth": [I, P, P, ctypes.c_long, I, ctypes.POINTER(D)],
        "aar_hip_pinv_noalloc": [P, P, P, I, P, P, P],
        "aar_hip_set_stencil": [P, I],
        "aar_hip_kspcg": [G, P, P, P, I, D, I, S],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = I
    L.aar_hip_last_error.argtypes = []
    L.aar_hip_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def load_for_symbols():
    """Load without requiring a GPU (symbol checks on a CPU-only host)."""
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built")
    import torch  # noqa: F401
    return ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)


class AarHipError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        msg = _lib.aar_hip_last_error().decode() if _lib is not None else ""
        raise AarHipError(f"libaar_hip error {rc}: {msg}")
