"""ctypes access to the parity checkers.  TEST INFRASTRUCTURE ONLY.

* ``liboracle.so`` -- the C restatement of the reference MPI_REAL path
  (``aar_oracle.c``), bit-faithful to the reference on one rank.
* ``_ref/ref_harness`` -- the real reference solvers compiled from the
  reference tree (``Makefile``), driven through their public entry points.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg
use this module.  Arrays are flat float64 in the reference layout
``idx = i + j*nx + k*nx*ny`` (MPI_REAL/system.c:82); numpy arrays of shape
(nz, ny, nx) in C order have exactly that layout.
"""
from __future__ import annotations

import ctypes
import os
import re
import shutil
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")
MPIRUN = "/opt/conda/bin/mpirun"

_lib = None


class OrcStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("relres", ctypes.c_double),
        ("rhs_norm", ctypes.c_double),
        ("converged", ctypes.c_int),
        ("printed_converged", ctypes.c_int),
        ("printed_warning", ctypes.c_int),
        ("trace_len", ctypes.c_int),
        ("seconds", ctypes.c_double),
        ("status", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_D = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile liboracle.so (and the reference harness when the reference
    tree is present).  Called by __graft_entry__.build()."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "oracle"], check=True)
    if os.path.isdir(os.environ.get("AAR_REFERENCE", "/root/reference") + "/MPI_REAL"):
        subprocess.run(["make", "-s", "-C", HERE, "-B", "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C {HERE} oracle`")
        L = ctypes.CDLL(LIB_PATH)
        I, Dd = ctypes.c_int, ctypes.c_double
        L.orc_lap_coefficient.argtypes = [_D, I]
        L.orc_matvec.argtypes = [I, I, I, I, _D, Dd, Dd, _D, _D]
        L.orc_norm2.argtypes = [_D, ctypes.c_long]
        L.orc_norm2.restype = Dd
        L.orc_pinv_times_vec.argtypes = [_D, _D, _D, I, _D, _D]
        L.orc_pinv_times_vec.restype = I
        L.orc_pinv_times_vec_textbook.argtypes = [_D, _D, _D, I, _D, _D]
        L.orc_pinv_times_vec_gelsd.argtypes = [_D, _D, _D, I, _D, _D]
        L.orc_pinv_times_vec_gelsd.restype = I
        L.orc_pinv_times_vec_textbook.restype = I
        L.orc_pythag.argtypes = [Dd, Dd]
        L.orc_pythag.restype = Dd
        L.orc_anderson_extrapolation.argtypes = [_D, _D, _D, Dd, I, ctypes.c_long, _D, _D, _D]
        L.orc_anderson_extrapolation.restype = I
        L.orc_aar.argtypes = [I, I, I, I, Dd, Dd, _D, _D, Dd, Dd, I, I, I, Dd, I, _D, I,
                              ctypes.POINTER(OrcStats)]
        L.orc_pgr.argtypes = [I, I, I, I, Dd, Dd, _D, _D, Dd, I, I, I, Dd, I, I, _D, I,
                              ctypes.POINTER(OrcStats)]
        L.orc_cg.argtypes = [I, I, I, I, Dd, Dd, _D, _D, I, Dd, I, _D, I, ctypes.POINTER(OrcStats)]
        L.orc_aar_complex.argtypes = [I, I, I, I, Dd, _D, Dd, Dd, _D, _D, Dd, Dd, I, I, I, Dd, _D, I,
                                      ctypes.POINTER(OrcStats)]
        L.orc_matvec_complex.argtypes = [I, I, I, I, _D, Dd, _D, Dd, Dd, _D, _D]
        L.orc_matvec_vc.argtypes = [I, I, I, I, _D, Dd, _D, _D, _D]
        L.orc_aar_vc.argtypes = [I, I, I, I, Dd, _D, _D, _D, Dd, Dd, I, I, I, Dd, I, _D, I,
                                 ctypes.POINTER(OrcStats)]
        L.orc_pgr_vc.argtypes = [I, I, I, I, Dd, _D, _D, _D, Dd, I, I, I, Dd, I, I, _D, I,
                                 ctypes.POINTER(OrcStats)]
        L.orc_cg_vc.argtypes = [I, I, I, I, Dd, _D, _D, _D, I, Dd, I, _D, I, ctypes.POINTER(OrcStats)]
        L.orc_gmres.argtypes = [I, I, I, I, Dd, Dd, _D, _D, _D, I, I, Dd, I, _D, I, ctypes.POINTER(OrcStats)]
        L.orc_bicg.argtypes = [I, I, I, I, Dd, Dd, _D, _D, _D, I, Dd, I, _D, I, ctypes.POINTER(OrcStats), _D]
        L.orc_kspcg.argtypes = [I, I, I, I, Dd, Dd, _D, _D, _D, I, Dd, I, _D, I, ctypes.POINTER(OrcStats)]
        L.orc_ilu0_factor.argtypes = [I, I, I, I, Dd, Dd, _D, _D]
        L.orc_aar_petsc.argtypes = [I, I, I, I, Dd, Dd, _D, _D, _D, Dd, Dd, I, I, I, Dd, I, _D, I,
                                    ctypes.POINTER(OrcStats)]
        L.orc_gmres_complex.argtypes = [I, I, I, I, Dd, _D, Dd, Dd, _D, _D, I, I, Dd, I, _D, I,
                                        ctypes.POINTER(OrcStats)]
        L.orc_bicg_complex.argtypes = [I, I, I, I, Dd, _D, Dd, Dd, _D, _D, I, Dd, I, _D, I,
                                       ctypes.POINTER(OrcStats)]
        L.orc_lgmres_complex.argtypes = [I, I, I, I, Dd, _D, Dd, Dd, _D, _D, I, I, I, Dd, I, _D, I,
                                         ctypes.POINTER(OrcStats)]
        L.orc_ilu0_factor_complex.argtypes = [I, I, I, I, Dd, _D, Dd, Dd, _D]
        L.orc_ilu0_factor_complex.restype = I
        L.orc_ilu0_apply_complex.argtypes = [I, I, I, _D, _D, _D, I]
        L.orc_ilu0_apply_complex.restype = None
        L.orc_pgr_petsc.argtypes = [I, I, I, I, Dd, Dd, _D, _D, _D, Dd, I, I, I, Dd, I, I, _D, I,
                                    ctypes.POINTER(OrcStats)]
        L.orc_ilu0_apply.argtypes = [I, I, I, _D, _D, _D, I]
        L.orc_ilu0_apply.restype = None
        for fn in ("orc_aar", "orc_pgr", "orc_cg", "orc_aar_complex", "orc_aar_vc", "orc_pgr_vc", "orc_cg_vc",
                   "orc_gmres", "orc_bicg", "orc_ilu0_factor", "orc_aar_petsc", "orc_pgr_petsc",
                   "orc_gmres_complex", "orc_bicg_complex", "orc_lgmres_complex"):
            getattr(L, fn).restype = I
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_D) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def lap_coefficient(fdn=6):
    c = np.zeros(fdn + 1)
    lib().orc_lap_coefficient(_p(c), fdn)
    return c


def matvec(x, shape, a, sigma=0.0, fdn=6, coeff=None):
    """system.c:189-205 with the weights of order 2*fdn, or the caller's
    coeff[0..fdn]."""
    nx, ny, nz = shape
    x = _f64(x).reshape(-1)
    y = np.empty_like(x)
    c = lap_coefficient(fdn) if coeff is None else _f64(coeff).reshape(-1)[:fdn + 1].copy()
    lib().orc_matvec(nx, ny, nz, fdn, _p(c), a, sigma, _p(x), _p(y))
    return y


def norm2(v):
    v = _f64(v).reshape(-1)
    return lib().orc_norm2(_p(v), v.size)


def pinv_times_vec(A, b, textbook=False, gelsd=False):
    """tools.c:52-105 (textbook: the textbook left reflector; gelsd: the
    dgelsd rcond = -1 cutoff of the PETSc variant)."""
    A = _f64(A)
    m = A.shape[0]
    b = _f64(b)
    x = np.zeros(m)
    w = np.zeros(m)
    V = np.zeros((m, m))
    fn = (lib().orc_pinv_times_vec_gelsd if gelsd else
          lib().orc_pinv_times_vec_textbook if textbook else lib().orc_pinv_times_vec)
    st = fn(_p(A), _p(b), _p(x), m, _p(w), _p(V))
    if st:
        raise RuntimeError("SVD did not converge in 250 sweeps")
    return x, w, V


def anderson_extrapolation(DX, DF, f, beta):
    DX, DF, f = _f64(DX), _f64(DF), _f64(f)
    m, n = DX.shape
    am = np.zeros(n)
    G = np.zeros(m * m + m)
    s = np.zeros(m)
    lib().orc_anderson_extrapolation(_p(DX), _p(DF), _p(f), beta, m, n, _p(am), _p(G), _p(s))
    return am, G, s


def _run(fn, args, trace, trace_cap):
    st = OrcStats()
    tr = np.zeros(trace_cap) if trace else None
    rc = fn(*args, _p(tr), trace_cap if trace else 0, ctypes.byref(st))
    d = st.as_dict()
    d["rc"] = rc
    if trace:
        d["trace"] = tr[: st.trace_len].copy()
    return d


def aar(x0, rhs, shape, a, omega, beta, m, p, max_iter, tol, sigma=0.0, precond=True,
        fdn=6, trace=True):
    """AAR.c:37-177 on one rank.  Returns (x, stats-dict incl. 'trace')."""
    nx, ny, nz = shape
    x = _f64(x0).reshape(-1).copy()
    b = _f64(rhs).reshape(-1)
    d = _run(lib().orc_aar, (nx, ny, nz, fdn, a, sigma, _p(x), _p(b), omega, beta, m, p,
                             max_iter, tol, int(precond)), trace, max_iter + 8)
    return x, d


def matvec_vc(x, shape, a, c, fdn=6):
    """y = a*Lap(x) + c(x) x (variable-coefficient operator)."""
    nx, ny, nz = shape
    x = _f64(x).reshape(-1)
    cv = _f64(c).reshape(-1)
    y = np.empty_like(x)
    lib().orc_matvec_vc(nx, ny, nz, fdn, _p(lap_coefficient(fdn)), a, _p(cv), _p(x), _p(y))
    return y


def aar_vc(x0, rhs, shape, a, c, omega, beta, m, p, max_iter, tol, precond=True, fdn=6, trace=True):
    """AAR.c:37-177 on a*Lap + c(x) with the pointwise Jacobi diagonal."""
    nx, ny, nz = shape
    x = _f64(x0).reshape(-1).copy()
    b = _f64(rhs).reshape(-1)
    cv = _f64(c).reshape(-1)
    d = _run(lib().orc_aar_vc, (nx, ny, nz, fdn, a, _p(cv), _p(x), _p(b), omega, beta, m, p,
                                max_iter, tol, int(precond)), trace, max_iter + 8)
    return x, d


def pgr_vc(x0, rhs, shape, a, c, omega, m, p, max_iter, tol, precond=True, l2=False, fdn=6, trace=True):
    """PGR.c / PL2R.c on a*Lap + c(x) with the pointwise Jacobi diagonal."""
    nx, ny, nz = shape
    x = _f64(x0).reshape(-1).copy()
    b = _f64(rhs).reshape(-1)
    cv = _f64(c).reshape(-1)
    d = _run(lib().orc_pgr_vc, (nx, ny, nz, fdn, a, _p(cv), _p(x), _p(b), omega, m, p, max_iter, tol,
                                int(precond), int(l2)), trace, max_iter + 8)
    return x, d


def cg_vc(x0, rhs, shape, a, c, max_iter, tol, precond=True, fdn=6, trace=True):
    """CG.c on a*Lap + c(x) with the pointwise Jacobi diagonal."""
    nx, ny, nz = shape
    x = _f64(x0).reshape(-1).copy()
    b = _f64(rhs).reshape(-1)
    cv = _f64(c).reshape(-1)
    d = _run(lib().orc_cg_vc, (nx, ny, nz, fdn, a, _p(cv), _p(x), _p(b), max_iter, tol, int(precond)),
             trace, max_iter // 50 + 8)
    return x, d


def pgr(x0, rhs, shape, a, omega, m, p, max_iter, tol, sigma=0.0, precond=True, l2=False,
        fdn=6, trace=True):
    """PGR.c:35-218 (l2=False) / PL2R.c:35-218 (l2=True) on one rank."""
    nx, ny, nz = shape
    x = _f64(x0).reshape(-1).copy()
    b = _f64(rhs).reshape(-1)
    d = _run(lib().orc_pgr, (nx, ny, nz, fdn, a, sigma, _p(x), _p(b), omega, m, p, max_iter,
                             tol, int(precond), int(l2)), trace, max_iter + 8)
    return x, d


def cg(x0, rhs, shape, a, max_iter, tol, sigma=0.0, precond=True, fdn=6, trace=True):
    """CG.c:33-175 on one rank."""
    nx, ny, nz = shape
    x = _f64(x0).reshape(-1).copy()
    b = _f64(rhs).reshape(-1)
    d = _run(lib().orc_cg, (nx, ny, nz, fdn, a, sigma, _p(x), _p(b), max_iter, tol,
                            int(precond)), trace, max_iter // 50 + 8)
    return x, d


ILU_SLOTS = 37
PC_NONE, PC_JACOBI, PC_ILU0 = 0, 1, 2


def ilu0_factor(shape, a, sigma=0.0, c=None, fdn=6):
    """ILU(0) factors of the periodic FD12 matrix, (n, 37) per row (see
    aar_oracle.h for the slot layout)."""
    nx, ny, nz = shape
    n = nx * ny * nz
    F = np.zeros(n * ILU_SLOTS)
    cv = None if c is None else _f64(c).reshape(-1)
    rc = lib().orc_ilu0_factor(nx, ny, nz, fdn, a, sigma, _p(cv), _p(F))
    if rc:
        raise ValueError(f"orc_ilu0_factor failed ({rc})")
    return F.reshape(n, ILU_SLOTS)


def ilu0_apply(shape, F, r, trans=False):
    nx, ny, nz = shape
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1)
    r = _f64(r).reshape(-1)
    out = np.empty_like(r)
    lib().orc_ilu0_apply(nx, ny, nz, _p(F), _p(r), _p(out), int(bool(trans)))
    return out


def ilu0_factor_complex(shape, a, c, c0, eta, fdn=6):
    nx, ny, nz = shape
    n = nx * ny * nz
    F = np.zeros(n * ILU_SLOTS, dtype=np.complex128)
    cv = None if c is None else _f64(c).reshape(-1)
    rc = lib().orc_ilu0_factor_complex(nx, ny, nz, fdn, a, _p(cv), c0, eta, _p(F.view(np.float64)))
    if rc:
        raise ValueError(f"orc_ilu0_factor_complex failed ({rc})")
    return F.reshape(n, ILU_SLOTS)


def ilu0_apply_complex(shape, F, r, trans=0):
    nx, ny, nz = shape
    F = np.ascontiguousarray(F, dtype=np.complex128).reshape(-1)
    r = np.ascontiguousarray(r, dtype=np.complex128).reshape(-1)
    out = np.empty_like(r)
    lib().orc_ilu0_apply_complex(nx, ny, nz, _p(F.view(np.float64)), _p(r.view(np.float64)),
                                 _p(out.view(np.float64)), int(trans))
    return out


# KSPConvergedReason codes of the Krylov comparators
KSP_CONVERGED_RTOL, KSP_CONVERGED_ATOL = 2, 3
KSP_DIVERGED_ITS, KSP_DIVERGED_BREAKDOWN, KSP_DIVERGED_BREAKDOWN_BICG = -3, -5, -6


def gmres(x0, rhs, shape, a, max_iter, tol, restart=30, sigma=0.0, c=None, precond=True, fdn=6,
          trace=True):
    """PETSc KSPGMRES as PETSC_REAL/main.c:86-122 runs it (restatement of
    PETSc's published algorithm); precond: 0/False none, 1/True Jacobi,
    2 ILU(0).  stats['status'] is the KSPConvergedReason; 'trace' holds
    every convergence-test norm."""
    nx, ny, nz = shape
    x = _f64(x0).reshape(-1).copy()
    b = _f64(rhs).reshape(-1)
    cv = None if c is None else _f64(c).reshape(-1)
    cap = max_iter + max_iter // max(1, restart) + 8
    d = _run(lib().orc_gmres, (nx, ny, nz, fdn, a, sigma, _p(cv), _p(x), _p(b), restart, max_iter, tol,
                               int(precond)), trace, cap)
    return x, d


def aar_petsc(x0, rhs, shape, a, omega, beta, m, p, max_iter, tol, sigma=0.0, c=None, precond=2, fdn=6,
              trace=True):
    """AAR of the PETSc variant (PETSC_REAL/AAR_Real.c:32-199): precond 1
    Jacobi, 2 ILU(0) for its pc = 1 (PCICC), 0 none."""
    nx, ny, nz = shape
    x = _f64(x0).reshape(-1).copy()
    b = _f64(rhs).reshape(-1)
    cv = None if c is None else _f64(c).reshape(-1)
    d = _run(lib().orc_aar_petsc, (nx, ny, nz, fdn, a, sigma, _p(cv), _p(x), _p(b), omega, beta, m, p, max_iter,
                                   tol, int(precond)), trace, max_iter // max(1, p) + 8)
    return x, d


def pgr_petsc(x0, rhs, shape, a, omega, m, p, max_iter, tol, sigma=0.0, c=None, precond=2, l2=False, fdn=6,
              trace=True):
    """PGR (l2=False) / PL2R (l2=True) of the PETSc variant (PGR_Real.c,
    PL2R_Real.c); precond as aar_petsc."""
    nx, ny, nz = shape
    x = _f64(x0).reshape(-1).copy()
    b = _f64(rhs).reshape(-1)
    cv = None if c is None else _f64(c).reshape(-1)
    d = _run(lib().orc_pgr_petsc, (nx, ny, nz, fdn, a, sigma, _p(cv), _p(x), _p(b), omega, m, p, max_iter, tol,
                                   int(precond), int(bool(l2))), trace, max_iter // max(1, p) + 8)
    return x, d


def gmres_complex(u0, rhs, shape, a, c, c0, eta, max_iter, tol, restart=30, precond=True, fdn=6, trace=True):
    """GMRES of the complex PETSc variant on the Helmholtz operator
    (complex128 arrays; c may be None)."""
    nx, ny, nz = shape
    u = np.ascontiguousarray(u0, dtype=np.complex128).reshape(-1).copy()
    b = np.ascontiguousarray(rhs, dtype=np.complex128).reshape(-1)
    cc = None if c is None else _f64(c).reshape(-1)
    cap = max_iter + max_iter // max(1, restart) + 8
    d = _run(lib().orc_gmres_complex, (nx, ny, nz, fdn, a, _p(cc), c0, eta, _p(u.view(np.float64)),
                                       _p(b.view(np.float64)), restart, max_iter, tol, int(precond)), trace, cap)
    return u, d


def lgmres_complex(u0, rhs, shape, a, c, c0, eta, max_iter, tol, restart=30, aug=2, precond=True, fdn=6,
                   trace=True):
    """LGMRES of the complex PETSc variant (PETSC_COMPLEX/main.c:120-147):
    restart - aug Arnoldi steps and one step per stored error approximation
    each cycle; aug = 0 is gmres_complex."""
    nx, ny, nz = shape
    u = np.ascontiguousarray(u0, dtype=np.complex128).reshape(-1).copy()
    b = np.ascontiguousarray(rhs, dtype=np.complex128).reshape(-1)
    cc = None if c is None else _f64(c).reshape(-1)
    cap = max_iter + max_iter // max(1, restart - aug) + 8
    d = _run(lib().orc_lgmres_complex, (nx, ny, nz, fdn, a, _p(cc), c0, eta, _p(u.view(np.float64)),
                                        _p(b.view(np.float64)), restart, aug, max_iter, tol, int(precond)), trace, cap)
    return u, d


def bicg_complex(u0, rhs, shape, a, c, c0, eta, max_iter, tol, precond=True, fdn=6, trace=True):
    """BiCG of the complex PETSc variant on the Helmholtz operator."""
    nx, ny, nz = shape
    u = np.ascontiguousarray(u0, dtype=np.complex128).reshape(-1).copy()
    b = np.ascontiguousarray(rhs, dtype=np.complex128).reshape(-1)
    cc = None if c is None else _f64(c).reshape(-1)
    d = _run(lib().orc_bicg_complex, (nx, ny, nz, fdn, a, _p(cc), c0, eta, _p(u.view(np.float64)),
                                      _p(b.view(np.float64)), max_iter, tol, int(precond)), trace, max_iter + 8)
    return u, d


def kspcg(x0, rhs, shape, a, max_iter, tol, sigma=0.0, c=None, precond=True, fdn=6, trace=True):
    """PETSc KSPCG as PETSC_REAL/main.c:152-175 runs it (restatement);
    precond 1 Jacobi, 2 ILU(0) (its pc = 1), 0 none."""
    nx, ny, nz = shape
    x = _f64(x0).reshape(-1).copy()
    b = _f64(rhs).reshape(-1)
    cv = None if c is None else _f64(c).reshape(-1)
    st = OrcStats()
    cap = max_iter + 8
    tr = np.zeros(cap) if trace else None
    rc = lib().orc_kspcg(nx, ny, nz, fdn, a, sigma, _p(cv), _p(x), _p(b), max_iter, tol, int(precond),
                         _p(tr), cap if trace else 0, ctypes.byref(st))
    d = st.as_dict()
    d["rc"] = rc
    if trace:
        d["trace"] = tr[: st.trace_len].copy()
    return x, d


def bicg(x0, rhs, shape, a, max_iter, tol, sigma=0.0, c=None, precond=True, fdn=6, trace=True):
    """PETSc KSPBICG + PCJACOBI as PETSC_REAL/main.c:124-150 runs it
    (restatement).  stats['shadow_gap'] = max |r~ - r| over the iterations."""
    nx, ny, nz = shape
    x = _f64(x0).reshape(-1).copy()
    b = _f64(rhs).reshape(-1)
    cv = None if c is None else _f64(c).reshape(-1)
    gap = np.zeros(1)
    st = OrcStats()
    cap = max_iter + 8
    tr = np.zeros(cap) if trace else None
    rc = lib().orc_bicg(nx, ny, nz, fdn, a, sigma, _p(cv), _p(x), _p(b), max_iter, tol, int(precond),
                        _p(tr), cap if trace else 0, ctypes.byref(st), _p(gap))
    d = st.as_dict()
    d["rc"] = rc
    d["shadow_gap"] = float(gap[0])
    if trace:
        d["trace"] = tr[: st.trace_len].copy()
    return x, d


def aar_complex(u0, rhs, shape, a, c, c0, eta, omega, beta, m, p, max_iter, tol, fdn=6,
                trace=True):
    """Complex Helmholtz AAR (extension; complex128 arrays)."""
    nx, ny, nz = shape
    u = np.ascontiguousarray(u0, dtype=np.complex128).reshape(-1).copy()
    b = np.ascontiguousarray(rhs, dtype=np.complex128).reshape(-1)
    cc = None if c is None else _f64(c).reshape(-1)
    ur = u.view(np.float64)
    br = b.view(np.float64)
    d = _run(lib().orc_aar_complex, (nx, ny, nz, fdn, a, _p(cc), c0, eta, _p(ur), _p(br), omega,
                                     beta, m, p, max_iter, tol), trace, max_iter + 8)
    return u, d


def matvec_complex(u, shape, a, c, c0, eta, fdn=6):
    nx, ny, nz = shape
    u = np.ascontiguousarray(u, dtype=np.complex128).reshape(-1)
    y = np.empty_like(u)
    cc = None if c is None else _f64(c).reshape(-1)
    coeff = lap_coefficient(fdn)
    lib().orc_matvec_complex(nx, ny, nz, fdn, _p(coeff), a, _p(cc), c0, eta,
                             _p(u.view(np.float64)), _p(y.view(np.float64)))
    return y


# ---------------------------------------------------------------------------
# the real reference (oracle/_ref/ref_harness)
# ---------------------------------------------------------------------------

def ref_available():
    return os.path.exists(REF_HARNESS)


def write_problem(path, shape, rhs, x0, a, sigma=0.0, omega=0.6, beta=0.6, m=7, p=6,
                  max_iter=1000, tol=1e-6, precond=True, fdn=6, coeff=None):
    """coeff: stencil weights coeff[0..fdn] replacing the reference's own
    (written as the "AARPRB02" form)."""
    nx, ny, nz = shape
    with open(path, "wb") as fp:
        fp.write(b"AARPRB01" if coeff is None else b"AARPRB02")
        fp.write(struct.pack("<8i", nx, ny, nz, fdn, m, p, max_iter, int(precond)))
        fp.write(struct.pack("<5d", a, sigma, omega, beta, tol))
        fp.write(_f64(rhs).reshape(-1).tobytes())
        fp.write(_f64(x0).reshape(-1).tobytes())
        if coeff is not None:
            c = _f64(coeff).reshape(-1)
            fp.write(struct.pack("<i", c.size))
            fp.write(c.tobytes())


_ITER_RE = re.compile(r"(AAR|PGR|PL2R|CG)(?: converged!)?:\s+Iterations = (\d+), "
                      r"Relative Residual = ([-+0-9.eEnaif]+)")


def ref_solve(solver, shape, rhs, x0, a, sigma=0.0, omega=0.6, beta=0.6, m=7, p=6,
              max_iter=1000, tol=1e-6, precond=True, ranks=1, timeout=600, trace=True, fdn=6, coeff=None):
    """Run the reference solver (one of aar, pgr, pl2r, cg).  Returns
    (x, dict(iterations, relres_printed, trace, stdout, converged_line,
    warning_line, seconds)).  trace=False: no residual recording (an empty
    trace), so `seconds` times the reference solver alone."""
    if not ref_available():
        raise RuntimeError("reference harness not built (oracle/_ref/ref_harness)")
    n = int(np.prod(shape))
    tmp = tempfile.mkdtemp(prefix="aarref_")
    try:
        prob = os.path.join(tmp, "p.bin")
        write_problem(prob, shape, rhs, x0, a, sigma, omega, beta, m, p, max_iter, tol, precond, fdn, coeff)
        xo, to = os.path.join(tmp, "x.bin"), os.path.join(tmp, "t.bin")
        cmd = [REF_HARNESS, "solve", prob, solver, xo, to if trace else "-"]
        if ranks > 1:
            cmd = [MPIRUN, "-n", str(ranks)] + cmd
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, check=True).stdout
        x = np.fromfile(xo, dtype=np.float64)
        assert x.size == n
        tr = np.fromfile(to, dtype=np.float64) if trace else np.zeros(0)
        mm = _ITER_RE.search(out)
        hs = re.search(r"HARNESS .* seconds=([0-9.]+)", out)
        return x, {
            "iterations": int(mm.group(2)) if mm else None,
            "relres_printed": float(mm.group(3)) if mm else None,
            "trace": tr,
            "stdout": out,
            "converged_line": f"{solver.upper()} converged!" in out,
            "warning_line": "exceeded maximum iterations" in out,
            "seconds": float(hs.group(1)) if hs else None,
        }
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def ref_matvec(shape, x, a, sigma=0.0, ranks=1, fdn=6, coeff=None):
    tmp = tempfile.mkdtemp(prefix="aarref_")
    try:
        prob = os.path.join(tmp, "p.bin")
        write_problem(prob, shape, np.zeros(int(np.prod(shape))), x, a, sigma, fdn=fdn, coeff=coeff)
        yo = os.path.join(tmp, "y.bin")
        cmd = [REF_HARNESS, "matvec", prob, yo]
        if ranks > 1:
            cmd = [MPIRUN, "-n", str(ranks)] + cmd
        subprocess.run(cmd, capture_output=True, text=True, timeout=600, check=True)
        return np.fromfile(yo, dtype=np.float64)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def ref_init_rhs(shape, ranks=1):
    """The reference's own RHS (MPI_REAL/system.c:425-471: srand(g+1),
    2*rand()/RAND_MAX, global mean removed)."""
    tmp = tempfile.mkdtemp(prefix="aarref_")
    try:
        o = os.path.join(tmp, "rhs.bin")
        cmd = [REF_HARNESS, "init", str(shape[0]), str(shape[1]), str(shape[2]), o]
        if ranks > 1:
            cmd = [MPIRUN, "-n", str(ranks)] + cmd
        subprocess.run(cmd, capture_output=True, text=True, timeout=600, check=True)
        return np.fromfile(o, dtype=np.float64)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def ref_pinv(A, b):
    A = _f64(A)
    m = A.shape[0]
    tmp = tempfile.mkdtemp(prefix="aarref_")
    try:
        i, o = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(i, "wb") as fp:
            fp.write(struct.pack("<i", m))
            fp.write(A.tobytes())
            fp.write(_f64(b).tobytes())
        subprocess.run([REF_HARNESS, "pinv", i, o], capture_output=True, timeout=60, check=True)
        return np.fromfile(o, dtype=np.float64)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def ref_step(kind, DX, DF, f, beta=0.0):
    """The reference's stand-alone history steps: kind "anderson"
    (AAR.h:24-25) returns (am, allredvec); "galerkin" (PGR.h:24-25) and
    "l2" (PL2R.h:24-25) return (svec, allredvec)."""
    DX, DF, f = _f64(DX), _f64(DF), _f64(f)
    m, n = DF.shape
    tmp = tempfile.mkdtemp(prefix="aarref_")
    try:
        i, o = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(i, "wb") as fp:
            fp.write(struct.pack("<iid", m, n, beta))
            fp.write(DX.tobytes())
            fp.write(DF.tobytes())
            fp.write(f.tobytes())
        subprocess.run([REF_HARNESS, kind, i, o], capture_output=True, timeout=60, check=True)
        out = np.fromfile(o, dtype=np.float64)
        k = n if kind == "anderson" else m
        return out[:k], out[k:]
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def ref_anderson(DX, DF, f, beta):
    DX, DF, f = _f64(DX), _f64(DF), _f64(f)
    m, n = DX.shape
    tmp = tempfile.mkdtemp(prefix="aarref_")
    try:
        i, o = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(i, "wb") as fp:
            fp.write(struct.pack("<iid", m, n, beta))
            fp.write(DX.tobytes())
            fp.write(DF.tobytes())
            fp.write(f.tobytes())
        subprocess.run([REF_HARNESS, "anderson", i, o], capture_output=True, timeout=60, check=True)
        out = np.fromfile(o, dtype=np.float64)
        return out[:n], out[n:]
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
