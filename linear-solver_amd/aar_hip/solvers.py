"""Python mirror of the reference solver interface over libaar_hip.

Names and argument meaning follow MPI_REAL (AAR.h:18-22, PGR.h:18-22,
PL2R.h:18-22, CG.h:18-21, system.h:82): ``x`` is the initial guess and is
overwritten with the solution, ``rhs``/``b`` is the right-hand side, ``a``
scales the Laplacian (y = a*Lap(x) + sigma*x), Jacobi preconditioning with
the constant diagonal a*(3*coeff[0]) + sigma unless ``precond=False``
(the reference's Precondition == NULL).  Vectors are torch tensors on the
GPU in the reference layout (x fastest; a C-ordered (nz, ny, nx) tensor or
its flat view).  The grid is given as ``shape = (nx, ny, nz)``.

Each solver returns the statistics the reference prints (iterations,
relative residual, converged / warning conditions) plus GPU time and
algorithmic HBM bytes.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import Grid, Instance, Stats, check


def _L():
    return _lib.load()


def _bind_stream():
    import torch
    _L().aar_hip_set_stream(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


def _dev(t, dtype, name):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA/HIP tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _grid(shape, a, sigma=0.0, fdn=6):
    nx, ny, nz = (int(s) for s in shape)
    return Grid(nx, ny, nz, fdn, float(a), float(sigma))


def _check_len(t, shape, name):
    n = int(np.prod(shape))
    if t.numel() != n:
        raise ValueError(f"{name} has {t.numel()} elements, grid {tuple(shape)} needs {n}")


def Lap_Vec_mult(x, shape, a, sigma=0.0, out=None, fdn=6):
    """system.c:80-213 -- y = a*Lap(x) (+ sigma*x), FP64 or FP32."""
    import torch
    _check_len(x, shape, "x")
    y = torch.empty_like(x) if out is None else out
    g = _grid(shape, a, sigma, fdn)
    _bind_stream()
    if x.dtype == torch.float64:
        check(_L().aar_hip_matvec(ctypes.byref(g), _dev(x, torch.float64, "x"),
                                  _dev(y, torch.float64, "y")))
    elif x.dtype == torch.float32:
        check(_L().aar_hip_matvec_f32(ctypes.byref(g), _dev(x, torch.float32, "x"),
                                      _dev(y, torch.float32, "y")))
    else:
        raise TypeError("x must be float64 or float32")
    return y


def AAR(x, rhs, shape, a, omega, beta, m, p, max_iter, tol, sigma=0.0, precond=True, fdn=6):
    """AAR.c:37-177 (FP64 tensors), or the same iteration in FP32 (FP32 tensors;
    normal equations and the m x m solve stay FP64)."""
    import torch
    _check_len(x, shape, "x")
    _check_len(rhs, shape, "rhs")
    g = _grid(shape, a, sigma, fdn)
    st = Stats()
    _bind_stream()
    if x.dtype == torch.float64:
        check(_L().aar_hip_aar(ctypes.byref(g), _dev(x, torch.float64, "x"),
                               _dev(rhs, torch.float64, "rhs"), omega, beta, m, p, max_iter, tol,
                               int(bool(precond)), ctypes.byref(st)))
    else:
        check(_L().aar_hip_aar_f32(ctypes.byref(g), _dev(x, torch.float32, "x"),
                                   _dev(rhs, torch.float32, "rhs"), omega, beta, m, p, max_iter,
                                   tol, int(bool(precond)), ctypes.byref(st)))
    return st.as_dict()


def AAR_mixed(x, rhs, shape, a, omega, beta, m, p, max_iter, tol, sigma=0.0, precond=True, fdn=6):
    """AAR.c:37-177 in mixed precision (no reference counterpart): FP64 x, b,
    true residual and normal equations; FP32 Richardson sweeps and history."""
    import torch
    _check_len(x, shape, "x")
    _check_len(rhs, shape, "rhs")
    g = _grid(shape, a, sigma, fdn)
    st = Stats()
    _bind_stream()
    check(_L().aar_hip_aar_mixed(ctypes.byref(g), _dev(x, torch.float64, "x"),
                                 _dev(rhs, torch.float64, "rhs"), omega, beta, m, p, max_iter, tol,
                                 int(bool(precond)), ctypes.byref(st)))
    return st.as_dict()


def Lap_Vec_mult_vc(x, shape, a, c, out=None):
    """y = a*Lap(x) + c(x) x (no reference counterpart: the device operator
    behind libaar_hip's Lap_Vec_mult_vc slot).  c: float64 CUDA/HIP tensor,
    one value per grid point."""
    import torch
    _check_len(x, shape, "x")
    _check_len(c, shape, "c")
    y = torch.empty_like(x) if out is None else out
    g = _grid(shape, a, 0.0)
    _bind_stream()
    check(_L().aar_hip_matvec_vc(ctypes.byref(g), _dev(c, torch.float64, "c"), _dev(x, torch.float64, "x"),
                                 _dev(y, torch.float64, "y")))
    return y


def AAR_vc(x, rhs, shape, a, c, omega, beta, m, p, max_iter, tol, precond=True, mixed=False):
    """AAR.c:37-177 on A = a*Lap + c(x) with the pointwise Jacobi diagonal
    a*(3*coeff[0]) + c(x) (no reference counterpart); mixed: FP32 sweeps as
    AAR_mixed."""
    import torch
    _check_len(x, shape, "x")
    _check_len(rhs, shape, "rhs")
    _check_len(c, shape, "c")
    g = _grid(shape, a, 0.0)
    st = Stats()
    _bind_stream()
    check(_L().aar_hip_aar_vc(ctypes.byref(g), _dev(c, torch.float64, "c"), _dev(x, torch.float64, "x"),
                              _dev(rhs, torch.float64, "rhs"), omega, beta, m, p, max_iter, tol,
                              int(bool(precond)), int(bool(mixed)), ctypes.byref(st)))
    return st.as_dict()


FDN_USER = -1   # fdn value selecting the weights registered with set_stencil


def set_stencil(coeff):
    """Registers user 1D stencil weights coeff[0..fdn] (fdn = len - 1 <= 6)
    for calls made with fdn=FDN_USER: y = a*(3 c0 x + sum_o c_o S_o x);
    None clears them."""
    if coeff is None:
        check(_L().aar_hip_set_stencil(None, 0))
        return
    c = np.ascontiguousarray(coeff, dtype=np.float64)
    check(_L().aar_hip_set_stencil(c.ctypes.data_as(ctypes.c_void_p), int(c.size - 1)))


def set_coefficient(c):
    """Registers c(x) (a host numpy array, or None to clear) for the drop-in
    Lap_Vec_mult_vc / Precondition_vc slots."""
    import numpy as np
    if c is None:
        check(_L().aar_hip_set_coefficient(None, 0))
        return
    arr = np.ascontiguousarray(c, dtype=np.float64)
    check(_L().aar_hip_set_coefficient(arr.ctypes.data_as(ctypes.c_void_p), arr.size))


def AAR_batch(instances, omega, beta, m, p, max_iter, tol, precond=True, mixed=False):
    """Independent AAR solves run together (no reference counterpart).

    instances: sequence of (x, rhs, shape, a, sigma) with float64 CUDA
    tensors; every x is overwritten with its solution.  Returns one stats
    dict per instance (AAR.c:37-177 semantics each)."""
    import torch
    n = len(instances)
    arr = (Instance * n)()
    for i, (x, rhs, shape, a, sigma) in enumerate(instances):
        _check_len(x, shape, f"x[{i}]")
        _check_len(rhs, shape, f"rhs[{i}]")
        nx, ny, nz = (int(s) for s in shape)
        arr[i] = Instance(nx, ny, nz, float(a), float(sigma), _dev(x, torch.float64, f"x[{i}]").value,
                          _dev(rhs, torch.float64, f"rhs[{i}]").value)
    sts = (Stats * n)()
    _bind_stream()
    check(_L().aar_hip_aar_batch(arr, n, omega, beta, m, p, max_iter, tol, int(bool(precond)),
                                 int(bool(mixed)), sts))
    return [s.as_dict() for s in sts]


def PGR(x, rhs, shape, a, omega, m, p, max_iter, tol, sigma=0.0, precond=True, _l2=False, c=None, fdn=6):
    """PGR.c:35-173.  c: a float64 coefficient field (one value per point)
    for the variable-coefficient operator a*Lap + c(x) (then sigma must be 0)."""
    import torch
    _check_len(x, shape, "x")
    _check_len(rhs, shape, "rhs")
    g = _grid(shape, a, sigma, fdn)
    st = Stats()
    _bind_stream()
    if c is None:
        check(_L().aar_hip_pgr(ctypes.byref(g), _dev(x, torch.float64, "x"),
                               _dev(rhs, torch.float64, "rhs"), omega, m, p, max_iter, tol,
                               int(bool(precond)), int(bool(_l2)), ctypes.byref(st)))
    else:
        _check_len(c, shape, "c")
        check(_L().aar_hip_pgr_vc(ctypes.byref(g), _dev(c, torch.float64, "c"), _dev(x, torch.float64, "x"),
                                  _dev(rhs, torch.float64, "rhs"), omega, m, p, max_iter, tol,
                                  int(bool(precond)), int(bool(_l2)), ctypes.byref(st)))
    return st.as_dict()


def PL2R(x, rhs, shape, a, omega, m, p, max_iter, tol, sigma=0.0, precond=True, c=None, fdn=6):
    """PL2R.c:35-172 (c as in PGR)."""
    return PGR(x, rhs, shape, a, omega, m, p, max_iter, tol, sigma, precond, _l2=True, c=c, fdn=fdn)


def CG(x, b, shape, a, max_iter, tol, sigma=0.0, precond=True, c=None, fdn=6):
    """CG.c:33-175 (c as in PGR)."""
    import torch
    _check_len(x, shape, "x")
    _check_len(b, shape, "b")
    g = _grid(shape, a, sigma, fdn)
    st = Stats()
    _bind_stream()
    if c is None:
        check(_L().aar_hip_cg(ctypes.byref(g), _dev(x, torch.float64, "x"), _dev(b, torch.float64, "b"),
                              max_iter, tol, int(bool(precond)), ctypes.byref(st)))
    else:
        _check_len(c, shape, "c")
        check(_L().aar_hip_cg_vc(ctypes.byref(g), _dev(c, torch.float64, "c"), _dev(x, torch.float64, "x"),
                                 _dev(b, torch.float64, "b"), max_iter, tol, int(bool(precond)),
                                 ctypes.byref(st)))
    return st.as_dict()


ILU0 = 2     # precond value of GMRES/BiCG for the ILU(0) preconditioner


def _pc(precond):
    return ILU0 if precond == ILU0 and not isinstance(precond, bool) else int(bool(precond))


def GMRES(x, b, shape, a, max_iter=2000, tol=1e-6, restart=30, sigma=0.0, precond=True, c=None):
    """PETSc's KSPGMRES as PETSC_REAL/main.c:86-122 runs it (restart 30,
    classical Gram-Schmidt, left preconditioning, converged on the
    preconditioned residual norm); precond: False/0 none, True/1 Jacobi,
    2 ILU(0) (the PETSc variant's pc = 1).  Returns the stats; 'reason' is the
    KSPConvergedReason, 'relres' the true ||b - A x||/||b||."""
    import torch
    _check_len(x, shape, "x")
    _check_len(b, shape, "b")
    g = _grid(shape, a, sigma)
    st = Stats()
    _bind_stream()
    cp = None
    if c is not None:
        _check_len(c, shape, "c")
        cp = _dev(c, torch.float64, "c")
    check(_L().aar_hip_gmres(ctypes.byref(g), cp, _dev(x, torch.float64, "x"), _dev(b, torch.float64, "b"),
                             restart, max_iter, tol, _pc(precond), ctypes.byref(st)))
    return st.as_dict()


def AAR_petsc(x, b, shape, a, omega, beta, m, p, max_iter, tol, sigma=0.0, precond=ILU0, c=None):
    """The PETSc variant's AAR (PETSC_REAL/AAR_Real.c:32-199): precond
    ILU0 (its pc = 1, PCICC, as ILU(0)), True/1 Jacobi, False/0 none."""
    import torch
    _check_len(x, shape, "x")
    _check_len(b, shape, "b")
    g = _grid(shape, a, sigma)
    st = Stats()
    _bind_stream()
    cp = None
    if c is not None:
        _check_len(c, shape, "c")
        cp = _dev(c, torch.float64, "c")
    check(_L().aar_hip_aar_petsc(ctypes.byref(g), cp, _dev(x, torch.float64, "x"), _dev(b, torch.float64, "b"),
                                 omega, beta, m, p, max_iter, tol, _pc(precond), ctypes.byref(st)))
    return st.as_dict()


def PGR_petsc(x, b, shape, a, omega, m, p, max_iter, tol, sigma=0.0, precond=ILU0, c=None, l2=False):
    """The PETSc variant's PGR (l2=False, PGR_Real.c) / PL2R (l2=True,
    PL2R_Real.c); precond as AAR_petsc."""
    import torch
    _check_len(x, shape, "x")
    _check_len(b, shape, "b")
    g = _grid(shape, a, sigma)
    st = Stats()
    _bind_stream()
    cp = None
    if c is not None:
        _check_len(c, shape, "c")
        cp = _dev(c, torch.float64, "c")
    check(_L().aar_hip_pgr_petsc(ctypes.byref(g), cp, _dev(x, torch.float64, "x"), _dev(b, torch.float64, "b"),
                                 omega, m, p, max_iter, tol, _pc(precond), int(bool(l2)), ctypes.byref(st)))
    return st.as_dict()


def ilu0_apply(r, shape, a, sigma=0.0, c=None, trans=False):
    """z = M^{-1} r (trans: M^{-T} r) for the ILU(0) factors of a*Lap +
    sigma (+ c): the PETSc variant's pc = 1 preconditioner on one rank."""
    import torch
    _check_len(r, shape, "r")
    g = _grid(shape, a, sigma)
    _bind_stream()
    cp = None
    if c is not None:
        _check_len(c, shape, "c")
        cp = _dev(c, torch.float64, "c")
    z = torch.empty_like(r)
    check(_L().aar_hip_ilu0_apply(ctypes.byref(g), cp, _dev(r, torch.float64, "r"), _dev(z, torch.float64, "z"),
                                  int(bool(trans))))
    return z


def BiCG(x, b, shape, a, max_iter=2000, tol=1e-6, sigma=0.0, precond=True, c=None):
    """PETSc's KSPBICG as PETSC_REAL/main.c:124-150 runs it; precond as
    in GMRES (2: ILU(0), the shadow sequence then carried explicitly)."""
    import torch
    _check_len(x, shape, "x")
    _check_len(b, shape, "b")
    g = _grid(shape, a, sigma)
    st = Stats()
    _bind_stream()
    cp = None
    if c is not None:
        _check_len(c, shape, "c")
        cp = _dev(c, torch.float64, "c")
    check(_L().aar_hip_bicg(ctypes.byref(g), cp, _dev(x, torch.float64, "x"), _dev(b, torch.float64, "b"),
                            max_iter, tol, _pc(precond), ctypes.byref(st)))
    return st.as_dict()


def KSPCG(x, b, shape, a, max_iter=2000, tol=1e-6, sigma=0.0, precond=True, c=None, fdn=6):
    """PETSc's KSPCG as PETSC_REAL/main.c:152-175 runs it; precond as in
    GMRES (2: ILU(0), its pc = 1)."""
    import torch
    _check_len(x, shape, "x")
    _check_len(b, shape, "b")
    g = _grid(shape, a, sigma, fdn)
    st = Stats()
    _bind_stream()
    cp = None
    if c is not None:
        _check_len(c, shape, "c")
        cp = _dev(c, torch.float64, "c")
    check(_L().aar_hip_kspcg(ctypes.byref(g), cp, _dev(x, torch.float64, "x"), _dev(b, torch.float64, "b"),
                             max_iter, tol, _pc(precond), ctypes.byref(st)))
    return st.as_dict()


def AAR_refined(x, rhs, shape, a, omega, beta, m, p, tol, inner_tol=1e-5, max_outer=30,
                max_inner=100000, sigma=0.0, precond=True):
    """FP64 residual refinement around FP32 AAR (no reference counterpart)."""
    import torch
    _check_len(x, shape, "x")
    _check_len(rhs, shape, "rhs")
    g = _grid(shape, a, sigma)
    st = Stats()
    _bind_stream()
    check(_L().aar_hip_aar_refined(ctypes.byref(g), _dev(x, torch.float64, "x"),
                                   _dev(rhs, torch.float64, "rhs"), omega, beta, m, p, max_outer,
                                   max_inner, tol, inner_tol, int(bool(precond)), ctypes.byref(st)))
    return st.as_dict()


def pinv(A, b):
    """tools.c:52-105 on the GPU: x = V diag(1/w) V^T b, returns (x, w, V)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    m = A.shape[0]
    x = np.zeros(m)
    w = np.zeros(m)
    V = np.zeros((m, m))
    dp = lambda arr: arr.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    check(_L().aar_hip_pinv(dp(A), dp(b), dp(x), m, dp(w), dp(V)))
    return x, w, V


def last_system(n):
    """The last normal-equation system of the last solve (aar_hip_last_system):
    (A, rhs, w) with A n x n."""
    G = np.zeros(n * n + n)
    w = np.zeros(n)
    _L().aar_hip_last_system(int(n), G.ctypes.data_as(ctypes.c_void_p), w.ctypes.data_as(ctypes.c_void_p))
    A = G[: n * n].reshape(n, n).T.copy()      # A[i][j] = G[j*n + i]
    return A, G[n * n:].copy(), w


def history(cap=1 << 20):
    """Residual history ||b - A x_k|| of the last solve (see aar_hip.h)."""
    buf = np.zeros(cap)
    n = _L().aar_hip_history(buf.ctypes.data_as(ctypes.c_void_p), cap)
    return buf[:n].copy()


def set_tuning(lz=0, wy=8, py=4):
    """Stencil launch shape for every grid; (0, 8, 4) returns to the
    automatic (or tuned) shapes (aar_hip_set_tuning)."""
    check(_L().aar_hip_set_tuning(int(lz), int(wy), int(py)))


def autotune(shape, fp32=False, batch=False):
    """Tune the stencil launch shape for a grid (aar_hip_autotune); later
    solves on that grid use it.  Returns {lz, wy, py, us}."""
    nx, ny, nz = (int(s) for s in shape)
    lz, wy, py, us = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    _bind_stream()
    check(_L().aar_hip_autotune(nx, ny, nz, int(bool(fp32)), int(bool(batch)), ctypes.byref(lz),
                                ctypes.byref(wy), ctypes.byref(py), ctypes.byref(us)))
    return {"lz": lz.value, "wy": wy.value, "py": py.value, "us": us.value}


def set_option(name, value):
    """Tuning option of the history sweeps (see aar_hip_set_option)."""
    check(_L().aar_hip_set_option(name.encode(), int(value)))


def profile(enable=True):
    _L().aar_hip_profile(int(bool(enable)))


def profile_reset():
    _L().aar_hip_profile_reset()


def profile_query(name):
    c = ctypes.c_int()
    ms = ctypes.c_double()
    by = ctypes.c_double()
    _L().aar_hip_profile_query(name.encode(), ctypes.byref(c), ctypes.byref(ms), ctypes.byref(by))
    return {"count": c.value, "ms": ms.value, "bytes": by.value}


def stream_bandwidth(kind, nbytes=4 << 30, reps=10, dst=None, src=None):
    """Attainable HBM bandwidth (GB/s): the best of three launch shapes of a
    streaming kernel with 16-byte accesses, kind "read", "write" or "copy"
    over nbytes (buffers allocated here unless given as device tensors)."""
    import torch
    k = {"read": 0, "write": 1, "copy": 2}[kind]
    nbytes = int(nbytes) // 16 * 16
    if src is None and k != 1:
        src = torch.zeros(nbytes // 8, dtype=torch.float64, device="cuda")
    if dst is None and k != 0:
        dst = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    for t in (src, dst):
        if t is not None and (not t.is_cuda or t.numel() * t.element_size() < nbytes):
            raise ValueError("stream buffers must be device tensors of at least nbytes")
    _bind_stream()
    g = ctypes.c_double()
    check(_L().aar_hip_stream_bandwidth(k, ctypes.c_void_p(dst.data_ptr() if dst is not None else 0),
                                        ctypes.c_void_p(src.data_ptr() if src is not None else 0),
                                        nbytes, int(reps), ctypes.byref(g)))
    return g.value


def release():
    _L().aar_hip_release()
