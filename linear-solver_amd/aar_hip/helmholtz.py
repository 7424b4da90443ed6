"""Complex Helmholtz solves and their adjoint, over libaar_hip's C ABI.

    A(c) u = a*Lap_FD12(u) + (c(x) + i*eta) u = b,   periodic,

with complex128 torch tensors on the GPU (x fastest, as every vector of the
library) and a real coefficient field c (float64, N points) or a constant
c0.  ``HelmholtzSolve`` is a ``torch.autograd.Function``: its forward is one
AAR solve (aar_hip_helm_aar), its backward one adjoint AAR solve with the
same kernels on A^H (eta -> -eta) followed by the gradient kernel:

    A^H lambda = dL/du,   dL/db = lambda,   dL/dc = -Re(conj(lambda) u).

Gradients follow torch's convention for complex tensors (dL/dRe + i dL/dIm),
in which dL/du of L = 1/2 ||u - u*||^2 is u - u*.  Only u is kept for the
backward pass.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import Stats, check


class Helm(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("fdn", ctypes.c_int),
                ("a", ctypes.c_double), ("c0", ctypes.c_double), ("eta", ctypes.c_double),
                ("c", ctypes.c_void_p)]


def _L():
    return _lib.load()


def _bind_stream():
    import torch
    _L().aar_hip_set_stream(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


def _cplx(t, name):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.complex128:
        raise TypeError(f"{name} must be a complex128 CUDA/HIP tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _helm(shape, a, c, c0, eta):
    import torch
    nx, ny, nz = (int(s) for s in shape)
    cp = None
    if c is not None:
        if not isinstance(c, torch.Tensor) or not c.is_cuda or c.dtype != torch.float64 or not c.is_contiguous():
            raise TypeError("c must be a contiguous float64 CUDA/HIP tensor")
        if c.numel() != nx * ny * nz:
            raise ValueError("c must have one value per grid point")
        cp = c.data_ptr()
    return Helm(nx, ny, nz, 6, float(a), float(c0), float(eta), cp)


def helm_matvec(u, shape, a, c=None, c0=0.0, eta=0.0, out=None):
    """y = A(c) u."""
    import torch
    y = torch.empty_like(u) if out is None else out
    h = _helm(shape, a, c, c0, eta)
    _bind_stream()
    check(_L().aar_hip_helm_matvec(ctypes.byref(h), _cplx(u, "u"), _cplx(y, "y")))
    return y


def helm_aar(u, b, shape, a, c=None, c0=0.0, eta=0.0, omega=0.6, beta=0.6, m=7, p=6, max_iter=10000,
             tol=1e-10, adjoint=False):
    """AAR on A u = b (or A^H u = b): u is the initial guess, overwritten."""
    h = _helm(shape, a, c, c0, eta)
    st = Stats()
    _bind_stream()
    check(_L().aar_hip_helm_aar(ctypes.byref(h), _cplx(u, "u"), _cplx(b, "b"), float(omega), float(beta),
                                int(m), int(p), int(max_iter), float(tol), int(bool(adjoint)),
                                ctypes.byref(st)))
    return st.as_dict()


def _pcz(precond):
    """precond: False/0 none, True/1 Jacobi, 2 ILU(0) (the complex variant's pc = 1)."""
    return 2 if precond == 2 and not isinstance(precond, bool) else int(bool(precond))


def helm_ilu0_apply(r, shape, a, c=None, c0=0.0, eta=0.0, trans=0):
    """z = M^{-1} r (trans 0), M^{-T} r (1) or M^{-H} r (2) for the complex
    ILU(0) factors of the Helmholtz matrix."""
    import torch
    h = _helm(shape, a, c, c0, eta)
    _bind_stream()
    z = torch.empty_like(r)
    check(_L().aar_hip_helm_ilu0_apply(ctypes.byref(h), _cplx(r, "r"), _cplx(z, "z"), int(trans)))
    return z


def helm_gmres(u, b, shape, a, c=None, c0=0.0, eta=0.0, max_iter=2000, tol=1e-6, restart=30, precond=True):
    """GMRES of the complex PETSc variant (PETSC_COMPLEX/main.c:80-118) on
    A u = b: u is the initial guess, overwritten; 'reason' in the stats is
    PETSc's KSPConvergedReason."""
    h = _helm(shape, a, c, c0, eta)
    st = Stats()
    _bind_stream()
    check(_L().aar_hip_helm_gmres(ctypes.byref(h), _cplx(u, "u"), _cplx(b, "b"), int(restart), int(max_iter),
                                  float(tol), _pcz(precond), ctypes.byref(st)))
    return st.as_dict()


def helm_lgmres(u, b, shape, a, c=None, c0=0.0, eta=0.0, max_iter=2000, tol=1e-6, restart=30, aug=2,
                precond=True):
    """LGMRES of the complex PETSc variant (PETSC_COMPLEX/main.c:120-147,
    KSPLGMRES with PETSc's defaults restart = 30, aug = 2) on A u = b."""
    h = _helm(shape, a, c, c0, eta)
    st = Stats()
    _bind_stream()
    check(_L().aar_hip_helm_lgmres(ctypes.byref(h), _cplx(u, "u"), _cplx(b, "b"), int(restart), int(aug),
                                   int(max_iter), float(tol), _pcz(precond), ctypes.byref(st)))
    return st.as_dict()


def helm_bicg(u, b, shape, a, c=None, c0=0.0, eta=0.0, max_iter=2000, tol=1e-6, precond=True):
    """BiCG of the complex PETSc variant (PETSC_COMPLEX/main.c:150-175)."""
    h = _helm(shape, a, c, c0, eta)
    st = Stats()
    _bind_stream()
    check(_L().aar_hip_helm_bicg(ctypes.byref(h), _cplx(u, "u"), _cplx(b, "b"), int(max_iter), float(tol),
                                 _pcz(precond), ctypes.byref(st)))
    return st.as_dict()


def helm_grad_c(lam, u):
    """-Re(conj(lambda) u), float64."""
    import torch
    dc = torch.empty(u.numel(), dtype=torch.float64, device=u.device)
    _bind_stream()
    check(_L().aar_hip_helm_grad_c(int(u.numel()), _cplx(lam, "lambda"), _cplx(u, "u"),
                                   ctypes.c_void_p(dc.data_ptr())))
    return dc


class HelmholtzSolve:
    """u = A(c)^{-1} b as a differentiable torch op (see the module docstring).

    ``HelmholtzSolve.apply(b, c, shape, a, eta, opts)``: b complex128 (N),
    c float64 (N), opts a dict of AAR parameters (omega, beta, m, p,
    max_iter, tol).  The last solve statistics are in ``HelmholtzSolve.last``.
    """

    last = {}

    @staticmethod
    def _fn():
        import torch

        class _F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, b, c, shape, a, eta, opts):
                u = torch.zeros_like(b)
                st = helm_aar(u, b.contiguous(), shape, a, c.contiguous(), 0.0, eta, adjoint=False, **opts)
                if not st["converged"]:
                    raise RuntimeError(f"Helmholtz solve did not converge: {st}")
                HelmholtzSolve.last = {"forward": st}
                ctx.save_for_backward(u, c)
                ctx.shape, ctx.a, ctx.eta, ctx.opts = shape, a, eta, opts
                return u

            @staticmethod
            def backward(ctx, gu):
                u, c = ctx.saved_tensors
                lam = torch.zeros_like(u)
                st = helm_aar(lam, gu.contiguous(), ctx.shape, ctx.a, c, 0.0, ctx.eta, adjoint=True, **ctx.opts)
                if not st["converged"]:
                    raise RuntimeError(f"adjoint solve did not converge: {st}")
                HelmholtzSolve.last["adjoint"] = st
                gb = lam if ctx.needs_input_grad[0] else None
                gc = helm_grad_c(lam, u) if ctx.needs_input_grad[1] else None
                return gb, gc, None, None, None, None

        return _F

    _cached = None

    @classmethod
    def apply(cls, b, c, shape, a, eta, opts=None):
        if cls._cached is None:
            cls._cached = cls._fn()
        return cls._cached.apply(b, c, tuple(int(s) for s in shape), float(a), float(eta), dict(opts or {}))


def coefficient_field(shape, seed, modes=8):
    """c(x) = 1 + 0.5 * (sum of `modes` seeded Fourier modes, max |.| = 1)."""
    from .problems import helmholtz_coefficient
    return np.ascontiguousarray(helmholtz_coefficient(shape, seed, modes))
