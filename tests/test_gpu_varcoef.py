"""GPU variable-coefficient operator A = a*Lap_FD12 + c(x) against the oracle.

The matvec is bit-exact.  AAR with the pointwise Jacobi diagonal
a*(3*coeff[0]) + c(x) (or none) gives the oracle's iteration counts
(orc_aar_vc: the reference's AAR.c control flow on this operator; it equals
the reference-pinned screened AAR bit for bit when c is constant and the
dense solve when not), solutions within 1e-8 (1e-6 for an iterate stopped
at max_iter) and residual histories within 1e-6, in FP64 and in mixed
precision.  The drop-in slots Lap_Vec_mult_vc / Precondition_vc run through
the reference's own AAR() signature with a registered c(x).
"""
import ctypes

import numpy as np
import pytest

from aar_hip import problems as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def dev(v):
    import torch
    return torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to("cuda")


@pytest.mark.parametrize("shape", [(24, 24, 24), (20, 12, 16), (70, 33, 17), (64, 64, 40)])
def test_vc_matvec_bitwise(gpu, oracle, shape):
    say("vc matvec", shape)
    n = int(np.prod(shape))
    x = np.random.default_rng(12).standard_normal(n)
    c = P.helmholtz_coefficient(shape, 7)
    y = gpu.Lap_Vec_mult_vc(dev(x), shape, -4.0, dev(c)).cpu().numpy()
    assert np.array_equal(y, oracle.matvec_vc(x, shape, -4.0, c))


CASES = [
    # shape, h, m, p, max_iter, tol, precond, omega
    ((16, 16, 16), 0.5, 7, 6, 1000, 1e-8, True, 0.6),
    ((24, 24, 24), 0.5, 7, 6, 1000, 1e-10, True, 0.6),
    ((20, 12, 16), 0.5, 3, 5, 1000, 1e-8, True, 0.6),
    ((16, 16, 16), 0.5, 4, 1, 1000, 1e-8, True, 0.6),
    ((16, 16, 16), 0.5, 7, 6, 13, 1e-8, True, 0.6),
    ((16, 16, 16), 0.5, 7, 6, 600, 1e-6, False, 0.015),
    ((32, 32, 32), 0.25, 7, 6, 2000, 1e-8, True, 0.6),
]


def _cid(c):
    return "x".join(map(str, c[0])) + f"_m{c[2]}p{c[3]}i{c[4]}" + ("" if c[6] else "_noprec")


@pytest.mark.parametrize("case", CASES, ids=_cid)
@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
def test_vc_aar_parity(gpu, oracle, case, mixed):
    shape, h, m, p, max_iter, tol, precond, om = case
    say("vc aar", shape, m, p, max_iter, precond, "mixed" if mixed else "fp64")
    n = int(np.prod(shape))
    a = P.poisson_a(h)
    c = P.helmholtz_coefficient(shape, 1007)
    b = P.rhs_normal(shape, 1007)
    xr, dr = oracle.aar_vc(np.zeros(n), b, shape, a, c, om, om, m, p, max_iter, tol, precond=precond)
    xd = dev(np.zeros(n))
    st = gpu.AAR_vc(xd, dev(b), shape, a, dev(c), om, om, m, p, max_iter, tol, precond=precond, mixed=mixed)
    assert st["iterations"] == dr["iterations"]
    assert bool(st["printed_converged"]) == bool(dr["printed_converged"])
    x = xd.cpu().numpy()
    tol_x = 1e-8 if dr["printed_converged"] else 1e-6
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= tol_x
    if not mixed:
        h_gpu = gpu.history()
        tr = dr["trace"]
        assert len(h_gpu) == len(tr)
        assert np.all(np.abs(h_gpu - tr) <= 1e-6 * tr + 1e-10 * tr[0])


def test_vc_constant_field_matches_screened_solve(gpu):
    """c = sigma everywhere: the same iteration count and solution (to the
    rounding of the two recurrence forms) as the screened-operator solver."""
    import torch
    shape = (32, 32, 32)
    n = int(np.prod(shape))
    a, sigma = P.poisson_a(0.5), 0.5
    b = dev(P.rhs_normal(shape, 1008))
    x1 = torch.zeros(n, dtype=torch.float64, device="cuda")
    s1 = gpu.AAR(x1, b, shape, a, 0.6, 0.6, 7, 6, 2000, 1e-10, sigma=sigma)
    x2 = torch.zeros_like(x1)
    s2 = gpu.AAR_vc(x2, b, shape, a, torch.full_like(x1, sigma), 0.6, 0.6, 7, 6, 2000, 1e-10)
    assert s1["iterations"] == s2["iterations"] and s1["converged"]
    assert float(torch.linalg.norm(x1 - x2) / torch.linalg.norm(x1)) < 1e-12


class POISSON(ctypes.Structure):
    """MPI_REAL/system.h:26-63 under MPICH (MPI_Comm is an int)."""
    I, P_, D = ctypes.c_int, ctypes.c_void_p, ctypes.c_double
    _fields_ = [("FDn", I), ("ssize", I * 3), ("psize", I * 3), ("np", I * 3), ("coords", I * 3),
                ("rem", I * 3), ("send_neighs", P_), ("rec_neighs", P_), ("send_counts", P_),
                ("rec_counts", P_), ("send_layers", I * 6), ("rec_layers", I * 6), ("scounts", P_),
                ("sdispls", P_), ("rcounts", P_), ("rdispls", P_), ("n_in", I), ("n_out", I), ("phi", P_),
                ("rhs", P_), ("Lap_phi", P_), ("solver_tol", D), ("solver_maxiter", I), ("m", I), ("p", I),
                ("omega", D), ("beta", D), ("coeff_lap", P_), ("comm_laplacian", I), ("cart", I)]


def _dropin(L):
    I, D, V = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    L.AAR.argtypes = [ctypes.POINTER(POISSON), V, V, D, V, V, D, D, I, I, I, D, I, I]
    L.AAR.restype = None
    L.PGR.argtypes = [ctypes.POINTER(POISSON), V, V, D, V, V, D, I, I, I, D, I, I]
    L.PGR.restype = None
    L.Lap_Vec_mult_vc.argtypes = [ctypes.POINTER(POISSON), D, V, V, I]
    L.Lap_Vec_mult_vc.restype = None
    L.Precondition_vc.argtypes = [D, V, I]
    L.Precondition_vc.restype = None
    return {k: ctypes.cast(getattr(L, k), V).value for k in
            ("Lap_Vec_mult", "Precondition", "Lap_Vec_mult_vc", "Precondition_vc")}


def _system(shape, coeff):
    s = POISSON()
    s.FDn = 6
    for d in range(3):
        s.ssize[d] = s.psize[d] = shape[d]
        s.np[d] = 1
    s.coeff_lap = coeff.ctypes.data_as(ctypes.c_void_p)
    return s


def test_vc_dropin_slots(gpu, oracle):
    """The reference's AAR() entry with Lap_Vec_mult_vc / Precondition_vc and
    a registered c(x): the oracle's iteration count and solution; the host
    slots themselves are bit-exact; mismatched or unsupported uses are
    refused without touching x."""
    from aar_hip import _lib
    L = _lib.load()
    fp = _dropin(L)
    shape = (24, 20, 16)
    n = int(np.prod(shape))
    a = P.poisson_a(0.5)
    coeff = oracle.lap_coefficient(6)
    c = P.helmholtz_coefficient(shape, 1009)
    b = P.rhs_normal(shape, 1009)
    sysd = _system(shape, coeff)
    gpu.set_coefficient(c)
    # the slots on host arrays
    u = np.random.default_rng(14).standard_normal(n)
    y = np.empty(n)
    L.Lap_Vec_mult_vc(ctypes.byref(sysd), a, u.ctypes.data, y.ctypes.data, 0)
    assert np.array_equal(y, oracle.matvec_vc(u, shape, a, c))
    r = u.copy()
    diag = a * (3 * coeff[0])
    L.Precondition_vc(diag, r.ctypes.data, n)
    assert np.array_equal(r, u / (diag + c))
    # AAR() through the reference signature
    x = np.zeros(n)
    bb = b.copy()
    L.AAR(ctypes.byref(sysd), fp["Lap_Vec_mult_vc"], fp["Precondition_vc"], a, x.ctypes.data, bb.ctypes.data,
          0.6, 0.6, 7, 6, 1000, 1e-8, n, 0)
    st = _lib.Stats()
    L.aar_hip_last_stats(ctypes.byref(st))
    xr, dr = oracle.aar_vc(np.zeros(n), b, shape, a, c, 0.6, 0.6, 7, 6, 1000, 1e-8)
    assert st.iterations == dr["iterations"]
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= 1e-8
    # the variable-coefficient operator with the constant-diagonal
    # preconditioner, and PGR on it, are refused; x is left alone
    x2 = np.zeros(n)
    L.AAR(ctypes.byref(sysd), fp["Lap_Vec_mult_vc"], fp["Precondition"], a, x2.ctypes.data, bb.ctypes.data,
          0.6, 0.6, 7, 6, 1000, 1e-8, n, 0)
    assert not x2.any()
    L.PGR(ctypes.byref(sysd), fp["Lap_Vec_mult_vc"], fp["Precondition_vc"], a, x2.ctypes.data, bb.ctypes.data,
          0.6, 7, 6, 1000, 1e-8, n, 0)
    assert not x2.any()
    gpu.set_coefficient(None)
    L.AAR(ctypes.byref(sysd), fp["Lap_Vec_mult_vc"], fp["Precondition_vc"], a, x2.ctypes.data, bb.ctypes.data,
          0.6, 0.6, 7, 6, 1000, 1e-8, n, 0)
    assert not x2.any()
