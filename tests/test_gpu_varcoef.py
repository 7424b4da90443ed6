"""GPU variable-coefficient operator A = a*Lap_FD12 + c(x) against the oracle.

The matvec is bit-exact.  AAR with the pointwise Jacobi diagonal
a*(3*coeff[0]) + c(x) (or none) gives the oracle's iteration counts
(orc_aar_vc: the reference's AAR.c control flow on this operator; it equals
the reference-pinned screened AAR bit for bit when c is constant and the
dense solve when not), solutions within 1e-8 (1e-6 for an iterate stopped
at max_iter) and residual histories within 1e-6, in FP64 and in mixed
precision; PGR, PL2R and CG (orc_pgr_vc, orc_cg_vc, pinned the same way)
likewise.  The drop-in slots Lap_Vec_mult_vc / Precondition_vc run through
the reference's own AAR(), PGR() and CG() signatures with a registered c(x).
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


PGR_CASES = [
    # solver, shape, m, p, max_iter, tol, precond, omega
    ("pgr", (16, 16, 16), 7, 6, 1000, 1e-8, True, 0.6),
    ("pl2r", (16, 16, 16), 7, 6, 1000, 1e-8, True, 0.6),
    ("pgr", (20, 12, 16), 3, 5, 1000, 1e-8, True, 0.6),
    ("pl2r", (24, 24, 24), 7, 6, 17, 1e-8, True, 0.6),
    ("pgr", (16, 16, 16), 7, 6, 600, 1e-6, False, 0.015),
    ("cg", (16, 16, 16), 0, 0, 1000, 1e-8, True, 0.0),
    ("cg", (20, 12, 16), 0, 0, 1000, 1e-10, True, 0.0),
    ("cg", (24, 24, 24), 0, 0, 37, 1e-10, True, 0.0),
    ("cg", (16, 16, 16), 0, 0, 1000, 1e-8, False, 0.0),
]


@pytest.mark.parametrize("case", PGR_CASES, ids=lambda c: f"{c[0]}_" + "x".join(map(str, c[1])) +
                         f"_i{c[4]}" + ("" if c[6] else "_noprec"))
def test_vc_pgr_cg_parity(gpu, oracle, case):
    """PGR / PL2R / CG on a*Lap + c(x): the oracle's iteration counts and
    printed conditions, solutions within 1e-8 (CG 1e-7; 1e-6 at a max_iter
    stop) and residual traces within 1e-6."""
    solver, shape, m, p, max_iter, tol, precond, om = case
    say("vc", solver, shape, max_iter, precond)
    n = int(np.prod(shape))
    a = P.poisson_a(0.5)
    c = P.helmholtz_coefficient(shape, 1011)
    b = P.rhs_normal(shape, 1011)
    xd = dev(np.zeros(n))
    if solver == "cg":
        xr, dr = oracle.cg_vc(np.zeros(n), b, shape, a, c, max_iter, tol, precond=precond)
        st = gpu.CG(xd, dev(b), shape, a, max_iter, tol, precond=precond, c=dev(c))
    else:
        l2 = solver == "pl2r"
        xr, dr = oracle.pgr_vc(np.zeros(n), b, shape, a, c, om, m, p, max_iter, tol, precond=precond, l2=l2)
        fn = gpu.PL2R if l2 else gpu.PGR
        st = fn(xd, dev(b), shape, a, om, m, p, max_iter, tol, precond=precond, c=dev(c))
    assert st["iterations"] == dr["iterations"]
    assert bool(st["printed_converged"]) == bool(dr["printed_converged"])
    x = xd.cpu().numpy()
    tol_x = (1e-7 if solver == "cg" else 1e-8) if dr["printed_converged"] else 1e-6
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= tol_x
    h_gpu = gpu.history()
    tr = dr["trace"]
    if solver == "cg":
        # CG's trace holds the explicit residuals: the start and every 50th
        # iteration (CG.c:109); the history has one entry per iteration
        h_gpu = np.concatenate([h_gpu[:1], h_gpu[1::50]])
    k = min(len(h_gpu), len(tr))
    assert k == len(tr)
    assert np.all(np.abs(h_gpu[:k] - tr) <= 1e-6 * tr + 1e-10 * tr[0])


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
    """The reference's AAR(), PGR() and CG() entries with Lap_Vec_mult_vc /
    Precondition_vc and a registered c(x): the oracle's iteration counts and
    solutions; the host slots themselves are bit-exact; a mismatched pair or
    a missing registration is refused without touching x."""
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
    # PGR() and CG() through the reference signatures
    xp = np.zeros(n)
    L.PGR(ctypes.byref(sysd), fp["Lap_Vec_mult_vc"], fp["Precondition_vc"], a, xp.ctypes.data, bb.ctypes.data,
          0.6, 7, 6, 1000, 1e-8, n, 0)
    L.aar_hip_last_stats(ctypes.byref(st))
    xr, dr = oracle.pgr_vc(np.zeros(n), b, shape, a, c, 0.6, 7, 6, 1000, 1e-8)
    assert st.iterations == dr["iterations"]
    assert np.linalg.norm(xp - xr) / np.linalg.norm(xr) <= 1e-8
    L.CG.argtypes = [ctypes.POINTER(POISSON), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                     ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    L.CG.restype = None
    xc = np.zeros(n)
    L.CG(ctypes.byref(sysd), fp["Lap_Vec_mult_vc"], fp["Precondition_vc"], a, xc.ctypes.data, bb.ctypes.data,
         1000, 1e-8, n, 0)
    L.aar_hip_last_stats(ctypes.byref(st))
    xr, dr = oracle.cg_vc(np.zeros(n), b, shape, a, c, 1000, 1e-8)
    assert st.iterations == dr["iterations"]
    assert np.linalg.norm(xc - xr) / np.linalg.norm(xr) <= 1e-7
    # the variable-coefficient operator with the constant-diagonal
    # preconditioner is refused; x is left alone
    x2 = np.zeros(n)
    L.AAR(ctypes.byref(sysd), fp["Lap_Vec_mult_vc"], fp["Precondition"], a, x2.ctypes.data, bb.ctypes.data,
          0.6, 0.6, 7, 6, 1000, 1e-8, n, 0)
    assert not x2.any()
    gpu.set_coefficient(None)
    L.AAR(ctypes.byref(sysd), fp["Lap_Vec_mult_vc"], fp["Precondition_vc"], a, x2.ctypes.data, bb.ctypes.data,
          0.6, 0.6, 7, 6, 1000, 1e-8, n, 0)
    assert not x2.any()
