"""Pins of the oracle's Krylov comparators (orc_gmres, orc_bicg) -- CPU only.

They restate PETSc's KSPGMRES / KSPBICG with PCJACOBI as the PETSc variant
of the reference runs them (PETSC_REAL/main.c:86-150).  PETSc is not in
the reference tree and not installed, so nothing pins them to PETSc's own
rounding; they are pinned to what the mathematics fixes:
* GMRES: the residual estimate after k Arnoldi steps is the minimum of
  ||r0 - B A V y|| over the k-dimensional Krylov space (dense least squares
  on a small grid), and the solution built at a cycle's end has exactly
  that preconditioned residual;
* BiCG with r~0 = r0 on a symmetric operator is preconditioned CG (a numpy
  PCG run step by step), with the shadow sequence equal to the primal one;
* both converge to the FFT solution; the KSPConvergedReason codes at the
  edges (max_it reached, zero residual).
"""
import numpy as np
import pytest


def fft_solve(b, shape, a, coeff, sigma):
    """Spectral solve of a*Lap(x) + sigma x = b on the periodic grid
    (sigma > 0: the operator is nonsingular)."""
    nx, ny, nz = shape
    def s1(n):
        th = 2 * np.pi * np.arange(n) / n
        return coeff[0] + 2 * sum(coeff[p] * np.cos(p * th) for p in range(1, len(coeff)))
    lam = a * (s1(nz)[:, None, None] + s1(ny)[None, :, None] + s1(nx)[None, None, :]) + sigma
    return np.real(np.fft.ifftn(np.fft.fftn(b.reshape(nz, ny, nx)) / lam)).reshape(-1)


def dense_operator(oracle, shape, a, sigma):
    n = int(np.prod(shape))
    A = np.empty((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[:] = 0.0
        e[j] = 1.0
        A[:, j] = oracle.matvec(e, shape, a, sigma=sigma)
    return A


def diag_of(a, sigma):
    from oracle import oracle as O
    return a * 3 * O.lap_coefficient(6)[0] + sigma


def rhs(shape, seed):
    rng = np.random.default_rng(seed)
    b = rng.uniform(-0.5, 0.5, int(np.prod(shape)))
    return b - b.mean()


@pytest.mark.parametrize("sigma,precond", [(0.5, 1), (0.0, 1), (0.5, 0)])
def test_gmres_estimate_is_krylov_minimum(oracle, sigma, precond):
    shape = (8, 8, 8)
    a = -4.0
    b = rhs(shape, 11)
    x0 = np.ones(b.size)
    A = dense_operator(oracle, shape, a, sigma)
    d = diag_of(a, sigma)
    Bd = (1.0 / d) if precond else 1.0
    M = Bd * A
    r0 = Bd * (b - A @ x0)
    _, st = oracle.gmres(x0, b, shape, a, 12, 1e-14, restart=30, sigma=sigma, precond=precond)
    tr = st["trace"]
    assert st["iterations"] == 12 and st["status"] == oracle.KSP_DIVERGED_ITS
    assert tr[0] == pytest.approx(np.linalg.norm(r0), rel=1e-13)
    # orthonormal Krylov basis by twice-repeated Gram-Schmidt
    Q = [r0 / np.linalg.norm(r0)]
    for k in range(1, 13):
        v = M @ Q[-1]
        for _ in range(2):
            for q in Q:
                v = v - (q @ v) * q
        Q.append(v / np.linalg.norm(v))
        K = np.array(Q[:k]).T
        y, *_ = np.linalg.lstsq(M @ K, r0, rcond=None)
        best = np.linalg.norm(r0 - M @ K @ y)
        assert tr[k] == pytest.approx(best, rel=1e-7), k


def test_gmres_cycle_solution_has_the_estimated_residual(oracle):
    """BuildSoln: after one cycle of 6 steps, ||B(b - A x)|| equals the
    last estimate; the next cycle starts from that true residual."""
    shape = (8, 8, 8)
    a, sigma = -4.0, 0.5
    b = rhs(shape, 12)
    A = dense_operator(oracle, shape, a, sigma)
    d = diag_of(a, sigma)
    x, st = oracle.gmres(np.ones(b.size), b, shape, a, 6, 1e-14, restart=6, sigma=sigma)
    true = np.linalg.norm((b - A @ x) / d)
    assert st["trace"][6] == pytest.approx(true, rel=1e-9)
    x2, st2 = oracle.gmres(np.ones(b.size), b, shape, a, 7, 1e-14, restart=6, sigma=sigma)
    assert len(st2["trace"]) == 9          # start, 6 steps, restart, 1 step
    assert st2["trace"][7] == pytest.approx(true, rel=1e-9)


def test_gmres_and_bicg_converge_to_fft_solution(oracle):
    shape = (16, 16, 16)
    a, sigma = -4.0, 0.5
    b = rhs(shape, 13)
    ref = fft_solve(b, shape, a, oracle.lap_coefficient(6), sigma)
    for fn in (oracle.gmres, oracle.bicg):
        x, st = fn(np.ones(b.size), b, shape, a, 2000, 1e-10, sigma=sigma)
        assert st["status"] == oracle.KSP_CONVERGED_RTOL
        assert np.abs(x - ref).max() <= 1e-7 * np.abs(ref).max()


def test_bicg_is_pcg_on_the_symmetric_operator(oracle):
    shape = (12, 12, 12)
    a, sigma = -4.0, 0.25
    b = rhs(shape, 14)
    A = dense_operator(oracle, shape, a, sigma)
    dinv = 1.0 / diag_of(a, sigma)
    x0 = np.ones(b.size)
    x, st = oracle.bicg(x0, b, shape, a, 2000, 1e-6, sigma=sigma)
    assert st["shadow_gap"] == 0.0
    # textbook PCG with z = B r, stopped on ||z|| <= rtol ||B b||
    xx = x0.copy()
    r = b - A @ xx
    z = dinv * r
    p = z.copy()
    rz = r @ z
    ttol = 1e-6 * np.linalg.norm(dinv * b)
    norms = [np.linalg.norm(z)]
    its = 0
    while norms[-1] > ttol:
        q = A @ p
        al = rz / (p @ q)
        xx = xx + al * p
        r = r - al * q
        z = dinv * r
        norms.append(np.linalg.norm(z))
        its += 1
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    assert st["iterations"] == its
    np.testing.assert_allclose(st["trace"], norms, rtol=1e-6)
    assert np.abs(x - xx).max() <= 1e-9 * np.abs(xx).max()


def test_ksp_reason_codes(oracle):
    shape = (8, 8, 8)
    a = -4.0
    b = rhs(shape, 15)
    n = b.size
    _, st = oracle.gmres(np.zeros(n), b, shape, a, 5, 1e-12, restart=30)
    assert st["status"] == oracle.KSP_DIVERGED_ITS and st["iterations"] == 5
    _, st = oracle.bicg(np.zeros(n), b, shape, a, 5, 1e-12)
    assert st["status"] == oracle.KSP_DIVERGED_ITS and st["iterations"] == 5
    # zero residual at the start: converged on the absolute tolerance
    x, st = oracle.gmres(np.zeros(n), np.zeros(n), shape, a, 100, 1e-6)
    assert st["status"] == oracle.KSP_CONVERGED_ATOL and st["iterations"] == 0 and not x.any()
    x, st = oracle.bicg(np.zeros(n), np.zeros(n), shape, a, 100, 1e-6)
    assert st["status"] == oracle.KSP_CONVERGED_ATOL and st["iterations"] == 0 and not x.any()
    # max_it = 0: no step; PETSc's BiCG loop runs one step before its test
    _, st = oracle.gmres(np.zeros(n), b, shape, a, 0, 1e-6)
    assert st["status"] == oracle.KSP_DIVERGED_ITS and st["iterations"] == 0
    _, st = oracle.bicg(np.zeros(n), b, shape, a, 0, 1e-6)
    assert st["status"] == oracle.KSP_DIVERGED_ITS and st["iterations"] == 1


def test_gmres_variable_coefficient_reduces_to_screened(oracle):
    """c(x) = sigma everywhere gives the screened run bit for bit."""
    shape = (10, 10, 10)
    a, sigma = -4.0, 0.5
    b = rhs(shape, 16)
    c = np.full(b.size, sigma)
    x1, s1 = oracle.gmres(np.ones(b.size), b, shape, a, 300, 1e-8, sigma=sigma)
    x2, s2 = oracle.gmres(np.ones(b.size), b, shape, a, 300, 1e-8, c=c)
    assert s1["iterations"] == s2["iterations"]
    assert np.array_equal(x1, x2) and np.array_equal(s1["trace"], s2["trace"])
    x1, s1 = oracle.bicg(np.ones(b.size), b, shape, a, 300, 1e-8, sigma=sigma)
    x2, s2 = oracle.bicg(np.ones(b.size), b, shape, a, 300, 1e-8, c=c)
    assert np.array_equal(x1, x2) and np.array_equal(s1["trace"], s2["trace"])
