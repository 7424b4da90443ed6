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

from aar_hip import problems as P


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


# ---------------------------------------------------------------------------
# ILU(0) (the pc = 1 preconditioner of the PETSc variant, one block)
# ---------------------------------------------------------------------------
def slot(axis, d):
    return 1 + 12 * axis + (d + 6 if d < 0 else d + 5)


def pdiff(cj, c, n):
    d = cj - c
    if d > n // 2:
        d -= n
    if d < -(n // 2):
        d += n
    return d


def star_rows(shape, a, sigma, c=None):
    """The assembled matrix as {column: value} per row (PETSC_REAL/system.c:243-273)."""
    from oracle import oracle as O
    nx, ny, nz = shape
    cf = O.lap_coefficient(6)
    rows = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = len(rows)
                d = {r: a * (3 * cf[0]) + sigma + (0.0 if c is None else c[r])}
                for l in range(1, 7):
                    for q in (((i - l) % nx) + j * nx + k * nx * ny, ((i + l) % nx) + j * nx + k * nx * ny,
                              i + ((j - l) % ny) * nx + k * nx * ny, i + ((j + l) % ny) * nx + k * nx * ny,
                              i + j * nx + ((k - l) % nz) * nx * ny, i + j * nx + ((k + l) % nz) * nx * ny):
                        d[q] = d.get(q, 0.0) + a * cf[l]
                rows.append(d)
    return rows


def generic_ilu0(rows):
    """Textbook IKJ ILU(0) on a general sparse pattern (columns ascending;
    the pivot used through its reciprocal); returns per-row dicts of L
    (below the diagonal) and U entries and the reciprocal pivots."""
    n = len(rows)
    U = [dict(r) for r in rows]
    inv = [0.0] * n
    for i in range(n):
        ri = U[i]
        for kk in sorted(q for q in ri if q < i):
            l = ri[kk] * inv[kk]
            ri[kk] = l
            for jj in sorted(q for q in U[kk] if q > kk):
                if jj in ri:
                    ri[jj] = ri[jj] - l * U[kk][jj]
        inv[i] = 1.0 / ri[i]
    return U, inv


def to_slots(shape, U, inv):
    nx, ny, nz = shape
    n = len(U)
    F = np.zeros((n, 37))
    dims, strides = (nx, ny, nz), (1, nx, nx * ny)
    for i in range(n):
        cr = (i % nx, (i // nx) % ny, i // (nx * ny))
        F[i, 0] = inv[i]
        for col, v in U[i].items():
            if col == i:
                continue
            cc = (col % nx, (col // nx) % ny, col // (nx * ny))
            ax = [q for q in range(3) if cc[q] != cr[q]]
            assert len(ax) == 1
            ax = ax[0]
            F[i, slot(ax, pdiff(cc[ax], cr[ax], dims[ax]))] = v
    return F


@pytest.mark.parametrize("shape,sigma,vc", [((13, 13, 13), 0.0, False), ((14, 13, 15), 0.5, False),
                                            ((13, 14, 13), 0.0, True)])
def test_ilu0_factor_equals_generic_ikj(oracle, shape, sigma, vc):
    a = -4.0
    n = int(np.prod(shape))
    c = np.random.default_rng(3).uniform(0.5, 1.5, n) if vc else None
    F = oracle.ilu0_factor(shape, a, sigma=sigma, c=c)
    U, inv = generic_ilu0(star_rows(shape, a, sigma, c))
    assert np.array_equal(F, to_slots(shape, U, inv))


def test_ilu0_defining_identity_and_solves(oracle):
    """(L U)_ij = A_ij on the pattern; the apply is the dense triangular
    solves with L and U, and its transpose."""
    import scipy.linalg as sl
    shape = (13, 13, 13)
    a, sigma = -4.0, 0.25
    rows = star_rows(shape, a, sigma)
    F = oracle.ilu0_factor(shape, a, sigma=sigma)
    U, inv = generic_ilu0(rows)
    n = len(rows)
    Ld = np.eye(n)
    Ud = np.zeros((n, n))
    for i in range(n):
        for col, v in U[i].items():
            if col < i:
                Ld[i, col] = v
            else:
                Ud[i, col] = v
    LU = Ld @ Ud
    for i in range(0, n, 7):
        for col, v in rows[i].items():
            assert LU[i, col] == pytest.approx(v, rel=1e-12, abs=1e-12)
    r = np.random.default_rng(4).standard_normal(n)
    z = sl.solve_triangular(Ud, sl.solve_triangular(Ld, r, lower=True, unit_diagonal=True), lower=False)
    np.testing.assert_allclose(oracle.ilu0_apply(shape, F, r), z, rtol=1e-11, atol=1e-14)
    zt = sl.solve_triangular(Ld.T, sl.solve_triangular(Ud.T, r, lower=True), lower=False, unit_diagonal=True)
    np.testing.assert_allclose(oracle.ilu0_apply(shape, F, r, trans=True), zt, rtol=1e-11, atol=1e-14)
    # in exact arithmetic ILU(0) of a symmetric matrix is symmetric (= ICC(0))
    M = LU
    assert np.abs(M - M.T).max() <= 1e-11 * np.abs(M).max()


def test_gmres_ilu0_estimate_is_krylov_minimum(oracle):
    shape = (13, 13, 13)
    a, sigma = -4.0, 0.5
    b = rhs(shape, 21)
    x0 = np.ones(b.size)
    A = dense_operator(oracle, shape, a, sigma)
    F = oracle.ilu0_factor(shape, a, sigma=sigma)
    Minv = lambda v: oracle.ilu0_apply(shape, F, v)  # noqa: E731
    r0 = Minv(b - A @ x0)
    _, st = oracle.gmres(x0, b, shape, a, 10, 1e-14, restart=30, sigma=sigma, precond=2)
    tr = st["trace"]
    assert tr[0] == pytest.approx(np.linalg.norm(r0), rel=1e-13)
    Q = [r0 / np.linalg.norm(r0)]
    MA = np.array([Minv(A[:, j]) for j in range(A.shape[1])]).T
    for k in range(1, 11):
        v = MA @ Q[-1]
        for _ in range(2):
            for q in Q:
                v = v - (q @ v) * q
        Q.append(v / np.linalg.norm(v))
        K = np.array(Q[:k]).T
        y, *_ = np.linalg.lstsq(MA @ K, r0, rcond=None)
        assert tr[k] == pytest.approx(np.linalg.norm(r0 - MA @ K @ y), rel=1e-7), k


def test_bicg_ilu0_is_pcg_with_the_ilu_preconditioner(oracle):
    shape = (13, 13, 13)
    a, sigma = -4.0, 0.25
    b = rhs(shape, 22)
    A = dense_operator(oracle, shape, a, sigma)
    F = oracle.ilu0_factor(shape, a, sigma=sigma)
    Minv = lambda v: oracle.ilu0_apply(shape, F, v)  # noqa: E731
    x0 = np.ones(b.size)
    x, st = oracle.bicg(x0, b, shape, a, 500, 1e-8, sigma=sigma, precond=2)
    assert st["status"] == oracle.KSP_CONVERGED_RTOL
    assert st["shadow_gap"] < 1e-10 * np.abs(b).max()      # M^-T = M^-1 up to rounding
    xx = x0.copy()
    r = b - A @ xx
    z = Minv(r)
    p = z.copy()
    rz = r @ z
    norms = [np.linalg.norm(z)]
    ttol = 1e-8 * np.linalg.norm(Minv(b))
    while norms[-1] > ttol:
        q = A @ p
        al = rz / (p @ q)
        xx = xx + al * p
        r = r - al * q
        z = Minv(r)
        norms.append(np.linalg.norm(z))
        rzn = r @ z
        p = z + (rzn / rz) * p
        rz = rzn
    assert st["iterations"] == len(norms) - 1
    np.testing.assert_allclose(st["trace"], norms, rtol=1e-5)
    ref = fft_solve(b, shape, a, oracle.lap_coefficient(6), sigma)
    assert np.abs(x - ref).max() <= 1e-6 * np.abs(ref).max()


# ---------------------------------------------------------------------------
# AAR of the PETSc variant (AAR_Real.c) with Jacobi or ILU(0)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("m,p", [(7, 6), (5, 4), (3, 3), (4, 2)])
def test_aar_petsc_jacobi_matches_the_reference_pinned_aar(oracle, m, p):
    """With p >= 2 the PETSc variant's AAR (Jacobi) runs the control flow of
    MPI_REAL/AAR.c, which the golden vectors pin: same iteration count,
    solutions equal to rounding (reciprocal vs division, update order).
    (m = 3, p = 2 is left out: its Gram is ill-conditioned enough that the
    rounding differences grow ~10x every few checks and move the count.)"""
    shape = (16, 16, 16)
    a, sigma = -4.0, 0.0
    b = rhs(shape, 31)
    x1, d1 = oracle.aar_petsc(np.ones(b.size), b, shape, a, 0.6, 0.6, m, p, 1000, 1e-8, precond=1)
    x2, d2 = oracle.aar(np.ones(b.size), b, shape, a, 0.6, 0.6, m, p, 1000, 1e-8)
    assert d1["iterations"] == d2["iterations"] and d1["converged"]
    assert np.abs(x1 - x2).max() <= 1e-10 * np.abs(x2).max()
    np.testing.assert_allclose(d1["trace"], d2["trace"][p::p][:len(d1["trace"])], rtol=1e-6,
                               atol=1e-10 * d2["trace"][0])


def test_aar_petsc_first_steps(oracle):
    """p = 1: the Anderson step at iter = 1 has an empty history, so x1 =
    x0 + beta B r0 (AAR_Real.c:116 has no iter > 1 guard); pure Richardson
    (p > max_iter) is x_{k+1} = x_k + omega B (b - A x_k) with B = M^{-1}."""
    shape = (13, 13, 13)
    a, sigma = -4.0, 0.5
    b = rhs(shape, 32)
    A = dense_operator(oracle, shape, a, sigma)
    F = oracle.ilu0_factor(shape, a, sigma=sigma)
    Minv = lambda v: oracle.ilu0_apply(shape, F, v)  # noqa: E731
    x0 = np.ones(b.size)
    x1, _ = oracle.aar_petsc(x0, b, shape, a, 0.6, 0.9, 4, 1, 1, 1e-14, sigma=sigma, precond=2)
    np.testing.assert_allclose(x1, x0 + 0.9 * Minv(b - A @ x0), rtol=1e-12, atol=1e-14)
    xr, _ = oracle.aar_petsc(x0, b, shape, a, 0.7, 0.6, 4, 100, 5, 1e-14, sigma=sigma, precond=2)
    xx = x0.copy()
    for _ in range(5):
        xx = xx + 0.7 * Minv(b - A @ xx)
    np.testing.assert_allclose(xr, xx, rtol=1e-11, atol=1e-13)


def test_aar_petsc_ilu0_converges_to_fft_solution(oracle):
    shape = (16, 16, 16)
    a, sigma = -4.0, 0.5
    b = rhs(shape, 33)
    ref = fft_solve(b, shape, a, oracle.lap_coefficient(6), sigma)
    x, d = oracle.aar_petsc(np.ones(b.size), b, shape, a, 0.6, 0.6, 7, 6, 1000, 1e-10, sigma=sigma, precond=2)
    assert d["converged"] and d["iterations"] % 6 == 0
    assert np.abs(x - ref).max() <= 1e-7 * np.abs(ref).max()
    _, dj = oracle.aar_petsc(np.ones(b.size), b, shape, a, 0.6, 0.6, 7, 6, 1000, 1e-10, sigma=sigma, precond=1)
    assert d["iterations"] < dj["iterations"]


@pytest.mark.parametrize("l2", [False, True])
@pytest.mark.parametrize("m,p", [(7, 6), (5, 4), (4, 3)])
def test_pgr_petsc_jacobi_matches_the_reference_pinned_pgr(oracle, l2, m, p):
    """The PETSc variant's PGR/PL2R with Jacobi runs the control flow of
    MPI_REAL/PGR.c / PL2R.c (pinned by the golden vectors): same counts,
    solutions equal to rounding."""
    shape = (16, 16, 16)
    a = -4.0
    b = rhs(shape, 41)
    x1, d1 = oracle.pgr_petsc(np.ones(b.size), b, shape, a, 0.6, m, p, 1000, 1e-8, precond=1, l2=l2)
    x2, d2 = oracle.pgr(np.ones(b.size), b, shape, a, 0.6, m, p, 1000, 1e-8, l2=l2)
    assert d1["iterations"] == d2["iterations"] and d1["converged"]
    assert np.abs(x1 - x2).max() <= 1e-9 * np.abs(x2).max()


@pytest.mark.parametrize("l2", [False, True])
def test_pgr_petsc_ilu0_first_step_and_solution(oracle, l2):
    """One iteration is x0 + omega B r0 (the history step alone, p > 1);
    converged solves reach the FFT solution."""
    shape = (13, 13, 13)
    a, sigma = -4.0, 0.5
    b = rhs(shape, 42)
    A = dense_operator(oracle, shape, a, sigma)
    F = oracle.ilu0_factor(shape, a, sigma=sigma)
    x0 = np.ones(b.size)
    x1, _ = oracle.pgr_petsc(x0, b, shape, a, 0.6, 4, 3, 1, 1e-14, sigma=sigma, precond=2, l2=l2)
    np.testing.assert_allclose(x1, x0 + 0.6 * oracle.ilu0_apply(shape, F, b - A @ x0), rtol=1e-12, atol=1e-14)
    x, d = oracle.pgr_petsc(x0, b, shape, a, 0.6, 7, 6, 1000, 1e-10, sigma=sigma, precond=2, l2=l2)
    ref = fft_solve(b, shape, a, oracle.lap_coefficient(6), sigma)
    assert d["converged"] and np.abs(x - ref).max() <= 1e-7 * np.abs(ref).max()


# ---------------------------------------------------------------------------
# GMRES of the complex PETSc variant on the Helmholtz operator
# ---------------------------------------------------------------------------
def test_gmres_complex_estimate_is_krylov_minimum(oracle):
    shape = (8, 8, 8)
    n = 512
    a, c0, eta = -4.0, 0.3, 0.5
    rng = np.random.default_rng(51)
    c = rng.uniform(0.5, 1.5, n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    u0 = np.ones(n, dtype=complex)
    A = np.empty((n, n), dtype=complex)
    e = np.zeros(n, dtype=complex)
    for j in range(n):
        e[:] = 0
        e[j] = 1
        A[:, j] = oracle.matvec_complex(e, shape, a, c, c0, eta)
    D = a * 3 * oracle.lap_coefficient(6)[0] + c + 1j * eta
    M = A / D[:, None]
    r0 = (b - A @ u0) / D
    _, st = oracle.gmres_complex(u0, b, shape, a, c, c0, eta, 10, 1e-14)
    tr = st["trace"]
    assert tr[0] == pytest.approx(np.linalg.norm(r0), rel=1e-12)
    Q = [r0 / np.linalg.norm(r0)]
    for k in range(1, 11):
        v = M @ Q[-1]
        for _ in range(2):
            for q in Q:
                v = v - (np.vdot(q, v)) * q
        Q.append(v / np.linalg.norm(v))
        K = np.array(Q[:k]).T
        y, *_ = np.linalg.lstsq(M @ K, r0, rcond=None)
        assert tr[k] == pytest.approx(np.linalg.norm(r0 - M @ K @ y), rel=1e-7), k


def test_gmres_complex_reduces_to_real_and_solves(oracle):
    """eta = 0, constant c, real b: the real screened GMRES with sigma = c
    (same count, norms to rounding: the Jacobi reciprocal differs in its
    last bit); eta > 0: converges to the FFT solution."""
    shape = (12, 12, 12)
    n = int(np.prod(shape))
    a = -4.0
    b = rhs(shape, 52)
    xr, dr = oracle.gmres(np.ones(n), b, shape, a, 500, 1e-8, sigma=0.7)
    xc, dc = oracle.gmres_complex(np.ones(n, dtype=complex), b.astype(complex), shape, a, None, 0.7, 0.0, 500, 1e-8)
    assert dc["iterations"] == dr["iterations"]
    np.testing.assert_allclose(dc["trace"], dr["trace"], rtol=1e-9, atol=1e-11 * dr["trace"][0])
    assert np.abs(xc.imag).max() == 0.0 and np.abs(xc.real - xr).max() <= 1e-10 * np.abs(xr).max()
    eta, c0 = 0.5, 0.3
    bc = np.random.default_rng(53).standard_normal(n) + 1j * np.random.default_rng(54).standard_normal(n)
    u, d = oracle.gmres_complex(np.ones(n, dtype=complex), bc, shape, a, None, c0, eta, 2000, 1e-10)
    cf = oracle.lap_coefficient(6)
    nx, ny, nz = shape
    def s1(m):
        th = 2 * np.pi * np.arange(m) / m
        return cf[0] + 2 * sum(cf[q] * np.cos(q * th) for q in range(1, 7))
    lam = a * (s1(nz)[:, None, None] + s1(ny)[None, :, None] + s1(nx)[None, None, :]) + c0 + 1j * eta
    ref = np.fft.ifftn(np.fft.fftn(bc.reshape(nz, ny, nx)) / lam).reshape(-1)
    assert d["status"] == oracle.KSP_CONVERGED_RTOL
    assert np.abs(u - ref).max() <= 1e-7 * np.abs(ref).max()


def test_bicg_complex_reduces_to_real_and_solves(oracle):
    shape = (12, 12, 12)
    n = int(np.prod(shape))
    a = -4.0
    b = rhs(shape, 56)
    xr, dr = oracle.bicg(np.ones(n), b, shape, a, 500, 1e-8, sigma=0.7)
    xc, dc = oracle.bicg_complex(np.ones(n, dtype=complex), b.astype(complex), shape, a, None, 0.7, 0.0, 500, 1e-8)
    assert dc["iterations"] == dr["iterations"]
    np.testing.assert_allclose(dc["trace"], dr["trace"], rtol=1e-9, atol=1e-11 * dr["trace"][0])
    assert np.abs(xc.imag).max() == 0.0 and np.abs(xc.real - xr).max() <= 1e-10 * np.abs(xr).max()
    eta, c0 = 0.5, 0.3
    bc = np.random.default_rng(57).standard_normal(n) + 1j * np.random.default_rng(58).standard_normal(n)
    u, d = oracle.bicg_complex(np.ones(n, dtype=complex), bc, shape, a, None, c0, eta, 2000, 1e-10)
    ug, dg = oracle.gmres_complex(np.ones(n, dtype=complex), bc, shape, a, None, c0, eta, 2000, 1e-10)
    cf = oracle.lap_coefficient(6)
    nx, ny, nz = shape
    def s1(m):
        th = 2 * np.pi * np.arange(m) / m
        return cf[0] + 2 * sum(cf[q] * np.cos(q * th) for q in range(1, 7))
    lam = a * (s1(nz)[:, None, None] + s1(ny)[None, :, None] + s1(nx)[None, None, :]) + c0 + 1j * eta
    ref = np.fft.ifftn(np.fft.fftn(bc.reshape(nz, ny, nx)) / lam).reshape(-1)
    assert d["status"] == oracle.KSP_CONVERGED_RTOL
    assert np.abs(u - ref).max() <= 1e-7 * np.abs(ref).max()


# ---------------------------------------------------------------------------
# LGMRES of the complex PETSc variant (PETSC_COMPLEX/main.c:120-147)
# ---------------------------------------------------------------------------
def _helm_dense(oracle, shape, a, c, c0, eta):
    n = int(np.prod(shape))
    A = np.empty((n, n), dtype=complex)
    e = np.zeros(n, dtype=complex)
    for j in range(n):
        e[:] = 0
        e[j] = 1
        A[:, j] = oracle.matvec_complex(e, shape, a, c, c0, eta)
    D = a * 3 * oracle.lap_coefficient(6)[0] + (c0 if c is None else c) + 1j * eta
    return A, D


def _krylov(M, r, k):
    Q = [r / np.linalg.norm(r)]
    while len(Q) < k:
        v = M @ Q[-1]
        for _ in range(2):
            for q in Q:
                v = v - np.vdot(q, v) * q
        Q.append(v / np.linalg.norm(v))
    return np.array(Q).T


def test_lgmres_complex_cycles_minimise_over_the_augmented_space(oracle):
    """restart 6, aug 2: 4 Arnoldi steps per cycle, then one step per stored
    error approximation (newest first).  Stopping after a whole cycle (max_it
    = 4, 9, 15) or after the first augmentation step (14), the iterate is
    x_prev + W y with y minimising ||M (x_prev + W y) - B b||, W = [K_4(M,
    r_prev), z newest, ...] and z the earlier cycles' corrections; the last
    trace entry is that minimum.  A wrong B A z, a dropped or misordered z
    or a wrong combination fails it."""
    shape = (8, 8, 8)
    n = 512
    a, c0, eta = -4.0, 0.3, 0.5
    rng = np.random.default_rng(71)
    c = rng.uniform(0.5, 1.5, n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    u0 = np.ones(n, dtype=complex)
    A, D = _helm_dense(oracle, shape, a, c, c0, eta)
    M = A / D[:, None]

    def run(its):
        u, d = oracle.lgmres_complex(u0, b, shape, a, c, c0, eta, its, 1e-14, restart=6, aug=2)
        assert d["iterations"] == its and d["status"] == oracle.KSP_DIVERGED_ITS
        return u, d

    xs = [u0] + [run(k)[0] for k in (4, 9)]
    for k, prev, zs in ((4, 0, []), (9, 1, [0]), (15, 2, [1, 0]), (14, 2, [1])):
        x, d = run(k)
        xp = xs[prev]
        r = (b - A @ xp) / D
        W = _krylov(M, r, 4)
        for j in zs:
            W = np.column_stack([W, xs[j + 1] - xs[j]])
        y, *_ = np.linalg.lstsq(M @ W, r, rcond=None)
        pred = xp + W @ y
        assert np.linalg.norm(x - pred) <= 1e-8 * np.linalg.norm(pred), k
        assert d["trace"][-1] == pytest.approx(np.linalg.norm(r - M @ W @ y), rel=1e-7), k
    # one trace entry per cycle start and per step: 1+4, 1+5, 1+6
    assert len(run(15)[1]["trace"]) == 18


def test_lgmres_complex_aug0_is_gmres_and_solves(oracle):
    """aug = 0 is GMRES(restart) bit for bit; with PETSc's defaults (30, 2)
    LGMRES converges to the FFT solution with Jacobi and with ILU(0)."""
    shape = (14, 14, 14)
    n = int(np.prod(shape))
    a, c0, eta = -4.0, 0.3, 0.5
    bc = np.random.default_rng(72).standard_normal(n) + 1j * np.random.default_rng(73).standard_normal(n)
    ug, dg = oracle.gmres_complex(np.ones(n, dtype=complex), bc, shape, a, None, c0, eta, 300, 1e-9, restart=12)
    ul, dl = oracle.lgmres_complex(np.ones(n, dtype=complex), bc, shape, a, None, c0, eta, 300, 1e-9, restart=12,
                                   aug=0)
    assert np.array_equal(ug, ul) and np.array_equal(dg["trace"], dl["trace"])
    cf = oracle.lap_coefficient(6)
    def s1(m):
        th = 2 * np.pi * np.arange(m) / m
        return cf[0] + 2 * sum(cf[q] * np.cos(q * th) for q in range(1, 7))
    lam = a * (s1(14)[:, None, None] + s1(14)[None, :, None] + s1(14)[None, None, :]) + c0 + 1j * eta
    ref = np.fft.ifftn(np.fft.fftn(bc.reshape(14, 14, 14)) / lam).reshape(-1)
    for pc in (1, 2):
        u, d = oracle.lgmres_complex(np.ones(n, dtype=complex), bc, shape, a, None, c0, eta, 2000, 1e-10, precond=pc)
        assert d["status"] == oracle.KSP_CONVERGED_RTOL
        assert np.abs(u - ref).max() <= 1e-7 * np.abs(ref).max()


def generic_ilu0_complex(rows):
    """generic_ilu0 for complex entries; the pivot reciprocal by the naive
    formula (re - i im) / (re^2 + im^2), the oracle's."""
    n = len(rows)
    U = [dict(r) for r in rows]
    inv = [0j] * n
    for i in range(n):
        ri = U[i]
        for kk in sorted(q for q in ri if q < i):
            l = ri[kk] * inv[kk]
            ri[kk] = l
            for jj in sorted(q for q in U[kk] if q > kk):
                if jj in ri:
                    ri[jj] = ri[jj] - l * U[kk][jj]
        d = ri[i]
        den = d.real * d.real + d.imag * d.imag
        inv[i] = complex(d.real / den, -d.imag / den)
    return U, inv


def test_ilu0_complex_equals_generic_ikj_and_solves(oracle):
    import scipy.linalg as sl
    shape = (13, 14, 13)
    a, c0, eta = -4.0, 0.0, 0.5
    n = int(np.prod(shape))
    c = np.random.default_rng(61).uniform(0.5, 1.5, n)
    rows = star_rows(shape, a, 0.0, c)
    for i, r in enumerate(rows):
        r[i] = r[i] + 1j * eta
    F = oracle.ilu0_factor_complex(shape, a, c, c0, eta)
    U, inv = generic_ilu0_complex(rows)
    G = np.zeros((n, 37), dtype=complex)
    nx, ny, nz = shape
    dims = (nx, ny, nz)
    for i in range(n):
        cr = (i % nx, (i // nx) % ny, i // (nx * ny))
        G[i, 0] = inv[i]
        for col, v in U[i].items():
            if col == i:
                continue
            cc = (col % nx, (col // nx) % ny, col // (nx * ny))
            ax = [q for q in range(3) if cc[q] != cr[q]][0]
            G[i, slot(ax, pdiff(cc[ax], cr[ax], dims[ax]))] = v
    assert np.array_equal(F, G)
    # solves against dense triangular factors: M^-1, M^-T, M^-H
    Ld = np.eye(n, dtype=complex)
    Ud = np.zeros((n, n), dtype=complex)
    for i in range(n):
        for col, v in U[i].items():
            if col < i:
                Ld[i, col] = v
            else:
                Ud[i, col] = v
    r = np.random.default_rng(62).standard_normal(n) + 1j * np.random.default_rng(63).standard_normal(n)
    z = sl.solve_triangular(Ud, sl.solve_triangular(Ld, r, lower=True, unit_diagonal=True), lower=False)
    np.testing.assert_allclose(oracle.ilu0_apply_complex(shape, F, r, 0), z, rtol=1e-11, atol=1e-14)
    M = Ld @ Ud
    np.testing.assert_allclose(oracle.ilu0_apply_complex(shape, F, r, 1), np.linalg.solve(M.T, r), rtol=1e-9,
                               atol=1e-12)
    np.testing.assert_allclose(oracle.ilu0_apply_complex(shape, F, r, 2), np.linalg.solve(M.conj().T, r),
                               rtol=1e-9, atol=1e-12)


def test_complex_krylov_ilu0_converge_to_fft_solution(oracle):
    shape = (14, 14, 14)
    n = int(np.prod(shape))
    a, c0, eta = -4.0, 0.3, 0.5
    bc = np.random.default_rng(64).standard_normal(n) + 1j * np.random.default_rng(65).standard_normal(n)
    cf = oracle.lap_coefficient(6)
    def s1(m):
        th = 2 * np.pi * np.arange(m) / m
        return cf[0] + 2 * sum(cf[q] * np.cos(q * th) for q in range(1, 7))
    lam = a * (s1(14)[:, None, None] + s1(14)[None, :, None] + s1(14)[None, None, :]) + c0 + 1j * eta
    ref = np.fft.ifftn(np.fft.fftn(bc.reshape(14, 14, 14)) / lam).reshape(-1)
    for fn in (oracle.gmres_complex, oracle.bicg_complex):
        u, d = fn(np.ones(n, dtype=complex), bc, shape, a, None, c0, eta, 2000, 1e-10, precond=2)
        _, dj = fn(np.ones(n, dtype=complex), bc, shape, a, None, c0, eta, 2000, 1e-10, precond=1)
        assert d["status"] == oracle.KSP_CONVERGED_RTOL and d["iterations"] < dj["iterations"]
        assert np.abs(u - ref).max() <= 1e-7 * np.abs(ref).max()


def test_icc0_equals_ilu0_on_the_symmetric_matrix(oracle):
    """PCICC (the variant's pc = 1 for AAR/PGR/PL2R) is an incomplete
    Cholesky factorization; it is realised here as ILU(0).  A textbook IC(0)
    in LDL^T form on the same pattern gives L equal to ILU(0)'s unit lower
    factor and U = D L^T, to rounding: the same preconditioner."""
    shape = (13, 13, 13)
    a, sigma = -4.0, 0.25
    rows = star_rows(shape, a, sigma)
    n = len(rows)
    lower = [sorted(q for q in r if q < i) for i, r in enumerate(rows)]
    L = [dict() for _ in range(n)]
    d = [0.0] * n
    for i in range(n):
        li = L[i]
        for k in lower[i]:
            s = rows[i][k]
            lk = L[k]
            for j, v in li.items():           # columns j < k already computed in row i
                if j in lk:
                    s -= v * d[j] * lk[j]
            li[k] = s / d[k]
        d[i] = rows[i][i] - sum(v * v * d[k] for k, v in li.items())
    U, inv = generic_ilu0(rows)
    err_l = max(abs(U[i][k] - L[i][k]) / max(abs(L[i][k]), 1e-300) for i in range(n) for k in lower[i])
    err_u = max(abs(U[i][j] - d[i] * L[j][i]) / max(abs(U[i][j]), 1e-300)
                for i in range(n) for j in rows[i] if j > i)
    err_d = max(abs(1.0 / inv[i] - d[i]) / abs(d[i]) for i in range(n))
    assert err_l < 1e-11 and err_u < 1e-11 and err_d < 1e-12


def _np_gmres_estimates(M, r0, k, drop_first=False):
    """Residual estimates of k Arnoldi/Givens steps of left-preconditioned
    GMRES on M (dense), classical Gram-Schmidt; drop_first plants a mistake:
    the projection on v_0 is skipped after the first step."""
    n = r0.size
    V = np.zeros((k + 1, n))
    H = np.zeros((k + 1, k))
    beta = np.linalg.norm(r0)
    V[0] = r0 / beta
    g = np.zeros(k + 1)
    g[0] = beta
    cs, sn = np.zeros(k), np.zeros(k)
    est = [beta]
    for j in range(k):
        w = M @ V[j]
        h = V[: j + 1] @ w
        if drop_first and j > 0:
            h[0] = 0.0
        w = w - h @ V[: j + 1]
        H[: j + 1, j] = h
        H[j + 1, j] = np.linalg.norm(w)
        V[j + 1] = w / H[j + 1, j]
        for i in range(j):
            t = H[i, j]
            H[i, j] = cs[i] * t + sn[i] * H[i + 1, j]
            H[i + 1, j] = cs[i] * H[i + 1, j] - sn[i] * t
        r = np.hypot(H[j, j], H[j + 1, j])
        cs[j], sn[j] = H[j, j] / r, H[j + 1, j] / r
        g[j + 1] = -sn[j] * g[j]
        g[j] = cs[j] * g[j]
        est.append(abs(g[j + 1]))
    return np.array(est)


def test_krylov_checks_catch_planted_mistakes(oracle):
    """The Krylov-minimum check passes a correct GMRES restatement and fails
    one that drops a Gram-Schmidt projection; the PCG check fails a BiCG
    whose beta has the wrong sign."""
    shape = (8, 8, 8)
    a, sigma = -4.0, 0.5
    b = rhs(shape, 71)
    A = dense_operator(oracle, shape, a, sigma)
    M = A / diag_of(a, sigma)
    r0 = (b - A @ np.ones(b.size)) / diag_of(a, sigma)
    Q = [r0 / np.linalg.norm(r0)]
    best = [np.linalg.norm(r0)]
    for k in range(1, 11):
        v = M @ Q[-1]
        for _ in range(2):
            for q in Q:
                v = v - (q @ v) * q
        Q.append(v / np.linalg.norm(v))
        K = np.array(Q[:k]).T
        y, *_ = np.linalg.lstsq(M @ K, r0, rcond=None)
        best.append(np.linalg.norm(r0 - M @ K @ y))
    good = _np_gmres_estimates(M, r0, 10)
    bad = _np_gmres_estimates(M, r0, 10, drop_first=True)
    np.testing.assert_allclose(good, best, rtol=1e-7)
    assert np.max(np.abs(bad - best) / np.array(best)) > 1e-3
    # BiCG with a flipped beta sign is no longer PCG
    dinv = 1.0 / diag_of(a, sigma)
    def pcg_norms(flip):
        x = np.ones(b.size)
        r = b - A @ x
        z = dinv * r
        p = z.copy()
        rz = r @ z
        out = [np.linalg.norm(z)]
        for _ in range(8):
            q = A @ p
            al = rz / (p @ q)
            r = r - al * q
            z = dinv * r
            out.append(np.linalg.norm(z))
            rzn = r @ z
            p = z + (-1.0 if flip else 1.0) * (rzn / rz) * p
            rz = rzn
        return np.array(out)
    _, st = oracle.bicg(np.ones(b.size), b, shape, a, 8, 1e-14, sigma=sigma)
    np.testing.assert_allclose(st["trace"], pcg_norms(False), rtol=1e-8)
    assert np.max(np.abs(st["trace"] - pcg_norms(True)) / st["trace"]) > 1e-3


def test_gelsd_cutoff_restatement(oracle):
    """The PETSc variant's least squares (LAPACKE_dgelsd with rcond = -1,
    PETSC_REAL/AAR_Real.c:223) drops singular values s <= 2^-53 s_max, where
    tools.c drops s < m s_max 2.22e-16.  On diagonal systems (whose SVD is
    exact) a singular value between the two cutoffs is kept by the dgelsd
    restatement and by scipy's gelsd driver, and dropped by tools.c; one at
    exactly 2^-53 s_max is dropped by both.  On a well-conditioned SPD system
    the restatement equals scipy's gelsd to rounding."""
    import scipy.linalg as sl
    b = np.array([1.0, 1.0, 1.0])
    A = np.diag([1.0, 0.5, 5e-16])
    xg, wg, _ = oracle.pinv_times_vec(A, b, gelsd=True)
    xt, wt, _ = oracle.pinv_times_vec(A, b)
    xs = sl.lstsq(A, b, cond=-1, lapack_driver="gelsd")[0]
    assert abs(xg[2] / 2e15 - 1.0) < 1e-14 and xt[2] == 0.0
    assert np.allclose(xg, xs, rtol=1e-14)
    A2 = np.diag([1.0, 2.0 ** -53, 0.25])
    xg2, _, _ = oracle.pinv_times_vec(A2, b, gelsd=True)
    assert xg2[1] == 0.0 and np.allclose(xg2[[0, 2]], [1.0, 4.0], rtol=1e-15)
    rng = np.random.default_rng(5)
    F = rng.standard_normal((40, 7))
    G = F.T @ F
    g = rng.standard_normal(7)
    xg3, _, _ = oracle.pinv_times_vec(G, g, gelsd=True)
    assert np.allclose(xg3, sl.lstsq(G, g, cond=-1, lapack_driver="gelsd")[0], rtol=1e-11, atol=0)


@pytest.mark.parametrize("precond", [0, 1, 2])
def test_kspcg_restatement_pins(oracle, precond):
    """KSPCG (PETSc's KSPSolve_CG) on these symmetric operators with a
    symmetric preconditioner is BiCG with its shadow sequence equal to the
    primal one: the oracle's KSPCG reproduces its KSPBICG restatement's
    iterates and counts (bit for bit with Jacobi or none; ILU(0)'s M^{-T}
    differs from M^{-1} in rounding), and converges to the dense solution.
    A planted error in either recurrence would break the equality."""
    shape = (14, 13, 13)
    a = P.poisson_a(0.5)
    b = P.rhs_uniform(shape, 1001)
    x0 = np.ones(b.size)
    x, d = oracle.kspcg(x0, b, shape, a, 2000, 1e-10, sigma=0.3, precond=precond)
    xb, db = oracle.bicg(x0, b, shape, a, 2000, 1e-10, sigma=0.3, precond=precond)
    assert d["iterations"] == db["iterations"] and d["status"] == db["status"] == 2
    if precond == 2:
        assert np.linalg.norm(x - xb) <= 1e-12 * np.linalg.norm(xb)
    else:
        assert np.array_equal(x, xb)
    n = b.size
    M = np.zeros((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        M[:, j] = oracle.matvec(e, shape, a, 0.3)
        e[j] = 0.0
    xs = np.linalg.solve(M, b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8


def test_kspcg_indefinite_exit(oracle):
    """An indefinite operator (screened with sigma = -60) stops KSPCG with
    a divergence reason (indefinite matrix or preconditioner), not a
    convergence."""
    shape = (12, 12, 12)
    b = P.rhs_uniform(shape, 3)
    _, d = oracle.kspcg(np.ones(b.size), b, shape, P.poisson_a(0.5), 500, 1e-8, sigma=-60.0)
    assert d["status"] in (-8, -10)
