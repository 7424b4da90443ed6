"""Pins of the CPU oracle (oracle/aar_oracle.c) -- CPU only.

The oracle is the parity checker for the GPU path, so it is itself pinned:
* bit-for-bit against fixtures written by the real reference solvers
  (tests/golden, made by tests/golden/make_golden.py from oracle/_ref);
* against what the mathematics fixes: the closed-form FD12 weights, the
  Fourier symbol of the periodic stencil, a dense brute-force operator,
  numpy's SVD for the pseudo-inverse, and the FFT spectral solution of the
  periodic system;
* and each check is shown to catch a planted mistake (tests at the end).
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from aar_hip import problems as P

C12 = [Fraction(-5369, 1800), Fraction(12, 7), Fraction(-15, 56), Fraction(10, 189),
       Fraction(-1, 112), Fraction(2, 1925), Fraction(-1, 16632)]


def fd_weights_exact(fdn):
    """Central 2nd-derivative weights of order 2*fdn from the closed form
    c_p = 2 (-1)^(p+1) (fdn!)^2 / ((fdn-p)! (fdn+p)!) / p^2, c0 = -2 sum 1/p^2."""
    c = [Fraction(0)] * (fdn + 1)
    for p in range(1, fdn + 1):
        c[p] = Fraction(2 * (-1) ** (p + 1) * math.factorial(fdn) ** 2,
                        math.factorial(fdn - p) * math.factorial(fdn + p) * p * p)
        c[0] -= Fraction(2, p * p)
    return c


def symbol(shape, a, coeff, sigma=0.0):
    """Eigenvalues of y = a*Lap(x) + sigma x on the periodic grid, shaped (nz, ny, nx)."""
    nx, ny, nz = shape
    def s1(n):
        th = 2 * np.pi * np.arange(n) / n
        return coeff[0] + 2 * sum(coeff[p] * np.cos(p * th) for p in range(1, len(coeff)))
    lam = s1(nz)[:, None, None] + s1(ny)[None, :, None] + s1(nx)[None, None, :]
    return a * lam + sigma


def fft_solve(b, shape, a, coeff, sigma=0.0, mean=0.0):
    nx, ny, nz = shape
    lam = symbol(shape, a, coeff, sigma)
    bh = np.fft.fftn(b.reshape(nz, ny, nx))
    with np.errstate(divide="ignore", invalid="ignore"):
        xh = np.where(np.abs(lam) > 1e-300, bh / lam, 0.0)
    if sigma == 0.0:
        xh[0, 0, 0] = mean * b.size
    return np.real(np.fft.ifftn(xh)).reshape(-1)


def dense_operator(shape, a, coeff):
    nx, ny, nz = shape
    n = nx * ny * nz
    M = np.zeros((n, n))
    w = [a * c for c in coeff]
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = i + j * nx + k * nx * ny
                M[r, r] += 3 * w[0]
                for o in range(1, len(coeff)):
                    for (ii, jj, kk) in ((i - o, j, k), (i + o, j, k), (i, j - o, k), (i, j + o, k),
                                         (i, j, k - o), (i, j, k + o)):
                        M[r, (ii % nx) + (jj % ny) * nx + (kk % nz) * nx * ny] += w[o]
    return M


# ---------------------------------------------------------------------------
# operator
# ---------------------------------------------------------------------------

def test_fd12_weights_closed_form(oracle):
    c = oracle.lap_coefficient(6)
    for p in range(7):
        assert c[p] == pytest.approx(float(C12[p]), rel=2e-16, abs=0)
        assert C12[p] == fd_weights_exact(6)[p]
    assert abs(c[0] + 2 * c[1:].sum()) < 1e-14          # zero row sum


@pytest.mark.parametrize("fdn,exact", [(1, [-2, 1]), (2, [Fraction(-5, 2), Fraction(4, 3), Fraction(-1, 12)])])
def test_low_order_weights(oracle, fdn, exact):
    c = oracle.lap_coefficient(fdn)
    assert np.allclose(c, [float(e) for e in exact], rtol=1e-15, atol=0)


def test_matvec_fourier_symbol(oracle):
    shape = (16, 12, 10)
    nx, ny, nz = shape
    coeff = oracle.lap_coefficient(6)
    a = -3.7
    k = np.arange(nz)[:, None, None]
    j = np.arange(ny)[None, :, None]
    i = np.arange(nx)[None, None, :]
    lam = symbol(shape, a, coeff)
    for (mx, my, mz) in [(1, 0, 0), (0, 2, 0), (0, 0, 3), (3, 5, 4), (8, 6, 5)]:
        x = np.cos(2 * np.pi * (mx * i / nx + my * j / ny + mz * k / nz)).reshape(-1)
        y = oracle.matvec(x, shape, a)
        assert np.allclose(y, lam[mz, my, mx] * x, rtol=0, atol=1e-12 * np.abs(lam).max())


def test_matvec_dense_bruteforce(oracle):
    shape = (8, 6, 7)                     # sides below 2*FDn+1: wraps overlap, as in the reference
    coeff = oracle.lap_coefficient(6)
    M = dense_operator(shape, -4.0, coeff)
    x = np.random.default_rng(3).standard_normal(M.shape[0])
    assert np.allclose(oracle.matvec(x, shape, -4.0), M @ x, rtol=0, atol=1e-11)
    assert np.allclose(M, M.T)                             # symmetric
    assert np.allclose(M.sum(axis=1), 0.0, atol=1e-12)     # zero row sums


def test_matvec_screened(oracle):
    shape = (12, 12, 12)
    x = np.random.default_rng(4).standard_normal(12 ** 3)
    y0 = oracle.matvec(x, shape, -4.0)
    y1 = oracle.matvec(x, shape, -4.0, sigma=0.5)
    assert np.allclose(y1, y0 + 0.5 * x, rtol=0, atol=1e-13)


# ---------------------------------------------------------------------------
# pseudo-inverse
# ---------------------------------------------------------------------------

def numpy_pinv_ref(A, b):
    """tools.c semantics from numpy: V diag(1/w) V^T b with w < n w_max eps zeroed."""
    U, w, Vt = np.linalg.svd(A)
    n = A.shape[0]
    w = w.copy()
    w[w < n * w.max() * 2.22044605e-16] = 0.0
    inv = np.where(w > 0, 1.0 / np.where(w > 0, w, 1.0), 0.0)
    V = Vt.T
    return V @ (inv * (V.T @ b)), w


@pytest.mark.parametrize("m", [1, 2, 4, 7, 8, 13, 20, 26, 32])
def test_pinv_matches_numpy_svd(oracle, m):
    rng = np.random.default_rng(10 + m)
    F = rng.standard_normal((40, m))
    A = F.T @ F
    b = rng.standard_normal(m)
    x, w, V = oracle.pinv_times_vec(A, b)
    xr, wr = numpy_pinv_ref(A, b)
    assert np.allclose(np.sort(w), np.sort(wr), rtol=1e-12, atol=0)
    assert np.allclose(x, xr, rtol=1e-9, atol=1e-12)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-8)
    assert np.allclose(V.T @ V, np.eye(m), atol=1e-12)


def test_pinv_rank_deficient(oracle):
    rng = np.random.default_rng(11)
    F = rng.standard_normal((40, 7))
    F[:, 5:] = 0.0                                   # never-written history columns
    A = F.T @ F
    b = F.T @ rng.standard_normal(40)
    x, w, _ = oracle.pinv_times_vec(A, b)
    assert (w == 0).sum() == 2
    assert np.allclose(x, np.linalg.pinv(A) @ b, rtol=1e-9, atol=1e-12)
    assert np.allclose(x[5:], 0.0, atol=1e-14)


def test_pinv_nonsymmetric_uses_v_only(oracle):
    """The reference replaces U by V (tools.c:324-329): for a non-symmetric
    matrix the result is (A^T A)^(-1/2) b, not pinv(A) b."""
    rng = np.random.default_rng(12)
    A = rng.standard_normal((5, 5)) + 5 * np.eye(5)
    b = rng.standard_normal(5)
    x, _, _ = oracle.pinv_times_vec(A, b)
    ev, Q = np.linalg.eigh(A.T @ A)
    assert np.allclose(x, Q @ ((Q.T @ b) / np.sqrt(ev)), rtol=1e-10)
    assert not np.allclose(x, np.linalg.solve(A, b), rtol=1e-3)


def test_pinv_golden_bitwise(oracle, golden):
    meta, g = golden
    ms, A, b, xg = g["pinv_m"], g["pinv_A"], g["pinv_b"], g["pinv_x"]
    ia = ib = 0
    for m in ms:
        x, _, _ = oracle.pinv_times_vec(A[ia:ia + m * m].reshape(m, m), b[ib:ib + m])
        assert np.array_equal(x, xg[ib:ib + m]), f"m={m}"
        ia += m * m
        ib += m


def test_anderson_golden_bitwise(oracle, golden):
    _, g = golden
    am, red, _ = oracle.anderson_extrapolation(g["and_DX"], g["and_DF"], g["and_f"], 0.6)
    assert np.array_equal(am, g["and_am"])
    assert np.array_equal(red, g["and_red"])


# ---------------------------------------------------------------------------
# solvers vs the reference's own outputs
# ---------------------------------------------------------------------------

def case_inputs(meta, name):
    c = meta["cases"][name]["params"]
    shape = tuple(c["shape"])
    n = int(np.prod(shape))
    if c["rhs"] == "uniform":
        b = P.rhs_uniform(shape, c["seed"])
    elif c["rhs"] == "normal":
        b = P.rhs_normal(shape, c["seed"])
    elif c["rhs"] == "zero":
        b = np.zeros(n)
    else:
        import os
        from conftest import GOLDEN_DIR
        b = np.load(os.path.join(GOLDEN_DIR, f"refrhs_{shape[0]}.npy"))
    x0 = np.full(n, float(c["x0"]))
    return c, shape, b, x0


def run_oracle(O, solver, c, shape, b, x0):
    if solver == "aar":
        return O.aar(x0, b, shape, c["a"], c["omega"], c["beta"], c["m"], c["p"], c["max_iter"],
                     c["tol"], sigma=c["sigma"], precond=bool(c["precond"]))
    if solver == "cg":
        return O.cg(x0, b, shape, c["a"], c["max_iter"], c["tol"], sigma=c["sigma"],
                    precond=bool(c["precond"]))
    return O.pgr(x0, b, shape, c["a"], c["omega"], c["m"], c["p"], c["max_iter"], c["tol"],
                 sigma=c["sigma"], precond=bool(c["precond"]), l2=(solver == "pl2r"))


SMALL_CASES = ["c1_16", "c1_24", "screened_24", "aniso_20x12x16", "maxiter13_16", "maxiter12_16",
               "p1_16", "m1_16", "m3p5_16", "noprec_16", "zero_rhs_16", "refdrv_16", "m16p3_16"]


@pytest.mark.parametrize("name", SMALL_CASES + ["c1_48", "refdrv_48", "readme_48"])
@pytest.mark.parametrize("solver", ["aar", "pgr", "pl2r", "cg"])
def test_oracle_bitwise_vs_reference(oracle, golden, name, solver):
    meta, g = golden
    c, shape, b, x0 = case_inputs(meta, name)
    assert float(np.sum(b)) == meta["cases"][name]["rhs_sum"]
    x, d = run_oracle(oracle, solver, c, shape, b, x0)
    ref = meta["cases"][name]["solvers"][solver]
    assert d["iterations"] == ref["iterations"]
    assert bool(d["printed_converged"]) == ref["converged_line"]
    assert bool(d["printed_warning"]) == ref["warning_line"]
    assert np.array_equal(d["trace"], g[f"{name}__{solver}__trace"])
    if f"{name}__{solver}__x" in g:
        assert np.array_equal(x, g[f"{name}__{solver}__x"])
    else:
        assert np.array_equal(x[::97], g[f"{name}__{solver}__xs"])
        assert float(np.linalg.norm(x)) == ref["x_norm"]
    if ref["relres_printed"] is not None and np.isfinite(ref["relres_printed"]):
        assert float(f"{d['relres']:g}") == ref["relres_printed"]


DEEP_CASES = ["m20p6_24", "m32p4_24", "m20p24_24", "m32p11_28x24x20"]


@pytest.mark.parametrize("name", DEEP_CASES)
@pytest.mark.parametrize("solver", ["aar", "pgr", "pl2r"])
def test_oracle_bitwise_vs_reference_deep(oracle, golden_deep, name, solver):
    """The deep histories (m = 20, 32; p below and above m + 1): counts,
    traces and solutions bit-identical to the reference build's."""
    meta, g = golden_deep
    c = meta["cases"][name]["params"]
    shape = tuple(c["shape"])
    b = P.rhs_uniform(shape, c["seed"])
    x, d = run_oracle(oracle, solver, c, shape, b, np.zeros(b.size))
    ref = meta["cases"][name]["solvers"][solver]
    assert d["iterations"] == ref["iterations"] and d["iterations"] > c["m"]
    assert bool(d["printed_converged"]) == ref["converged_line"]
    assert np.array_equal(d["trace"], g[f"{name}__{solver}__trace"])
    assert np.array_equal(x, g[f"{name}__{solver}__x"])


def test_pinv_deep_golden_bitwise(oracle, golden_deep):
    _, g = golden_deep
    ia = ib = 0
    for m in g["pinv_m"]:
        A = g["pinv_A"][ia:ia + m * m].reshape(m, m)
        x, _, _ = oracle.pinv_times_vec(A, g["pinv_b"][ib:ib + m])
        assert np.array_equal(x, g["pinv_x"][ib:ib + m]), m
        ia += m * m
        ib += m


# ---------------------------------------------------------------------------
# closed forms and invariants of the converged solutions
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("solver", ["aar", "pgr", "pl2r", "cg"])
@pytest.mark.parametrize("sigma", [0.0, 0.5])
def test_oracle_solution_vs_fft(oracle, solver, sigma):
    shape = (16, 16, 16)
    a = P.poisson_a(0.5)
    b = P.rhs_uniform(shape, 77) if sigma == 0.0 else P.rhs_normal(shape, 78)
    x0 = np.zeros(b.size)
    c = dict(a=a, omega=0.6, beta=0.6, m=7, p=6, max_iter=5000, tol=1e-11, sigma=sigma, precond=1)
    x, d = run_oracle(oracle, solver, c, shape, b, x0)
    xf = fft_solve(b, shape, a, oracle.lap_coefficient(6), sigma, mean=0.0)
    assert d["converged"]
    assert np.linalg.norm(x - xf) / np.linalg.norm(xf) < 1e-8
    if solver in ("aar", "pgr", "pl2r"):
        assert d["iterations"] % 6 == 0          # residual checked on projection steps only
    if sigma == 0.0:
        assert abs(x.mean()) < 1e-12 * np.abs(x).max()   # mean(x) = mean(x0) for periodic Poisson


# ---------------------------------------------------------------------------
# each check catches a planted mistake (mutations of the oracle's arithmetic,
# emulated in numpy on the CPU)
# ---------------------------------------------------------------------------

def _numpy_matvec(x, shape, a, coeff, drop=None, sign_flip=None):
    nx, ny, nz = shape
    X = x.reshape(nz, ny, nx)
    w = [a * c for c in coeff]
    y = X * 3 * w[0]
    for o in range(1, len(coeff)):
        terms = [np.roll(X, o, 2), np.roll(X, -o, 2), np.roll(X, o, 1), np.roll(X, -o, 1),
                 np.roll(X, o, 0), np.roll(X, -o, 0)]
        if drop is not None and o == drop[0]:
            terms[drop[1]] = 0 * terms[drop[1]]
        s = sum(terms)
        y = y + s * (-w[o] if sign_flip == o else w[o])
    return y.reshape(-1)


def test_symbol_check_catches_dropped_or_flipped_term(oracle):
    shape = (16, 12, 10)
    coeff = oracle.lap_coefficient(6)
    x = np.cos(2 * np.pi * 3 * np.arange(16) / 16)[None, None, :] * np.ones((10, 12, 1))
    x = x.reshape(-1)
    good = _numpy_matvec(x, shape, -3.7, coeff)
    assert np.allclose(good, oracle.matvec(x, shape, -3.7), atol=1e-12)
    lam = symbol(shape, -3.7, coeff)[0, 0, 3]
    for bad in (_numpy_matvec(x, shape, -3.7, coeff, drop=(2, 1)),
                _numpy_matvec(x, shape, -3.7, coeff, sign_flip=4)):
        assert not np.allclose(bad, lam * x, atol=1e-9)


def test_pinv_check_catches_transpose(oracle):
    rng = np.random.default_rng(5)
    F = rng.standard_normal((30, 6))
    A = F.T @ F
    b = rng.standard_normal(6)
    xr, _ = numpy_pinv_ref(A, b)
    U, w, Vt = np.linalg.svd(A)
    wrong = Vt @ ((U.T @ b) / w)                 # V^T in place of V
    assert not np.allclose(wrong, xr, rtol=1e-6)


def test_generator_numpy_vs_torch_bits():
    import torch
    u_np = P.uniform01(5000, 1234, 3)
    u_t = P.uniform01_torch(5000, 1234, 3, device="cpu").numpy()
    assert np.array_equal(u_np, u_t)
    assert 0.0 <= u_np.min() and u_np.max() < 1.0
    assert abs(u_np.mean() - 0.5) < 0.02
    assert not np.array_equal(u_np, P.uniform01(5000, 1234, 4))
    b = P.rhs_uniform((8, 8, 8), 1)
    assert abs(b.mean()) < 1e-16
    del torch


# ---------------------------------------------------------------------------
# the complex Helmholtz restatement: A u = a Lap u + (c + i eta) u
# ---------------------------------------------------------------------------

def fft_solve_complex(b, shape, a, coeff, c0, eta):
    nx, ny, nz = shape
    lam = symbol(shape, a, coeff) + c0 + 1j * eta
    return np.fft.ifftn(np.fft.fftn(b.reshape(nz, ny, nx)) / lam).reshape(-1)


def test_oracle_complex_matvec_symbol(oracle):
    """Every grid Fourier mode is an eigenvector with eigenvalue
    a*sym(theta) + c0 + i*eta (constant c); a variable c adds c(x) u
    pointwise."""
    shape = (12, 10, 8)
    nx, ny, nz = shape
    a, c0, eta = -4.0, 1.3, 0.5
    coeff = oracle.lap_coefficient(6)
    lam = symbol(shape, a, coeff) + c0 + 1j * eta
    rng = np.random.default_rng(3)
    for _ in range(6):
        kx, ky, kz = rng.integers(0, nx), rng.integers(0, ny), rng.integers(0, nz)
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        e = np.exp(2j * np.pi * (kx * i / nx + ky * j / ny + kz * k / nz)).reshape(-1)
        y = oracle.matvec_complex(e, shape, a, None, c0, eta)
        assert np.allclose(y, lam[kz, ky, kx] * e, rtol=0, atol=1e-12 * abs(lam).max())
    # variable c: A(c) u - A(0) u = c u
    u = rng.standard_normal(nx * ny * nz) + 1j * rng.standard_normal(nx * ny * nz)
    c = rng.uniform(0.5, 1.5, nx * ny * nz)
    d = oracle.matvec_complex(u, shape, a, c, 0.0, eta) - oracle.matvec_complex(u, shape, a, None, 0.0, eta)
    assert np.allclose(d, c * u, rtol=1e-12, atol=1e-12)


def test_oracle_complex_aar_vs_fft(oracle):
    """Constant c: the converged complex AAR solution is the FFT solution."""
    shape = (16, 16, 16)
    a, c0, eta = P.poisson_a(0.5), 1.0, 0.5
    b = P.rhs_complex_normal(shape, 1005)
    u, d = oracle.aar_complex(np.zeros_like(b), b, shape, a, None, c0, eta, 0.6, 0.6, 7, 6, 2000, 1e-12)
    assert d["converged"] and d["iterations"] % 6 == 0
    uf = fft_solve_complex(b, shape, a, oracle.lap_coefficient(6), c0, eta)
    assert np.linalg.norm(u - uf) / np.linalg.norm(uf) < 1e-10


def test_oracle_complex_reduces_to_real_aar(oracle):
    """eta = 0, c = sigma constant and a real b: the complex AAR is the real
    (reference-pinned) AAR on the screened operator -- the same iteration
    count and solution, with a zero imaginary part."""
    shape = (16, 16, 16)
    a, sigma = P.poisson_a(0.5), 0.5
    b = P.rhs_normal(shape, 78)
    xr, dr = oracle.aar(np.zeros(b.size), b, shape, a, 0.6, 0.6, 7, 6, 1000, 1e-8, sigma=sigma)
    u, dc = oracle.aar_complex(np.zeros(b.size, complex), b.astype(complex), shape, a, None, sigma, 0.0,
                               0.6, 0.6, 7, 6, 1000, 1e-8)
    assert dc["iterations"] == dr["iterations"]
    assert np.abs(u.imag).max() == 0.0
    assert np.linalg.norm(u.real - xr) / np.linalg.norm(xr) < 1e-10


def test_oracle_complex_adjoint_identity(oracle):
    """<A^H v, u> = <v, A u> for a variable c (A^H: eta -> -eta)."""
    shape = (8, 8, 8)
    rng = np.random.default_rng(5)
    n = 512
    u = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    c = rng.uniform(0.5, 1.5, n)
    Au = oracle.matvec_complex(u, shape, -4.0, c, 0.0, 0.5)
    AHv = oracle.matvec_complex(v, shape, -4.0, c, 0.0, -0.5)
    assert abs(np.vdot(AHv, u) - np.vdot(v, Au)) < 1e-10 * abs(np.vdot(v, Au))


@pytest.mark.parametrize("m", [2, 7, 14, 16, 26, 32])
def test_pinv_textbook_matches_numpy(oracle, m):
    """The textbook Golub-Reinsch variant (the complex solver's m x m solve)
    is the pseudo-inverse: V diag(1/w) V^T b against numpy's SVD on
    symmetric PSD and rank-deficient matrices."""
    rng = np.random.default_rng(m)
    B = rng.standard_normal((m, m + 3))
    for A in (B @ B.T, (B[:, : m // 2 + 1] @ B[:, : m // 2 + 1].T)):
        b = rng.standard_normal(m)
        x, w, V = oracle.pinv_times_vec(A, b, textbook=True)
        U, S, Vt = np.linalg.svd(A)
        keep = S >= m * S[0] * 2.22044605e-16
        xr = Vt.T[:, keep] @ ((Vt[keep] @ b) / S[keep])
        assert np.linalg.norm(x - xr) <= 1e-10 * np.linalg.norm(xr)


def test_pinv_textbook_converges_where_reference_order_does_not(oracle):
    """tests/golden/pinv_embedded_*.npy: the embedded Hermitian Anderson
    system on which the tools.c:137 order (left reflector from column 1)
    does not converge in 250 sweeps; the textbook order solves it."""
    import os
    from conftest import GOLDEN_DIR
    A = np.load(os.path.join(GOLDEN_DIR, "pinv_embedded_A.npy"))
    r = np.load(os.path.join(GOLDEN_DIR, "pinv_embedded_r.npy"))
    with pytest.raises(RuntimeError):
        oracle.pinv_times_vec(A, r)
    x, w, V = oracle.pinv_times_vec(A, r, textbook=True)
    U, S, Vt = np.linalg.svd(A)
    keep = S >= 14 * S[0] * 2.22044605e-16
    xr = Vt.T[:, keep] @ ((Vt[keep] @ r) / S[keep])
    assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)


# ---------------------------------------------------------------------------
# variable-coefficient operator a*Lap + c(x) (extension)
# ---------------------------------------------------------------------------

def test_vc_constant_field_is_the_screened_oracle(oracle):
    """c = sigma everywhere: orc_matvec_vc and orc_aar_vc equal the
    (reference-pinned) screened orc_matvec / orc_aar bit for bit -- solution,
    residual trace and iteration count, with and without Jacobi."""
    shape = (16, 12, 10)
    n = int(np.prod(shape))
    a, sigma = P.poisson_a(0.5), 0.75
    c = np.full(n, sigma)
    x = np.random.default_rng(5).standard_normal(n)
    assert np.array_equal(oracle.matvec_vc(x, shape, a, c), oracle.matvec(x, shape, a, sigma))
    b = P.rhs_normal(shape, 81)
    for precond, om, mi in ((True, 0.6, 600), (False, 0.015, 300)):
        xs, ds = oracle.aar(np.zeros(n), b, shape, a, om, om, 7, 6, mi, 1e-9, sigma=sigma, precond=precond)
        xv, dv = oracle.aar_vc(np.zeros(n), b, shape, a, c, om, om, 7, 6, mi, 1e-9, precond=precond)
        assert dv["iterations"] == ds["iterations"]
        assert np.array_equal(xv, xs)
        assert np.array_equal(dv["trace"], ds["trace"])


def test_vc_matvec_dense_bruteforce(oracle):
    """a*Lap + diag(c) as a dense matrix (periodic wraps overlapping at sides
    below 13) against orc_matvec_vc; a dropped or misplaced c term fails."""
    shape = (8, 6, 7)
    n = int(np.prod(shape))
    coeff = oracle.lap_coefficient(6)
    c = P.helmholtz_coefficient(shape, 9)
    M = dense_operator(shape, -4.0, coeff) + np.diag(c)
    x = np.random.default_rng(6).standard_normal(n)
    y = oracle.matvec_vc(x, shape, -4.0, c)
    assert np.allclose(y, M @ x, rtol=0, atol=1e-11)
    assert not np.allclose(y, M @ x - c * x, rtol=0, atol=1e-3)
    assert not np.allclose(y, M @ x + np.roll(c, 1) * x - c * x, rtol=0, atol=1e-3)


@pytest.mark.parametrize("precond", [True, False])
def test_vc_aar_solution_dense(oracle, precond):
    """orc_aar_vc's converged solution against numpy's dense solve of the
    same operator (a 9 x 8 x 7 grid, c(x) in [0.5, 1.5])."""
    shape = (9, 8, 7)
    n = int(np.prod(shape))
    a = P.poisson_a(0.5)
    c = P.helmholtz_coefficient(shape, 10)
    M = dense_operator(shape, a, oracle.lap_coefficient(6)) + np.diag(c)
    b = P.rhs_normal(shape, 82)
    xd = np.linalg.solve(M, b)
    om, mi = (0.6, 3000) if precond else (0.015, 20000)
    x, d = oracle.aar_vc(np.zeros(n), b, shape, a, c, om, om, 7, 6, mi, 1e-12, precond=precond)
    assert d["converged"]
    assert d["iterations"] % 6 == 0
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9


def test_vc_real_oracle_vs_complex_oracle(oracle):
    """The real variable-coefficient AAR against the complex oracle with
    eta = 0 on the same c(x) and a real b: the same converged solution and a
    zero imaginary part (the two differ only in the order of their SVD)."""
    shape = (16, 16, 16)
    n = int(np.prod(shape))
    a = P.poisson_a(0.5)
    c = P.helmholtz_coefficient(shape, 11)
    b = P.rhs_normal(shape, 83)
    x, d = oracle.aar_vc(np.zeros(n), b, shape, a, c, 0.6, 0.6, 7, 6, 2000, 1e-12)
    u, dc = oracle.aar_complex(np.zeros(n, complex), b.astype(complex), shape, a, c, 0.0, 0.0,
                               0.6, 0.6, 7, 6, 2000, 1e-12)
    assert d["converged"] and dc["converged"]
    assert np.abs(u.imag).max() == 0.0
    assert np.linalg.norm(u.real - x) / np.linalg.norm(x) < 1e-9


def test_device_coefficient_generator_matches_host():
    """problems.helmholtz_coefficient_torch (the bench's 512^3 coefficient)
    is the host generator's formula."""
    import torch
    shape = (20, 12, 16)
    h = P.helmholtz_coefficient(shape, 1010)
    d = P.helmholtz_coefficient_torch(shape, 1010, "cpu").numpy()
    assert np.allclose(d, h, rtol=0, atol=1e-13)


@pytest.mark.parametrize("solver", ["pgr", "pl2r", "cg"])
def test_vc_pgr_cg_constant_field_is_the_screened_oracle(oracle, solver):
    """orc_pgr_vc / orc_cg_vc with c = sigma everywhere equal the
    reference-pinned screened orc_pgr / orc_cg bit for bit."""
    shape = (16, 12, 10)
    n = int(np.prod(shape))
    a, sigma = P.poisson_a(0.5), 0.75
    c = np.full(n, sigma)
    b = P.rhs_normal(shape, 84)
    for precond, om in ((True, 0.6), (False, 0.015)):
        if solver == "cg":
            xs, ds = oracle.cg(np.zeros(n), b, shape, a, 500, 1e-9, sigma=sigma, precond=precond)
            xv, dv = oracle.cg_vc(np.zeros(n), b, shape, a, c, 500, 1e-9, precond=precond)
        else:
            l2 = solver == "pl2r"
            xs, ds = oracle.pgr(np.zeros(n), b, shape, a, om, 7, 6, 400, 1e-9, sigma=sigma, precond=precond, l2=l2)
            xv, dv = oracle.pgr_vc(np.zeros(n), b, shape, a, c, om, 7, 6, 400, 1e-9, precond=precond, l2=l2)
        assert dv["iterations"] == ds["iterations"]
        assert np.array_equal(xv, xs)
        assert np.array_equal(dv["trace"], ds["trace"])


@pytest.mark.parametrize("solver", ["pgr", "pl2r", "cg"])
def test_vc_pgr_cg_solution_dense(oracle, solver):
    """orc_pgr_vc / orc_cg_vc converged solutions against numpy's dense solve
    (9 x 8 x 7 grid, c(x) in [0.5, 1.5])."""
    shape = (9, 8, 7)
    n = int(np.prod(shape))
    a = P.poisson_a(0.5)
    c = P.helmholtz_coefficient(shape, 12)
    M = dense_operator(shape, a, oracle.lap_coefficient(6)) + np.diag(c)
    b = P.rhs_normal(shape, 85)
    xd = np.linalg.solve(M, b)
    if solver == "cg":
        x, d = oracle.cg_vc(np.zeros(n), b, shape, a, c, 3000, 1e-12)
    else:
        x, d = oracle.pgr_vc(np.zeros(n), b, shape, a, c, 0.6, 7, 6, 3000, 1e-12, l2=solver == "pl2r")
    assert d["converged"]
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9


def _edge():
    import json
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    with open(os.path.join(here, "golden_edge.json")) as fp:
        meta = json.load(fp)
    return meta, np.load(os.path.join(here, "golden_edge.npz"), allow_pickle=False)


@pytest.mark.parametrize("fdn", [1, 2, 3, 4, 5, 6])
def test_oracle_matvec_fdn_matches_reference(oracle, fdn):
    """The oracle's stencil at every FD order the reference computes
    weights for (FDn = 1..6, system.c:480-498, loop of system.c:195-201) is
    bit-identical to the reference build's."""
    meta, g = _edge()
    shape = tuple(meta["fdn"]["matvec_shape"])
    y = oracle.matvec(g["fdn_matvec_x"], shape, -4.0, fdn=fdn)
    assert np.array_equal(y, g[f"fdn{fdn}_matvec_y"])


def test_oracle_matvec_user_weights_match_reference(oracle):
    """A caller's own weights in coeff_lap (the reference's extension point):
    the oracle's stencil with them equals the reference build's."""
    meta, g = _edge()
    shape = tuple(meta["fdn"]["matvec_shape"])
    c = g["user_coeff"]
    y = oracle.matvec(g["fdn_matvec_x"], shape, -4.0, fdn=c.size - 1, coeff=c)
    assert np.array_equal(y, g["user_matvec_y"])


@pytest.mark.parametrize("label,fdn", [("fdn2", 2), ("fdn4", 4)])
def test_oracle_solvers_fdn_match_reference(oracle, label, fdn):
    """AAR, PGR, PL2R and CG at FD orders 4 and 8: the reference's iteration
    counts and solutions (within rounding)."""
    meta, g = _edge()
    shape = tuple(meta["fdn"]["solve_shape"])
    b = g["fdn_solve_b"]
    x0 = np.zeros(b.size)
    for sv in ("aar", "pgr", "pl2r", "cg"):
        ref = meta["fdn"]["solves"][f"{label}_{sv}"]
        if sv == "aar":
            x, d = oracle.aar(x0, b, shape, -4.0, 0.6, 0.6, 7, 6, 1000, 1e-8, fdn=fdn, trace=False)
        elif sv == "cg":
            x, d = oracle.cg(x0, b, shape, -4.0, 1000, 1e-8, fdn=fdn, trace=False)
        else:
            x, d = oracle.pgr(x0, b, shape, -4.0, 0.6, 7, 6, 1000, 1e-8, l2=sv == "pl2r", fdn=fdn, trace=False)
        assert d["iterations"] == ref["iterations"], sv
        xr = g[f"{label}_{sv}_x"]
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-10, sv
