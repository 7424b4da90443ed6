"""Seeded synthetic problems for the FD12 Poisson / screened-Poisson /
Helmholtz systems.

The reference seeds its RHS with glibc ``srand(global_index + 1)`` so that
the problem does not depend on the rank count (MPI_REAL/system.c:452-460).
This module keeps that property with a counter-based generator of its own:
value ``i`` of stream ``s`` under seed ``S`` is

    key = mix64(S * G + s),   u_i = (mix64(key + (i + 1) * G) >> 11) * 2**-53

where ``mix64`` is the SplitMix64 finaliser (Steele, Lea, Flood, OOPSLA
2014) and G = 0x9E3779B97F4A7C15.  The same bits come out of the numpy and
the torch implementations, so a RHS can be generated directly in HBM.
Normal variates use Box-Muller on the pair (u_{2i}, u_{2i+1}).

Layout is x fastest: ``idx = i + j*nx + k*nx*ny`` (MPI_REAL/system.c:82),
i.e. a C-ordered array of shape (nz, ny, nx).
"""
from __future__ import annotations

import math

import numpy as np

GOLDEN = 0x9E3779B97F4A7C15
M1 = 0xBF58476D1CE4E5B9
M2 = 0x94D049BB133111EB
_MASK64 = (1 << 64) - 1


def _mix64_int(z: int) -> int:
    z &= _MASK64
    z = ((z ^ (z >> 30)) * M1) & _MASK64
    z = ((z ^ (z >> 27)) * M2) & _MASK64
    return z ^ (z >> 31)


def stream_key(seed: int, stream: int = 0) -> int:
    return _mix64_int((seed * GOLDEN + stream) & _MASK64)


def _mix64_np(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(M1)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(M2)
    return z ^ (z >> np.uint64(31))


def uniform01(n: int, seed: int, stream: int = 0, offset: int = 0, chunk: int = 1 << 24):
    """n doubles in [0, 1), counter-based (numpy)."""
    key = np.uint64(stream_key(seed, stream))
    out = np.empty(n, dtype=np.float64)
    with np.errstate(over="ignore"):
        for s in range(0, n, chunk):
            e = min(n, s + chunk)
            idx = np.arange(offset + s + 1, offset + e + 1, dtype=np.uint64)
            z = _mix64_np(key + idx * np.uint64(GOLDEN))
            out[s:e] = (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return out


def normal(n: int, seed: int, stream: int = 0):
    """n standard normals by Box-Muller (numpy)."""
    u = uniform01(2 * n, seed, stream)
    u1 = 1.0 - u[0::2]            # (0, 1]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


# ---------------------------------------------------------------------------
# torch (device) versions: identical bits to the numpy ones
# ---------------------------------------------------------------------------

def _signed(v: int) -> int:
    v &= _MASK64
    return v - (1 << 64) if v >= (1 << 63) else v


def _srl(torch, z, s: int):
    # logical right shift on int64 two's-complement values
    return (z >> s) & ((1 << (64 - s)) - 1)


def uniform01_torch(n: int, seed: int, stream: int = 0, device="cuda", chunk: int = 1 << 26):
    import torch
    key = _signed(stream_key(seed, stream))
    out = torch.empty(n, dtype=torch.float64, device=device)
    g, m1, m2 = _signed(GOLDEN), _signed(M1), _signed(M2)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        z = torch.arange(s + 1, e + 1, dtype=torch.int64, device=device) * g + key
        z = (z ^ _srl(torch, z, 30)) * m1
        z = (z ^ _srl(torch, z, 27)) * m2
        z = z ^ _srl(torch, z, 31)
        out[s:e] = _srl(torch, z, 11).to(torch.float64) * (2.0 ** -53)
    return out


def normal_torch(n: int, seed: int, stream: int = 0, device="cuda"):
    import torch
    u = uniform01_torch(2 * n, seed, stream, device)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    return torch.sqrt(-2.0 * torch.log(u1)) * torch.cos(2.0 * math.pi * u2)


# ---------------------------------------------------------------------------
# the workloads
# ---------------------------------------------------------------------------

def poisson_a(h: float) -> float:
    """Scale a of y = a*Lap(x) for -Laplace with spacing h (a = -1/h^2)."""
    return -1.0 / (h * h)


def rhs_uniform(shape, seed, stream=0, lo=-0.5, hi=0.5, remove_mean=True):
    """b_i ~ U[lo, hi] i.i.d.; mean removed for the singular periodic
    Poisson system (as MPI_REAL/system.c:462-466 does for its RHS)."""
    n = int(np.prod(shape))
    b = lo + (hi - lo) * uniform01(n, seed, stream)
    if remove_mean:
        b -= b.sum() / n
    return b


def rhs_uniform_torch(shape, seed, stream=0, lo=-0.5, hi=0.5, remove_mean=True, device="cuda"):
    n = int(np.prod(shape))
    b = lo + (hi - lo) * uniform01_torch(n, seed, stream, device)
    if remove_mean:
        b -= b.sum() / n
    return b


def rhs_normal(shape, seed, stream=0):
    return normal(int(np.prod(shape)), seed, stream)


def rhs_normal_torch(shape, seed, stream=0, device="cuda"):
    """Device version of rhs_normal (same uniforms; log/cos of the device
    library, so equal to the numpy values to rounding, not bitwise)."""
    return normal_torch(int(np.prod(shape)), seed, stream, device)


def rhs_complex_normal(shape, seed, stream=0):
    """complex128 b with real and imaginary parts i.i.d. N(0,1)."""
    n = int(np.prod(shape))
    v = normal(2 * n, seed, stream)
    return (v[0::2] + 1j * v[1::2]).astype(np.complex128)


def helmholtz_coefficient(shape, seed, modes=8, stream=100):
    """c(x) = 1 + 0.5 * (sum of `modes` seeded random Fourier modes,
    normalised to max |.| = 1) on the periodic unit grid."""
    nx, ny, nz = shape
    u = uniform01(5 * modes, seed, stream)
    k = np.arange(nz)[:, None, None] / nz
    j = np.arange(ny)[None, :, None] / ny
    i = np.arange(nx)[None, None, :] / nx
    acc = np.zeros((nz, ny, nx))
    for q in range(modes):
        kx, ky, kz = (1 + int(u[5 * q + r] * 4) for r in range(3))
        amp = 2.0 * u[5 * q + 3] - 1.0
        ph = 2.0 * math.pi * u[5 * q + 4]
        acc += amp * np.cos(2.0 * math.pi * (kx * i + ky * j + kz * k) + ph)
    acc /= np.abs(acc).max()
    return (1.0 + 0.5 * acc).reshape(-1)


def helmholtz_coefficient_torch(shape, seed, device, modes=8, stream=100):
    """helmholtz_coefficient evaluated on the device (float64; the same
    modes, amplitudes and phases, cosines from the device's libm), for
    grids where the host evaluation is slow."""
    import torch
    nx, ny, nz = shape
    u = uniform01(5 * modes, seed, stream)
    k = torch.arange(nz, dtype=torch.float64, device=device)[:, None, None] / nz
    j = torch.arange(ny, dtype=torch.float64, device=device)[None, :, None] / ny
    i = torch.arange(nx, dtype=torch.float64, device=device)[None, None, :] / nx
    acc = torch.zeros((nz, ny, nx), dtype=torch.float64, device=device)
    for q in range(modes):
        kx, ky, kz = (1 + int(u[5 * q + r] * 4) for r in range(3))
        amp = 2.0 * float(u[5 * q + 3]) - 1.0
        ph = 2.0 * math.pi * float(u[5 * q + 4])
        acc += amp * torch.cos(2.0 * math.pi * (kx * i + ky * j + kz * k) + ph)
    acc /= acc.abs().max()
    return (1.0 + 0.5 * acc).reshape(-1)


def lap_weights(fdn=6):
    """1D FD weights of order 2*fdn (MPI_REAL/system.c:480-498): c0 =
    -2 sum 1/p^2, c_p = 2 (-1)^(p+1) prod_{fdn-p+1..fdn} i / prod_{fdn+1..fdn+p} i
    / p^2, evaluated in that order (so the doubles equal the reference's)."""
    c = [0.0] * (fdn + 1)
    for p in range(1, fdn + 1):
        c[0] += -(2.0 / (p * p))
    for p in range(1, fdn + 1):
        num = den = 1.0
        for i in range(fdn - p + 1, fdn + 1):
            num *= i
        for i in range(fdn + 1, fdn + p + 1):
            den *= i
        c[p] = 2.0 * (1.0 if p % 2 else -1.0) * (num / den) / (p * p)
    return c


def spectral_solve_torch(b, shape, a, sigma=0.0, fdn=6):
    """The exact solution of (a*Lap_FD + sigma) u = b on the periodic grid,
    by a device FFT: every grid Fourier mode is an eigenvector of the
    operator, with eigenvalue a*(3 c0 + sum_d 2 sum_p c_p cos(p theta_d)) +
    sigma.  For sigma = 0 the zero mode is dropped (the solution of zero
    mean; b must then have zero mean).  b: float64 device tensor."""
    import torch
    nx, ny, nz = (int(s) for s in shape)
    coeff = lap_weights(fdn)
    dev = b.device

    def sym(n, half=False):
        th = 2.0 * math.pi * torch.arange(n // 2 + 1 if half else n, dtype=torch.float64, device=dev) / n
        s = torch.zeros_like(th)
        for p in range(1, fdn + 1):
            s += 2.0 * coeff[p] * torch.cos(p * th)
        return s

    sx, sy, sz = sym(nx, True), sym(ny), sym(nz)
    lam = a * (3.0 * coeff[0] + sx[None, None, :] + sy[None, :, None] + sz[:, None, None]) + sigma
    B = torch.fft.rfftn(b.reshape(nz, ny, nx))
    if sigma == 0.0:
        lam[0, 0, 0] = 1.0
        B[0, 0, 0] = 0.0
    B /= lam
    del lam
    return torch.fft.irfftn(B, s=(nz, ny, nx)).reshape(-1)


def batch_instances(seed=1004):
    """The uneven batch: 96 screened-Poisson instances whose cubic side is
    drawn from {24, 32, 48, 64, 96, 128} and sigma_i ~ U[0.1, 2], then 16
    right-hand sides sharing one 128^3 operator (one sigma).  Returns
    [(shape, sigma, rhs_stream)]; instance k's RHS is N(0,1) from stream
    100 + k of `seed`."""
    sides = (24, 32, 48, 64, 96, 128)
    u = uniform01(3 * 96 + 1, seed, 7)
    inst = []
    for i in range(96):
        side = sides[min(5, int(u[2 * i] * 6))]
        sigma = 0.1 + 1.9 * float(u[2 * i + 1])
        inst.append(((side,) * 3, sigma, 100 + i))
    sig16 = 0.1 + 1.9 * float(u[-1])
    for r in range(16):
        inst.append(((128,) * 3, sig16, 100 + 96 + r))
    return inst


# Parameter sets of the measured configurations (FD12, m=7, p=6 unless said).
C1 = dict(shape=(48, 48, 48), h=0.5, omega=0.6, beta=0.6, m=7, p=6, tol=1e-6, max_iter=1000,
          seed=1001)
C2 = dict(shape=(512, 512, 512), h=0.25, omega=0.6, beta=0.6, m=7, p=6, tol=1e-6,
          max_iter=20000, seed=1002)
