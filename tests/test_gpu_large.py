"""The largest grids one MI355X holds: 1024 x 1024 x 1000 FP64 (1.05 G points,
8.4 GB per vector, about 110 GB of AAR state).  Plane offsets come within a
factor of two of 2^31 here, so this is where 32-bit index arithmetic would
break.  Checked through the C ABI:

* the matvec at 2000 sampled points against the stencil evaluated on the host
  from the same input (the operator of system.c:189-205 with periodic wraps);
* 12 AAR iterations (two Anderson cycles): the residual the solver reports is
  ||b - A x|| of the x it returns (recomputed by a separate matvec);
* a grid of more than 2^31 points is refused with AAR_HIP_ERR_ARG.
"""
import ctypes

import numpy as np
import pytest

from aar_hip import problems as P

pytestmark = pytest.mark.gpu

SHAPE = (1024, 1024, 1000)


def say(*a):
    print(*a, flush=True)


def _host_stencil(x, shape, a, coeff, pts):
    """a*Lap(x) at flat indices pts, the reference order of system.c:195-201."""
    nx, ny, nz = shape
    i = pts % nx
    j = (pts // nx) % ny
    k = pts // (nx * ny)
    w = [a * c for c in coeff]

    def at(ii, jj, kk):
        return x[(ii % nx) + (jj % ny) * nx + (kk % nz) * nx * ny]

    y = at(i, j, k) * 3 * w[0]
    for o in range(1, 7):
        nb = at(i - o, j, k) + at(i + o, j, k)
        nb = nb + at(i, j - o, k)
        nb = nb + at(i, j + o, k)
        nb = nb + at(i, j, k - o)
        nb = nb + at(i, j, k + o)
        y = y + nb * w[o]
    return y


def test_large_grid_matvec_and_aar(gpu, oracle):
    import torch
    nx, ny, nz = SHAPE
    n = nx * ny * nz
    a = P.poisson_a(1.0 / 1024)
    say("large grid", SHAPE)
    b = P.rhs_uniform_torch(SHAPE, 1012, device="cuda")
    y = gpu.Lap_Vec_mult(b, SHAPE, a)
    rng = np.random.default_rng(13)
    pts = np.concatenate([rng.integers(0, n, 1990), [0, n - 1, nx - 1, nx * ny - 1, n - nx * ny,
                                                     nx * ny * (nz // 2), 5, n - 6, nx * 7 + 3, n // 2]])
    # the host evaluation needs x at the sampled points and their 36 neighbours
    nxy = nx * ny
    i, j, k = pts % nx, (pts // nx) % ny, pts // nxy
    need = [pts]
    for o in range(1, 7):
        need += [((i - o) % nx) + j * nx + k * nxy, ((i + o) % nx) + j * nx + k * nxy,
                 i + ((j - o) % ny) * nx + k * nxy, i + ((j + o) % ny) * nx + k * nxy,
                 i + j * nx + ((k - o) % nz) * nxy, i + j * nx + ((k + o) % nz) * nxy]
    idx = np.unique(np.concatenate(need))
    vals = b[torch.from_numpy(idx).to("cuda")].cpu().numpy()
    lookup = dict(zip(idx.tolist(), vals.tolist()))

    class Sparse:
        def __getitem__(self, q):
            return np.array([lookup[int(t)] for t in np.atleast_1d(q)])

    yh = _host_stencil(Sparse(), SHAPE, a, oracle.lap_coefficient(6), pts)
    yg = y[torch.from_numpy(pts).to("cuda")].cpu().numpy()
    assert np.array_equal(yg, yh)
    del y
    torch.cuda.empty_cache()

    x = torch.zeros_like(b)
    st = gpu.AAR(x, b, SHAPE, a, 0.6, 0.6, 7, 6, 12, 1e-30)
    say("large aar", st["iterations"], st["relres"], st["solve_ms"])
    assert st["iterations"] == 12 and np.isfinite(st["relres"])
    r = b - gpu.Lap_Vec_mult(x, SHAPE, a)
    rel = float(torch.linalg.norm(r) / torch.linalg.norm(b))
    assert abs(rel - st["relres"]) <= 1e-9 * rel
    assert float(x.mean()) == pytest.approx(0.0, abs=1e-12)      # the periodic mean is preserved
    del x, b, r
    gpu.release()                      # the library's 1024^3 rings (about 90 GB)
    torch.cuda.empty_cache()


def test_grid_beyond_32bit_points_refused(gpu):
    from aar_hip import _lib
    L = _lib.load()
    g = _lib.Grid(1300, 1300, 1300, 6, -1.0, 0.0)
    rc = L.aar_hip_matvec(ctypes.byref(g), None, None)
    assert rc == 1
    assert b"too large" in L.aar_hip_last_error()
