"""The device operator at every FD order the reference computes weights for
(FDn = 1..6, MPI_REAL/system.c:480-498 and the stencil loop of
system.c:195-201), and with a caller's own weights -- the reference's
extension point behind the Lap_Vec_mult / Precondition slots (README:25):
against the reference build's own outputs (tests/golden/golden_edge.*).

* the stencil, bit for bit, on a ragged grid;
* AAR, PGR, PL2R and CG at FD4, FD8 and with the user weights: the
  reference's iteration counts, solutions within 1e-8;
* the same user weights through the reference-signature entry points (the
  caller's POISSON holds them in coeff_lap), and registered for the device
  API with aar_hip_set_stencil;
* ILU(0) refuses a stencil other than FD12.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import dropin

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def edge():
    with open(os.path.join(HERE, "golden", "golden_edge.json")) as fp:
        meta = json.load(fp)
    return meta, np.load(os.path.join(HERE, "golden", "golden_edge.npz"), allow_pickle=False)


def dev(v):
    import torch
    return torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to("cuda")


@pytest.mark.parametrize("fdn", [1, 2, 3, 4, 5, 6])
def test_matvec_fdn_bitwise(gpu, oracle, edge, fdn):
    meta, g = edge
    shape = tuple(meta["fdn"]["matvec_shape"])
    x = g["fdn_matvec_x"]
    y = gpu.Lap_Vec_mult(dev(x), shape, -4.0, fdn=fdn).cpu().numpy()
    assert np.array_equal(y, g[f"fdn{fdn}_matvec_y"])
    assert np.array_equal(y, oracle.matvec(x, shape, -4.0, fdn=fdn))


def test_matvec_user_weights_bitwise(gpu, edge):
    meta, g = edge
    shape = tuple(meta["fdn"]["matvec_shape"])
    x = g["fdn_matvec_x"]
    gpu.set_stencil(g["user_coeff"])
    try:
        y = gpu.Lap_Vec_mult(dev(x), shape, -4.0, fdn=gpu.FDN_USER).cpu().numpy()
    finally:
        gpu.set_stencil(None)
    assert np.array_equal(y, g["user_matvec_y"])


def _solve(gpu, sv, x, b, shape, fdn):
    if sv == "aar":
        return gpu.AAR(x, b, shape, -4.0, 0.6, 0.6, 7, 6, 1000, 1e-8, fdn=fdn)
    if sv == "pgr":
        return gpu.PGR(x, b, shape, -4.0, 0.6, 7, 6, 1000, 1e-8, fdn=fdn)
    if sv == "pl2r":
        return gpu.PL2R(x, b, shape, -4.0, 0.6, 7, 6, 1000, 1e-8, fdn=fdn)
    return gpu.CG(x, b, shape, -4.0, 1000, 1e-8, fdn=fdn)


CASES = [(lab, sv) for lab in ("fdn2", "fdn4", "user") for sv in ("aar", "pgr", "pl2r", "cg")]


@pytest.mark.parametrize("label,solver", CASES, ids=[f"{a}-{b}" for a, b in CASES])
def test_solvers_fdn_vs_reference(gpu, edge, label, solver):
    meta, g = edge
    shape = tuple(meta["fdn"]["solve_shape"])
    b = dev(g["fdn_solve_b"])
    x = dev(np.zeros(b.numel()))
    if label == "user":
        gpu.set_stencil(g["user_coeff"])
        fdn = gpu.FDN_USER
    else:
        fdn = int(label[3:])
    try:
        st = _solve(gpu, solver, x, b, shape, fdn)
    finally:
        gpu.set_stencil(None)
    ref = meta["fdn"]["solves"][f"{label}_{solver}"]
    xr = g[f"{label}_{solver}_x"]
    rel = np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr)
    say(f"{label} {solver}: gpu {st['iterations']} reference {ref['iterations']} rel diff {rel:.2e}")
    assert st["iterations"] == ref["iterations"]
    assert bool(st["printed_converged"]) == ref["converged_line"]
    assert rel <= 1e-8


def test_user_weights_through_reference_signature(gpu, edge):
    """The reference's own way to change the operator: the caller's POISSON
    carries FDn and coeff_lap.  Lap_Vec_mult (host arrays) is bit-exact and
    AAR() / CG() give the reference's counts with those weights."""
    from aar_hip import _lib
    meta, g = edge
    L = _lib.load()
    fp = dropin.bind(L)
    c = np.ascontiguousarray(g["user_coeff"])
    shape = tuple(meta["fdn"]["matvec_shape"])
    sysd = dropin.system(shape, c)
    sysd.FDn = c.size - 1
    x = np.ascontiguousarray(g["fdn_matvec_x"])
    y = np.full(x.size, np.nan)
    L.Lap_Vec_mult(ctypes.byref(sysd), -4.0, x.ctypes.data, y.ctypes.data, 0)
    assert np.array_equal(y, g["user_matvec_y"])
    shape = tuple(meta["fdn"]["solve_shape"])
    sysd = dropin.system(shape, c)
    sysd.FDn = c.size - 1
    n = int(np.prod(shape))
    b = np.ascontiguousarray(g["fdn_solve_b"]).copy()
    st = _lib.Stats()
    for sv in ("aar", "cg"):
        xs = np.zeros(n)
        if sv == "aar":
            L.AAR(ctypes.byref(sysd), fp["Lap_Vec_mult"], fp["Precondition"], -4.0, xs.ctypes.data, b.ctypes.data,
                  0.6, 0.6, 7, 6, 1000, 1e-8, n, 0)
        else:
            L.CG(ctypes.byref(sysd), fp["Lap_Vec_mult"], fp["Precondition"], -4.0, xs.ctypes.data, b.ctypes.data,
                 1000, 1e-8, n, 0)
        L.aar_hip_last_stats(ctypes.byref(st))
        ref = meta["fdn"]["solves"][f"user_{sv}"]
        xr = g[f"user_{sv}_x"]
        assert st.iterations == ref["iterations"], sv
        assert np.linalg.norm(xs - xr) / np.linalg.norm(xr) <= 1e-8, sv


def test_stencil_order_limits(gpu):
    """fdn beyond the sweep's radius, an unregistered user stencil and ILU(0)
    on a stencil other than FD12 are refused with an argument error."""
    import torch
    from aar_hip import _lib
    shape = (16, 16, 16)
    x = torch.zeros(16 ** 3, dtype=torch.float64, device="cuda")
    for fdn in (0, 7, gpu.FDN_USER):
        with pytest.raises(_lib.AarHipError):
            gpu.Lap_Vec_mult(x, shape, -4.0, fdn=fdn)
    L = _lib.load()
    grid = _lib.Grid(16, 16, 16, 4, -4.0, 0.0)
    z = torch.zeros_like(x)
    rc = L.aar_hip_ilu0_apply(ctypes.byref(grid), None, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(z.data_ptr()), 0)
    assert rc == 1 and b"FD12" in L.aar_hip_last_error()
