"""The reference-signature entry points of libaar_hip on host arrays, each
against the reference's own output (tests/golden, written by the reference
build) or the oracle:

* Vector2Norm / VectorDotProduct (MPI_REAL/tools.c:18-43);
* pythag, SingularValueDecomp, PseudoInverseTimesVec (tools.c:52-352);
* Galerkin_Richardson (PGR.h:24-25) and L2_Richardson (PL2R.h:24-25);
* the host-array Lap_Vec_mult and Precondition (system.h:82-84);
* non-finite inputs through AAR(), PGR(), PL2R() and CG(): the reference's
  printed lines, its process exit (tools.c:270) and the returned x.
"""
import json
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import dropin
from aar_hip import problems as P

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def L(gpu):
    from aar_hip import _lib
    lib = _lib.load()
    dropin.bind(lib)
    return lib


@pytest.fixture(scope="module")
def edge():
    with open(os.path.join(HERE, "golden", "golden_edge.json")) as fp:
        meta = json.load(fp)
    return meta, np.load(os.path.join(HERE, "golden", "golden_edge.npz"), allow_pickle=False)


@pytest.mark.parametrize("n", [0, 1, 1000, 300001])
def test_vector_norm_and_dot(L, oracle, n):
    """tools.c:18-43: the sums (in a fixed order on the GPU) against the
    oracle's sequential sums, to rounding of a different summation order;
    len = 0 gives 0 as the reference's empty loop does."""
    import ctypes
    rng = np.random.default_rng(n + 5)
    u = rng.standard_normal(n)
    v = rng.standard_normal(n)
    r = ctypes.c_double(-1.0)
    L.Vector2Norm(u.ctypes.data, n, ctypes.byref(r), 0)
    nrm = oracle.norm2(u) if n else 0.0
    assert abs(r.value - nrm) <= 1e-14 * max(nrm, 1e-300)
    L.VectorDotProduct(u.ctypes.data, v.ctypes.data, n, ctypes.byref(r), 0)
    dot = float(np.dot(u, v)) if n else 0.0
    assert abs(r.value - dot) <= 1e-13 * max(np.linalg.norm(u) * np.linalg.norm(v), 1e-300)


def test_pythag(L):
    """tools.c:345-352: sqrt(a^2 + b^2) as the larger magnitude times
    sqrt(1 + (smaller^2 / larger^2)).  The reference squares both before
    dividing, so where both squares overflow (|a|, |b| > ~1e154) or both
    underflow to zero it returns NaN, and so must the drop-in."""
    assert L.pythag(3.0, 4.0) == 5.0
    assert L.pythag(-5.0, 12.0) == 13.0
    assert L.pythag(0.0, 0.0) == 0.0
    assert L.pythag(0.0, -2.5) == 2.5
    assert math.isnan(L.pythag(1e300, 1e300)) and math.isnan(L.pythag(1e300, 2e299))
    assert math.isnan(L.pythag(3e-300, 4e-300))
    assert L.pythag(1e300, 0.0) == 1e300
    rng = np.random.default_rng(3)
    for a_, b_ in rng.standard_normal((200, 2)) * 10.0 ** rng.integers(-100, 100, (200, 1)):
        h = L.pythag(float(a_), float(b_))
        assert abs(h - math.hypot(a_, b_)) <= 4e-16 * math.hypot(a_, b_)


def test_pinv_and_svd_reference_abi(L, oracle, golden):
    """PseudoInverseTimesVec (tools.c:52-105) on the reference's known-answer
    vectors, bit for bit; SingularValueDecomp (tools.c:113-337) returns the
    oracle's (reference-pinned) truncated singular values and V in a."""
    import ctypes
    _, g = golden
    ms, A, b, xg = g["pinv_m"], g["pinv_A"], g["pinv_b"], g["pinv_x"]
    ia = ib = 0
    for m in ms:
        m = int(m)
        Am = np.ascontiguousarray(A[ia:ia + m * m].reshape(m, m))
        bb = np.ascontiguousarray(b[ib:ib + m])
        x = np.zeros(m)
        L.PseudoInverseTimesVec(dropin.rows(Am), bb.ctypes.data, x.ctypes.data, m)
        assert np.array_equal(x, xg[ib:ib + m]), f"pinv m={m}"
        a2 = Am.copy()
        w = np.zeros(m)
        V = np.zeros((m, m))
        L.SingularValueDecomp(dropin.rows(a2), m, m, w.ctypes.data, dropin.rows(V))
        _, wo, vo = oracle.pinv_times_vec(Am, bb)
        assert np.array_equal(w, wo), f"svd w m={m}"
        assert np.array_equal(V, vo.reshape(m, m)) and np.array_equal(a2, V), f"svd V m={m}"
        ia += m * m
        ib += m


@pytest.mark.parametrize("kind", ["galerkin", "l2"])
@pytest.mark.parametrize("case", ["m7", "m13"])
def test_projection_steps_reference_abi(L, edge, kind, case):
    """Galerkin_Richardson (PGR.c:181-218) and L2_Richardson (PL2R.c:180-218)
    on host arrays against the reference's own svec and allredvec (m = 7
    with two never-written columns: a rank-deficient system; m = 13)."""
    _, g = edge
    DX = np.ascontiguousarray(g[f"step_{case}_DX"])
    DF = np.ascontiguousarray(g[f"step_{case}_DF"])
    f = np.ascontiguousarray(g[f"step_{case}_f"])
    m, n = DF.shape
    red = np.zeros(m * m + m)
    rhs = np.zeros(m)
    sv = np.zeros(m)
    M = np.zeros((m, m))
    if kind == "galerkin":
        L.Galerkin_Richardson(dropin.rows(DX), dropin.rows(DF), f.ctypes.data, m, n, dropin.rows(M),
                              red.ctypes.data, rhs.ctypes.data, sv.ctypes.data, 0)
    else:
        L.L2_Richardson(dropin.rows(DF), f.ctypes.data, m, n, dropin.rows(M), red.ctypes.data,
                        rhs.ctypes.data, sv.ctypes.data, 0)
    rr = g[f"step_{case}_{kind}_red"]
    rs = g[f"step_{case}_{kind}_svec"]
    # the sums in another order: rounding of N = 4096 / 3000 terms
    assert np.allclose(red, rr, rtol=1e-13, atol=1e-12 * np.abs(rr).max())
    assert np.array_equal(red[m * m:], rhs)
    assert np.allclose(sv, rs, rtol=1e-9, atol=1e-11 * np.abs(rs).max())


def test_host_lap_vec_mult_and_precondition(L, oracle):
    """Lap_Vec_mult (system.c:80-213) and Precondition (system.c:222-227)
    on host arrays: bit-identical to the oracle's stencil and to the
    division."""
    shape = (20, 12, 16)
    n = int(np.prod(shape))
    coeff = oracle.lap_coefficient(6)
    sysd = dropin.system(shape, coeff)
    import ctypes
    x = np.random.default_rng(12).standard_normal(n)
    y = np.full(n, np.nan)
    for a in (-4.0, -1.0 / (4 * math.pi)):
        L.Lap_Vec_mult(ctypes.byref(sysd), a, x.ctypes.data, y.ctypes.data, 0)
        assert np.array_equal(y, oracle.matvec(x, shape, a))
    r = x.copy()
    diag = -4.0 * (3 * coeff[0])
    L.Precondition(diag, r.ctypes.data, n)
    assert np.array_equal(r, x / diag)


def _mask(s):
    return re.sub(r"Time = [-0-9.na]+", "Time = <t>", s).strip()


NF = [(c, s) for c in ("rhs_nan", "rhs_inf", "x0_nan") for s in ("aar", "pgr", "pl2r", "cg")]


@pytest.mark.parametrize("case,solver", NF, ids=[f"{c}-{s}" for c, s in NF])
def test_nonfinite_inputs_dropin(gpu, edge, case, solver):
    """A NaN or Inf in rhs, or a NaN in x0, through the reference-signature
    solver entries: the reference's printed lines (the stop rule relres >
    tol is false for NaN, AAR.c:91; an Inf ||b|| reports relres 0), its exit
    status (the SVD of a NaN Gram does not converge and tools.c:270 ends the
    process with status 1) and the returned x (unchanged where the
    reference's loop never runs)."""
    meta, _ = edge
    ref = meta["nonfinite"][case][solver]
    say("nonfinite", case, solver, ref)
    r = subprocess.run([sys.executable, os.path.join(HERE, "dropin.py"), case, solver], capture_output=True,
                       text=True, timeout=120)
    say(r.stdout[-800:], r.stderr[-800:])
    assert r.returncode == ref["exit_status"]
    lines = [ln for ln in r.stdout.splitlines() if not ln.startswith("RESULT")]
    assert _mask("\n".join(lines)) == _mask(ref["stdout"])
    if ref["exit_status"] == 0:
        res = json.loads([ln for ln in r.stdout.splitlines() if ln.startswith("RESULT")][-1][7:])
        assert res["x_nan_count"] == ref["x_nan_count"]
        assert res["x_equals_x0"] == ref["x_equals_x0"]
