"""The streaming updates' element-by-element forms.  The anchor update, the
PGR/PL2R projection update and the CG update stream pairs of points; an odd
point count, or an iterate that is not on a 16-byte boundary, takes their
element-by-element walk instead.  Both are solved here against the oracle
(AAR.c:37-177, PGR.c:35-218, PL2R.c, CG.c:33-175 restated): the oracle's
iteration counts and printed conditions, solutions within 1e-8 (CG 1e-7,
the reference's own rank spread at these sizes), residual traces within
1e-6.  The mixed-precision anchor update's three walks (four points, pairs,
single points) against the FP64 oracle: its count within one Anderson
period, its solution within 1e-7 at tol 1e-10.
"""
import numpy as np
import pytest

from aar_hip import problems as P

pytestmark = pytest.mark.gpu

ODD = (15, 13, 13)          # 2535 points
EVEN = (16, 14, 12)


def say(*a):
    print(*a, flush=True)


def _dev_offset(v, offset):
    """v on the device as a view `offset` elements into its buffer (offset 1:
    the data starts 8 bytes past a 16-byte boundary)."""
    import torch
    buf = torch.zeros(v.size + offset, dtype=torch.float64, device="cuda")
    t = buf[offset:offset + v.size]
    t.copy_(torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)))
    return t


def _check(st, dr, x, xr, solver):
    assert st["iterations"] == dr["iterations"], (st["iterations"], dr["iterations"])
    assert bool(st["printed_converged"]) == bool(dr["printed_converged"])
    tol_x = 1e-7 if solver == "cg" else 1e-8
    rel = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    say(f"  {solver}: {st['iterations']} iterations, solution rel diff {rel:.2e}")
    assert rel <= tol_x


def _trace(gpu, dr, solver):
    h = gpu.history()
    tr = dr["trace"]
    if solver == "cg":
        h = np.concatenate([h[:1], h[1::50]])
    assert len(h) >= len(tr)
    h = h[:len(tr)]
    assert np.all(np.abs(h - tr) <= 1e-6 * tr + 1e-10 * tr[0])


# PGR/PL2R copy the caller's x into their own ring (and back), so only the
# odd point count reaches their update's element-by-element walk
CASES = [(s, ODD, 0) for s in ("aar", "pgr", "pl2r", "cg")] + [(s, EVEN, 1) for s in ("aar", "cg")]


@pytest.mark.parametrize("solver,shape,offset", CASES,
                         ids=[f"{c[0]}-{'odd_n' if c[2] == 0 else 'x_off_16B'}" for c in CASES])
def test_elementwise_walks_vs_oracle(gpu, oracle, solver, shape, offset):
    say("elementwise", solver, shape, "offset", offset)
    n = int(np.prod(shape))
    a = P.poisson_a(0.5)
    b = P.rhs_normal(shape, 2027)
    b = b - b.mean()
    x = _dev_offset(np.zeros(n), offset)
    bd = _dev_offset(b, 0)
    if solver == "aar":
        xr, dr = oracle.aar(np.zeros(n), b, shape, a, 0.6, 0.6, 7, 6, 3000, 1e-8)
        st = gpu.AAR(x, bd, shape, a, 0.6, 0.6, 7, 6, 3000, 1e-8)
    elif solver == "cg":
        xr, dr = oracle.cg(np.zeros(n), b, shape, a, 3000, 1e-8)
        st = gpu.CG(x, bd, shape, a, 3000, 1e-8)
    else:
        l2 = solver == "pl2r"
        xr, dr = oracle.pgr(np.zeros(n), b, shape, a, 0.6, 7, 6, 3000, 1e-8, l2=l2)
        fn = gpu.PL2R if l2 else gpu.PGR
        st = fn(x, bd, shape, a, 0.6, 7, 6, 3000, 1e-8)
    _check(st, dr, x.cpu().numpy(), xr, solver)
    _trace(gpu, dr, solver)


@pytest.mark.parametrize("offset", [0, 1], ids=["aligned", "x_off_16B"])
def test_cg_variable_coefficient_walks(gpu, oracle, offset):
    """CG on a*Lap + c(x) (the update reads c for the Jacobi divisor): odd n,
    and an iterate off a 16-byte boundary on an even grid."""
    shape = ODD if offset == 0 else EVEN
    say("elementwise cg vc", shape, "offset", offset)
    n = int(np.prod(shape))
    a = P.poisson_a(0.5)
    c = P.helmholtz_coefficient(shape, 1011)
    b = P.rhs_normal(shape, 1011)
    x = _dev_offset(np.zeros(n), offset)
    xr, dr = oracle.cg_vc(np.zeros(n), b, shape, a, c, 3000, 1e-8)
    st = gpu.CG(x, _dev_offset(b, 0), shape, a, 3000, 1e-8, c=_dev_offset(c, 0))
    _check(st, dr, x.cpu().numpy(), xr, "cg")
    _trace(gpu, dr, "cg")


# mixed precision: the anchor update streams four points per thread when the
# point count is a multiple of 4 and x is on a 16-byte boundary; otherwise
# pairs (n % 4 == 2) or single points (odd n, x off the boundary)
MIXED_CASES = [((16, 14, 12), 0, "quads"), ((14, 13, 13), 0, "pairs_n_mod4_2"), (ODD, 0, "odd_n"),
               ((16, 14, 12), 1, "x_off_16B")]


@pytest.mark.parametrize("shape,offset,walk", MIXED_CASES, ids=[c[2] for c in MIXED_CASES])
def test_mixed_anchor_walks_vs_oracle(gpu, oracle, shape, offset, walk):
    """FP32 sweeps and history with FP64 true residuals, on each walk of the
    anchor update: the FP64 oracle's count within one Anderson period p,
    and its solution within 1e-7 at tol 1e-10 (the bar of the deep-history
    mixed test)."""
    say("elementwise mixed", walk, shape, "offset", offset)
    n = int(np.prod(shape))
    a = P.poisson_a(0.5)
    b = P.rhs_normal(shape, 2029)
    b = b - b.mean()
    x = _dev_offset(np.zeros(n), offset)
    xr, dr = oracle.aar(np.zeros(n), b, shape, a, 0.6, 0.6, 7, 6, 3000, 1e-10, trace=False)
    st = gpu.AAR_mixed(x, _dev_offset(b, 0), shape, a, 0.6, 0.6, 7, 6, 3000, 1e-10)
    rel = np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr)
    say(f"  mixed {st['iterations']} oracle {dr['iterations']} rel diff {rel:.2e}")
    assert st["converged"] and abs(st["iterations"] - dr["iterations"]) <= 6
    assert rel <= 1e-7
