"""The deep Anderson histories, 16 < m <= 32, against the reference build
(tests/golden/golden_deep.*).  The reference takes any history depth
(MPI_REAL/AAR.c:51-73, PGR.c:50-72, PL2R.c:50-72); libaar_hip takes m <= 32,
the largest m x m system its one-wavefront SVD solves.

* AAR, PGR and PL2R at m = 20 and 32 with p below and above m + 1: the
  reference's iteration counts and printed conditions; solutions within
  1e-8, or 10x the reference's own spread over 1, 2, 4 and 8 ranks where
  that is larger; residual traces within 1e-6 (1e-4 for the entries that
  are not true residuals at omega = 1.0 > 0.843, as in the shallow
  fixtures) or 10x that spread, entry by entry.  A deep history's m x m
  system is ill-conditioned enough that the summation order of its dots
  moves the late residuals beyond rounding in the reference itself: with
  p > m + 1 (every column from a run of plain Richardson steps) its own 1-
  and 8-rank solutions differ by up to 1e-7.  The fixtures keep only cases
  whose counts are the same on every rank count;
* mixed precision at the same histories: the FP64 counts, solutions within
  the tolerance's reach;
* the pseudo-inverse at n = 24 and 32, bit for bit against tools.c;
* m = 33 refused.
"""
import numpy as np
import pytest

from aar_hip import problems as P

pytestmark = pytest.mark.gpu

CASES = ["m20p6_24", "m32p4_24", "m20p24_24", "m32p11_28x24x20"]


def say(*a):
    print(*a, flush=True)


def dev(v):
    import torch
    return torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to("cuda")


def _solve(gpu, solver, c, x, b, mixed=False):
    shape = tuple(c["shape"])
    if solver == "aar":
        fn = gpu.AAR_mixed if mixed else gpu.AAR
        return fn(x, b, shape, c["a"], c["omega"], c["beta"], c["m"], c["p"], c["max_iter"], c["tol"])
    fn = gpu.PGR if solver == "pgr" else gpu.PL2R
    return fn(x, b, shape, c["a"], c["omega"], c["m"], c["p"], c["max_iter"], c["tol"])


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("solver", ["aar", "pgr", "pl2r"])
def test_deep_history_vs_reference(gpu, golden_deep, name, solver):
    meta, g = golden_deep
    c = meta["cases"][name]["params"]
    shape = tuple(c["shape"])
    say("deep", name, solver)
    b = P.rhs_uniform(shape, c["seed"])
    x = dev(np.zeros(b.size))
    st = _solve(gpu, solver, c, x, dev(b))
    ref = meta["cases"][name]["solvers"][solver]
    xr = g[f"{name}__{solver}__x"]
    rel = np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr)
    say(f"  gpu {st['iterations']} reference {ref['iterations']} solution rel diff {rel:.2e}")
    assert st["iterations"] == ref["iterations"]
    assert bool(st["printed_converged"]) == ref["converged_line"]
    assert rel <= max(1e-8, 10 * ref["rank_spread_rel"])
    h = gpu.history()
    tr = g[f"{name}__{solver}__trace"]
    spread = g[f"{name}__{solver}__trace_spread"]
    assert len(h) == len(tr)
    rtol = np.full(len(tr), 1e-6)
    if c["omega"] > 0.843:
        if solver == "aar":
            rtol[np.arange(len(tr)) % c["p"] != 0] = 1e-4
        else:
            rtol[:] = 1e-4
    err = np.abs(h - tr)
    bar = np.maximum(rtol * np.abs(tr), 10 * spread) + 1e-10 * tr[0]
    assert np.all(err <= bar), (int(np.argmax(err / bar)), float(np.max(err / np.abs(tr))))


@pytest.mark.parametrize("name", ["m20p6_24", "m32p4_24"])
def test_deep_history_mixed(gpu, golden_deep, name):
    """FP32 sweeps and history with FP64 true residuals at m = 20 and 32:
    the FP64 reference's count, and the solution as close to the
    reference's as the tolerance allows (both stop at relres 1e-10)."""
    meta, g = golden_deep
    c = meta["cases"][name]["params"]
    shape = tuple(c["shape"])
    say("deep mixed", name)
    b = P.rhs_uniform(shape, c["seed"])
    x = dev(np.zeros(b.size))
    st = _solve(gpu, "aar", c, x, dev(b), mixed=True)
    ref = meta["cases"][name]["solvers"]["aar"]
    xr = g[f"{name}__aar__x"]
    rel = np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr)
    say(f"  mixed {st['iterations']} reference {ref['iterations']} rel diff {rel:.2e}")
    assert st["converged"] and abs(st["iterations"] - ref["iterations"]) <= c["p"]
    assert rel <= 1e-7


def test_deep_pinv_bitwise(gpu, golden_deep):
    _, g = golden_deep
    ia = ib = 0
    for m in g["pinv_m"]:
        A = g["pinv_A"][ia:ia + m * m].reshape(m, m)
        x, _, _ = gpu.pinv(A, g["pinv_b"][ib:ib + m])
        assert np.array_equal(x, g["pinv_x"][ib:ib + m]), m
        ia += m * m
        ib += m


def test_history_limit(gpu):
    from aar_hip._lib import AarHipError
    shape = (16, 16, 16)
    b = dev(P.rhs_uniform(shape, 1001))
    x = dev(np.zeros(b.numel()))
    for fn in (lambda: gpu.AAR(x, b, shape, -4.0, 0.6, 0.6, 33, 6, 100, 1e-6),
               lambda: gpu.PGR(x, b, shape, -4.0, 0.6, 33, 6, 100, 1e-6)):
        with pytest.raises(AarHipError):
            fn()
