"""Batched AAR: uneven grids, screened operators and several right-hand
sides solved together (one graph-captured Anderson cycle per replay).

Every instance must give the reference's iteration count (the oracle's,
bit-faithful to MPI_REAL/AAR.c), in FP64 and in mixed precision.  FP64
solutions agree with the oracle's within 1e-8 when converged (1e-6 at a
max_iter stop).  A mixed solution stores the residual history in FP32: with
a deep history on a slowly converging instance it moves up to ~2e-7 from
the FP64 iterate, as the single-solve mixed path does on the same
instance.  Its bar: its true residual meets the tolerance, and it is no
farther from the FP64 oracle's iterate than the two residuals allow,
(||r|| + ||r_64||)/lambda_min.  The FP64 batch must agree with the
single-solve path on the same instance.
"""
import numpy as np
import pytest

from aar_hip import problems as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def dev(v):
    import torch
    return torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to("cuda")


def instances():
    """(shape, h, sigma, rhs): uneven grids, Poisson and screened, and three
    right-hand sides sharing one operator."""
    out = [((16, 16, 16), 0.5, 0.5, P.rhs_normal((16, 16, 16), 21)),
           ((24, 24, 24), 0.5, 0.0, P.rhs_uniform((24, 24, 24), 22)),
           ((20, 12, 16), 0.5, 1.7, P.rhs_normal((20, 12, 16), 23)),
           ((32, 32, 32), 0.25, 0.1, P.rhs_normal((32, 32, 32), 24))]
    for s in range(3):
        out.append(((24, 24, 24), 0.5, 0.8, P.rhs_normal((24, 24, 24), 30, stream=s)))
    return out


SCHEDULES = [  # m, p, max_iter, tol
    (7, 6, 1000, 1e-8),
    (4, 1, 1000, 1e-8),
    (3, 4, 1000, 1e-8),
    (7, 6, 13, 1e-8),
    # the READMEs' deep history, m = 16, and p > m + 1 (residuals folded
    # into the anchor between Anderson steps), with a max_iter stop inside
    # a cycle's folds
    (13, 11, 1000, 1e-8),
    (16, 5, 1000, 1e-8),
    (3, 8, 1000, 1e-8),
    (4, 9, 40, 1e-12),
]


@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: f"m{s[0]}p{s[1]}i{s[2]}")
@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
def test_batch_vs_oracle(gpu, oracle, sched, mixed):
    m, p, max_iter, tol = sched
    say("batch", sched, "mixed" if mixed else "fp64")
    inst = instances()
    xs = [dev(np.zeros(b.size)) for (_, _, _, b) in inst]
    bs = [dev(b) for (_, _, _, b) in inst]
    args = [(xs[i], bs[i], shape, P.poisson_a(h), sig) for i, (shape, h, sig, _) in enumerate(inst)]
    sts = gpu.AAR_batch(args, 0.6, 0.6, m, p, max_iter, tol, mixed=mixed)
    for i, (shape, h, sig, b) in enumerate(inst):
        xr, dr = oracle.aar(np.zeros(b.size), b, shape, P.poisson_a(h), 0.6, 0.6, m, p, max_iter, tol,
                            sigma=sig, trace=False)
        say(f"  instance {i} {shape} sigma={sig}: gpu {sts[i]['iterations']} oracle {dr['iterations']}")
        assert sts[i]["iterations"] == dr["iterations"]
        assert bool(sts[i]["printed_converged"]) == bool(dr["printed_converged"])
        x = xs[i].cpu().numpy()
        if not mixed:
            lim = 1e-8 if dr["printed_converged"] else 1e-6
            assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= lim
        elif not dr["printed_converged"]:
            # a max_iter stop: the same iteration's FP64 iterate (the
            # single-solve mixed path measures 1.8e-6 on the m = 3, p = 8 stop)
            assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= 1e-5
        else:
            # two converged iterates of this symmetric operator differ by at
            # most what their true residuals allow: (||r|| + ||r64||)/lambda_min
            a = P.poisson_a(h)
            r = np.linalg.norm(b - oracle.matvec(x, shape, a, sigma=sig))
            r64 = np.linalg.norm(b - oracle.matvec(xr, shape, a, sigma=sig))
            cf = oracle.lap_coefficient(6)
            th = 2 * np.pi / max(shape)
            lmin = sig if sig > 0 else abs(a * (cf[0] + 2 * sum(cf[q] * np.cos(q * th) for q in range(1, 7))))
            d = np.linalg.norm((x - x.mean()) - (xr - xr.mean())) if sig == 0 else np.linalg.norm(x - xr)
            say(f"    mixed vs fp64 {d / np.linalg.norm(xr):.2e}, residual bound {(r + r64) / lmin / np.linalg.norm(xr):.2e}")
            assert r <= tol * np.linalg.norm(b) * (1 + 1e-12)
            assert d <= (r + r64) / lmin


def test_batch_matches_single_solves(gpu):
    """The FP64 batch and the single-solve path compute the same iterates."""
    inst = instances()
    xs = [dev(np.zeros(b.size)) for (_, _, _, b) in inst]
    bs = [dev(b) for (_, _, _, b) in inst]
    args = [(xs[i], bs[i], shape, P.poisson_a(h), sig) for i, (shape, h, sig, _) in enumerate(inst)]
    sts = gpu.AAR_batch(args, 0.6, 0.6, 7, 6, 1000, 1e-10)
    for i, (shape, h, sig, b) in enumerate(inst):
        x1 = dev(np.zeros(b.size))
        st = gpu.AAR(x1, bs[i], shape, P.poisson_a(h), 0.6, 0.6, 7, 6, 1000, 1e-10, sigma=sig)
        assert st["iterations"] == sts[i]["iterations"]
        r = float(((x1 - xs[i]).norm() / x1.norm()).cpu())
        assert r < 1e-10, (i, r)


def test_autotune_shapes_keep_parity(gpu, oracle):
    """Tuned launch shapes are valid candidates and the stencil and the
    solver give the same results under them (bit-exact matvec, the
    reference's iteration count)."""
    shape = (40, 36, 44)
    for fp32 in (False, True):
        t = gpu.autotune(shape, fp32=fp32)
        say("tuned", shape, "fp32" if fp32 else "fp64", t)
        assert (t["wy"], t["py"]) in {(8, 4), (4, 8), (8, 2), (4, 4), (4, 2)}
        assert min(44, 16) <= t["lz"] <= 44 and t["us"] > 0
    tb = gpu.autotune(shape, batch=True)
    assert (tb["wy"], tb["py"]) == (8, 4)
    x = np.random.default_rng(9).standard_normal(int(np.prod(shape)))
    y = gpu.Lap_Vec_mult(dev(x), shape, -4.0).cpu().numpy()
    assert np.array_equal(y, oracle.matvec(x, shape, -4.0))
    b = P.rhs_uniform(shape, 41)
    xr, dr = oracle.aar(np.zeros(b.size), b, shape, -4.0, 0.6, 0.6, 7, 6, 1000, 1e-8, trace=False)
    xg = dev(np.zeros(b.size))
    st = gpu.AAR(xg, dev(b), shape, -4.0, 0.6, 0.6, 7, 6, 1000, 1e-8)
    assert st["iterations"] == dr["iterations"]
    assert np.linalg.norm(xg.cpu().numpy() - xr) / np.linalg.norm(xr) <= 1e-8
    xb = dev(np.zeros(b.size))
    sts = gpu.AAR_batch([(xb, dev(b), shape, -4.0, 0.0)], 0.6, 0.6, 7, 6, 1000, 1e-8)
    assert sts[0]["iterations"] == dr["iterations"]


@pytest.mark.parametrize("lz", [8, 20, 64], ids=lambda v: f"lz{v}")
def test_batch_fixed_slab_depth_vs_oracle(gpu, oracle, lz):
    """A slab depth fixed by set_tuning reaches the batch (clamped to each
    instance's nz; 64 is one slab per column for every instance here): the
    reference's iteration count and the oracle's iterate within 1e-8."""
    inst = instances()
    xs = [dev(np.zeros(b.size)) for (_, _, _, b) in inst]
    bs = [dev(b) for (_, _, _, b) in inst]
    args = [(xs[i], bs[i], shape, P.poisson_a(h), sig) for i, (shape, h, sig, _) in enumerate(inst)]
    say("batch with lz", lz)
    gpu.set_tuning(lz, 8, 4)
    try:
        sts = gpu.AAR_batch(args, 0.6, 0.6, 7, 6, 1000, 1e-8)
    finally:
        gpu.set_tuning(0, 8, 4)
    for i, (shape, h, sig, b) in enumerate(inst):
        xr, dr = oracle.aar(np.zeros(b.size), b, shape, P.poisson_a(h), 0.6, 0.6, 7, 6, 1000, 1e-8,
                            sigma=sig, trace=False)
        assert sts[i]["iterations"] == dr["iterations"], (i, shape)
        x = xs[i].cpu().numpy()
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= 1e-8, (i, shape)
