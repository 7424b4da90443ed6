"""Benchmark: AAR solves/s on the 512^3 FD12 periodic Poisson system.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512]

A step is one AAR solve (MPI_REAL/AAR.c semantics) from x0 = 0 to
||b - A x|| / ||b|| <= 1e-6 on the 512^3 grid, h = 0.25 (a = -16),
b ~ U[-0.5, 0.5] (seed 1002, mean removed), omega = beta = 0.6, m = 7,
p = 6, with b and x already resident in HBM.  `value` is FP64 AAR solves/s.
Reported on the same JSON line: the mixed-precision AAR (FP32 sweeps and
history inside an FP64 true-residual loop) on the same solve and, against
FP64 AAR, on b ~ N(0,1) to 1e-10; the stencil matvec bandwidth; the
roofline of the dominant kernel (the residual-recurrence stencil sweep,
timed with HIP events on the library stream); a CPU baseline (the reference
solver on the host's cores, bounded sample).

--mode kernels runs a short profiling workload (a few matvecs, one plain
copy of known size, 24 AAR iterations in FP64 and in mixed precision) for
rocprofv3.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import shutil
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "linear-solver_amd"))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--mode", default="bench", choices=["bench", "kernels"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-ranks", type=int, default=0, help="MPI ranks of the all-core CPU run (0: min(16, cores))")
    ap.add_argument("--no-helm", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-vc", action="store_true")
    ap.add_argument("--no-petsc", action="store_true")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fp:
            for line in fp:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(gpu_rates, n_gpu, gpu_iterations, ranks_all=None):
    """Full reference AAR solves (oracle/_ref: the reference MPI_REAL sources
    built from their own files) on the host cores, timed around the AAR()
    call as MPI_REAL/main.c:67-69 does: the C2 problem at 256^3 with one MPI
    rank per core, and at 128^3 with one rank.  Each is a whole solve from
    x0 = 0 to relres 1e-6.  `value` states the 256^3 all-core rate in the
    metric's unit (solves/s of the n_gpu^3 solve) by the point ratio and the
    iteration ratio; the measured rates and the GPU's at the same sizes are
    reported beside it.  A failed run is recorded as such."""
    import numpy as np
    from aar_hip import problems as P
    from oracle import oracle as O
    cores = ranks_all or min(16, len(os.sched_getaffinity(0)))
    out = {"unit": "solves/s", "kind": "reference", "cores": cores, "cpu": cpu_model(),
           "host_threads": os.cpu_count(), "runs": {}}
    if not O.ref_available():
        out["error"] = "oracle/_ref/ref_harness is not built (make -C oracle ref)"
        out["value"] = None
        return out
    a = P.poisson_a(0.25)
    for side, ranks in ((256, cores), (128, 1)):
        shape = (side,) * 3
        b = P.rhs_uniform(shape, 1002)
        key = f"{side}^3_{ranks}rank" + ("s" if ranks > 1 else "")
        try:
            _, d = O.ref_solve("aar", shape, b, np.zeros(b.size), a, 0.0, 0.6, 0.6, 7, 6, 20000, 1e-6, True,
                               ranks=ranks, timeout=900, trace=False)
            rate = 1.0 / d["seconds"]
            out["runs"][key] = {"ranks": ranks, "seconds": d["seconds"], "iterations": d["iterations"],
                                "solves_per_s": rate, "relres": d["relres_printed"],
                                "gpu_fp64_solves_per_s": gpu_rates.get(side),
                                "gpu_over_cpu": (gpu_rates[side] / rate) if gpu_rates.get(side) else None}
            log(f"cpu baseline {key}: {d['iterations']} its in {d['seconds']:.1f} s")
        except Exception as e:  # noqa: BLE001
            out["runs"][key] = {"ranks": ranks, "error": str(e)[-400:]}
            log(f"cpu baseline {key} failed: {e}")
    r = out["runs"].get(f"256^3_{cores}ranks", {})
    if "seconds" in r:
        scale = (n_gpu / 256) ** 3 * gpu_iterations / r["iterations"]
        out["value"] = 1.0 / (r["seconds"] * scale)
        out["sample"] = (f"full reference AAR solve of the same problem at 256^3 on {cores} MPI ranks "
                         f"({r['iterations']} iterations, {r['seconds']:.1f} s), stated at {n_gpu}^3 by "
                         f"x{(n_gpu / 256) ** 3:.0f} points and x{gpu_iterations / r['iterations']:.2f} "
                         f"iterations ({gpu_iterations} at {n_gpu}^3); the full solves and the GPU's rates at "
                         "the same sizes are under runs")
    else:
        out["value"] = None
        out["error"] = "the 256^3 reference solve did not complete"
    return out


def helmholtz_case(n, dev, a, directions=20, hstep=1e-6, tol=1e-13):
    """One forward solve and its adjoint (autograd backward) of
    L = 1/2 ||u(b, c) - u*||^2 on the n^3 complex Helmholtz problem, then
    the gradients checked by central differences."""
    import numpy as np
    import torch
    import aar_hip
    from aar_hip import problems as P
    shape = (n,) * 3
    N = n ** 3
    eta = 0.5
    c = torch.from_numpy(P.helmholtz_coefficient(shape, 1005)).to(dev)

    def crhs(stream):
        v = P.normal_torch(2 * N, 1005, stream, dev)
        return torch.complex(v[0::2].contiguous(), v[1::2].contiguous())

    b, bs = crhs(0), crhs(1)
    opts = dict(omega=0.6, beta=0.6, m=7, p=6, max_iter=20000, tol=tol)
    ustar = torch.zeros_like(b)
    aar_hip.helm_aar(ustar, bs, shape, a, c, 0.0, eta, **opts)

    def loss(b_, c_):
        u = aar_hip.HelmholtzSolve.apply(b_, c_, shape, a, eta, opts)
        return 0.5 * torch.sum(torch.abs(u - ustar) ** 2)

    bb = b.clone().requires_grad_(True)
    cc = c.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L = loss(bb, cc)
    L.backward()
    torch.cuda.synchronize()
    t_fb = time.perf_counter() - t0
    last = aar_hip.HelmholtzSolve.last
    gb, gc = bb.grad.detach(), cc.grad.detach()
    gen = torch.Generator(device=dev)
    gen.manual_seed(2024)
    eb, ec = [], []

    def solve(b_, c_):
        return aar_hip.HelmholtzSolve.apply(b_, c_, shape, a, eta, opts)

    def central(up, um):
        # (L(+) - L(-)) / 2h with L(+) - L(-) = 1/2 (||u+ - u*||^2 - ||u- - u*||^2)
        # = Re <u+ - u-, (u+ + u-)/2 - u*>, exactly; summed in this form the
        # difference does not cancel against L itself
        return float(torch.sum((torch.conj(up - um) * (0.5 * (up + um) - ustar)).real)) / (2 * hstep)

    with torch.no_grad():
        for _ in range(directions):
            db = torch.complex(torch.randn(N, dtype=torch.float64, device=dev, generator=gen),
                               torch.randn(N, dtype=torch.float64, device=dev, generator=gen))
            dc = torch.randn(N, dtype=torch.float64, device=dev, generator=gen)
            fd = central(solve(b + hstep * db, c), solve(b - hstep * db, c))
            ad = float(torch.sum((gb.conj() * db).real))
            eb.append(abs(fd - ad) / abs(ad))
            fd = central(solve(b, c + hstep * dc), solve(b, c - hstep * dc))
            ad = float(torch.sum(gc * dc))
            ec.append(abs(fd - ad) / abs(ad))
    f, adj = last["forward"], last["adjoint"]
    # the complex PETSc variant's Krylov comparators on the same system
    # (PCJACOBI, rtol 1e-6, max_it 2000 as PETSC_COMPLEX/main.c:96,127,155)
    # against AAR to the same tolerance
    kry = {}
    runs = [("aar", lambda u: aar_hip.helm_aar(u, b, shape, a, c, 0.0, eta, **dict(opts, tol=1e-6))),
            ("gmres30", lambda u: aar_hip.helm_gmres(u, b, shape, a, c, 0.0, eta, 2000, 1e-6)),
            ("bicg", lambda u: aar_hip.helm_bicg(u, b, shape, a, c, 0.0, eta, 2000, 1e-6))]
    if n <= 128:
        # the variant's pc = 1 (complex ILU(0), level-scheduled: 3n - 2 launches per triangular solve)
        runs += [("gmres30_ilu0", lambda u: aar_hip.helm_gmres(u, b, shape, a, c, 0.0, eta, 2000, 1e-6, precond=2)),
                 ("bicg_ilu0", lambda u: aar_hip.helm_bicg(u, b, shape, a, c, 0.0, eta, 2000, 1e-6, precond=2))]
    for name, fn in runs:
        u = torch.zeros_like(b)
        st = fn(u)
        kry[name] = {"iterations": st["iterations"], "solve_ms": st["solve_ms"], "relres": st["relres"]}
    return {"grid": [n, n, n], "h": 0.25, "eta": eta, "tol": opts["tol"],
            "rtol_1e-6_solvers": kry,
            "forward_iterations": f["iterations"], "forward_ms": f["solve_ms"],
            "adjoint_iterations": adj["iterations"], "adjoint_ms": adj["solve_ms"],
            "solves_per_s": 1e3 / f["solve_ms"], "forward_backward_s": t_fb,
            "fd_directions": directions, "fd_step": hstep,
            "fd_max_rel_err_b": float(np.max(eb)), "fd_max_rel_err_c": float(np.max(ec))}


def varcoef_case(dev, shape, a, b, tol):
    """AAR on the variable-coefficient operator a*Lap + c(x) (c = 1 + 0.5 *
    8 seeded Fourier modes, in [0.5, 1.5]) with the pointwise Jacobi
    diagonal: FP64 and mixed-precision solves, true residuals and the
    recurrence sweep's kernel time (3 algorithmic vectors per launch)."""
    import torch
    import aar_hip
    from aar_hip import problems as P
    c = P.helmholtz_coefficient_torch(shape, 1010, dev)
    x = torch.zeros_like(b)
    out = {"operator": "a*Lap_FD12 + c(x), c = 1 + 0.5*(8 seeded Fourier modes), Jacobi diag a*3*c0 + c(x)",
           "rhs": "the headline RHS"}
    for label, mixed in (("fp64", False), ("mixed", True)):
        x.zero_()
        aar_hip.AAR_vc(x, b, shape, a, c, 0.6, 0.6, 7, 6, 20000, tol, mixed=mixed)    # warm-up
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = aar_hip.AAR_vc(x, b, shape, a, c, 0.6, 0.6, 7, 6, 20000, tol, mixed=mixed)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        r = b - aar_hip.Lap_Vec_mult_vc(x, shape, a, c)
        rel = float(torch.linalg.norm(r) / torch.linalg.norm(b))
        x.zero_()
        aar_hip.profile(True)
        aar_hip.profile_reset()
        aar_hip.AAR_vc(x, b, shape, a, c, 0.6, 0.6, 7, 6, 20000, tol, mixed=mixed)
        aar_hip.profile(False)
        kr = aar_hip.profile_query("stencil_aar_rec")
        out[label] = {"solves_per_s": 1.0 / dt, "seconds": dt, "iterations": st["iterations"],
                      "converged": bool(st["converged"]), "relres_true": rel,
                      "algorithmic_GBps": st["bytes"] / st["solve_ms"] / 1e6,
                      "stencil_aar_rec_vc": {"avg_ms": kr["ms"] / kr["count"],
                                             "GBps": kr["bytes"] / kr["ms"] / 1e6}}
        log(f"varcoef {label}: {st['iterations']} its {dt * 1e3:.0f} ms relres {rel:.2e}")
    out["mixed_speedup_vs_fp64"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    del c, x
    return out


def petsc_variant_case(dev, tol, n=128):
    """The PETSc variant's solvers (PETSC_REAL/main.c with pc = 0 and 1) on
    the 128^3 Poisson system (h = 0.25, the same seeded RHS scheme): AAR
    with Jacobi and with ILU(0) (its PCICC), PGR and PL2R with ILU(0),
    GMRES(30) and BiCG with ILU(0)
    (its PCBJACOBI/PCILU), max_it 2000 for the Krylov methods.  The
    level-scheduled ILU(0) sweeps are launch-latency bound (n+n+n-2 levels
    per triangular solve)."""
    import torch
    import aar_hip
    from aar_hip import problems as P
    shape = (n, n, n)
    a = P.poisson_a(0.25)
    b = P.rhs_uniform_torch(shape, 1002, device=dev)
    x = torch.zeros_like(b)
    out = {"grid": list(shape)}
    runs = (("aar_jacobi", lambda: aar_hip.AAR_petsc(x, b, shape, a, 0.6, 0.6, 7, 6, 100000, tol, precond=True)),
            ("aar_ilu0", lambda: aar_hip.AAR_petsc(x, b, shape, a, 0.6, 0.6, 7, 6, 100000, tol,
                                                   precond=aar_hip.ILU0)),
            ("pgr_ilu0", lambda: aar_hip.PGR_petsc(x, b, shape, a, 0.6, 7, 6, 100000, tol, precond=aar_hip.ILU0)),
            ("pl2r_ilu0", lambda: aar_hip.PGR_petsc(x, b, shape, a, 0.6, 7, 6, 100000, tol, precond=aar_hip.ILU0,
                                                    l2=True)),
            ("gmres30_ilu0", lambda: aar_hip.GMRES(x, b, shape, a, 2000, tol, precond=aar_hip.ILU0)),
            ("bicg_ilu0", lambda: aar_hip.BiCG(x, b, shape, a, 2000, tol, precond=aar_hip.ILU0)))
    for name, fn in runs:
        x.zero_()
        st = fn()
        out[name] = {"iterations": st["iterations"], "converged": bool(st["converged"]), "relres": st["relres"],
                     "solve_ms": st["solve_ms"]}
        log(f"petsc variant {name}: {st['iterations']} its {st['solve_ms']:.0f} ms")
    # one ILU(0) factor + apply, timed on the device
    r = torch.randn_like(b)
    aar_hip.ilu0_apply(r, shape, a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    aar_hip.ilu0_apply(r, shape, a)
    e1.record()
    torch.cuda.synchronize()
    out["ilu0_factor_apply_ms"] = e0.elapsed_time(e1)
    return out


def batch_case(dev, seed=1004):
    """96 screened-Poisson instances, side drawn from {24,32,48,64,96,128},
    sigma_i ~ U[0.1, 2], b_i ~ N(0,1), plus 16 right-hand sides on one 128^3
    operator; h = 0.5.  FP64 and mixed batches (one graph-replayed Anderson
    cycle), slab depths autotuned per size; the same instances solved one
    by one on the single-solve path for comparison."""
    import numpy as np
    import torch
    import aar_hip
    from aar_hip import problems as P
    inst = P.batch_instances(seed)
    a = P.poisson_a(0.5)
    m, p, tol, max_iter = 7, 6, 1e-6, 5000
    t0 = time.perf_counter()
    tuned = {s_: aar_hip.autotune((s_,) * 3, batch=True)["lz"] for s_ in sorted({sh[0] for sh, _, _ in inst})}
    tuned.update({f"{s_}_f32": aar_hip.autotune((s_,) * 3, fp32=True, batch=True)["lz"]
                  for s_ in sorted({sh[0] for sh, _, _ in inst})})
    t_tune = time.perf_counter() - t0
    bs = [P.normal_torch(int(np.prod(sh)), seed, k, dev) for sh, _, k in inst]
    xs = [torch.zeros_like(b_) for b_ in bs]
    points = int(sum(np.prod(sh) for sh, _, _ in inst))
    out = {"instances": len(inst), "points": points, "sides": [sh[0] for sh, _, _ in inst],
           "sigma": [round(sg, 4) for _, sg, _ in inst], "h": 0.5, "m": m, "p": p, "tol": tol,
           "autotune_s": t_tune, "tuned_lz": tuned}
    for label, mixed in (("fp64", False), ("mixed", True)):
        for x_ in xs:
            x_.zero_()
        args = [(xs[i], bs[i], sh, a, sg) for i, (sh, sg, _) in enumerate(inst)]
        aar_hip.AAR_batch(args, 0.6, 0.6, m, p, max_iter, tol, mixed=mixed)       # warm-up
        for x_ in xs:
            x_.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sts = aar_hip.AAR_batch(args, 0.6, 0.6, m, p, max_iter, tol, mixed=mixed)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        out[label] = {"seconds": tb, "instance_solves_per_s": len(inst) / tb,
                      "iterations": [s_["iterations"] for s_ in sts],
                      "all_converged": all(bool(s_["converged"]) for s_ in sts)}
        log(f"batch {label}: {tb * 1e3:.0f} ms, {len(inst) / tb:.1f} instance-solves/s")
    # one by one (FP64 single-solve path) on the same instances
    its1 = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, (sh, sg, _) in enumerate(inst):
        x1 = torch.zeros_like(bs[i])
        its1.append(aar_hip.AAR(x1, bs[i], sh, a, 0.6, 0.6, m, p, max_iter, tol, sigma=sg)["iterations"])
    torch.cuda.synchronize()
    t1 = time.perf_counter() - t0
    out["single_fp64"] = {"seconds": t1, "instance_solves_per_s": len(inst) / t1,
                          "iterations_equal_batch": its1 == out["fp64"]["iterations"]}
    out["batch_speedup_vs_single"] = t1 / out["fp64"]["seconds"]
    log(f"single fp64: {t1 * 1e3:.0f} ms; iterations equal: {its1 == out['fp64']['iterations']}")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "libaar_hip runs on one GPU: the whole grid is resident in "
                                   "one GPU's HBM; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    import aar_hip
    from aar_hip import problems as P

    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    n = args.n
    shape = (n, n, n)
    N = n ** 3
    a = P.poisson_a(0.25)
    omega = beta = 0.6
    m, p, max_iter = 7, 6, 20000
    b = P.rhs_uniform_torch(shape, 1002, device=dev)
    x = torch.zeros_like(b)
    torch.cuda.synchronize()

    if args.mode == "kernels":
        y = torch.empty_like(b)
        for _ in range(5):
            aar_hip.Lap_Vec_mult(b, shape, a, out=y)
        y.copy_(b)                                     # plain copy of 2*8*N bytes (counter scale)
        for k in ("read", "write", "copy"):            # the library's stream kernels, 8*N bytes each way
            aar_hip.stream_bandwidth(k, 8 * N, 2, dst=y, src=b)
        aar_hip.AAR(x, b, shape, a, omega, beta, m, p, 24, 1e-30)
        x.zero_()
        aar_hip.AAR_mixed(x, b, shape, a, omega, beta, m, p, 24, 1e-30)
        torch.cuda.synchronize()
        log("kernels workload done")
        return

    def solve(fn, rhs, tol):
        x.zero_()
        return fn(x, rhs, shape, a, omega, beta, m, p, max_iter, tol)

    def timed(fn, rhs, tol, steps, label):
        st = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            st = solve(fn, rhs, tol)
            log(f"{label} step {k}: {st['iterations']} its {st['solve_ms']:.0f} ms")
        torch.cuda.synchronize()
        assert st["converged"], st
        return (time.perf_counter() - t0) / steps, st

    def true_relres(rhs):
        r = rhs - aar_hip.Lap_Vec_mult(x, shape, a)
        return float(torch.linalg.norm(r) / torch.linalg.norm(rhs))

    def diff_mod_mean(u, v):
        u = u - u.mean()
        v = v - v.mean()
        return float(torch.linalg.norm(u - v) / torch.linalg.norm(v))

    # warm-up (allocations, code objects)
    for _ in range(args.warmup):
        st = solve(aar_hip.AAR, b, args.tol)
        log(f"warmup fp64: {st['iterations']} its relres {st['relres']:.3e} {st['solve_ms']:.0f} ms")
        st = solve(aar_hip.AAR_mixed, b, args.tol)
        log(f"warmup mixed: {st['iterations']} its relres {st['relres']:.3e} {st['solve_ms']:.0f} ms")

    # timed FP64 solves (the headline value)
    t64, st64 = timed(aar_hip.AAR, b, args.tol, args.steps, "fp64")
    relres_true = true_relres(b)
    x64 = x.clone()
    # the same solve in mixed precision
    tmx, stmx = timed(aar_hip.AAR_mixed, b, args.tol, args.steps, "mixed")
    relres_mx = true_relres(b)
    diff_mx = diff_mod_mean(x, x64)

    # the host-array entry points (AAR() of AAR.h) add a PCIe round trip:
    # x and rhs from pageable host memory in, x back out
    b_host = b.cpu()
    x_host = torch.zeros_like(b_host)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xd = x_host.to(dev)
    bd = b_host.to(dev)
    aar_hip.AAR(xd, bd, shape, a, omega, beta, m, p, max_iter, args.tol)
    x_host = xd.cpu()
    t_pcie = time.perf_counter() - t0
    del b_host, x_host, xd, bd
    log(f"pcie-inclusive fp64 solve: {t_pcie * 1e3:.0f} ms")

    # mixed vs FP64 to 1e-10 on b ~ N(0,1): Poisson (mean removed) and
    # screened Poisson, sigma = 0.5
    def pair_1e10(sigma):
        bn = P.rhs_normal_torch(shape, 1003, device=dev)
        if sigma == 0.0:
            bn -= bn.mean()

        def f64(xx, rhs, *args):
            return aar_hip.AAR(xx, rhs, *args, sigma=sigma)

        def fmx(xx, rhs, *args):
            return aar_hip.AAR_mixed(xx, rhs, *args, sigma=sigma)

        # the exact solution of the discrete system (device FFT)
        ue = P.spectral_solve_torch(bn, shape, a, sigma)
        nue = float(torch.linalg.norm(ue))

        def err_exact():
            xx = x - x.mean() if sigma == 0.0 else x
            return float(torch.linalg.norm(xx - ue)) / nue

        t64t, st64t = timed(f64, bn, 1e-10, 1, f"fp64 1e-10 sigma={sigma}")
        x64t = x.clone()
        e64 = err_exact()
        tmxt, stmxt = timed(fmx, bn, 1e-10, 1, f"mixed 1e-10 sigma={sigma}")
        emx = err_exact()
        r = bn - aar_hip.Lap_Vec_mult(x, shape, a, sigma)
        rel = float(torch.linalg.norm(r) / torch.linalg.norm(bn))
        diff = diff_mod_mean(x, x64t) if sigma == 0.0 else float(torch.linalg.norm(x - x64t) /
                                                                 torch.linalg.norm(x64t))
        del ue, x64t
        log(f"1e-10 sigma={sigma}: error vs exact fp64 {e64:.3e} mixed {emx:.3e}")
        return {"rhs": "N(0,1) seed 1003" + (", mean removed" if sigma == 0.0 else ""), "sigma": sigma,
                "fp64_solves_per_s": 1.0 / t64t, "fp64_iterations": st64t["iterations"],
                "mixed_solves_per_s": 1.0 / tmxt, "mixed_iterations": stmxt["iterations"],
                "speedup_vs_fp64": t64t / tmxt, "mixed_relres_true": rel,
                "fp64_err_vs_exact": e64, "mixed_err_vs_exact": emx,
                "mixed_rel_diff_vs_fp64" + ("_mod_mean" if sigma == 0.0 else ""): diff}

    tol10 = pair_1e10(0.0)
    tol10_screened = pair_1e10(0.5)

    # the comparison solvers on the same 512^3 system (one solve each; GPU
    # time of the solve from the library's events)
    others = {}
    for name, fn in (("pgr", lambda: aar_hip.PGR(x, b, shape, a, omega, m, p, max_iter, args.tol)),
                     ("pl2r", lambda: aar_hip.PL2R(x, b, shape, a, omega, m, p, max_iter, args.tol)),
                     ("cg", lambda: aar_hip.CG(x, b, shape, a, max_iter, args.tol)),
                     # the PETSc variant's Krylov comparators with its max_it 2000
                     # (PETSC_REAL/main.c:100,132)
                     ("bicg", lambda: aar_hip.BiCG(x, b, shape, a, 2000, args.tol)),
                     ("gmres30", lambda: aar_hip.GMRES(x, b, shape, a, 2000, args.tol, restart=30))):
        x.zero_()
        st = fn()
        others[name] = {"iterations": st["iterations"], "converged": bool(st["converged"]),
                        "relres": st["relres"], "solve_ms": st["solve_ms"],
                        "solves_per_s": 1e3 / st["solve_ms"],
                        "algorithmic_GBps": st["bytes"] / st["solve_ms"] / 1e6}
        if name in ("bicg", "gmres30"):
            others[name]["ksp_reason"] = st["reason"]
        log(f"{name}: {st['iterations']} its {st['solve_ms']:.0f} ms")

    # the variable-coefficient operator on the same grid and RHS
    petsc = None
    if not args.no_petsc:
        try:
            petsc = petsc_variant_case(dev, args.tol)
        except Exception as e:  # noqa: BLE001
            petsc = {"error": str(e)}

    vc = None
    if not args.no_vc:
        try:
            vc = varcoef_case(dev, shape, a, b, args.tol)
        except Exception as e:  # noqa: BLE001
            vc = {"error": str(e)}

    # differentiable complex Helmholtz (-Lap + c(x) + i eta) u = b, h = 0.25,
    # eta = 0.5: forward + adjoint solves through the autograd harness, and
    # the gradients against central differences (step 1e-6) on 20 directions
    helm = {}
    if not args.no_helm:
        for nh, tl in ((128, 1e-13), (256, 1e-13)):
            try:
                helm[str(nh)] = helmholtz_case(nh, dev, a, tol=tl)
            except Exception as e:  # noqa: BLE001
                helm[str(nh)] = {"error": str(e)}
            log(f"helmholtz {nh}^3: {helm[str(nh)]}")

    batch = None
    if not args.no_batch:
        try:
            batch = batch_case(dev)
        except Exception as e:  # noqa: BLE001
            batch = {"error": str(e)}

    # roofline of the dominant kernel: one profiled FP64 solve, HIP events
    # around every launch of the recurrence stencil sweep on the library stream
    aar_hip.profile(True)
    aar_hip.profile_reset()
    solve(aar_hip.AAR, b, args.tol)
    aar_hip.profile(False)
    names = ("stencil_aar_rec", "stencil_aar_true", "gram", "ls", "anchor_update", "dot")
    kern = {k: aar_hip.profile_query(k) for k in names}
    aar_hip.profile(True)
    aar_hip.profile_reset()
    solve(aar_hip.AAR_mixed, b, args.tol)
    aar_hip.profile(False)
    kern_mx = {k: aar_hip.profile_query(k) for k in names}
    kr = kern["stencil_aar_rec"]
    avg_ms = kr["ms"] / kr["count"]
    bytes_per_launch = kr["bytes"] / kr["count"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9

    # stencil matvec alone: y = a Lap x, 2*8*N algorithmic bytes
    y = torch.empty_like(b)
    aar_hip.profile(True)
    aar_hip.profile_reset()
    for _ in range(10):
        aar_hip.Lap_Vec_mult(b, shape, a, out=y)
    aar_hip.profile(False)
    mv = aar_hip.profile_query("stencil_matvec")
    mv_gbs = mv["bytes"] / mv["ms"] / 1e6
    del y

    traffic = None
    tf = os.path.join(REPO, "profiles", "traffic_stencil_aar_rec_512.json")
    if os.path.exists(tf) and n == 512:
        with open(tf) as fp:
            traffic = json.load(fp).get("bytes_per_launch")

    log(f"roofline: stencil_aar_rec {achieved:.0f} GB/s; matvec {mv_gbs:.0f} GB/s")

    # attainable HBM bandwidth on this device: streaming kernels with 16-byte
    # non-temporal accesses over 4 GiB (read-only, write-only, copy)
    del x64
    sb = torch.empty(4 << 27, dtype=torch.float64, device=dev)       # 4 GiB
    sb.zero_()
    sd = torch.empty_like(sb)
    attainable = {k: aar_hip.stream_bandwidth(k, 4 << 30, 10, dst=sd, src=sb) for k in ("read", "write", "copy")}
    del sb, sd
    log(f"attainable (stream kernels, 4 GiB): {attainable}")

    # the whole FP64 solve against the roofline: the library's algorithmic
    # byte count of the solve (per Anderson cycle at m = 7, p = 6: 5
    # recurrence sweeps x 2 vectors, the Gram 8, the anchor update 12 (x read
    # and written, 8 residual slots, one correction read and one written),
    # the true residual 3; 33 vectors) over the solve's wall time
    solve_gbs = st64["bytes"] / t64 / 1e9

    # GPU FP64 AAR at the CPU baseline's sizes (same problem scheme)
    gpu_rates = {}
    for side in (128, 256):
        sh = (side,) * 3
        bs_ = P.rhs_uniform_torch(sh, 1002, device=dev)
        xs_ = torch.zeros_like(bs_)
        aar_hip.AAR(xs_, bs_, sh, a, omega, beta, m, p, max_iter, args.tol)
        ts = []
        for _ in range(3):
            xs_.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            aar_hip.AAR(xs_, bs_, sh, a, omega, beta, m, p, max_iter, args.tol)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        gpu_rates[side] = 1.0 / sorted(ts)[1]
        del bs_, xs_
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(gpu_rates, n, st64["iterations"], args.cpu_ranks or None)
        except Exception as e:  # noqa: BLE001
            log(f"cpu baseline failed: {e}")
            cpu = {"error": str(e), "value": None}

    def ktab(kk):
        return {k: {"count": v["count"], "avg_ms": (v["ms"] / v["count"]) if v["count"] else 0.0,
                    "GBps": (v["bytes"] / v["ms"] / 1e6) if v["ms"] > 0 else 0.0}
                for k, v in kk.items()}

    out = {
        "metric": "AAR solves/s at 512^3 to rel-res 1e-6: FP32-refined vs FP64; matvec HBM GB/s",
        "value": 1.0 / t64,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t64 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded U[-0.5,0.5] RHS, mean removed)",
        "config": {"workload": f"FD12 periodic Poisson {n}^3, h=0.25, AAR m=7 p=6 omega=beta=0.6, "
                               f"x0=0, tol {args.tol:g}", "grid": [n, n, n], "points": N},
        "iterations": st64["iterations"],
        "relres": st64["relres"],
        "relres_true": relres_true,
        "pcie_inclusive": {"solves_per_s": 1.0 / t_pcie, "seconds": t_pcie,
                           "note": "x and rhs copied in from pageable host memory and x out, as the "
                                   "reference-signature AAR() does; not the headline value"},
        "fp32_refined": {
            "solver": "mixed-precision AAR: FP32 Richardson sweeps, residual history and Gram reads; "
                      "FP64 true residual b-Ax every Anderson cycle, FP64 iterate and normal equations",
            "solves_per_s": 1.0 / tmx, "ms_per_step": tmx * 1e3, "speedup_vs_fp64": t64 / tmx,
            "iterations": stmx["iterations"], "relres_true": relres_mx,
            "rel_diff_vs_fp64_mod_mean": diff_mx,
            "tol_1e-10": tol10,
            "tol_1e-10_screened": tol10_screened,
            "kernels": ktab(kern_mx),
        },
        "other_solvers": others,
        "variable_coefficient": vc,
        "petsc_variant": petsc,
        "helmholtz": helm,
        "batch": batch,
        "matvec": {"GBps": mv_gbs, "frac_of_peak": mv_gbs / HBM_PEAK_GBS,
                   "avg_ms": mv["ms"] / mv["count"], "bytes_per_launch": 16.0 * N},
        "kernels": ktab(kern),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "attainable": attainable,
                     "frac_of_attainable_read": achieved / attainable["read"],
                     "kernel": "stencil_aar_rec (FD12 stencil with the Richardson step folded in: "
                               "g' = g - (omega/d) A g; algorithmic 2x8 B/pt)",
                     "avg_ms": avg_ms, "bytes_per_launch": bytes_per_launch},
        "solve_roofline": {"algorithmic_bytes": st64["bytes"], "seconds": t64, "GBps": solve_gbs,
                           "frac": solve_gbs / HBM_PEAK_GBS,
                           "byte_model": "per Anderson cycle (m=7, p=6): 5 recurrence sweeps x 2 vectors, "
                                         "Gram 8, anchor update 12, true residual 3 = 33 vectors of 8 B/pt"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
