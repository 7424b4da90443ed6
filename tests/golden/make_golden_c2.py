"""Write tests/golden/golden_c2.{json,npz}: the REAL reference solvers on the
C2 problem at 128^3 and 256^3.

The problem is FD12 periodic Poisson with h = 0.25 (a = -16), b ~ U[-0.5, 0.5]
from seed 1002 with its mean removed, x0 = 0, Jacobi, omega = beta = 0.6,
m = 7, p = 6, tol 1e-6.  oracle/_ref/ref_harness (the reference MPI_REAL
sources compiled by oracle/Makefile) runs each solver with 8 MPI ranks and,
at 128^3, also with 1 rank.  Stored per run: the printed iteration count and
relative residual, the residual trace the harness records
(MPI_REAL/AAR.c:119-123 evaluates b - A x on the same operands), the
solution's norm and sum, and strided solution samples.  The 1-rank vs
8-rank relative difference of the solution is the reference's own
rounding-noise floor at that size.

    python tests/golden/make_golden_c2.py

Only runs where the reference tree and MPICH exist (this build container).
The 256^3 runs take a few minutes each on 8 cores.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "linear-solver_amd"))

from oracle import oracle as O  # noqa: E402
from aar_hip import problems as P  # noqa: E402

PARAMS = dict(a=P.poisson_a(0.25), sigma=0.0, omega=0.6, beta=0.6, m=7, p=6, tol=1e-6,
              max_iter=20000, precond=1, seed=1002, rhs="uniform", x0=0.0)

# (side, solvers, rank counts); the one-rank runs give the noise floor
RUNS = [(128, ("aar", "pgr", "pl2r", "cg"), (8, 1)),
        (256, ("aar", "cg"), (8,))]


def stride(n):
    return 997 if n <= 128 ** 3 else 4099


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/ref_harness is not built (make -C oracle ref)")
    out_json = os.path.join(HERE, "golden_c2.json")
    meta = {"params": PARAMS, "cases": {}}
    arrays = {}
    for side, solvers, ranks in RUNS:
        shape = (side, side, side)
        b = P.rhs_uniform(shape, PARAMS["seed"])
        x0 = np.zeros(b.size)
        st = stride(b.size)
        name = f"c2_{side}"
        entry = {"shape": list(shape), "stride": st, "rhs_sum": float(np.sum(b)),
                 "rhs_norm": float(np.linalg.norm(b)), "solvers": {}}
        for sv in solvers:
            r = {}
            xs = {}
            for nr in ranks:
                t0 = time.time()
                x, d = O.ref_solve(sv, shape, b, x0, PARAMS["a"], PARAMS["sigma"], PARAMS["omega"],
                                   PARAMS["beta"], PARAMS["m"], PARAMS["p"], PARAMS["max_iter"],
                                   PARAMS["tol"], True, ranks=nr, timeout=7200)
                wall = time.time() - t0
                xs[nr] = x
                key = f"rank{nr}"
                r[key] = {"iterations": d["iterations"], "relres_printed": d["relres_printed"],
                          "converged_line": d["converged_line"], "warning_line": d["warning_line"],
                          "x_norm": float(np.linalg.norm(x)), "x_sum": float(np.sum(x)),
                          "x_mean": float(np.mean(x)), "seconds_solver": d["seconds"],
                          "seconds_wall": wall, "trace_len": int(d["trace"].size)}
                arrays[f"{name}__{sv}__rank{nr}__trace"] = d["trace"]
                arrays[f"{name}__{sv}__rank{nr}__xs"] = x[::st].copy()
                print(name, sv, nr, d["iterations"], d["relres_printed"], f"{wall:.1f}s", flush=True)
            if 1 in xs and 8 in xs:
                x1, x8 = xs[1], xs[8]
                r["rank_spread_rel"] = float(np.linalg.norm(x8 - x1) / np.linalg.norm(x1))
                print(name, sv, "rank spread", r["rank_spread_rel"], flush=True)
            entry["solvers"][sv] = r
            meta["cases"][name] = entry
            # written after every solver, so a long run keeps what it has
            np.savez_compressed(os.path.join(HERE, "golden_c2.npz"), **arrays)
            with open(out_json, "w") as fp:
                json.dump(meta, fp, indent=1, sort_keys=True)
    print("wrote", out_json)


if __name__ == "__main__":
    main()
