"""Write tests/golden/golden_deep.{json,npz}: the deep Anderson histories
(16 < m <= 32) from the REAL reference solvers.

The reference takes any history depth m (MPI_REAL/AAR.c:51-73,
PGR.c:50-72, PL2R.c:50-72).  This runs oracle/_ref/ref_harness (the
reference sources compiled by oracle/Makefile) for AAR, PGR and PL2R at
m = 20 and 32, with p below and above m + 1, on seeded inputs, and stores
iteration counts, printed conditions, residual traces and solutions, with
the same solves on 2, 4 and 8 ranks (other dot summation orders) as the
reference's own rounding spread; each case's counts are the same on every
rank count.  It also stores tools.c pseudo-inverse known answers at n = 24
and 32.

    python tests/golden/make_golden_deep.py

Only runs where the reference tree and MPICH exist (this build container).
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "linear-solver_amd"))

from oracle import oracle as O  # noqa: E402
from aar_hip import problems as P  # noqa: E402

SOLVERS = ("aar", "pgr", "pl2r")


def cases():
    base = dict(omega=0.6, beta=0.6, sigma=0.0, precond=1, max_iter=2000, seed=1001)
    return {
        "m20p6_24": dict(base, shape=[24, 24, 24], m=20, p=6, tol=1e-10),
        "m32p4_24": dict(base, shape=[24, 24, 24], m=32, p=4, tol=1e-10),
        "m20p24_24": dict(base, shape=[24, 24, 24], m=20, p=24, tol=1e-8),
        "m32p11_28x24x20": dict(base, shape=[28, 24, 20], m=32, p=11, tol=1e-10, omega=1.0, beta=0.9),
    }


def case_inputs(c):
    shape = tuple(c["shape"])
    b = P.rhs_uniform(shape, c["seed"])
    return shape, b, np.zeros(b.size)


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/ref_harness is not built (make -C oracle ref)")
    meta = {"cases": {}, "kats": {}}
    arrays = {}
    for name, c in cases().items():
        shape, b, x0 = case_inputs(c)
        a = P.poisson_a(0.5)
        entry = {"params": dict(c, a=a), "rhs_norm": float(np.linalg.norm(b)), "solvers": {}}
        for sv in SOLVERS:
            x, d = O.ref_solve(sv, shape, b, x0, a, c["sigma"], c["omega"], c["beta"], c["m"], c["p"],
                               c["max_iter"], c["tol"], bool(c["precond"]))
            # the reference's own rounding spread: on 2, 4 and 8 ranks every
            # dot is summed in another order (MPI_Allreduce of per-rank
            # partials); the largest deviation from the 1-rank run, entry by
            # entry for the trace
            sol_spread, tr_spread = 0.0, np.zeros_like(d["trace"])
            for ranks in (2, 4, 8):
                xr, dr = O.ref_solve(sv, shape, b, x0, a, c["sigma"], c["omega"], c["beta"], c["m"], c["p"],
                                     c["max_iter"], c["tol"], bool(c["precond"]), ranks=ranks)
                assert dr["iterations"] == d["iterations"], (name, sv, ranks, dr["iterations"], d["iterations"])
                sol_spread = max(sol_spread, float(np.linalg.norm(xr - x) / np.linalg.norm(x)))
                tr_spread = np.maximum(tr_spread, np.abs(dr["trace"] - d["trace"]))
            entry["solvers"][sv] = {"iterations": d["iterations"], "relres_printed": d["relres_printed"],
                                    "converged_line": d["converged_line"], "warning_line": d["warning_line"],
                                    "rank_spread_rel": sol_spread}
            arrays[f"{name}__{sv}__x"] = x
            arrays[f"{name}__{sv}__trace"] = d["trace"]
            arrays[f"{name}__{sv}__trace_spread"] = tr_spread
            print(name, sv, d["iterations"], d["relres_printed"], "rank spread", sol_spread,
                  "max trace spread", float(np.max(tr_spread / d["trace"])), flush=True)
        meta["cases"][name] = entry
    # tools.c pseudo-inverse known answers at the deep sizes: an SPD normal
    # matrix and a rank-deficient one (zero row and column)
    rng = np.random.default_rng(24032)
    ms, As, bs, xs = [], [], [], []
    for m in (24, 32):
        F = rng.standard_normal((3 * m, m))
        G = F.T @ F
        for A in (G, np.where((np.arange(m)[:, None] == m - 2) | (np.arange(m)[None, :] == m - 2), 0.0, G)):
            bb = rng.standard_normal(m)
            ms.append(m)
            As.append(A.reshape(-1))
            bs.append(bb)
            xs.append(O.ref_pinv(A, bb))
    arrays["pinv_m"] = np.array(ms, dtype=np.int64)
    arrays["pinv_A"] = np.concatenate(As)
    arrays["pinv_b"] = np.concatenate(bs)
    arrays["pinv_x"] = np.concatenate(xs)
    with open(os.path.join(HERE, "golden_deep.json"), "w") as fp:
        json.dump(meta, fp, indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "golden_deep.npz"), **arrays)


if __name__ == "__main__":
    main()
