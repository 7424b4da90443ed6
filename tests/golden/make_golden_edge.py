"""Write tests/golden/golden_edge.{json,npz}: reference behaviour on the edges
of the interface, from the REAL reference (oracle/_ref, built from the
reference's own sources by oracle/Makefile).

* The stand-alone projection steps Galerkin_Richardson (PGR.h:24-25) and
  L2_Richardson (PL2R.h:24-25): svec and allredvec for seeded DX, DF, f
  (m = 7 with two never-written columns, m = 13).
* Non-finite inputs (SURVEY pitfall: the stop rule `relres > tol` of
  AAR.c:91 is false for NaN): a NaN and an Inf in the right-hand side, and
  a NaN in the initial guess, for AAR, PGR, PL2R and CG on 16^3 -- the
  printed outcome, the reported iteration count (None: nothing printed)
  and the returned x.
* A wrong argument count of the reference driver (system.c:37-39): its
  stdout (the time line masked) and exit status, for several argc.

    python tests/golden/make_golden_edge.py

Only runs where the reference tree and MPICH exist (this build container).
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "linear-solver_amd"))

from oracle import oracle as O  # noqa: E402
from aar_hip import problems as P  # noqa: E402

REF_MAIN = os.path.join(REPO, "oracle", "_ref", "ref_main")


def mask_time(out):
    return re.sub(r"Starting time: .*", "Starting time: <masked>", out)


def nonfinite_cases():
    shape = (16, 16, 16)
    n = 16 ** 3
    b = P.rhs_uniform(shape, 1001)
    out = {}
    bn = b.copy()
    bn[100] = np.nan
    out["rhs_nan"] = (bn, np.zeros(n))
    bi = b.copy()
    bi[7] = np.inf
    out["rhs_inf"] = (bi, np.zeros(n))
    x0 = np.zeros(n)
    x0[5] = np.nan
    out["x0_nan"] = (b.copy(), x0)
    return shape, out


def user_weights():
    """A caller's own radius-3 stencil: the mean of the FD2 and FD6 weights
    (a convex combination of two Laplacians, so the operator stays negative
    semi-definite with a zero row sum) -- not one of the reference's orders."""
    c1 = O.lap_coefficient(1)
    c3 = O.lap_coefficient(3)
    return [0.5 * (c3[o] + (c1[o] if o <= 1 else 0.0)) for o in range(4)]


def main():
    if not O.ref_available() or not os.path.exists(REF_MAIN):
        raise SystemExit("build the reference first: make -C oracle ref ref_main")
    meta = {"steps": {}, "nonfinite": {}, "argc": {}}
    arrays = {}
    rng = np.random.default_rng(77)
    for m, n in ((7, 4096), (13, 3000)):
        DX = rng.standard_normal((m, n))
        DF = rng.standard_normal((m, n))
        if m == 7:
            DF[5:] = 0.0                       # two never-written columns (rank-deficient system)
            DX[5:] = 0.0
        f = rng.standard_normal(n)
        key = f"m{m}"
        arrays[f"step_{key}_DX"], arrays[f"step_{key}_DF"], arrays[f"step_{key}_f"] = DX, DF, f
        for kind in ("galerkin", "l2"):
            sv, red = O.ref_step(kind, DX, DF, f)
            arrays[f"step_{key}_{kind}_svec"] = sv
            arrays[f"step_{key}_{kind}_red"] = red
            print("step", key, kind, sv[:3], flush=True)
    meta["steps"] = {"cases": ["m7", "m13"]}

    shape, cases = nonfinite_cases()
    a = P.poisson_a(0.5)
    for name, (b, x0) in cases.items():
        arrays[f"nf_{name}_b"] = b
        arrays[f"nf_{name}_x0"] = x0
        ent = {}
        for sv in ("aar", "pgr", "pl2r", "cg"):
            try:
                x, d = O.ref_solve(sv, shape, b, x0, a, 0.0, 0.6, 0.6, 7, 6, 1000, 1e-6, True)
            except subprocess.CalledProcessError as e:
                # the reference ended the process (tools.c:270 exit(1) when the
                # SVD of a non-finite Gram does not converge)
                ent[sv] = {"exit_status": e.returncode, "stdout": e.stdout}
                print("nonfinite", name, sv, "exit", e.returncode, repr(e.stdout), flush=True)
                continue
            lines = [ln for ln in d["stdout"].splitlines() if not ln.startswith("HARNESS")]
            ent[sv] = {"exit_status": 0, "iterations": d["iterations"], "converged_line": d["converged_line"],
                       "warning_line": d["warning_line"], "stdout": "\n".join(lines),
                       "x_nan_count": int(np.isnan(x).sum()), "x_equals_x0": bool(np.array_equal(x, x0,
                                                                                               equal_nan=True))}
            arrays[f"nf_{name}_{sv}_x"] = x
            print("nonfinite", name, sv, ent[sv]["iterations"], ent[sv]["x_nan_count"], repr(ent[sv]["stdout"]),
                  flush=True)
        meta["nonfinite"][name] = ent

    # FD orders 2..12 (FDn = 1..6, system.c:480-498) and a caller's own
    # weights in coeff_lap (the reference's extension point, README:25): the
    # reference's stencil on a ragged grid, and its solvers on 16^3
    rng2 = np.random.default_rng(91)
    mshape = (13, 10, 14)
    xm = rng2.standard_normal(int(np.prod(mshape)))
    arrays["fdn_matvec_x"] = xm
    user = user_weights()
    arrays["user_coeff"] = np.array(user)
    meta["fdn"] = {"matvec_shape": list(mshape), "solve_shape": [16, 16, 16], "user_fdn": len(user) - 1,
                   "solves": {}}
    for fdn in range(1, 7):
        arrays[f"fdn{fdn}_matvec_y"] = O.ref_matvec(mshape, xm, -4.0, fdn=fdn)
    arrays["user_matvec_y"] = O.ref_matvec(mshape, xm, -4.0, fdn=len(user) - 1, coeff=user)
    sshape = (16, 16, 16)
    bs = P.rhs_uniform(sshape, 1001)
    arrays["fdn_solve_b"] = bs
    for label, fdn, coeff in (("fdn2", 2, None), ("fdn4", 4, None), ("user", len(user) - 1, user)):
        for sv in ("aar", "pgr", "pl2r", "cg"):
            x, d = O.ref_solve(sv, sshape, bs, np.zeros(bs.size), -4.0, 0.0, 0.6, 0.6, 7, 6, 1000, 1e-8, True,
                               fdn=fdn, coeff=coeff)
            meta["fdn"]["solves"][f"{label}_{sv}"] = {"iterations": d["iterations"],
                                                      "converged_line": d["converged_line"]}
            arrays[f"{label}_{sv}_x"] = x
            print("fdn", label, sv, d["iterations"], flush=True)

    with tempfile.TemporaryDirectory() as tmp:
        for args in ([], ["1e-6"], ["1e-6", "7", "6", "0.6"], ["1e-6", "7", "6", "0.6", "0.6", "1"]):
            r = subprocess.run([REF_MAIN] + args, capture_output=True, text=True, timeout=120, cwd=tmp)
            meta["argc"][str(len(args) + 1)] = {"args": args, "stdout": mask_time(r.stdout), "rc": r.returncode}
            print("argc", len(args) + 1, r.returncode, repr(mask_time(r.stdout)[-60:]), flush=True)

    np.savez_compressed(os.path.join(HERE, "golden_edge.npz"), **arrays)
    with open(os.path.join(HERE, "golden_edge.json"), "w") as fp:
        json.dump(meta, fp, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "golden_edge.json"))


if __name__ == "__main__":
    main()
