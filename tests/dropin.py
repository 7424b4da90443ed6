"""ctypes bindings of libaar_hip's reference-signature entry points
(include/AAR.h, PGR.h, PL2R.h, CG.h, system.h, tools.h) for the tests, and
a runner that calls one solver on host arrays in a child process, so that
the reference's printed lines and its process exit (tools.c:270 ends the
process when the SVD does not converge) can be compared as they are.

    python tests/dropin.py <golden_edge case> <aar|pgr|pl2r|cg>
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
for p_ in (REPO, os.path.join(REPO, "linear-solver_amd")):
    if p_ not in sys.path:
        sys.path.insert(0, p_)

I, D, V = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
DP = ctypes.POINTER(ctypes.c_double)
DPP = ctypes.POINTER(DP)


class POISSON(ctypes.Structure):
    """MPI_REAL/system.h:26-63 under MPICH (MPI_Comm is an int)."""
    P_ = ctypes.c_void_p
    _fields_ = [("FDn", I), ("ssize", I * 3), ("psize", I * 3), ("np", I * 3), ("coords", I * 3),
                ("rem", I * 3), ("send_neighs", P_), ("rec_neighs", P_), ("send_counts", P_),
                ("rec_counts", P_), ("send_layers", I * 6), ("rec_layers", I * 6), ("scounts", P_),
                ("sdispls", P_), ("rcounts", P_), ("rdispls", P_), ("n_in", I), ("n_out", I), ("phi", P_),
                ("rhs", P_), ("Lap_phi", P_), ("solver_tol", D), ("solver_maxiter", I), ("m", I), ("p", I),
                ("omega", D), ("beta", D), ("coeff_lap", P_), ("comm_laplacian", I), ("cart", I)]


def bind(L):
    """argtypes of the reference-signature entry points; returns the
    addresses of the library's operator / preconditioner slots."""
    S = ctypes.POINTER(POISSON)
    sigs = {
        "AAR": [S, V, V, D, V, V, D, D, I, I, I, D, I, I],
        "PGR": [S, V, V, D, V, V, D, I, I, I, D, I, I],
        "PL2R": [S, V, V, D, V, V, D, I, I, I, D, I, I],
        "CG": [S, V, V, D, V, V, I, D, I, I],
        "Lap_Vec_mult": [S, D, V, V, I],
        "Precondition": [D, V, I],
        "Vector2Norm": [V, I, DP, I],
        "VectorDotProduct": [V, V, I, DP, I],
        "SingularValueDecomp": [DPP, I, I, V, DPP],
        "PseudoInverseTimesVec": [DPP, V, V, I],
        "Galerkin_Richardson": [DPP, DPP, V, I, I, DPP, V, V, V, I],
        "L2_Richardson": [DPP, V, I, I, DPP, V, V, V, I],
    }
    for k, a in sigs.items():
        fn = getattr(L, k)
        fn.argtypes = a
        fn.restype = None
    L.pythag.argtypes = [D, D]
    L.pythag.restype = D
    return {k: ctypes.cast(getattr(L, k), V).value for k in ("Lap_Vec_mult", "Precondition")}


def system(shape, coeff):
    s = POISSON()
    s.FDn = 6
    for d in range(3):
        s.ssize[d] = s.psize[d] = shape[d]
        s.np[d] = 1
    s.coeff_lap = coeff.ctypes.data_as(ctypes.c_void_p)
    return s


def rows(M):
    """double** over the rows of a C-ordered 2-D array (kept alive by M)."""
    r = (DP * M.shape[0])()
    for i in range(M.shape[0]):
        r[i] = M[i].ctypes.data_as(DP)
    return r


def main():
    case, solver = sys.argv[1], sys.argv[2]
    from aar_hip import _lib, problems as P
    L = _lib.load()
    fp = bind(L)
    g = np.load(os.path.join(HERE, "golden", "golden_edge.npz"), allow_pickle=False)
    b = np.ascontiguousarray(g[f"nf_{case}_b"])
    x = np.ascontiguousarray(g[f"nf_{case}_x0"]).copy()
    x0 = x.copy()
    shape = (16, 16, 16)
    n = x.size
    coeff = np.array(P.lap_weights(6))
    sysd = system(shape, coeff)
    a = P.poisson_a(0.5)
    args = (ctypes.byref(sysd), fp["Lap_Vec_mult"], fp["Precondition"], a, x.ctypes.data, b.ctypes.data)
    if solver == "aar":
        L.AAR(*args, 0.6, 0.6, 7, 6, 1000, 1e-6, n, 0)
    elif solver == "pgr":
        L.PGR(*args, 0.6, 7, 6, 1000, 1e-6, n, 0)
    elif solver == "pl2r":
        L.PL2R(*args, 0.6, 7, 6, 1000, 1e-6, n, 0)
    else:
        L.CG(*args, 1000, 1e-6, n, 0)
    sys.stdout.flush()
    libc = ctypes.CDLL(None)
    libc.fflush(None)
    print("RESULT " + json.dumps({"x_nan_count": int(np.isnan(x).sum()),
                                  "x_equals_x0": bool(np.array_equal(x, x0, equal_nan=True))}), flush=True)


if __name__ == "__main__":
    main()
