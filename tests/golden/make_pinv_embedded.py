"""Regenerate pinv_embedded_{A,r}.npy on a GPU host (not part of the tests).

The 14 x 14 real embedding [Gr -Gi; Gi Gr] of the Hermitian Anderson system
(m = 7) met by the complex Helmholtz solver at 256^3 (h = 0.25, eta = 0.5,
c = helmholtz_coefficient(seed 1005), tol 1e-12) for the right-hand side
b + 1e-6 db of the fourth finite-difference direction (torch generator seed
2024).  On it the Golub-Reinsch SVD in the order of MPI_REAL/tools.c:137
does not converge in 250 sweeps, the textbook order does.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "linear-solver_amd"))


def main():
    import torch
    import aar_hip
    from aar_hip import problems as P
    n = 256
    shape = (n,) * 3
    N = n ** 3
    c = torch.from_numpy(P.helmholtz_coefficient(shape, 1005)).to("cuda")
    v = P.normal_torch(2 * N, 1005, 0, "cuda")
    b = torch.complex(v[0::2].contiguous(), v[1::2].contiguous())
    gen = torch.Generator(device="cuda")
    gen.manual_seed(2024)
    for _ in range(4):
        db = torch.complex(torch.randn(N, dtype=torch.float64, device="cuda", generator=gen),
                           torch.randn(N, dtype=torch.float64, device="cuda", generator=gen))
        torch.randn(N, dtype=torch.float64, device="cuda", generator=gen)
    u = torch.zeros_like(b)
    aar_hip.set_option("complex_pinv_reference_order", 1)   # the tools.c order, to meet the failure
    try:
        aar_hip.helm_aar(u, b + 1e-6 * db, shape, P.poisson_a(0.25), c, 0.0, 0.5, 0.6, 0.6, 7, 6, 20000, 1e-12)
    except Exception as e:  # noqa: BLE001
        print("solver stopped:", e)
    A, r, _ = aar_hip.last_system(14)
    np.save(os.path.join(HERE, "pinv_embedded_A.npy"), A)
    np.save(os.path.join(HERE, "pinv_embedded_r.npy"), r)


if __name__ == "__main__":
    main()
