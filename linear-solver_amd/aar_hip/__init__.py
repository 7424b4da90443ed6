"""aar_hip -- MI355X-native Anderson-accelerated Richardson solvers (AAR, PGR,
PL2R; CG as comparator) for the FD12 periodic Poisson family, over the C ABI
of libaar_hip.so.  See solvers.py for the reference-shaped interface."""
from .solvers import (AAR, CG, GMRES, ILU0, KSPCG, PGR, PL2R, AAR_batch, AAR_petsc, BiCG, PGR_petsc, ilu0_apply, AAR_mixed, AAR_refined, AAR_vc, Lap_Vec_mult,  # noqa: F401
                      Lap_Vec_mult_vc, autotune, history, last_system, pinv, profile, profile_query, profile_reset,
                      release, set_coefficient, set_option, set_stencil, set_tuning, stream_bandwidth, FDN_USER)
from . import problems  # noqa: F401
from .helmholtz import (HelmholtzSolve, helm_aar, helm_bicg, helm_gmres, helm_grad_c, helm_ilu0_apply, helm_lgmres,  # noqa: F401
                        helm_matvec)
