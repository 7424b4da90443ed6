/*
 * PL2R.h -- drop-in replacement for MPI_REAL/PL2R.h, served by libaar_hip.
 * Same callback rules and device residency as AAR() (AAR.h).
 */
#ifndef AAR_HIP_PL2R_H
#define AAR_HIP_PL2R_H

#include "system.h"
#include "tools.h"

#ifdef __cplusplus
extern "C" {
#endif

/* MPI_REAL/PL2R.h:18-22 */
void PL2R(POISSON *system,
          void (*Lap_Vec_mult)(POISSON *, double, double *, double *, MPI_Comm),
          void (*Precondition)(double, double *, int), double a,
          double *x, double *rhs, double omega, int m, int p,
          int max_iter, double tol, int Np, MPI_Comm comm_dist_graph_cart);

/* MPI_REAL/PL2R.h:24-25 -- svec = pinv(DF^T DF) (DF^T f) from host arrays */
void L2_Richardson(double **DF, double *f, int m, int Np, double **FtF,
                   double *allredvec, double *Ftf, double *svec, MPI_Comm comm);

#ifdef __cplusplus
}
#endif
#endif
