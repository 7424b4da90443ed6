/*
 * PGR.h -- drop-in replacement for MPI_REAL/PGR.h, served by libaar_hip.
 * Same callback rules and device residency as AAR() (AAR.h).
 */
#ifndef AAR_HIP_PGR_H
#define AAR_HIP_PGR_H

#include "system.h"
#include "tools.h"

#ifdef __cplusplus
extern "C" {
#endif

/* MPI_REAL/PGR.h:18-22 */
void PGR(POISSON *system,
         void (*Lap_Vec_mult)(POISSON *, double, double *, double *, MPI_Comm),
         void (*Precondition)(double, double *, int), double a,
         double *x, double *rhs, double omega, int m, int p,
         int max_iter, double tol, int Np, MPI_Comm comm_dist_graph_cart);

/* MPI_REAL/PGR.h:24-25 -- svec = pinv(DX^T DF) (DX^T f) from host arrays */
void Galerkin_Richardson(double **DX, double **DF, double *f, int m, int Np,
                         double **XtF, double *allredvec, double *Xtf, double *svec, MPI_Comm comm);

#ifdef __cplusplus
}
#endif
#endif
