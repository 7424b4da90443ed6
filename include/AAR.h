/*
 * AAR.h -- drop-in replacement for MPI_REAL/AAR.h, served by libaar_hip.
 *
 * AAR() runs the whole Alternating Anderson-Richardson solve on the GPU:
 * x and rhs are copied to HBM on entry and x back on exit.  The operator
 * and preconditioner slots must hold libaar_hip's own Lap_Vec_mult and
 * Precondition (or NULL for no preconditioner); any other callback is
 * rejected with a message and x is left unchanged (no host execution).
 * Reports are printed as the reference prints them (AAR.c:151-159); the
 * numbers are also available from aar_hip_last_stats() (aar_hip.h).
 */
#ifndef AAR_HIP_AAR_H
#define AAR_HIP_AAR_H

#include "system.h"
#include "tools.h"

#ifdef __cplusplus
extern "C" {
#endif

/* MPI_REAL/AAR.h:18-22 */
void AAR(POISSON *system,
         void (*Lap_Vec_mult)(POISSON *, double, double *, double *, MPI_Comm),
         void (*Precondition)(double, double *, int), double a,
         double *x, double *rhs, double omega, double beta, int m, int p,
         int max_iter, double tol, int Np, MPI_Comm comm);

/* MPI_REAL/AAR.h:24-25 -- host arrays DX[m][N], DF[m][N], f[N]; the Gram,
 * the m x m solve and am_vec are computed on the GPU.  FtF, allredvec,
 * Ftf, svec receive the same contents as in the reference. */
void AndersonExtrapolation(double **DX, double **DF, double *f, double beta_mix, int m,
                           int N, double *am_vec, double **FtF, double *allredvec, double *Ftf,
                           double *svec, MPI_Comm comm);

#ifdef __cplusplus
}
#endif
#endif
