/*
 * CG.h -- drop-in replacement for MPI_REAL/CG.h, served by libaar_hip.
 * Jacobi-preconditioned CG with an explicit residual every 50 iterations,
 * all on the GPU.  Same callback rules as AAR() (AAR.h).
 */
#ifndef AAR_HIP_CG_H
#define AAR_HIP_CG_H

#include "system.h"
#include "tools.h"

#ifdef __cplusplus
extern "C" {
#endif

/* MPI_REAL/CG.h:18-21 */
void CG(POISSON *system,
        void (*Lap_Vec_mult)(POISSON *, double, double *, double *, MPI_Comm),
        void (*Precondition)(double, double *, int), double a,
        double *x, double *b, int max_iter, double tol, int Np, MPI_Comm comm);

#ifdef __cplusplus
}
#endif
#endif
