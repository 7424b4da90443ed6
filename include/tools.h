/*
 * tools.h -- drop-in replacement for MPI_REAL/tools.h, served by libaar_hip.
 * Norms and dots run as deterministic device reductions; the pseudo-inverse
 * runs the Golub-Reinsch SVD on the GPU (one wavefront), bit-faithful to
 * MPI_REAL/tools.c:52-352.
 */
#ifndef AAR_HIP_TOOLS_H
#define AAR_HIP_TOOLS_H

#include <mpi.h>
#include <math.h>
#include <stdlib.h>
#include <stdio.h>

#ifndef SIGN
#define SIGN(a, b) ((b) >= 0.0 ? fabs(a) : -fabs(a))
#endif
#ifndef max
#define max(a, b) ((a) > (b) ? (a) : (b))
#endif
#ifndef min
#define min(a, b) ((a) < (b) ? (a) : (b))
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* MPI_REAL/tools.h:24 -- sqrt(a^2 + b^2) without destructive over/underflow */
double pythag(double a, double b);

/* MPI_REAL/tools.h:25 -- square m == n <= 16: a <- V, w <- truncated singular values, v <- V */
void SingularValueDecomp(double **a, int m, int n, double *w, double **v);

/* MPI_REAL/tools.h:26 -- x = V diag(1/w) V^T b, m <= 16 */
void PseudoInverseTimesVec(double **A, double *b, double *x, int m);

/* MPI_REAL/tools.h:27 -- host vector in, GPU reduction */
void Vector2Norm(double *Vec, int len, double *ResVal, MPI_Comm comm);

/* MPI_REAL/tools.h:28 */
void VectorDotProduct(double *Vec1, double *Vec2, int len, double *ResVal, MPI_Comm comm);

#ifdef __cplusplus
}
#endif
#endif
