/*
 * system.h -- drop-in replacement for MPI_REAL/system.h of the reference,
 * served by libaar_hip.  A caller written against the reference (its
 * main.c) compiles unchanged against this header and links libaar_hip in
 * place of system.o/tools.o/AAR.o/PGR.o/PL2R.o/CG.o.
 *
 * The POISSON descriptor keeps the reference's field order and types
 * (MPI_REAL/system.h:26-63) because callers allocate and fill it
 * themselves; libaar_hip reads FDn, psize/ssize and coeff_lap from it.
 * libaar_hip is single-rank and single-GPU: the MPI fields are carried but
 * never used by the library (it makes no MPI calls), the decomposition
 * routines produce the one-rank layout, and the whole grid is resident in
 * HBM during a solve.
 */
#ifndef AAR_HIP_SYSTEM_H
#define AAR_HIP_SYSTEM_H

#include <math.h>
#include <time.h>
#include <stdio.h>
#include <stdlib.h>
#include <mpi.h>
#include <assert.h>

#ifndef Min
#define Min(a, b) ((a) < (b) ? (a) : (b))
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* MPI_REAL/system.h:26-63 */
typedef struct {
    int FDn;                 /* stencil half-order: 6 for the 12th-order Laplacian */
    int ssize[3];            /* global grid points per axis */
    int psize[3];            /* this rank's block (== ssize on one rank) */

    int np[3];               /* process grid (1,1,1) */
    int coords[3];
    int rem[3];

    int *send_neighs;
    int *rec_neighs;
    int *send_counts;
    int *rec_counts;
    int send_layers[6];
    int rec_layers[6];

    int *scounts;
    int *sdispls;
    int *rcounts;
    int *rdispls;
    int n_in;
    int n_out;

    double *phi;             /* solution / initial guess */
    double *rhs;             /* right-hand side */
    double *Lap_phi;         /* operator output scratch */

    double solver_tol;
    int solver_maxiter;
    int m;
    int p;
    double omega;
    double beta;

    double *coeff_lap;       /* FD weights c_0..c_FDn */
    MPI_Comm comm_laplacian;
    MPI_Comm cart;
} POISSON;

/* MPI_REAL/system.h:65 -- argv: tol m p omega beta; exits on a wrong count */
void CheckInputs(POISSON *system, int argc, char **argv);

/* MPI_REAL/system.h:67 -- one rank: np = (1,1,1), psize = ssize */
void Processor_Domain(int ssize[3], int psize[3], int np[3], int coords[3], int rem[3], MPI_Comm comm, MPI_Comm *cart);

/* MPI_REAL/system.h:69-70 -- one rank: no neighbours, no ghost traffic */
void Comm_topologies(POISSON *system, int FDn, int psize[3], int coords[3], int rem[3], int np[3], MPI_Comm cart, MPI_Comm *comm_laplacian,
    int *send_neighs, int *rec_neighs, int *send_counts, int *rec_counts, int send_layers[6], int rec_layers[6], int *n_in, int *n_out);

/* MPI_REAL/system.h:72 */
void Max_layer(int ssize[3], int np[3], int FDn, int *max_layer);

/* MPI_REAL/system.h:74 -- allocations, FD weights, seeded RHS (system.c:425-471) */
void Initialize(POISSON *system, int max_layer);

/* MPI_REAL/system.h:76 -- also releases libaar_hip's cached device buffers */
void Deallocate_memory(POISSON *system);

/* MPI_REAL/system.h:78 */
void Vec_copy(int *a, int *b, int n);

/* MPI_REAL/system.h:80 -- weights of the 2*FDn-order central second derivative */
void Lap_coefficient(double *coeff_lap, int FDn);

/* MPI_REAL/system.h:82 -- Lap_phi = a * Lap * phi on the GPU (host arrays in and out) */
void Lap_Vec_mult(POISSON *system, double a, double *phi, double *Lap_phi, MPI_Comm comm_laplacian);

/* MPI_REAL/system.h:84 -- res /= diag on the GPU (host array) */
void Precondition(double diag, double *res, int Np);

/* Extension, no reference counterpart: the variable-coefficient operator
 * a*Lap + c(x) and its pointwise Jacobi step res /= (diag + c(x)), with the
 * c(x) registered by aar_hip_set_coefficient (aar_hip.h).  Same signatures
 * as Lap_Vec_mult / Precondition (MPI_REAL/system.h:82-84), so a caller
 * passes them to AAR(), PGR(), PL2R() or CG() in their place (README:25's
 * callback slot), which then solve A(c) x = rhs on the GPU. */
void Lap_Vec_mult_vc(POISSON *system, double a, double *phi, double *Lap_phi, MPI_Comm comm_laplacian);
void Precondition_vc(double diag, double *res, int Np);

/* MPI_REAL/system.h:86 */
void Find_size_dir(int rem, int coords, int psize, int *small, int *large);

/* MPI_REAL/system.h:88 */
void Get_block_origin_global_coords(int coords[3], int rem[3], int psize[3], int g_origin[3]);

#ifdef __cplusplus
}
#endif
#endif
