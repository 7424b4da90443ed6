// poisson_abi.h -- the POISSON descriptor as libaar_hip sees it.
//
// Callers compile against include/system.h, which takes MPI_Comm from
// their MPI's <mpi.h>.  libaar_hip makes no MPI calls and is built without
// MPI headers; it only needs the struct layout.  With MPICH (the MPI of
// this image, MPICH 3.3.2) MPI_Comm is an int handle, so the layout below
// is identical; tests/test_abi.py compiles both and compares sizeof and
// every offsetof.
#pragma once

extern "C" {

typedef int aar_mpi_comm;                    // MPICH: typedef int MPI_Comm
#define AAR_MPI_COMM_NULL ((aar_mpi_comm)0x04000000)   // MPICH MPI_COMM_NULL

// MPI_REAL/system.h:26-63
typedef struct {
    int FDn;
    int ssize[3];
    int psize[3];
    int np[3];
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
    double *phi;
    double *rhs;
    double *Lap_phi;
    double solver_tol;
    int solver_maxiter;
    int m;
    int p;
    double omega;
    double beta;
    double *coeff_lap;
    aar_mpi_comm comm_laplacian;
    aar_mpi_comm cart;
} POISSON;

typedef void (*aar_op_cb)(POISSON *, double, double *, double *, aar_mpi_comm);
typedef void (*aar_pc_cb)(double, double *, int);

}  // extern "C"
