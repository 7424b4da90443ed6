/*
 * aar_oracle.h -- CPU restatement of the reference MPI_REAL solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * HIP product in linear-solver_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  It is never linked into
 * libaar_hip and is never a fallback for it.
 *
 * Each function restates one reference routine (cited file:line, paths
 * relative to the reference tree) on one rank, with the reference's
 * floating-point operation order kept so that results are bit-identical
 * to the reference built with gcc -O3 -funroll-loops on x86-64.
 *
 * Extensions with no reference counterpart (sigma != 0 screened operator,
 * the residual trace, the complex Helmholtz solver) are marked as such.
 */
#ifndef AAR_ORACLE_H
#define AAR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int iterations;     /* as printed by the reference: iter-1 (AAR/PGR/PL2R), iter_count (CG) */
    double relres;      /* last checked ||b-Ax|| / ||b|| */
    double rhs_norm;    /* ||b|| */
    int converged;      /* relres <= tol*||b|| at exit */
    int printed_converged; /* the reference's "converged!" line condition */
    int printed_warning;   /* the reference's "exceeded maximum iterations" condition */
    int trace_len;      /* number of entries written to trace */
    double seconds;     /* wall time of the iteration loop */
    int status;         /* 0 ok, 1 SVD failed to converge */
} orc_stats;

/* system.c:480-498 Lap_coefficient */
void orc_lap_coefficient(double *coeff, int fdn);

/* system.c:189-205 (1-rank branch of Lap_Vec_mult): y = a*Lap(x), periodic.
 * sigma != 0 adds sigma*x afterwards (screened operator; no reference
 * counterpart, applied the way a user operator callback would). */
void orc_matvec(int nx, int ny, int nz, int fdn, const double *coeff,
                double a, double sigma, const double *x, double *y);

/* tools.c:18-43 Vector2Norm / VectorDotProduct (sequential sums) */
double orc_norm2(const double *v, long n);
double orc_dot(const double *u, const double *v, long n);

/* tools.c:52-105 PseudoInverseTimesVec, tools.c:113-337 SingularValueDecomp.
 * A is m*m row-major (A[i*m+j] is the reference's A[i][j]).  Returns 0, or
 * 1 when the QR sweep exceeds 250 iterations (the reference exits(1)).
 * w_out, v_out (optional) receive the truncated singular values and V. */
/* The same with LAPACK dgelsd's rcond = -1 cutoff (singular values
 * s_j <= 2^-53 * s_max dropped), the least squares of the PETSc variant
 * (PETSC_REAL/AAR_Real.c:223, PGR_Real.c:216, PL2R_Real.c:220). */
int orc_pinv_times_vec_gelsd(const double *Ain, const double *b, double *x, int m,
                             double *w_out, double *v_out);
/* The same with the textbook Golub-Reinsch left reflector (columns
 * i+1..n-1); the complex solver's m x m solves use it. */
int orc_pinv_times_vec_textbook(const double *A, const double *b, double *x, int m,
                                double *w_out, double *v_out);
int orc_pinv_times_vec(const double *A, const double *b, double *x, int m,
                       double *w_out, double *v_out);
double orc_pythag(double a, double b);

/* AAR.c:185-232 AndersonExtrapolation.  DX, DF are m columns of length n
 * (column j at DX + j*n).  G_out (m*m + m, optional) receives allredvec. */
int orc_anderson_extrapolation(const double *DX, const double *DF, const double *f,
                               double beta, int m, long n, double *am,
                               double *G_out, double *s_out);

/* AAR.c:37-177.  precond: 1 = Jacobi with diag = a*(3*coeff[0]) + sigma,
 * 0 = none (the reference's Precondition == NULL).  trace (optional,
 * trace_cap entries) receives ||rhs - A x|| each time the operator is
 * applied to the iterate (initial residual first). */
int orc_aar(int nx, int ny, int nz, int fdn, double a, double sigma,
            double *x, const double *rhs, double omega, double beta,
            int m, int p, int max_iter, double tol, int precond,
            double *trace, int trace_cap, orc_stats *st);

/* Variable-coefficient operator (no MPI_REAL counterpart; the operator a
 * caller would pass through the reference's Lap_Vec_mult slot, README:25):
 * y = a*Lap(x) + c(x) x, the term added after the stencil as orc_matvec
 * adds sigma*x.  orc_aar_vc / orc_pgr_vc / orc_cg_vc are orc_aar / orc_pgr /
 * orc_cg on that operator with the Jacobi diagonal a*(3*coeff[0]) + c(x) per
 * point; with c = sigma everywhere each equals its screened counterpart
 * with that sigma bit for bit. */
void orc_matvec_vc(int nx, int ny, int nz, int fdn, const double *coeff,
                   double a, const double *c, const double *x, double *y);
int orc_aar_vc(int nx, int ny, int nz, int fdn, double a, const double *c,
               double *x, const double *rhs, double omega, double beta,
               int m, int p, int max_iter, double tol, int precond,
               double *trace, int trace_cap, orc_stats *st);

int orc_pgr_vc(int nx, int ny, int nz, int fdn, double a, const double *c,
               double *x, const double *rhs, double omega, int m, int p,
               int max_iter, double tol, int precond, int l2,
               double *trace, int trace_cap, orc_stats *st);
int orc_cg_vc(int nx, int ny, int nz, int fdn, double a, const double *c,
              double *x, const double *b, int max_iter, double tol, int precond,
              double *trace, int trace_cap, orc_stats *st);

/* PGR.c:35-218 (l2 = 0) and PL2R.c:35-218 (l2 = 1). */
int orc_pgr(int nx, int ny, int nz, int fdn, double a, double sigma,
            double *x, const double *rhs, double omega, int m, int p,
            int max_iter, double tol, int precond, int l2,
            double *trace, int trace_cap, orc_stats *st);

/* CG.c:33-175.  trace receives ||b - A x|| at each explicit residual
 * evaluation (start, and every 50 iterations). */
int orc_cg(int nx, int ny, int nz, int fdn, double a, double sigma,
           double *x, const double *b, int max_iter, double tol, int precond,
           double *trace, int trace_cap, orc_stats *st);

/* Complex Helmholtz AAR (no MPI_REAL counterpart; semantics of
 * PETSC_COMPLEX/AAR_Complex.c:103-246 with a Jacobi preconditioner):
 * A u = a*Lap(u) + (c(x) + i*eta) u, complex128 interleaved (re, im).
 * c may be NULL (c = c0 everywhere).  The Anderson LS uses the Hermitian
 * Gram DF^H DF and DF^H f solved through the real 2m x 2m embedding with
 * the pinv of tools.c in its textbook form (orc_pinv_times_vec_textbook).
 * Parity with PETSc/zgelsd is unpinned; the solver is pinned by the FFT
 * solution, the adjoint identity and its reduction to the real AAR. */
int orc_aar_complex(int nx, int ny, int nz, int fdn, double a,
                    const double *c, double c0, double eta,
                    double *u, const double *rhs, double omega, double beta,
                    int m, int p, int max_iter, double tol,
                    double *trace, int trace_cap, orc_stats *st);

void orc_matvec_complex(int nx, int ny, int nz, int fdn, const double *coeff,
                        double a, const double *c, double c0, double eta,
                        const double *u, double *y);

/* Krylov comparators of the PETSc variant (PETSC_REAL/main.c:86-150):
 * PETSc's KSPGMRES (restart `restart`, classical Gram-Schmidt, left
 * preconditioning) and KSPBICG (shadow residual r0), both with PCJACOBI
 * (precond = 1: multiply by 1/diag, diag = a*(3*coeff[0]) + sigma, or
 * + c(x) when c != NULL) or no preconditioner (precond = 0), converged by
 * KSPConvergedDefault on the preconditioned residual norm with
 * ttol = rtol*||B b||.  PETSc is absent: these restate its published
 * algorithms (aar_oracle.c lists the sources and defaults).  trace
 * receives the norm of every convergence test (GMRES: the start of each
 * cycle and each Arnoldi step's estimate); stats: iterations = PETSc's
 * its, relres = ||b - A x||/||b|| of the returned x (main.c:112-118),
 * status = the KSPConvergedReason below.  orc_bicg's shadow_gap (may be
 * NULL) receives max |r~ - r| over the iterations, 0 on a symmetric
 * operator. */
#define ORC_GMRES_MAXK 32
#define ORC_GMRES_HAPTOL 1e-30
#define ORC_LGMRES_MAXAUG 8
#define ORC_KSP_ATOL 1e-50
#define ORC_KSP_DTOL 1e5
#define ORC_KSP_CONVERGED_RTOL 2
#define ORC_KSP_CONVERGED_ATOL 3
#define ORC_KSP_DIVERGED_NULL (-2)
#define ORC_KSP_DIVERGED_ITS (-3)
#define ORC_KSP_DIVERGED_DTOL (-4)
#define ORC_KSP_DIVERGED_BREAKDOWN (-5)
#define ORC_KSP_DIVERGED_BREAKDOWN_BICG (-6)
#define ORC_KSP_DIVERGED_INDEFINITE_PC (-8)
#define ORC_KSP_DIVERGED_NANORINF (-9)
#define ORC_KSP_DIVERGED_INDEFINITE_MAT (-10)
/* precond for orc_gmres / orc_bicg: 0 none, 1 Jacobi, 2 ILU(0) (below). */

/* ILU(0) of the periodic FD12 star matrix a*Lap + sigma (+ c(x)), the
 * PETSc variant's pc = 1 preconditioner on one rank (PCBJACOBI -> PCILU,
 * PETSC_REAL/main.c:88-91; PCICC of AAR_Real.c:58-60 is the same
 * preconditioner in exact arithmetic on this symmetric matrix).  F holds
 * ORC_ILU_SLOTS doubles per row: [0] = 1/u_ii, [1 + 12*axis + k] the entry
 * at periodic offset d along axis (0 x, 1 y, 2 z), k = d + 6 (d < 0) or
 * d + 5 (d > 0); l_ij below the row, u_ij above.  Every side >= 13.
 * Returns 0, -1 (bad size) or -2 (zero pivot).  orc_ilu0_apply: out =
 * M^{-1} in (trans = 0) or M^{-T} in (trans = 1), in place allowed. */
#define ORC_ILU_SLOTS 37
int orc_ilu0_factor(int nx, int ny, int nz, int fdn, double a, double sigma, const double *c, double *F);
void orc_ilu0_apply(int nx, int ny, int nz, const double *F, const double *in, double *out, int trans);
/* AAR of the PETSc variant, PETSC_REAL/AAR_Real.c:32-199 (its pc = 0/1:
 * precond 1 Jacobi by the reciprocal diagonal, 2 ILU(0) standing in for
 * PCICC, 0 none); its dgelsd least squares restated with dgelsd's rcond =
 * -1 cutoff on the Golub-Reinsch SVD (orc_pinv_times_vec_gelsd).  trace: ||b - A x|| at every check (iter % p == 0); stats as
 * orc_aar (iterations = iter - 1). */
int orc_aar_petsc(int nx, int ny, int nz, int fdn, double a, double sigma, const double *c,
                  double *x, const double *b, double omega, double beta, int m, int p, int max_iter,
                  double tol, int precond, double *trace, int trace_cap, orc_stats *st);
/* PGR (l2 = 0) / PL2R (l2 = 1) of the PETSc variant, PGR_Real.c:32-195 /
 * PL2R_Real.c:32-197, preconditioner and least squares as orc_aar_petsc. */
int orc_pgr_petsc(int nx, int ny, int nz, int fdn, double a, double sigma, const double *c,
                  double *x, const double *b, double omega, int m, int p, int max_iter, double tol,
                  int precond, int l2, double *trace, int trace_cap, orc_stats *st);
/* GMRES of the complex PETSc variant (PETSC_COMPLEX/main.c:80-118) on
 * the Helmholtz operator of orc_aar_complex (c may be NULL: c0), u and rhs
 * complex128 interleaved; precond 1 = Jacobi 1/D, 0 = none.  Complex
 * operations by explicit formulas (aar_oracle.c); stats/trace as
 * orc_gmres. */
int orc_gmres_complex(int nx, int ny, int nz, int fdn, double a, const double *c, double c0, double eta,
                      double *u, const double *rhs, int restart, int max_it, double rtol, int precond,
                      double *trace, int trace_cap, orc_stats *st);
/* LGMRES of the complex PETSc variant (PETSC_COMPLEX/main.c:120-147,
 * KSPLGMRES; PETSc's defaults restart = 30, aug = 2): restart - aug
 * Arnoldi steps and one step per stored error approximation each cycle
 * (aar_oracle.c); 1 <= restart <= 32, 0 <= aug < restart, aug <= 8; aug
 * = 0 is orc_gmres_complex. */
int orc_lgmres_complex(int nx, int ny, int nz, int fdn, double a, const double *c, double c0, double eta,
                       double *u, const double *rhs, int restart, int aug, int max_it, double rtol, int precond,
                       double *trace, int trace_cap, orc_stats *st);
/* BiCG of the complex PETSc variant (PETSC_COMPLEX/main.c:150-175); the
 * shadow sequence with A^H (eta -> -eta) and B^H = conj(1/D). */
int orc_bicg_complex(int nx, int ny, int nz, int fdn, double a, const double *c, double c0, double eta,
                     double *u, const double *rhs, int max_it, double rtol, int precond,
                     double *trace, int trace_cap, orc_stats *st);
/* ILU(0) of the complex Helmholtz matrix (PETSC_COMPLEX pc = 1), F with
 * ORC_ILU_SLOTS complex entries per row; apply trans 0: M^{-1}, 1: M^{-T},
 * 2: M^{-H} = conj(M^{-T} conj(r)).  orc_gmres_complex / orc_bicg_complex
 * take precond = 2 for it. */
int orc_ilu0_factor_complex(int nx, int ny, int nz, int fdn, double a, const double *c, double c0, double eta,
                            double *F);
void orc_ilu0_apply_complex(int nx, int ny, int nz, const double *F, const double *in, double *out, int trans);
int orc_gmres(int nx, int ny, int nz, int fdn, double a, double sigma, const double *c,
              double *x, const double *b, int restart, int max_it, double rtol, int precond,
              double *trace, int trace_cap, orc_stats *st);
/* CG of the PETSc variant, PETSC_REAL/main.c:152-175 (KSPCG with the
 * preconditioner of its pc: 1 Jacobi, 2 ILU(0), 0 none; rtol, max_it
 * 2000, nonzero initial guess): PETSc's KSPSolve_CG with the preconditioned
 * residual norm and KSPConvergedDefault, its beta = 0, indefinite-PC and
 * indefinite-matrix exits; trace and stats as orc_bicg. */
int orc_kspcg(int nx, int ny, int nz, int fdn, double a, double sigma, const double *c,
              double *x, const double *b, int max_it, double rtol, int precond,
              double *trace, int trace_cap, orc_stats *st);
int orc_bicg(int nx, int ny, int nz, int fdn, double a, double sigma, const double *c,
             double *x, const double *b, int max_it, double rtol, int precond,
             double *trace, int trace_cap, orc_stats *st, double *shadow_gap);

#ifdef __cplusplus
}
#endif
#endif
