/*
 * aar_hip.h -- extension entry points of libaar_hip (no MPI dependency).
 *
 * The reference's solvers take host arrays and a POISSON descriptor
 * (AAR.h, PGR.h, PL2R.h, CG.h in this directory keep those signatures).
 * The entry points below are their device-resident counterparts: every
 * vector argument is a DEVICE pointer into HBM (hipMalloc'd or a framework
 * tensor), laid out x fastest (idx = i + j*nx + k*nx*ny), and the whole
 * solve runs on the current GPU on the library stream (aar_hip_set_stream).
 * Each is the GPU form of the reference routine named in its comment; the
 * screened term sigma, the FP32-refined solver and the diagnostics have no
 * reference counterpart and say so.
 *
 * All functions return 0 on success or an AAR_HIP_ERR_* code; the text of
 * the last error is returned by aar_hip_last_error().
 */
#ifndef AAR_HIP_H
#define AAR_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define AAR_HIP_OK 0
#define AAR_HIP_ERR_ARG 1        /* invalid argument (sizes, m, fdn, null pointers) */
#define AAR_HIP_ERR_HIP 2        /* HIP runtime error (allocation, launch) */
#define AAR_HIP_ERR_SVD 3        /* SVD did not converge in 250 sweeps (tools.c:270 exits) */
#define AAR_HIP_ERR_CALLBACK 4   /* operator/preconditioner callback is not libaar_hip's */
#define AAR_HIP_ERR_NONFINITE 5  /* residual became NaN */

/* Grid and operator: y = a*Lap_FD(2*fdn)(x) + sigma*x, periodic.
 * fdn = 1..6: the central FD weights of order 2*fdn that Lap_coefficient
 * computes (system.c:480-498; MPI_REAL/main.c:35 uses 6, the 12th-order
 * radius-6 stencil), or AAR_HIP_FDN_USER: the weights registered with
 * aar_hip_set_stencil.  The sweeps have radius 6; a smaller fdn runs them
 * with zero weights beyond fdn (the same sums, at the cost of FD12).
 * sigma = 0 is the reference operator (system.c:189-205); sigma != 0 is
 * the screened operator (extension).  The Jacobi diagonal is
 * a*(3*coeff[0]) + sigma (AAR.c:94 for sigma = 0). */
#define AAR_HIP_FDN_USER (-1)
typedef struct {
    int nx, ny, nz;
    int fdn;
    double a;
    double sigma;
} aar_hip_grid;

typedef struct {
    int iterations;         /* AAR/PGR/PL2R: iter-1 (AAR.c:153); CG: iter_count (CG.c:159) */
    double relres;          /* last checked ||b - A x|| / ||b|| */
    double rhs_norm;        /* ||b|| */
    int converged;          /* relres <= tol at exit */
    int printed_converged;  /* the reference's "converged!" condition (AAR.c:152) */
    int printed_warning;    /* the reference's "exceeded maximum iterations" condition */
    int status;             /* 0 or an AAR_HIP_ERR_* code */
    int hist_len;           /* entries available from aar_hip_history() */
    double solve_ms;        /* GPU time of the solve (hipEvents on the library stream) */
    double bytes;           /* algorithmic HBM bytes moved by the solve */
    int outer_steps;        /* FP32-refined solver: FP64 residual-refinement steps */
    int inner_iterations;   /* FP32-refined solver: total FP32 AAR iterations */
    int reason;             /* Krylov comparators: PETSc's KSPConvergedReason (2 rtol,
                               3 atol, -3 max_it, -4 dtol, -5/-6 breakdown, -9 nan) */
} aar_hip_stats;

/* system.c:189-205 -- y = a*Lap(x) + sigma*x (device pointers, FP64) */
int aar_hip_matvec(const aar_hip_grid *g, const double *x, double *y);
/* same in FP32 (the arithmetic of the refined solver's inner sweeps) */
int aar_hip_matvec_f32(const aar_hip_grid *g, const float *x, float *y);

/* AAR.c:37-177 -- x: initial guess in, solution out; history m <= 32 (also
 * for PGR/PL2R, their FP32/mixed forms and the stand-alone
 * Anderson/Galerkin/L2 steps; the reference takes any m, AAR.c:51-73, and
 * 32 is the largest m x m system the one-wavefront SVD solves) */
int aar_hip_aar(const aar_hip_grid *g, double *x, const double *b, double omega, double beta,
                int m, int p, int max_iter, double tol, int precond, aar_hip_stats *st);

/* AAR.c:37-177 in FP32 (normal equations and the m x m solve stay FP64) */
int aar_hip_aar_f32(const aar_hip_grid *g, float *x, const float *b, double omega, double beta,
                    int m, int p, int max_iter, double tol, int precond, aar_hip_stats *st);

/* AAR.c:37-177 in mixed precision (extension, no reference counterpart):
 * x and b FP64; the true residual b - A x of every post-Anderson iterate,
 * the iterate itself and the m x m normal equations in FP64; the Richardson
 * sweeps (stencil), the residual history and the Gram reads in FP32.  The
 * FP64 true residual replaces the FP32 recurrence every p iterations, so
 * convergence is checked on b - A x in FP64 as in the reference. */
int aar_hip_aar_mixed(const aar_hip_grid *g, double *x, const double *b, double omega, double beta,
                      int m, int p, int max_iter, double tol, int precond, aar_hip_stats *st);

/* Variable-coefficient operator (extension; no MPI_REAL counterpart -- the
 * device-side operator behind the reference's Lap_Vec_mult callback slot,
 * README:25, and MPI_REAL/system.h:82-84's Lap_Vec_mult/Precondition):
 *     A = a*Lap_FD12 + c(x),  periodic,  Jacobi diagonal D(x) = a*(3*coeff[0]) + c(x).
 * c is a DEVICE array of N doubles; g->sigma must be 0 (a constant shift
 * belongs in c).  With c = sigma everywhere every result equals the screened
 * operator's bit for bit. */
int aar_hip_matvec_vc(const aar_hip_grid *g, const double *c, const double *x, double *y);
/* AAR.c:37-177 on A(c) with the pointwise Jacobi step r/D(x) (precond = 1)
 * or none; mixed = 1: the FP32 recurrence sweeps of aar_hip_aar_mixed. */
int aar_hip_aar_vc(const aar_hip_grid *g, const double *c, double *x, const double *b, double omega,
                   double beta, int m, int p, int max_iter, double tol, int precond, int mixed,
                   aar_hip_stats *st);
/* PGR.c:35-173 (l2 = 0) / PL2R.c:35-172 (l2 = 1) and CG.c:33-175 on A(c), with
 * the same pointwise Jacobi step. */
int aar_hip_pgr_vc(const aar_hip_grid *g, const double *c, double *x, const double *b, double omega, int m,
                   int p, int max_iter, double tol, int precond, int l2, aar_hip_stats *st);
int aar_hip_cg_vc(const aar_hip_grid *g, const double *c, double *x, const double *b, int max_iter, double tol,
                  int precond, aar_hip_stats *st);
/* The device operator behind the reference's callback slots (README:25):
 * registers a user's 1D stencil weights coeff[0..fdn] (HOST array, fdn =
 * 1..6), used as y = a*(3*coeff[0]*x + sum_{o=1..fdn} coeff[o]*S_o(x)) with
 * S_o the sum of the six neighbours at distance o (the form of
 * system.c:195-201), Jacobi diagonal a*3*coeff[0], by every call whose
 * grid (or Helmholtz operator) says fdn = AAR_HIP_FDN_USER.  coeff = NULL
 * clears them.  The reference-signature entry points take the weights the
 * caller's POISSON holds (FDn, coeff_lap), which may equally be a user's.
 * ILU(0) (precond = 2) needs the full FD12 stencil. */
int aar_hip_set_stencil(const double *coeff, int fdn);

/* Registers c(x) -- a HOST array of n doubles, copied to the GPU -- for the
 * drop-in slots Lap_Vec_mult_vc / Precondition_vc (system.h), which the
 * reference's AAR(), PGR(), PL2R() and CG() then run on the GPU; c = NULL
 * clears it.  Kept until
 * cleared, aar_hip_release() or Deallocate_memory(). */
int aar_hip_set_coefficient(const double *c, long n);

/* One solve of a batch: its own grid, operator a*Lap + sigma and vectors
 * (device pointers; x: initial guess in, solution out). */
typedef struct {
    int nx, ny, nz;
    double a, sigma;
    double *x;
    const double *b;
} aar_hip_instance;

/* Extension (no reference counterpart): `count` independent AAR solves
 * (AAR.c:37-177 semantics each, FD12, fdn = 6) sharing omega, beta, m, p,
 * max_iter and tol, run together: every batched kernel covers all
 * instances through a work table and one Anderson cycle is replayed as a
 * captured hipGraph.  mixed = 0: FP64; mixed = 1: the mixed-precision
 * iteration of aar_hip_aar_mixed.  m <= 16, any p >= 1.  stats: `count`
 * entries (solve_ms is the whole batch's).  The batch runs on the library's
 * own stream, ordered with the stream set by aar_hip_set_stream. */
int aar_hip_aar_batch(const aar_hip_instance *inst, int count, double omega, double beta, int m, int p,
                      int max_iter, double tol, int precond, int mixed, aar_hip_stats *stats);

/* PGR.c:35-173 (l2 = 0) and PL2R.c:35-172 (l2 = 1) */
int aar_hip_pgr(const aar_hip_grid *g, double *x, const double *b, double omega, int m, int p,
                int max_iter, double tol, int precond, int l2, aar_hip_stats *st);

/* CG.c:33-175 */
int aar_hip_cg(const aar_hip_grid *g, double *x, const double *b, int max_iter, double tol,
               int precond, aar_hip_stats *st);

/* The Krylov comparators of the PETSc variant, PETSC_REAL/main.c:86-122
 * (KSPGMRES) and 124-150 (KSPBICG): PETSc's algorithms with its defaults
 * (GMRES: restart 30, classical Gram-Schmidt, left preconditioning; BiCG:
 * shadow residual r0), the Jacobi preconditioner of PCJACOBI (precond = 1:
 * multiply by 1/diag; 0: none) and KSPConvergedDefault on the
 * preconditioned residual norm, ttol = tol*||B b|| for the nonzero guess.
 * c: a variable-coefficient field (device, a*Lap + c(x), diagonal
 * a*3*c0 + c(x); sigma must then be 0) or NULL.  restart <= 32.  stats:
 * iterations = PETSc's its, relres = ||b - A x||/||b|| of the returned x
 * (main.c:112-118), reason = the KSPConvergedReason; aar_hip_history()
 * returns the norm of every convergence test. */
#define AAR_HIP_GMRES_MAXK 32
#define AAR_HIP_LGMRES_MAXAUG 8
/* precond = 2 (both solvers): the pc = 1 preconditioner of the PETSc
 * variant on one rank, PCBJACOBI -> PCILU at 0 levels (main.c:88-91) = the
 * ILU(0) factors of the whole periodic matrix, computed on the device in
 * each call; every grid side must be >= 13.  BiCG then carries the shadow
 * sequence with the transposed factors. */
int aar_hip_gmres(const aar_hip_grid *g, const double *c, double *x, const double *b, int restart,
                  int max_iter, double tol, int precond, aar_hip_stats *st);
int aar_hip_bicg(const aar_hip_grid *g, const double *c, double *x, const double *b, int max_iter,
                 double tol, int precond, aar_hip_stats *st);
/* The CG of the PETSc variant, PETSC_REAL/main.c:152-175 (KSPCG with the
 * preconditioner of its pc: precond 1 = PCJACOBI, 2 = PCBJACOBI -> ILU(0)
 * as above, 0 = none): PETSc's KSPSolve_CG with the preconditioned residual
 * norm, KSPConvergedDefault (ttol = tol*||B b|| for the nonzero guess), its
 * beta = 0 (converged), indefinite-preconditioner and indefinite-matrix
 * exits (reasons 3, -8, -10).  stats as aar_hip_gmres. */
int aar_hip_kspcg(const aar_hip_grid *g, const double *c, double *x, const double *b, int max_iter, double tol,
                  int precond, aar_hip_stats *st);
/* PCApply / PCApplyTranspose of that ILU(0) preconditioner: factors the
 * matrix of g (+ c, may be NULL) and sets z = M^{-1} r (trans = 0) or
 * M^{-T} r (trans = 1); device pointers, z may equal r.  Level-scheduled
 * (x + y + z = const planes), bit-identical to a sequential IKJ ILU(0). */
int aar_hip_ilu0_apply(const aar_hip_grid *g, const double *c, const double *r, double *z, int trans);

/* PETSC_REAL/AAR_Real.c:32-199 -- the PETSc variant's AAR with its
 * preconditioner choice (AAR_Real.c:57-66): precond 1 = PCJACOBI (multiply
 * by 1/diag), 2 = its pc = 1, PCICC on the one block, realised as the ILU(0)
 * above (the same preconditioner in exact arithmetic on this symmetric
 * matrix), 0 = none.  Its semantics where they differ from AAR.c: the
 * Anderson step runs whenever iter % p == 0 (iter = 1 included), the
 * update is x += sum_i (-s_i)(beta DF_i + DX_i) column by column, then
 * x += beta f.  Its LAPACKE_dgelsd least squares (rcond = -1,
 * AAR_Real.c:223) is the minimum-norm solve over the singular values above
 * 2^-53 s_max, by the Golub-Reinsch SVD of the other solvers.  m <= 32; ||b - A x|| is checked every p iterations
 * (aar_hip_history); iterations = iter - 1 as AAR_Real.c:162-166 prints. */
int aar_hip_aar_petsc(const aar_hip_grid *g, const double *c, double *x, const double *b, double omega,
                      double beta, int m, int p, int max_iter, double tol, int precond, aar_hip_stats *st);

/* PETSC_REAL/PGR_Real.c:32-195 (l2 = 0) and PL2R_Real.c:32-197 (l2 = 1) --
 * the variant's PGR/PL2R with the same preconditioner choice (PCICC for
 * pc = 1, PGR_Real.c:55-62), the DX^T DF (DF^T DF) system solved with
 * dgelsd's rcond = -1 cutoff as in aar_hip_aar_petsc (dgelsd's column-major
 * call takes the transpose, equal in exact arithmetic for these symmetric
 * operators). */
int aar_hip_pgr_petsc(const aar_hip_grid *g, const double *c, double *x, const double *b, double omega, int m,
                      int p, int max_iter, double tol, int precond, int l2, aar_hip_stats *st);

/* Extension (no reference counterpart): FP64 residual refinement around
 * FP32 AAR.  Outer: r = b - A x in FP64, stop at ||r|| <= tol*||b||;
 * inner: FP32 AAR on A e = r from e = 0 to relative tolerance
 * max(inner_tol, 0.5*tol*||b||/||r||); x += e in FP64. */
int aar_hip_aar_refined(const aar_hip_grid *g, double *x, const double *b, double omega,
                        double beta, int m, int p, int max_outer, int max_inner, double tol,
                        double inner_tol, int precond, aar_hip_stats *st);

/* Complex Helmholtz (extension, no MPI_REAL counterpart; the complex AAR of
 * PETSC_COMPLEX/AAR_Complex.c:103-246 with a Jacobi preconditioner):
 *   A u = a*Lap_FD12(u) + (c(x) + i*eta) u,  periodic,
 * vectors complex128 interleaved (re, im) in device memory, N = nx*ny*nz
 * points.  c is a device array of N doubles, or NULL for c = c0
 * everywhere.  Jacobi diagonal D = (a*3*coeff[0] + c(x)) + i*eta. */
typedef struct {
    int nx, ny, nz;
    int fdn;            /* 6 */
    double a;
    double c0;
    double eta;
    const double *c;
} aar_hip_helm;

/* y = A u */
int aar_hip_helm_matvec(const aar_hip_helm *h, const double *u, double *y);

/* AAR on A u = b (adjoint = 0) or on A^H u = b (adjoint = 1: eta -> -eta),
 * the control flow of AAR.c:37-177 with the Hermitian Gram DF^H DF, DF^H f
 * solved through the real 2m x 2m embedding by the pinv of tools.c;
 * m <= 16.  u: initial guess in, solution out. */
int aar_hip_helm_aar(const aar_hip_helm *h, double *u, const double *b, double omega, double beta,
                     int m, int p, int max_iter, double tol, int adjoint, aar_hip_stats *st);

/* GMRES of the complex PETSc variant, PETSC_COMPLEX/main.c:80-118 (KSPGMRES
 * with PCJACOBI on the Helmholtz operator): aar_hip_gmres's algorithm in
 * complex arithmetic, <x, y> = sum x conj(y), complex Givens rotations;
 * precond 1 = Jacobi 1/D, 0 = none; restart <= 32.  u, b complex128
 * interleaved (device).  stats as aar_hip_gmres. */
int aar_hip_helm_gmres(const aar_hip_helm *h, double *u, const double *b, int restart, int max_iter, double tol,
                       int precond, aar_hip_stats *st);

/* LGMRES of the complex PETSc variant, PETSC_COMPLEX/main.c:120-147
 * (KSPLGMRES with PCJACOBI or, precond = 2, the complex ILU(0); PETSc's
 * defaults restart = 30, aug = 2): each cycle takes restart - aug Arnoldi
 * steps, then one step per stored error approximation z (newest first,
 * its vector the stored B A z); every step is an iteration and is tested.
 * The cycle's correction becomes the newest z (the oldest is dropped once
 * aug are stored), with B A z = V H y from the unrotated Hessenberg, both
 * scaled by 1/||z||.  1 <= restart <= 32, 0 <= aug < restart, aug <= 8;
 * aug = 0 is aar_hip_helm_gmres.  Stats as aar_hip_helm_gmres. */
int aar_hip_helm_lgmres(const aar_hip_helm *h, double *u, const double *b, int restart, int aug, int max_iter,
                        double tol, int precond, aar_hip_stats *st);

/* BiCG of the complex PETSc variant, PETSC_COMPLEX/main.c:150-175 (KSPBICG
 * with PCJACOBI): the shadow sequence with A^H (eta -> -eta; the operator
 * is complex symmetric) and B^H = conj(1/D). */
int aar_hip_helm_bicg(const aar_hip_helm *h, double *u, const double *b, int max_iter, double tol, int precond,
                      aar_hip_stats *st);
/* precond = 2 for aar_hip_helm_gmres / aar_hip_helm_bicg: the complex
 * variant's pc = 1, PCBJACOBI -> PCILU at 0 levels on one block = ILU(0)
 * of the complex matrix (every grid side >= 13); BiCG's shadow then uses
 * M^{-H} = conj(M^{-T} conj(r)).  aar_hip_helm_ilu0_apply: that
 * preconditioner alone, z = M^{-1} r (trans 0), M^{-T} r (1), M^{-H} r (2);
 * complex128 interleaved device arrays, bit-identical to a sequential
 * restatement. */
int aar_hip_helm_ilu0_apply(const aar_hip_helm *h, const double *r, double *z, int trans);

/* Gradient of a loss through u = A(c)^{-1} b with respect to the diagonal
 * coefficient field c, given the adjoint solution lambda of A^H lambda =
 * dL/du:  dc[k] = -Re(conj(lambda[k]) u[k]).  (dL/db = lambda.) */
int aar_hip_helm_grad_c(long n, const double *lambda, const double *u, double *dc);

/* tools.c:52-105 on the GPU: x = pinv(A) b as the reference defines it,
 * A row-major m x m (A[i*m+j] = A[i][j]), m <= 32.  HOST pointers.
 * w_out, v_out may be NULL. */
int aar_hip_pinv(const double *A, const double *b, double *x, int m, double *w_out, double *v_out);

/* No counterpart in the reference: aar_hip_pinv for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops).  DEVICE pointers: A (m x m, row-major), b and x (m doubles), and
 * optionally w (m: truncated singular values), V (m x m) and status (one
 * int: 0, or 1 if the SVD did not converge in 250 sweeps); m <= 32.  The
 * work is enqueued on the library stream (the one given to
 * aar_hip_set_stream, or the null stream before the library is set up);
 * the call allocates nothing, copies nothing to the host and does not wait.
 * Its work area is one static device buffer per GPU (18 KB, held until the
 * process ends), so calls that overlap on one GPU race on it. */
int aar_hip_pinv_noalloc(const double *A, const double *b, double *x, int m, double *w, double *V, int *status);

/* Residual history of the last solve: ||b - A x_k|| for every iterate
 * (AAR: entry k is the residual of the iterate entering iteration k+1;
 * PGR/PL2R: of the Richardson iterate before projection; CG: of r_k).
 * Copies min(cap, hist_len) entries to host memory; returns the count. */
int aar_hip_history(double *out, int cap);

/* Statistics of the last solve started by any entry point, including the
 * reference-signature ones in AAR.h/PGR.h/PL2R.h/CG.h. */
int aar_hip_last_stats(aar_hip_stats *st);

/* Stream for all work (a hipStream_t; NULL is the default stream).
 * Without this call the library uses a stream of its own that is ordered
 * with the default stream. */
int aar_hip_set_stream(void *stream);

/* Stencil launch shape: lz planes per block (0 = automatic), wy
 * wavefronts per block stacked in y (4 or 8), py rows per thread (2 or 4;
 * 8 with wy = 4);
 * a block owns a 64 x (wy*py) column.  Applies to subsequent calls and to
 * every grid; (0, 8, 4) returns to the automatic shapes, tuned ones where
 * a grid has been tuned.  Batched solves (aar_hip_aar_batch) take lz,
 * clamped to each instance's nz, and keep their 8 x 4 launch shape. */
int aar_hip_set_tuning(int lz, int wy, int py);

/* Autotuner (extension): times the recurrence stencil sweep on an nx x ny x
 * nz grid (fp32 = 1: the FP32 sweep of the mixed solver) over the launch
 * shapes (wy, py) and z-slab depths lz and keeps the fastest for that grid;
 * later solves on the grid use it.  batch_shape = 1 tunes only lz for the
 * batched launch (aar_hip_aar_batch).  Returns the shape and its time per
 * sweep in us.  Allocates two temporary vectors of the grid. */
int aar_hip_autotune(int nx, int ny, int nz, int fp32, int batch_shape, int *lz, int *wy, int *py, double *us);

/* Tuning options of the history sweeps (apply to subsequent solves):
 *   "ring_pad"    extra elements per ring slot, a multiple of 32 (default
 *                 512), so that slots whose size is a large power of two
 *                 do not alias onto the same HBM channels;
 *   "hist_blocks" grid of the reducing sweeps (0 = automatic, <= 2048);
 *   "two_pass"    1 (default): reductions finish in a one-block second
 *                 kernel; 0: in the last block of the reducing kernel;
 *   "autotune"    1: a grid's stencil shape is tuned on its first solve
 *                 (aar_hip_autotune); 0 (default): automatic shapes;
 *   "complex_pinv_reference_order"  1: the complex solver's m x m SVD runs
 *                 the left reflector in the order of tools.c:137 (from
 *                 column 1); 0 (default): the textbook order, which
 *                 converges on systems where that order does not. */
int aar_hip_set_option(const char *name, long value);

/* Per-kernel GPU timing with hipEvents (off by default).  Kernel names:
 * "stencil_matvec", "stencil_aar_true", "stencil_aar_rec", "stencil_resid",
 * "stencil_pgr_rich", "stencil_cg_q", "stencil_cg_resid", "stencil_matvec_res",
 * "gram", "ls", "anchor_update", "fold", "dot", "pgr_update", "cg_update". */
int aar_hip_profile(int enable);
int aar_hip_profile_reset(void);
int aar_hip_profile_query(const char *name, int *count, double *total_ms, double *bytes);

/* Diagnostics: the last m x m (complex: 2m x 2m) normal-equation system
 * of the last solve, in the reference's allredvec layout (matrix A[i][j]
 * at G[j*n + i], right-hand side at G[n*n + i]), and its truncated
 * singular values.  n is the system size passed in (the solver's m, or
 * 2m for the complex solver).  Returns 0. */
int aar_hip_last_system(int n, double *G, double *w);

/* Extension (no reference counterpart): the attainable HBM bandwidth, by
 * plain streaming kernels with 16-byte accesses over `bytes` (a positive
 * multiple of 16; 16-byte aligned device buffers): kind 0 reads src, 1
 * writes dst, 2 copies src to dst.  Three launch shapes (block size,
 * accesses per thread, chunk per block, store form) are each given one
 * warm-up launch and `reps` launches timed with hipEvents on the library
 * stream; *gbps = the best shape's bytes moved (a copy counts its read and
 * its write) / time, in 1e9 B/s. */
int aar_hip_stream_bandwidth(int kind, void *dst, const void *src, long bytes, int reps, double *gbps);

/* Release cached device workspace. */
int aar_hip_release(void);

const char *aar_hip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
