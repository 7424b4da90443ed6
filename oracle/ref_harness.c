/*
 * ref_harness.c -- driver that runs the REAL reference MPI_REAL solvers
 * (compiled from the reference tree by oracle/Makefile into oracle/_ref/)
 * on inputs written by the tests, to make golden fixtures and to pin the
 * C restatement in aar_oracle.c.  TEST INFRASTRUCTURE ONLY.
 *
 * It calls the reference through its public entry points only
 * (MPI_REAL/AAR.h:18-22, PGR.h:18-22, PL2R.h:18-22, CG.h:18-21,
 * system.h:65-88, tools.h:24-28) and plugs a recording operator into the
 * reference's operator slot (the extension point of MPI_REAL/README:25):
 * it forwards to the reference Lap_Vec_mult, adds sigma*x for a screened
 * operator, and records ||rhs - A x|| whenever the operand is the
 * solver's iterate.
 *
 * Usage (any rank count, e.g. mpirun -n 8):
 *   ref_harness solve  <problem.bin> <aar|pgr|pl2r|cg> <x_out|-> <trace_out|->
 *   ref_harness matvec <problem.bin> <y_out>
 *   ref_harness init   <nx> <ny> <nz> <rhs_out>     (the reference RHS of system.c:425-471)
 *   ref_harness pinv   <in.bin> <out.bin>           (rank 0 only)
 *   ref_harness anderson <in.bin> <out.bin>         (rank 0 only)
 *   ref_harness galerkin|l2 <in.bin> <out.bin>      (rank 0 only)
 *
 * problem.bin: "AARPRB01", int32 nx ny nz fdn m p max_iter precond,
 *              float64 a sigma omega beta tol, float64 rhs[N], x0[N].
 */
#include "system.h"
#include "AAR.h"
#include "PGR.h"
#include "PL2R.h"
#include "CG.h"
#include "tools.h"

#include <string.h>

typedef struct {
    int n[3], fdn, m, p, max_iter, precond;
    double a, sigma, omega, beta, tol;
} prob_hdr;

static double g_sigma = 0.0;
static const double *g_rhs_local = NULL;
static const double *g_iterate = NULL;
static int g_np = 0;
static double *g_trace = NULL;
static int g_trace_len = 0, g_trace_cap = 0;

static void rec_op(POISSON *sys, double a, double *x, double *y, MPI_Comm comm)
{
    Lap_Vec_mult(sys, a, x, y, comm);
    if (g_sigma != 0.0)
        for (int i = 0; i < g_np; i++) y[i] = y[i] + g_sigma * x[i];
    if (x == g_iterate && g_trace) {
        double s = 0.0, tot = 0.0;
        for (int i = 0; i < g_np; i++) {
            double r = g_rhs_local[i] - y[i];
            s = s + r * r;
        }
        MPI_Allreduce(&s, &tot, 1, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
        if (g_trace_len < g_trace_cap) g_trace[g_trace_len++] = sqrt(tot);
    }
}

static void rec_precond(double diag, double *r, int np)
{
    double d = diag;
    if (g_sigma != 0.0) d += g_sigma;
    for (int i = 0; i < np; i++) r[i] /= d;
}

/* "AARPRB02": the same, followed by int32 ncoef and float64 coeff[ncoef]:
 * a caller's stencil weights that replace the system's coeff_lap[0..FDn]
 * (the reference reads its weights only from there, system.c:102-104,
 * AAR.c:94) */
static int g_ncoef = 0;
static double g_coef[32];

static int read_problem(const char *path, prob_hdr *h, double **rhs, double **x0)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    char magic[8];
    int iv[8];
    double dv[5];
    if (fread(magic, 1, 8, fp) != 8) { fclose(fp); return -2; }
    const int v2 = memcmp(magic, "AARPRB02", 8) == 0;
    if (!v2 && memcmp(magic, "AARPRB01", 8) != 0) { fclose(fp); return -2; }
    if (fread(iv, sizeof(int), 8, fp) != 8 || fread(dv, sizeof(double), 5, fp) != 5) { fclose(fp); return -3; }
    h->n[0] = iv[0]; h->n[1] = iv[1]; h->n[2] = iv[2]; h->fdn = iv[3];
    h->m = iv[4]; h->p = iv[5]; h->max_iter = iv[6]; h->precond = iv[7];
    h->a = dv[0]; h->sigma = dv[1]; h->omega = dv[2]; h->beta = dv[3]; h->tol = dv[4];
    long n = (long)h->n[0] * h->n[1] * h->n[2];
    *rhs = (double *)malloc(sizeof(double) * n);
    *x0 = (double *)malloc(sizeof(double) * n);
    if (fread(*rhs, sizeof(double), n, fp) != (size_t)n || fread(*x0, sizeof(double), n, fp) != (size_t)n) {
        fclose(fp);
        return -4;
    }
    g_ncoef = 0;
    if (v2) {
        if (fread(&g_ncoef, sizeof(int), 1, fp) != 1 || g_ncoef < 1 || g_ncoef > 32 ||
            fread(g_coef, sizeof(double), g_ncoef, fp) != (size_t)g_ncoef) {
            fclose(fp);
            return -5;
        }
    }
    fclose(fp);
    return 0;
}

static POISSON *setup_system(const int n[3], int fdn)
{
    POISSON *sys = (POISSON *)calloc(1, sizeof(POISSON));
    int max_layer = 0;
    sys->ssize[0] = n[0]; sys->ssize[1] = n[1]; sys->ssize[2] = n[2];
    sys->FDn = fdn;
    Processor_Domain(sys->ssize, sys->psize, sys->np, sys->coords, sys->rem, MPI_COMM_WORLD, &sys->cart);
    Max_layer(sys->ssize, sys->np, sys->FDn, &max_layer);
    Initialize(sys, max_layer);
    Comm_topologies(sys, sys->FDn, sys->psize, sys->coords, sys->rem, sys->np, sys->cart, &sys->comm_laplacian,
                    sys->send_neighs, sys->rec_neighs, sys->send_counts, sys->rec_counts,
                    sys->send_layers, sys->rec_layers, &sys->n_in, &sys->n_out);
    for (int o = 0; o < g_ncoef && o <= fdn; o++) sys->coeff_lap[o] = g_coef[o];
    return sys;
}

/* copy global field -> local block */
static void scatter_local(POISSON *sys, const double *glob, double *loc)
{
    int o[3];
    Get_block_origin_global_coords(sys->coords, sys->rem, sys->psize, o);
    const int *ps = sys->psize, *ss = sys->ssize;
    for (int k = 0; k < ps[2]; k++)
        for (int j = 0; j < ps[1]; j++)
            for (int i = 0; i < ps[0]; i++)
                loc[i + j * ps[0] + k * ps[0] * ps[1]] =
                    glob[(o[0] + i) + (long)(o[1] + j) * ss[0] + (long)(o[2] + k) * ss[0] * ss[1]];
}

/* gather local blocks -> global field on rank 0 */
static void gather_global(POISSON *sys, const double *loc, double *glob)
{
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    int meta[6];
    Get_block_origin_global_coords(sys->coords, sys->rem, sys->psize, meta);
    meta[3] = sys->psize[0]; meta[4] = sys->psize[1]; meta[5] = sys->psize[2];
    const int *ss = sys->ssize;
    if (rank == 0) {
        for (int src = 0; src < size; src++) {
            int mt[6];
            const double *buf = loc;
            double *tmp = NULL;
            if (src == 0) {
                memcpy(mt, meta, sizeof(mt));
            } else {
                MPI_Recv(mt, 6, MPI_INT, src, 1, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
                long cnt = (long)mt[3] * mt[4] * mt[5];
                tmp = (double *)malloc(sizeof(double) * cnt);
                MPI_Recv(tmp, (int)cnt, MPI_DOUBLE, src, 2, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
                buf = tmp;
            }
            for (int k = 0; k < mt[5]; k++)
                for (int j = 0; j < mt[4]; j++)
                    for (int i = 0; i < mt[3]; i++)
                        glob[(mt[0] + i) + (long)(mt[1] + j) * ss[0] + (long)(mt[2] + k) * ss[0] * ss[1]] =
                            buf[i + j * mt[3] + k * mt[3] * mt[4]];
            free(tmp);
        }
    } else {
        long cnt = (long)meta[3] * meta[4] * meta[5];
        MPI_Send(meta, 6, MPI_INT, 0, 1, MPI_COMM_WORLD);
        MPI_Send((void *)loc, (int)cnt, MPI_DOUBLE, 0, 2, MPI_COMM_WORLD);
    }
}

static int write_doubles(const char *path, const double *v, long n)
{
    if (!path || strcmp(path, "-") == 0) return 0;
    FILE *fp = fopen(path, "wb");
    if (!fp) return -1;
    fwrite(v, sizeof(double), n, fp);
    fclose(fp);
    return 0;
}

static int mode_solve(int argc, char **argv, int rank)
{
    if (argc != 6) return 2;
    prob_hdr h;
    double *rhs_g, *x0_g;
    if (read_problem(argv[2], &h, &rhs_g, &x0_g)) { if (!rank) printf("HARNESS error reading %s\n", argv[2]); return 3; }
    POISSON *sys = setup_system(h.n, h.fdn);
    int np = sys->psize[0] * sys->psize[1] * sys->psize[2];
    scatter_local(sys, rhs_g, sys->rhs);
    scatter_local(sys, x0_g, sys->phi);
    g_sigma = h.sigma; g_rhs_local = sys->rhs; g_iterate = sys->phi; g_np = np;
    g_trace_cap = h.max_iter + 8;
    /* trace_out "-": no recording, so the timed solve is the reference's alone */
    g_trace = strcmp(argv[5], "-") ? (double *)calloc((size_t)g_trace_cap, sizeof(double)) : NULL;
    void (*pc)(double, double *, int) = h.precond ? rec_precond : NULL;
    const char *sv = argv[3];
    double t0 = MPI_Wtime();
    if (!strcmp(sv, "aar"))
        AAR(sys, rec_op, pc, h.a, sys->phi, sys->rhs, h.omega, h.beta, h.m, h.p, h.max_iter, h.tol, np, sys->comm_laplacian);
    else if (!strcmp(sv, "pgr"))
        PGR(sys, rec_op, pc, h.a, sys->phi, sys->rhs, h.omega, h.m, h.p, h.max_iter, h.tol, np, sys->comm_laplacian);
    else if (!strcmp(sv, "pl2r"))
        PL2R(sys, rec_op, pc, h.a, sys->phi, sys->rhs, h.omega, h.m, h.p, h.max_iter, h.tol, np, sys->comm_laplacian);
    else if (!strcmp(sv, "cg"))
        CG(sys, rec_op, pc, h.a, sys->phi, sys->rhs, h.max_iter, h.tol, np, sys->comm_laplacian);
    else { if (!rank) printf("HARNESS unknown solver %s\n", sv); return 2; }
    double t1 = MPI_Wtime();
    long n = (long)h.n[0] * h.n[1] * h.n[2];
    double *xg = rank == 0 ? (double *)malloc(sizeof(double) * n) : NULL;
    gather_global(sys, sys->phi, xg);
    if (rank == 0) {
        write_doubles(argv[4], xg, n);
        write_doubles(argv[5], g_trace, g_trace_len);
        int size;
        MPI_Comm_size(MPI_COMM_WORLD, &size);
        printf("HARNESS solver=%s ranks=%d trace_len=%d seconds=%.6f\n", sv, size, g_trace_len, t1 - t0);
        fflush(stdout);
    }
    free(xg);
    return 0;
}

static int mode_matvec(int argc, char **argv, int rank)
{
    if (argc != 4) return 2;
    prob_hdr h;
    double *rhs_g, *x0_g;
    if (read_problem(argv[2], &h, &rhs_g, &x0_g)) return 3;
    POISSON *sys = setup_system(h.n, h.fdn);
    int np = sys->psize[0] * sys->psize[1] * sys->psize[2];
    scatter_local(sys, x0_g, sys->phi);
    g_sigma = h.sigma; g_np = np;
    rec_op(sys, h.a, sys->phi, sys->Lap_phi, sys->comm_laplacian);
    long n = (long)h.n[0] * h.n[1] * h.n[2];
    double *yg = rank == 0 ? (double *)malloc(sizeof(double) * n) : NULL;
    gather_global(sys, sys->Lap_phi, yg);
    if (rank == 0) write_doubles(argv[3], yg, n);
    return 0;
}

static int mode_init(int argc, char **argv, int rank)
{
    if (argc != 6) return 2;
    int n[3] = {atoi(argv[2]), atoi(argv[3]), atoi(argv[4])};
    POISSON *sys = setup_system(n, 6);
    long ng = (long)n[0] * n[1] * n[2];
    double *g = rank == 0 ? (double *)malloc(sizeof(double) * ng) : NULL;
    gather_global(sys, sys->rhs, g);
    if (rank == 0) write_doubles(argv[5], g, ng);
    return 0;
}

/* pinv in.bin: int32 m, float64 A[m*m] (row-major A[i][j]), b[m];
 * out.bin: x[m] */
static int mode_pinv(int argc, char **argv)
{
    if (argc != 4) return 2;
    FILE *fp = fopen(argv[2], "rb");
    if (!fp) return 3;
    int m;
    if (fread(&m, sizeof(int), 1, fp) != 1) { fclose(fp); return 3; }
    double **A = (double **)malloc(sizeof(double *) * m);
    double *b = (double *)malloc(sizeof(double) * m), *x = (double *)malloc(sizeof(double) * m);
    for (int i = 0; i < m; i++) {
        A[i] = (double *)malloc(sizeof(double) * m);
        if (fread(A[i], sizeof(double), m, fp) != (size_t)m) { fclose(fp); return 3; }
    }
    if (fread(b, sizeof(double), m, fp) != (size_t)m) { fclose(fp); return 3; }
    fclose(fp);
    PseudoInverseTimesVec(A, b, x, m);
    return write_doubles(argv[3], x, m);
}

/* anderson in.bin: int32 m, int32 n, float64 beta, DX[m][n], DF[m][n], f[n];
 * out.bin: am[n], allredvec[m*m+m].
 * galerkin / l2 (the same input; beta unused, DX unused by l2): the PGR
 * and PL2R steps (PGR.h:24-25, PL2R.h:24-25); out.bin: svec[m],
 * allredvec[m*m+m]. */
static int mode_anderson(int argc, char **argv, int kind)
{
    if (argc != 4) return 2;
    FILE *fp = fopen(argv[2], "rb");
    if (!fp) return 3;
    int m, n;
    double beta;
    if (fread(&m, sizeof(int), 1, fp) != 1 || fread(&n, sizeof(int), 1, fp) != 1 ||
        fread(&beta, sizeof(double), 1, fp) != 1) { fclose(fp); return 3; }
    double **DX = (double **)malloc(sizeof(double *) * m), **DF = (double **)malloc(sizeof(double *) * m);
    double **FtF = (double **)malloc(sizeof(double *) * m);
    for (int i = 0; i < m; i++) {
        DX[i] = (double *)malloc(sizeof(double) * n);
        if (fread(DX[i], sizeof(double), n, fp) != (size_t)n) { fclose(fp); return 3; }
    }
    for (int i = 0; i < m; i++) {
        DF[i] = (double *)malloc(sizeof(double) * n);
        FtF[i] = (double *)calloc((size_t)m, sizeof(double));
        if (fread(DF[i], sizeof(double), n, fp) != (size_t)n) { fclose(fp); return 3; }
    }
    double *f = (double *)malloc(sizeof(double) * n);
    if (fread(f, sizeof(double), n, fp) != (size_t)n) { fclose(fp); return 3; }
    fclose(fp);
    double *am = (double *)calloc((size_t)n, sizeof(double));
    double *red = (double *)calloc((size_t)(m * m + m), sizeof(double));
    double *Ftf = (double *)calloc((size_t)m, sizeof(double));
    double *svec = (double *)calloc((size_t)m, sizeof(double));
    if (kind == 0)
        AndersonExtrapolation(DX, DF, f, beta, m, n, am, FtF, red, Ftf, svec, MPI_COMM_WORLD);
    else if (kind == 1)
        Galerkin_Richardson(DX, DF, f, m, n, FtF, red, Ftf, svec, MPI_COMM_WORLD);
    else
        L2_Richardson(DF, f, m, n, FtF, red, Ftf, svec, MPI_COMM_WORLD);
    FILE *fo = fopen(argv[3], "wb");
    if (!fo) return 3;
    if (kind == 0) fwrite(am, sizeof(double), n, fo);
    else fwrite(svec, sizeof(double), m, fo);
    fwrite(red, sizeof(double), m * m + m, fo);
    fclose(fo);
    return 0;
}

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    int rank, rc = 2;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    if (argc >= 2) {
        if (!strcmp(argv[1], "solve")) rc = mode_solve(argc, argv, rank);
        else if (!strcmp(argv[1], "matvec")) rc = mode_matvec(argc, argv, rank);
        else if (!strcmp(argv[1], "init")) rc = mode_init(argc, argv, rank);
        else if (!strcmp(argv[1], "pinv")) rc = rank == 0 ? mode_pinv(argc, argv) : 0;
        else if (!strcmp(argv[1], "anderson")) rc = rank == 0 ? mode_anderson(argc, argv, 0) : 0;
        else if (!strcmp(argv[1], "galerkin")) rc = rank == 0 ? mode_anderson(argc, argv, 1) : 0;
        else if (!strcmp(argv[1], "l2")) rc = rank == 0 ? mode_anderson(argc, argv, 2) : 0;
    }
    if (rc && !rank) fprintf(stderr, "ref_harness: mode failed (rc=%d)\n", rc);
    MPI_Finalize();
    return rc;
}
