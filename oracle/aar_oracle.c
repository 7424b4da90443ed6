/*
 * aar_oracle.c -- CPU restatement of the reference MPI_REAL solver path
 * (one rank, no MPI).  TEST INFRASTRUCTURE: the parity checker for the HIP
 * product, never a part of it.  See aar_oracle.h for the contract.
 *
 * Every routine keeps the reference's floating-point operation order
 * (left-to-right sums, no fused multiply-add, division where the
 * reference divides) so that, built with gcc -O3 -funroll-loops
 * -ffp-contract=off, its results are bit-identical to the reference's on
 * one rank.  Citations are file:line in the reference tree (MPI_REAL/).
 */
#include "aar_oracle.h"

#define ILU_SLOTS ORC_ILU_SLOTS

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_seconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------ */
/* Finite-difference weights: system.c:480-498.                        */
/* c0 = sum_p -2/p^2; c_p = 2 (-1)^(p+1) [prod_{FDn-p+1..FDn} i /      */
/* prod_{FDn+1..FDn+p} i] / p^2, evaluated in that order.              */
/* ------------------------------------------------------------------ */
void orc_lap_coefficient(double *coeff, int fdn)
{
    double acc = 0.0;
    for (int p = 1; p <= fdn; p++)
        acc += -(2.0 / (p * p));
    coeff[0] = acc;
    for (int p = 1; p <= fdn; p++) {
        double num = 1.0, den = 1.0;
        for (int i = fdn - p + 1; i <= fdn; i++) num *= i;
        for (int i = fdn + 1; i <= fdn + p; i++) den *= i;
        double ratio = num / den;
        double sgn = (p % 2 == 1) ? 1.0 : -1.0;   /* pow(-1, p+1), exact */
        coeff[p] = 2.0 * sgn * ratio / (p * p);
    }
}

/* ------------------------------------------------------------------ */
/* Stencil: system.c:189-205.  y = (x*3)*w0, then for o = 1..FDn       */
/* y += (x[i-o] + x[i+o] + x[j-o] + x[j+o] + x[k-o] + x[k+o]) * w_o    */
/* with w_o = a*coeff[o] and periodic wrap; x fastest (system.c:82).   */
/* ------------------------------------------------------------------ */
static inline int wrap(int i, int n) { i %= n; return i < 0 ? i + n : i; }

void orc_matvec(int nx, int ny, int nz, int fdn, const double *coeff,
                double a, double sigma, const double *x, double *y)
{
    double w[32];
    for (int o = 0; o <= fdn; o++) w[o] = a * coeff[o];
    const long sxy = (long)nx * ny;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const long c = i + (long)j * nx + k * sxy;
                double acc = x[c] * 3 * w[0];
                for (int o = 1; o <= fdn; o++) {
                    double nb = x[wrap(i - o, nx) + (long)j * nx + k * sxy];
                    nb = nb + x[wrap(i + o, nx) + (long)j * nx + k * sxy];
                    nb = nb + x[i + (long)wrap(j - o, ny) * nx + k * sxy];
                    nb = nb + x[i + (long)wrap(j + o, ny) * nx + k * sxy];
                    nb = nb + x[i + (long)j * nx + wrap(k - o, nz) * sxy];
                    nb = nb + x[i + (long)j * nx + wrap(k + o, nz) * sxy];
                    acc += nb * w[o];
                }
                y[c] = acc;
            }
    if (sigma != 0.0) {
        const long n = sxy * nz;
        for (long c = 0; c < n; c++) y[c] = y[c] + sigma * x[c];
    }
}

/* tools.c:18-28 and tools.c:35-43: plain sequential sums. */
double orc_norm2(const double *v, long n)
{
    double s = 0.0;
    for (long k = 0; k < n; k++) s = s + v[k] * v[k];
    return sqrt(s);
}

double orc_dot(const double *u, const double *v, long n)
{
    double s = 0.0;
    for (long k = 0; k < n; k++) s = s + u[k] * v[k];
    return s;
}

/* ------------------------------------------------------------------ */
/* SVD and pseudo-inverse: tools.c:52-352.                             */
/* Golub-Reinsch: Householder bidiagonalisation, accumulation of the   */
/* right transforms into V, implicit-shift QR on the bidiagonal (at    */
/* most 250 sweeps per singular value).  Only w and V are formed: the  */
/* reference overwrites U with V before use (tools.c:324-329), and no  */
/* U-only step feeds w, the superdiagonal or V.  The left Householder  */
/* step applies its reflector to columns 1..n-1 (tools.c:137), which   */
/* re-scales the current reflector before the trailing columns are    */
/* updated; that order is kept because it moves the result by an ulp.  */
/* ------------------------------------------------------------------ */
double orc_pythag(double a, double b)
{
    double absa = fabs(a), absb = fabs(b);
    if (absa > absb) return absa * sqrt(1.0 + absb * absb / (absa * absa));
    return absb == 0.0 ? 0.0 : absb * sqrt(1.0 + absa * absa / (absb * absb));
}

static inline double sign_of(double mag, double s) { return s >= 0.0 ? fabs(mag) : -fabs(mag); }
static inline double dmax(double a, double b) { return a > b ? a : b; }

/* a: n*n row-major (overwritten); w: n; v: n*n row-major.  Returns 0/1.
 * textbook = 1 runs the left reflector over columns i+1..n-1 (the textbook
 * Golub-Reinsch step) instead of tools.c's 1..n-1; used by the complex
 * solver, which has no MPI_REAL counterpart: on some of its embedded
 * Hermitian systems the tools.c order does not converge in 250 sweeps. */
static int gr_svd_wv(double *a, int n, double *w, double *v, int textbook)
{
#define A(r, c) a[(r) * n + (c)]
#define V(r, c) v[(r) * n + (c)]
    const int m = n;
    double *rv1 = (double *)calloc((size_t)n, sizeof(double));
    double g = 0.0, scale = 0.0, anorm = 0.0, s, f, h;
    int l = 0;

    for (int i = 0; i < n; i++) {
        l = i + 1;
        rv1[i] = scale * g;
        g = s = scale = 0.0;
        if (i < m) {
            for (int k = i; k < m; k++) scale += fabs(A(k, i));
            if (scale != 0.0) {
                for (int k = i; k < m; k++) {
                    A(k, i) /= scale;
                    s += A(k, i) * A(k, i);
                }
                f = A(i, i);
                g = -sign_of(sqrt(s), f);
                h = f * g - s;
                A(i, i) = f - g;
                for (int j = (textbook & 1) ? l : 1; j < n; j++) {   /* tools.c:137 starts at 1 */
                    double t = 0.0;
                    for (int k = i; k < m; k++) t += A(k, i) * A(k, j);
                    double fac = t / h;
                    for (int k = i; k < m; k++) A(k, j) += fac * A(k, i);
                }
                for (int k = i; k < m; k++) A(k, i) *= scale;
            }
        }
        w[i] = scale * g;
        g = s = scale = 0.0;
        if (i <= m - 1 && i != n - 1) {
            for (int k = l; k < n; k++) scale += fabs(A(i, k));
            if (scale != 0.0) {
                for (int k = l; k < n; k++) {
                    A(i, k) /= scale;
                    s += A(i, k) * A(i, k);
                }
                f = A(i, l);
                g = -sign_of(sqrt(s), f);
                h = f * g - s;
                A(i, l) = f - g;
                for (int k = l; k < n; k++) rv1[k] = A(i, k) / h;
                for (int j = l; j < m; j++) {
                    double t = 0.0;
                    for (int k = l; k < n; k++) t += A(j, k) * A(i, k);
                    for (int k = l; k < n; k++) A(j, k) += t * rv1[k];
                }
                for (int k = l; k < n; k++) A(i, k) *= scale;
            }
        }
        anorm = dmax(anorm, fabs(w[i]) + fabs(rv1[i]));
    }

    /* right-hand transforms -> V */
    for (int i = n - 1; i >= 0; i--) {
        if (i < n - 1) {
            if (g != 0.0) {
                for (int j = l; j < n; j++) V(j, i) = (A(i, j) / A(i, l)) / g;
                for (int j = l; j < n; j++) {
                    double t = 0.0;
                    for (int k = l; k < n; k++) t += A(i, k) * V(k, j);
                    for (int k = l; k < n; k++) V(k, j) += t * V(k, i);
                }
            }
            for (int j = l; j < n; j++) V(i, j) = V(j, i) = 0.0;
        }
        V(i, i) = 1.0;
        g = rv1[i];
        l = i;
    }

    /* implicit-shift QR on the bidiagonal (w, rv1) */
    int status = 0;
    for (int k = n - 1; k >= 0; k--) {
        for (int its = 0; its <= 250; its++) {
            int cancel = 1, nm = 0;
            for (l = k; l >= 0; l--) {
                nm = l - 1;
                if ((double)(fabs(rv1[l]) + anorm) == anorm) { cancel = 0; break; }
                if ((double)(fabs(w[nm]) + anorm) == anorm) break;
            }
            if (cancel) {
                double c = 0.0;
                s = 1.0;
                for (int i = l; i <= k; i++) {
                    f = s * rv1[i];
                    rv1[i] = c * rv1[i];
                    if ((double)(fabs(f) + anorm) == anorm) break;
                    g = w[i];
                    h = orc_pythag(f, g);
                    w[i] = h;
                    h = 1.0 / h;
                    c = g * h;
                    s = -f * h;
                }
            }
            double z = w[k];
            if (l == k) {
                if (z < 0.0) {
                    w[k] = -z;
                    for (int j = 0; j < n; j++) V(j, k) = -V(j, k);
                }
                break;
            }
            if (its == 250) { status = 1; break; }
            double x = w[l];
            nm = k - 1;
            double y = w[nm];
            g = rv1[nm];
            h = rv1[k];
            f = ((y - z) * (y + z) + (g - h) * (g + h)) / (2.0 * h * y);
            g = orc_pythag(f, 1.0);
            f = ((x - z) * (x + z) + h * ((y / (f + sign_of(g, f))) - h)) / x;
            double c = 1.0;
            s = 1.0;
            for (int j = l; j <= nm; j++) {
                int i = j + 1;
                g = rv1[i];
                y = w[i];
                h = s * g;
                g = c * g;
                z = orc_pythag(f, h);
                rv1[j] = z;
                c = f / z;
                s = h / z;
                f = x * c + g * s;
                g = g * c - x * s;
                h = y * s;
                y *= c;
                for (int jj = 0; jj < n; jj++) {
                    double vx = V(jj, j), vz = V(jj, i);
                    V(jj, j) = vx * c + vz * s;
                    V(jj, i) = vz * c - vx * s;
                }
                z = orc_pythag(f, h);
                w[j] = z;
                if (z != 0.0) {
                    z = 1.0 / z;
                    c = f * z;
                    s = h * z;
                }
                f = c * g + s * y;
                x = c * y - s * g;
            }
            rv1[l] = 0.0;
            rv1[k] = f;
            w[k] = x;
        }
        if (status) break;
    }
    free(rv1);

    double wmax = 0.0;
    for (int j = 0; j < n; j++) if (w[j] > wmax) wmax = w[j];
    if (textbook & 2) {
        /* LAPACK dgelsd with rcond < 0 (PETSC_REAL/AAR_Real.c:223): singular
         * values s_j <= eps * s_max are dropped, eps = dlamch('E') = 2^-53 */
        const double tol = 1.1102230246251565e-16 * wmax;
        for (int j = 0; j < n; j++) if (w[j] <= tol) w[j] = 0.0;
    } else {
        /* tools.c:331-335: zero w_j < n * w_max * 2.22044605e-16 */
        double wmin = n * wmax * (2.22044605e-16);
        for (int j = 0; j < n; j++) if (w[j] < wmin) w[j] = 0.0;
    }
    return status;
#undef A
#undef V
}

static int pinv_times_vec(const double *Ain, const double *b, double *x, int m, double *w_out, double *v_out,
                          int textbook);

int orc_pinv_times_vec(const double *Ain, const double *b, double *x, int m,
                       double *w_out, double *v_out)
{
    return pinv_times_vec(Ain, b, x, m, w_out, v_out, 0);
}

int orc_pinv_times_vec_textbook(const double *Ain, const double *b, double *x, int m,
                                double *w_out, double *v_out)
{
    return pinv_times_vec(Ain, b, x, m, w_out, v_out, 1);
}

/* The least-squares solve of the PETSc variant (LAPACKE_dgelsd with
 * rcond = -1, PETSC_REAL/AAR_Real.c:223, PGR_Real.c:216, PL2R_Real.c:220)
 * on its symmetric positive semi-definite normal matrices: the minimum-norm
 * solution V diag(1/s) V^T b over the singular values above eps * s_max,
 * computed by the same Golub-Reinsch SVD (dgelsd's own bidiagonal
 * divide-and-conquer is restated only through its cutoff). */
int orc_pinv_times_vec_gelsd(const double *Ain, const double *b, double *x, int m,
                             double *w_out, double *v_out)
{
    return pinv_times_vec(Ain, b, x, m, w_out, v_out, 2);
}

static int pinv_times_vec(const double *Ain, const double *b, double *x, int m, double *w_out, double *v_out,
                          int textbook)
{
    double *a = (double *)malloc(sizeof(double) * m * m);
    double *v = (double *)calloc((size_t)m * m, sizeof(double));
    double *w = (double *)calloc((size_t)m, sizeof(double));
    double *tmp = (double *)calloc((size_t)m, sizeof(double));
    memcpy(a, Ain, sizeof(double) * m * m);
    int st = gr_svd_wv(a, m, w, v, textbook);
    /* tools.c:76-94: x = V diag(1/w) V^T b, with U replaced by V */
    for (int j = 0; j < m; j++) {
        double s = 0.0;
        if (w[j] != 0.0) {
            for (int i = 0; i < m; i++) s += v[i * m + j] * b[i];
            s /= w[j];
        }
        tmp[j] = s;
    }
    for (int j = 0; j < m; j++) {
        double s = 0.0;
        for (int jj = 0; jj < m; jj++) s += v[j * m + jj] * tmp[jj];
        x[j] = s;
    }
    if (w_out) memcpy(w_out, w, sizeof(double) * m);
    if (v_out) memcpy(v_out, v, sizeof(double) * m * m);
    free(a); free(v); free(w); free(tmp);
    return st;
}

/* ------------------------------------------------------------------ */
/* AAR.c:185-232 AndersonExtrapolation (one rank: the allreduce of     */
/* AAR.c:212 is the identity).                                         */
/* ------------------------------------------------------------------ */
int orc_anderson_extrapolation(const double *DX, const double *DF, const double *f,
                               double beta, int m, long n, double *am,
                               double *G_out, double *s_out)
{
    double *red = (double *)calloc((size_t)(m * m + m), sizeof(double));
    double *G = (double *)calloc((size_t)m * m, sizeof(double));
    double *g = (double *)calloc((size_t)m, sizeof(double));
    double *s = (double *)calloc((size_t)m, sizeof(double));
    for (int j = 0; j < m; j++)
        for (int i = 0; i <= j; i++) {
            double t = orc_dot(DF + (long)i * n, DF + (long)j * n, n);
            red[j * m + i] = t;
            red[i * m + j] = t;
        }
    for (int j = 0; j < m; j++) red[m * m + j] = orc_dot(DF + (long)j * n, f, n);
    /* FtF[i][j] = red[j*m + i]; Ftf[j] = red[m*m + j] (AAR.c:214-221) */
    for (int j = 0; j < m; j++)
        for (int i = 0; i < m; i++) G[i * m + j] = red[j * m + i];
    for (int j = 0; j < m; j++) g[j] = red[m * m + j];
    int st = orc_pinv_times_vec(G, g, s, m, NULL, NULL);
    for (long k = 0; k < n; k++) {
        double t = 0.0;
        for (int j = 0; j < m; j++)
            t = t + (DX[(long)j * n + k] + beta * DF[(long)j * n + k]) * s[j];
        am[k] = t;
    }
    if (G_out) memcpy(G_out, red, sizeof(double) * (m * m + m));
    if (s_out) memcpy(s_out, s, sizeof(double) * m);
    free(red); free(G); free(g); free(s);
    return st;
}

static void finish_stats(orc_stats *st, int iters, double relres, double rhs_norm,
                         double tol_abs, int loop_iter, int max_iter, int trace_len,
                         double secs, int status)
{
    if (!st) return;
    st->iterations = iters;
    st->relres = relres / rhs_norm;
    st->rhs_norm = rhs_norm;
    st->converged = relres <= tol_abs;
    st->printed_converged = (loop_iter < max_iter && relres <= tol_abs);
    st->printed_warning = (loop_iter >= max_iter);
    st->trace_len = trace_len;
    st->seconds = secs;
    st->status = status;
}

/* ------------------------------------------------------------------ */
/* AAR.c:37-177.                                                       */
/* ------------------------------------------------------------------ */
/* y = a*Lap(x) + c(x) x: the reference operator (system.c:189-205) plus a
 * pointwise term added after it, as the screened term of orc_matvec. */
void orc_matvec_vc(int nx, int ny, int nz, int fdn, const double *coeff,
                   double a, const double *c, const double *x, double *y)
{
    orc_matvec(nx, ny, nz, fdn, coeff, a, 0.0, x, y);
    const long n = (long)nx * ny * nz;
    for (long i = 0; i < n; i++) y[i] = y[i] + c[i] * x[i];
}

/* the operator of the *_impl solvers: a*Lap + sigma, or a*Lap + c(x) */
#define OPV(xin, yout)                                                      \
    do {                                                                    \
        if (cv) orc_matvec_vc(nx, ny, nz, fdn, coeff, a, cv, xin, yout);    \
        else orc_matvec(nx, ny, nz, fdn, coeff, a, sigma, xin, yout);       \
    } while (0)

/* AAR.c:37-177 for the operator a*Lap + sigma (c == NULL) or a*Lap + c(x);
 * with c the Jacobi step of Precondition (system.c:222-227) divides by the
 * pointwise diagonal a*(3*coeff[0]) + c(x). */
static int aar_impl(int nx, int ny, int nz, int fdn, double a, double sigma, const double *cv,
                    double *x, const double *rhs, double omega, double beta,
                    int m, int p, int max_iter, double tol, int precond,
                    double *trace, int trace_cap, orc_stats *st)
{
    const long n = (long)nx * ny * nz;
    double coeff[32] = {0};
    orc_lap_coefficient(coeff, fdn);
    double diag = a * (3 * coeff[0]);
    if (sigma != 0.0) diag += sigma;

    double *f = (double *)calloc((size_t)n, sizeof(double));
    double *x_old = (double *)calloc((size_t)n, sizeof(double));
    double *f_old = (double *)calloc((size_t)n, sizeof(double));
    double *am = (double *)calloc((size_t)n, sizeof(double));
    double *DX = (double *)calloc((size_t)n * m, sizeof(double));
    double *DF = (double *)calloc((size_t)n * m, sizeof(double));
    int tlen = 0, status = 0;

    int iter = 1;
    double relres = tol + 1;                 /* AAR.c:76, before tol is scaled */
    double rhs_norm = orc_norm2(rhs, n);
    tol *= rhs_norm;
    OPV(x, f);
    for (long i = 0; i < n; i++) f[i] = rhs[i] - f[i];
    if (trace && tlen < trace_cap) trace[tlen++] = orc_norm2(f, n);

    double t0 = now_seconds();
    while (relres > tol && iter <= max_iter) {
        if (precond) {
            if (cv) for (long i = 0; i < n; i++) f[i] /= (diag + cv[i]);
            else for (long i = 0; i < n; i++) f[i] /= diag;
        }
        if (iter > 1) {
            const int col = (iter - 2) % m;
            double *dx = DX + (long)col * n, *df = DF + (long)col * n;
            for (long i = 0; i < n; i++) {
                dx[i] = x[i] - x_old[i];
                df[i] = f[i] - f_old[i];
            }
        }
        memcpy(x_old, x, sizeof(double) * n);
        memcpy(f_old, f, sizeof(double) * n);
        double rn = -1.0;
        if (iter % p == 0 && iter > 1) {
            if (orc_anderson_extrapolation(DX, DF, f, beta, m, n, am, NULL, NULL)) { status = 1; break; }
            for (long i = 0; i < n; i++) x[i] = x_old[i] + beta * f[i] - am[i];
            OPV(x, f);
            for (long i = 0; i < n; i++) f[i] = rhs[i] - f[i];
            relres = orc_norm2(f, n);
            rn = relres;
        } else {
            for (long i = 0; i < n; i++) x[i] = x_old[i] + omega * f[i];
            OPV(x, f);
            for (long i = 0; i < n; i++) f[i] = rhs[i] - f[i];
        }
        if (trace && tlen < trace_cap) trace[tlen++] = rn >= 0.0 ? rn : orc_norm2(f, n);
        iter = iter + 1;
    }
    double t1 = now_seconds();
    finish_stats(st, iter - 1, relres, rhs_norm, tol, iter, max_iter, tlen, t1 - t0, status);
    free(f); free(x_old); free(f_old); free(am); free(DX); free(DF);
    return status;
}

int orc_aar(int nx, int ny, int nz, int fdn, double a, double sigma,
            double *x, const double *rhs, double omega, double beta,
            int m, int p, int max_iter, double tol, int precond,
            double *trace, int trace_cap, orc_stats *st)
{
    return aar_impl(nx, ny, nz, fdn, a, sigma, NULL, x, rhs, omega, beta, m, p, max_iter, tol, precond,
                    trace, trace_cap, st);
}

int orc_aar_vc(int nx, int ny, int nz, int fdn, double a, const double *c,
               double *x, const double *rhs, double omega, double beta,
               int m, int p, int max_iter, double tol, int precond,
               double *trace, int trace_cap, orc_stats *st)
{
    return aar_impl(nx, ny, nz, fdn, a, 0.0, c, x, rhs, omega, beta, m, p, max_iter, tol, precond,
                    trace, trace_cap, st);
}

/* ------------------------------------------------------------------ */
/* PGR.c:35-218 and PL2R.c:35-218.                                     */
/* ------------------------------------------------------------------ */
/* PGR.c:35-218 / PL2R.c:35-218 for a*Lap + sigma (cv == NULL) or
 * a*Lap + c(x) with the pointwise Jacobi diagonal. */
static int pgr_impl(int nx, int ny, int nz, int fdn, double a, double sigma, const double *cv,
                    double *x, const double *rhs, double omega, int m, int p,
                    int max_iter, double tol, int precond, int l2,
                    double *trace, int trace_cap, orc_stats *st)
{
    const long n = (long)nx * ny * nz;
    double coeff[32] = {0};
    orc_lap_coefficient(coeff, fdn);
    double diag = a * (3 * coeff[0]);
    if (sigma != 0.0) diag += sigma;

    double *Ax = (double *)calloc((size_t)n, sizeof(double));
    double *Ax_old = (double *)calloc((size_t)n, sizeof(double));
    double *f = (double *)calloc((size_t)n, sizeof(double));
    double *x_old = (double *)calloc((size_t)n, sizeof(double));
    double *DX = (double *)calloc((size_t)n * m, sizeof(double));
    double *DF = (double *)calloc((size_t)n * m, sizeof(double));
    double *red = (double *)calloc((size_t)(m * m + m), sizeof(double));
    double *G = (double *)calloc((size_t)m * m, sizeof(double));
    double *g = (double *)calloc((size_t)m, sizeof(double));
    double *s = (double *)calloc((size_t)m, sizeof(double));
    int tlen = 0, status = 0;

    int iter = 1;
    double relres = tol + 1;
    double rhs_norm = orc_norm2(rhs, n);
    tol *= rhs_norm;
    OPV(x, Ax);
    for (long i = 0; i < n; i++) f[i] = rhs[i] - Ax[i];
    if (trace && tlen < trace_cap) trace[tlen++] = orc_norm2(f, n);

    double t0 = now_seconds();
    while (relres > tol && iter <= max_iter) {
        if (precond) {
            if (cv) for (long i = 0; i < n; i++) f[i] /= (diag + cv[i]);
            else for (long i = 0; i < n; i++) f[i] /= diag;
        }
        for (long i = 0; i < n; i++) {
            x_old[i] = x[i];
            Ax_old[i] = Ax[i];
            x[i] += (omega * f[i]);
        }
        OPV(x, Ax);
        for (long i = 0; i < n; i++) f[i] = rhs[i] - Ax[i];
        if (trace && tlen < trace_cap) trace[tlen++] = orc_norm2(f, n);

        const int col = (iter - 2 + m) % m;
        double *dx = DX + (long)col * n, *df = DF + (long)col * n;
        for (long i = 0; i < n; i++) {
            dx[i] = x[i] - x_old[i];
            df[i] = Ax[i] - Ax_old[i];
        }
        if (iter % p == 0) {
            /* PGR.c:181-218 (DX^T DF, DX^T f) or PL2R.c:180-218 (DF^T DF, DF^T f) */
            if (!l2) {
                for (int j = 0; j < m; j++)
                    for (int i = 0; i < m; i++)
                        red[j * m + i] = orc_dot(DX + (long)i * n, DF + (long)j * n, n);
                for (int j = 0; j < m; j++) red[m * m + j] = orc_dot(DX + (long)j * n, f, n);
            } else {
                for (int j = 0; j < m; j++)
                    for (int i = 0; i <= j; i++) {
                        double t = orc_dot(DF + (long)i * n, DF + (long)j * n, n);
                        red[j * m + i] = t;
                        red[i * m + j] = t;
                    }
                for (int j = 0; j < m; j++) red[m * m + j] = orc_dot(DF + (long)j * n, f, n);
            }
            for (int j = 0; j < m; j++)
                for (int i = 0; i < m; i++) G[i * m + j] = red[j * m + i];
            for (int j = 0; j < m; j++) g[j] = red[m * m + j];
            if (orc_pinv_times_vec(G, g, s, m, NULL, NULL)) { status = 1; break; }
            for (int i = 0; i < m; i++) {
                const double *dxi = DX + (long)i * n, *dfi = DF + (long)i * n;
                for (long j = 0; j < n; j++) {
                    x[j] = x[j] + dxi[j] * s[i];
                    Ax[j] = Ax[j] + dfi[j] * s[i];
                }
            }
            for (long i = 0; i < n; i++) {
                dx[i] = x[i] - x_old[i];
                df[i] = Ax[i] - Ax_old[i];
            }
        }
        for (long i = 0; i < n; i++) f[i] = rhs[i] - Ax[i];
        if (iter % p == 0) relres = orc_norm2(f, n);
        iter = iter + 1;
    }
    double t1 = now_seconds();
    finish_stats(st, iter - 1, relres, rhs_norm, tol, iter, max_iter, tlen, t1 - t0, status);
    free(Ax); free(Ax_old); free(f); free(x_old); free(DX); free(DF);
    free(red); free(G); free(g); free(s);
    return status;
}

int orc_pgr(int nx, int ny, int nz, int fdn, double a, double sigma,
            double *x, const double *rhs, double omega, int m, int p,
            int max_iter, double tol, int precond, int l2,
            double *trace, int trace_cap, orc_stats *st)
{
    return pgr_impl(nx, ny, nz, fdn, a, sigma, NULL, x, rhs, omega, m, p, max_iter, tol, precond, l2,
                    trace, trace_cap, st);
}

int orc_pgr_vc(int nx, int ny, int nz, int fdn, double a, const double *c,
               double *x, const double *rhs, double omega, int m, int p,
               int max_iter, double tol, int precond, int l2,
               double *trace, int trace_cap, orc_stats *st)
{
    return pgr_impl(nx, ny, nz, fdn, a, 0.0, c, x, rhs, omega, m, p, max_iter, tol, precond, l2,
                    trace, trace_cap, st);
}

/* ------------------------------------------------------------------ */
/* CG.c:33-175.                                                        */
/* ------------------------------------------------------------------ */
/* CG.c:33-175 for a*Lap + sigma (cv == NULL) or a*Lap + c(x) with the
 * pointwise Jacobi diagonal. */
static int cg_impl(int nx, int ny, int nz, int fdn, double a, double sigma, const double *cv,
                   double *x, const double *b, int max_iter, double tol, int precond,
                   double *trace, int trace_cap, orc_stats *st)
{
    const long n = (long)nx * ny * nz;
    double coeff[32] = {0};
    orc_lap_coefficient(coeff, fdn);
    double diag = a * (3 * coeff[0]);
    if (sigma != 0.0) diag += sigma;
    double *r = (double *)malloc(sizeof(double) * n);
    double *d = (double *)malloc(sizeof(double) * n);
    double *s = (double *)malloc(sizeof(double) * n);
    double *q = (double *)malloc(sizeof(double) * n);
    int tlen = 0;

    double b_2norm = orc_norm2(b, n);
    OPV(x, r);
    for (long i = 0; i < n; i++) {
        r[i] = b[i] - r[i];
        d[i] = r[i];
    }
    if (precond) {
        if (cv) for (long i = 0; i < n; i++) d[i] /= (diag + cv[i]);
        else for (long i = 0; i < n; i++) d[i] /= diag;
    }
    double delta_new = orc_dot(r, d, n), delta_old;
    double relres = precond ? orc_norm2(r, n) : sqrt(delta_new);
    if (trace && tlen < trace_cap) trace[tlen++] = orc_norm2(r, n);
    tol *= b_2norm;

    int iter_count = 0;
    double t0 = now_seconds();
    while (iter_count < max_iter && relres > tol) {
        OPV(d, q);
        double alpha = orc_dot(d, q, n);
        alpha = delta_new / alpha;
        for (long i = 0; i < n; i++) x[i] = x[i] + alpha * d[i];
        if ((iter_count % 50) == 0) {
            OPV(x, r);
            for (long i = 0; i < n; i++) {
                r[i] = b[i] - r[i];
                s[i] = r[i];
            }
            if (trace && tlen < trace_cap) trace[tlen++] = orc_norm2(r, n);
        } else {
            for (long i = 0; i < n; i++) {
                r[i] = r[i] - alpha * q[i];
                s[i] = r[i];
            }
        }
        if (precond) {
            if (cv) for (long i = 0; i < n; i++) s[i] /= (diag + cv[i]);
            else for (long i = 0; i < n; i++) s[i] /= diag;
        }
        delta_old = delta_new;
        delta_new = orc_dot(r, s, n);
        relres = precond ? orc_norm2(r, n) : sqrt(delta_new);
        double beta = delta_new / delta_old;
        for (long i = 0; i < n; i++) d[i] = s[i] + beta * d[i];
        iter_count++;
    }
    double t1 = now_seconds();
    finish_stats(st, iter_count, relres, b_2norm, tol, iter_count, max_iter, tlen, t1 - t0, 0);
    free(r); free(d); free(s); free(q);
    return 0;
}

int orc_cg(int nx, int ny, int nz, int fdn, double a, double sigma,
           double *x, const double *b, int max_iter, double tol, int precond,
           double *trace, int trace_cap, orc_stats *st)
{
    return cg_impl(nx, ny, nz, fdn, a, sigma, NULL, x, b, max_iter, tol, precond, trace, trace_cap, st);
}

int orc_cg_vc(int nx, int ny, int nz, int fdn, double a, const double *c,
              double *x, const double *b, int max_iter, double tol, int precond,
              double *trace, int trace_cap, orc_stats *st)
{
    return cg_impl(nx, ny, nz, fdn, a, 0.0, c, x, b, max_iter, tol, precond, trace, trace_cap, st);
}

/* ------------------------------------------------------------------ */
/* Complex Helmholtz (extension).  Operator order per point:           */
/*   lap = a*Lap(u) per component in the order of system.c:189-205,   */
/*   y_re = lap_re + (c*u_re - eta*u_im), y_im = lap_im + (c*u_im +    */
/*   eta*u_re).  Jacobi divides by d = (a*(3*c0_lap) + c) + i*eta with */
/*   f = ((r_re d_re + r_im d_im) + i (r_im d_re - r_re d_im)) / |d|^2.*/
/* ------------------------------------------------------------------ */
void orc_matvec_complex(int nx, int ny, int nz, int fdn, const double *coeff,
                        double a, const double *c, double c0, double eta,
                        const double *u, double *y)
{
    const long n = (long)nx * ny * nz;
    double *ur = (double *)malloc(sizeof(double) * n);
    double *ui = (double *)malloc(sizeof(double) * n);
    double *yr = (double *)malloc(sizeof(double) * n);
    double *yi = (double *)malloc(sizeof(double) * n);
    for (long k = 0; k < n; k++) { ur[k] = u[2 * k]; ui[k] = u[2 * k + 1]; }
    orc_matvec(nx, ny, nz, fdn, coeff, a, 0.0, ur, yr);
    orc_matvec(nx, ny, nz, fdn, coeff, a, 0.0, ui, yi);
    for (long k = 0; k < n; k++) {
        double ck = c ? c[k] : c0;
        y[2 * k] = yr[k] + (ck * ur[k] - eta * ui[k]);
        y[2 * k + 1] = yi[k] + (ck * ui[k] + eta * ur[k]);
    }
    free(ur); free(ui); free(yr); free(yi);
}

static double cnorm2(const double *v, long n)
{
    double s = 0.0;
    for (long k = 0; k < n; k++) s = s + (v[2 * k] * v[2 * k] + v[2 * k + 1] * v[2 * k + 1]);
    return sqrt(s);
}

/* <u, v> = sum conj(u) v */
static void cdot(const double *u, const double *v, long n, double *re, double *im)
{
    double sr = 0.0, si = 0.0;
    for (long k = 0; k < n; k++) {
        sr = sr + (u[2 * k] * v[2 * k] + u[2 * k + 1] * v[2 * k + 1]);
        si = si + (u[2 * k] * v[2 * k + 1] - u[2 * k + 1] * v[2 * k]);
    }
    *re = sr;
    *im = si;
}

int orc_aar_complex(int nx, int ny, int nz, int fdn, double a,
                    const double *c, double c0, double eta,
                    double *x, const double *rhs, double omega, double beta,
                    int m, int p, int max_iter, double tol,
                    double *trace, int trace_cap, orc_stats *st)
{
    const long n = (long)nx * ny * nz;
    const long n2 = 2 * n;
    double coeff[32] = {0};
    orc_lap_coefficient(coeff, fdn);
    const double dlap = a * (3 * coeff[0]);
    double *f = (double *)calloc((size_t)n2, sizeof(double));
    double *x_old = (double *)calloc((size_t)n2, sizeof(double));
    double *f_old = (double *)calloc((size_t)n2, sizeof(double));
    double *DX = (double *)calloc((size_t)n2 * m, sizeof(double));
    double *DF = (double *)calloc((size_t)n2 * m, sizeof(double));
    const int m2 = 2 * m;
    double *E = (double *)calloc((size_t)m2 * m2, sizeof(double));
    double *e = (double *)calloc((size_t)m2, sizeof(double));
    double *se = (double *)calloc((size_t)m2, sizeof(double));
    double *Gr = (double *)calloc((size_t)m * m, sizeof(double));
    double *Gi = (double *)calloc((size_t)m * m, sizeof(double));
    int tlen = 0, status = 0;

    int iter = 1;
    double relres = tol + 1;
    double rhs_norm = cnorm2(rhs, n);
    tol *= rhs_norm;
    orc_matvec_complex(nx, ny, nz, fdn, coeff, a, c, c0, eta, x, f);
    for (long i = 0; i < n2; i++) f[i] = rhs[i] - f[i];
    if (trace && tlen < trace_cap) trace[tlen++] = cnorm2(f, n);

    double t0 = now_seconds();
    while (relres > tol && iter <= max_iter) {
        for (long k = 0; k < n; k++) {
            const double dr = dlap + (c ? c[k] : c0), di = eta;
            const double den = dr * dr + di * di;
            const double rr = f[2 * k], ri = f[2 * k + 1];
            f[2 * k] = (rr * dr + ri * di) / den;
            f[2 * k + 1] = (ri * dr - rr * di) / den;
        }
        if (iter > 1) {
            const int col = (iter - 2) % m;
            double *dx = DX + (long)col * n2, *df = DF + (long)col * n2;
            for (long i = 0; i < n2; i++) {
                dx[i] = x[i] - x_old[i];
                df[i] = f[i] - f_old[i];
            }
        }
        memcpy(x_old, x, sizeof(double) * n2);
        memcpy(f_old, f, sizeof(double) * n2);
        double rn = -1.0;
        if (iter % p == 0 && iter > 1) {
            /* Hermitian Gram G_ij = <DF_i, DF_j>, g_i = <DF_i, f>
             * (PETSC_COMPLEX/AAR_Complex.c:213-236), solved through the real
             * embedding [Gr -Gi; Gi Gr][sr; si] = [gr; gi]. */
            for (int j = 0; j < m; j++)
                for (int i = 0; i <= j; i++) {
                    double re, im;
                    cdot(DF + (long)i * n2, DF + (long)j * n2, n, &re, &im);
                    Gr[i * m + j] = re; Gi[i * m + j] = im;
                    Gr[j * m + i] = re; Gi[j * m + i] = -im;
                }
            for (int i = 0; i < m; i++) cdot(DF + (long)i * n2, f, n, &e[i], &e[m + i]);
            for (int i = 0; i < m; i++)
                for (int j = 0; j < m; j++) {
                    E[i * m2 + j] = Gr[i * m + j];
                    E[i * m2 + m + j] = -Gi[i * m + j];
                    E[(m + i) * m2 + j] = Gi[i * m + j];
                    E[(m + i) * m2 + m + j] = Gr[i * m + j];
                }
            if (orc_pinv_times_vec_textbook(E, e, se, m2, NULL, NULL)) { status = 1; break; }
            for (long k = 0; k < n; k++) {
                double tr = 0.0, ti = 0.0;
                for (int j = 0; j < m; j++) {
                    const double *dx = DX + (long)j * n2, *df = DF + (long)j * n2;
                    const double vr = dx[2 * k] + beta * df[2 * k];
                    const double vi = dx[2 * k + 1] + beta * df[2 * k + 1];
                    tr = tr + (vr * se[j] - vi * se[m + j]);
                    ti = ti + (vr * se[m + j] + vi * se[j]);
                }
                x[2 * k] = x_old[2 * k] + beta * f[2 * k] - tr;
                x[2 * k + 1] = x_old[2 * k + 1] + beta * f[2 * k + 1] - ti;
            }
            orc_matvec_complex(nx, ny, nz, fdn, coeff, a, c, c0, eta, x, f);
            for (long i = 0; i < n2; i++) f[i] = rhs[i] - f[i];
            relres = cnorm2(f, n);
            rn = relres;
        } else {
            for (long i = 0; i < n2; i++) x[i] = x_old[i] + omega * f[i];
            orc_matvec_complex(nx, ny, nz, fdn, coeff, a, c, c0, eta, x, f);
            for (long i = 0; i < n2; i++) f[i] = rhs[i] - f[i];
        }
        if (trace && tlen < trace_cap) trace[tlen++] = rn >= 0.0 ? rn : cnorm2(f, n);
        iter = iter + 1;
    }
    double t1 = now_seconds();
    finish_stats(st, iter - 1, relres, rhs_norm, tol, iter, max_iter, tlen, t1 - t0, status);
    free(f); free(x_old); free(f_old); free(DX); free(DF);
    free(E); free(e); free(se); free(Gr); free(Gi);
    return status;
}

/* ------------------------------------------------------------------ */
/* Krylov comparators of the PETSc variant (PETSC_REAL/main.c:86-150): */
/* KSPGMRES and KSPBICG with PCJACOBI, rtol 1e-6, max_it 2000 and a    */
/* nonzero initial guess (main.c:97-101, 131-133).  PETSc (3.11.3 in   */
/* the reference's README) is neither vendored nor installed; these    */
/* follow its published algorithms and defaults:                       */
/*  - KSPConvergedDefault (src/ksp/ksp/interface/iterativ.c): at n = 0 */
/*    ttol = max(rtol*||B b||, atol) for a nonzero guess, rnorm0 =     */
/*    rnorm; converged when rnorm <= ttol (ATOL if rnorm < atol, else  */
/*    RTOL); diverged when rnorm >= dtol*rnorm0; atol 1e-50, dtol 1e5; */
/*  - PCJACOBI (src/ksp/pc/impls/jacobi/jacobi.c): y = x * (1/diag),   */
/*    the reciprocal taken once (the MPI_REAL Precondition divides);   */
/*  - KSPGMRES (src/ksp/ksp/impls/gmres/gmres.c, borthog.c): restart   */
/*    30, left preconditioning, the preconditioned residual norm,      */
/*    classical Gram-Schmidt without refinement, Givens rotations,     */
/*    happy-breakdown tolerance 1e-30;                                 */
/*  - KSPBICG (src/ksp/ksp/impls/bicg/bicg.c): left-preconditioned     */
/*    BiCG with the shadow residual r~0 = r0, preconditioned norm.     */
/* Parity with PETSc itself is unpinned (rounding order of its BLAS    */
/* kernels); the tests pin these routines by what the mathematics      */
/* fixes (GMRES minimises the preconditioned residual over the Krylov */
/* space; BiCG on a symmetric operator is PCG).                        */
/* ------------------------------------------------------------------ */
typedef struct {
    double ttol, rnorm0, rtol, bpnorm;
} ksp_conv;

static int ksp_converged(ksp_conv *c, int n, double rnorm)
{
    if (n == 0) {
        c->ttol = c->rtol * c->bpnorm;
        if (c->ttol < ORC_KSP_ATOL) c->ttol = ORC_KSP_ATOL;
        c->rnorm0 = rnorm;
    }
    if (isnan(rnorm) || isinf(rnorm)) return ORC_KSP_DIVERGED_NANORINF;
    if (rnorm <= c->ttol) return rnorm < ORC_KSP_ATOL ? ORC_KSP_CONVERGED_ATOL : ORC_KSP_CONVERGED_RTOL;
    if (rnorm >= ORC_KSP_DTOL * c->rnorm0) return ORC_KSP_DIVERGED_DTOL;
    return 0;
}

/* the Jacobi reciprocal 1/diag per point (1 without a preconditioner) */
static double *ksp_dinv(long n, double diag, const double *cv, int precond)
{
    double *d = (double *)malloc(sizeof(double) * n);
    for (long i = 0; i < n; i++) d[i] = precond ? 1.0 / (cv ? diag + cv[i] : diag) : 1.0;
    return d;
}

/* the preconditioner B: none (kind 0), Jacobi (1: y = x * (1/diag)) or
   ILU(0) (2: orc_ilu0_apply with the factors F) */
typedef struct {
    int kind;
    const double *dinv, *F;
    int nx, ny, nz;
} ksp_pc;

static void pc_apply(const ksp_pc *pc, const double *in, double *out, long n, int trans)
{
    if (pc->kind == 2) {
        orc_ilu0_apply(pc->nx, pc->ny, pc->nz, pc->F, in, out, trans);
        return;
    }
    for (long i = 0; i < n; i++) out[i] = pc->dinv[i] * in[i];
}

static int pc_setup(ksp_pc *pc, int precond, int nx, int ny, int nz, int fdn, double a, double sigma,
                    const double *cv, double diag, double **dinv, double **F)
{
    const long n = (long)nx * ny * nz;
    pc->kind = precond;
    pc->nx = nx;
    pc->ny = ny;
    pc->nz = nz;
    *dinv = ksp_dinv(n, diag, cv, precond == 1);
    *F = NULL;
    pc->dinv = *dinv;
    pc->F = NULL;
    if (precond == 2) {
        *F = (double *)malloc(sizeof(double) * n * ILU_SLOTS);
        if (orc_ilu0_factor(nx, ny, nz, fdn, a, sigma, cv, *F)) return -1;
        pc->F = *F;
    }
    return 0;
}

/* VecNormalize: scale by 1/norm unless the norm is 0 or 1 */
static double ksp_normalize(double *v, long n)
{
    double nrm = orc_norm2(v, n);
    if (nrm != 0.0 && nrm != 1.0) {
        const double s = 1.0 / nrm;
        for (long i = 0; i < n; i++) v[i] = s * v[i];
    }
    return nrm;
}

static void ksp_finish(orc_stats *st, int its, int reason, const double *x, const double *b, double *t,
                       int nx, int ny, int nz, int fdn, double a, double sigma, const double *cv,
                       const double *coeff, int tlen, double secs)
{
    if (!st) return;
    const long n = (long)nx * ny * nz;
    OPV(x, t);
    for (long i = 0; i < n; i++) t[i] = b[i] - t[i];
    const double bn = orc_norm2(b, n);
    st->iterations = its;
    st->relres = orc_norm2(t, n) / bn;      /* main.c:112-118: ||b - A x|| / ||b|| */
    st->rhs_norm = bn;
    st->converged = reason > 0;
    st->printed_converged = reason > 0;
    st->printed_warning = reason == ORC_KSP_DIVERGED_ITS;
    st->trace_len = tlen;
    st->seconds = secs;
    st->status = reason;
}

static int gmres_impl(int nx, int ny, int nz, int fdn, double a, double sigma, const double *cv,
                      double *x, const double *b, int restart, int max_it, double rtol, int precond,
                      double *trace, int trace_cap, orc_stats *st)
{
    if (restart < 1 || restart > ORC_GMRES_MAXK) return -1;
    const long n = (long)nx * ny * nz;
    const int K = restart;
    double coeff[32] = {0};
    orc_lap_coefficient(coeff, fdn);
    double diag = a * (3 * coeff[0]);
    if (sigma != 0.0) diag += sigma;
    double *dinv, *F;
    ksp_pc pc;
    if (pc_setup(&pc, precond, nx, ny, nz, fdn, a, sigma, cv, diag, &dinv, &F)) {
        free(dinv); free(F);
        return -2;
    }
    double *V = (double *)malloc(sizeof(double) * n * (K + 1));
    double *t = (double *)malloc(sizeof(double) * n);
    double HH[(ORC_GMRES_MAXK + 2) * (ORC_GMRES_MAXK + 1)];
    double cc[ORC_GMRES_MAXK + 1], ss[ORC_GMRES_MAXK + 1], grs[ORC_GMRES_MAXK + 2], nrs[ORC_GMRES_MAXK + 1];
    double h[ORC_GMRES_MAXK + 1];
#define HHE(r, c) HH[(c) * (K + 2) + (r)]
    int tlen = 0;
    ksp_conv cv_ = {0.0, 0.0, rtol, 0.0};
    pc_apply(&pc, b, t, n, 0);
    cv_.bpnorm = orc_norm2(t, n);                       /* ||B b|| */

    int its = 0, reason = 0;
    double t0 = now_seconds();
    while (!reason) {
        /* KSPInitialResidual: v0 = B (b - A x) */
        double *v0 = V;
        OPV(x, t);
        for (long i = 0; i < n; i++) t[i] = b[i] - t[i];
        pc_apply(&pc, t, v0, n, 0);
        double res = ksp_normalize(v0, n);
        grs[0] = res;
        int it = 0;
        if (trace && tlen < trace_cap) trace[tlen++] = res;
        if (res == 0.0) {
            reason = ORC_KSP_CONVERGED_ATOL;
            break;
        }
        reason = ksp_converged(&cv_, its, res);
        while (!reason && it < K && its < max_it) {
            double *vn = V + (long)(it + 1) * n;
            OPV(V + (long)it * n, t);
            pc_apply(&pc, t, vn, n, 0);
            /* classical Gram-Schmidt: h_j = <w, v_j>, then w -= sum_j h_j v_j */
            for (int j = 0; j <= it; j++) h[j] = orc_dot(vn, V + (long)j * n, n);
            for (long i = 0; i < n; i++) {
                double w = vn[i];
                for (int j = 0; j <= it; j++) w = w + (-h[j]) * V[(long)j * n + i];
                vn[i] = w;
            }
            for (int j = 0; j <= it; j++) HHE(j, it) = h[j];
            const double tt = ksp_normalize(vn, n);
            HHE(it + 1, it) = tt;
            double hapbnd = fabs(tt / grs[it]);
            if (hapbnd > ORC_GMRES_HAPTOL) hapbnd = ORC_GMRES_HAPTOL;
            const int hapend = tt < hapbnd;
            /* KSPGMRESUpdateHessenberg: earlier rotations, then a new one */
            double *hh = &HHE(0, it);
            for (int j = 1; j <= it; j++) {
                const double t1 = hh[j - 1];
                hh[j - 1] = cc[j - 1] * t1 + ss[j - 1] * hh[j];
                hh[j] = cc[j - 1] * hh[j] - ss[j - 1] * t1;
            }
            if (!hapend) {
                const double r = sqrt(hh[it] * hh[it] + hh[it + 1] * hh[it + 1]);
                if (r == 0.0) {
                    reason = ORC_KSP_DIVERGED_NULL;
                } else {
                    cc[it] = hh[it] / r;
                    ss[it] = hh[it + 1] / r;
                    grs[it + 1] = -(ss[it] * grs[it]);
                    grs[it] = cc[it] * grs[it];
                    hh[it] = cc[it] * hh[it] + ss[it] * hh[it + 1];
                    res = fabs(grs[it + 1]);
                }
            } else {
                res = 0.0;
            }
            it++;
            its++;
            if (trace && tlen < trace_cap) trace[tlen++] = res;
            if (!reason) reason = ksp_converged(&cv_, its, res);
            if (hapend) {
                if (!reason) reason = ORC_KSP_DIVERGED_BREAKDOWN;
                break;
            }
        }
        /* KSPGMRESBuildSoln: back-substitution, x += sum_j nrs_j v_j */
        if (it > 0) {
            const int k = it - 1;
            nrs[k] = HHE(k, k) != 0.0 ? grs[k] / HHE(k, k) : 0.0;
            for (int ii = 1; ii <= k; ii++) {
                const int kk = k - ii;
                double tt = grs[kk];
                for (int j = kk + 1; j <= k; j++) tt = tt - HHE(kk, j) * nrs[j];
                nrs[kk] = tt / HHE(kk, kk);
            }
            for (long i = 0; i < n; i++) {
                double s = 0.0;
                for (int j = 0; j <= k; j++) s = s + nrs[j] * V[(long)j * n + i];
                x[i] = x[i] + s;
            }
        }
        if (its >= max_it) {
            if (!reason) reason = ORC_KSP_DIVERGED_ITS;
            break;
        }
    }
#undef HHE
    double t1 = now_seconds();
    ksp_finish(st, its, reason, x, b, t, nx, ny, nz, fdn, a, sigma, cv, coeff, tlen, t1 - t0);
    free(dinv); free(F); free(V); free(t);
    return 0;
}

int orc_gmres(int nx, int ny, int nz, int fdn, double a, double sigma, const double *c,
              double *x, const double *b, int restart, int max_it, double rtol, int precond,
              double *trace, int trace_cap, orc_stats *st)
{
    return gmres_impl(nx, ny, nz, fdn, a, sigma, c, x, b, restart, max_it, rtol, precond, trace, trace_cap, st);
}

static int bicg_impl(int nx, int ny, int nz, int fdn, double a, double sigma, const double *cv,
                     double *x, const double *b, int max_it, double rtol, int precond,
                     double *trace, int trace_cap, orc_stats *st, double *shadow_gap)
{
    const long n = (long)nx * ny * nz;
    double coeff[32] = {0};
    orc_lap_coefficient(coeff, fdn);
    double diag = a * (3 * coeff[0]);
    if (sigma != 0.0) diag += sigma;
    double *dinv, *F;
    ksp_pc pc;
    if (pc_setup(&pc, precond, nx, ny, nz, fdn, a, sigma, cv, diag, &dinv, &F)) {
        free(dinv); free(F);
        return -2;
    }
    double *Rr = (double *)malloc(sizeof(double) * n), *Rl = (double *)malloc(sizeof(double) * n);
    double *Zr = (double *)malloc(sizeof(double) * n), *Zl = (double *)malloc(sizeof(double) * n);
    double *Pr = (double *)malloc(sizeof(double) * n), *Pl = (double *)malloc(sizeof(double) * n);
    int tlen = 0;
    ksp_conv cv_ = {0.0, 0.0, rtol, 0.0};
    pc_apply(&pc, b, Zr, n, 0);
    cv_.bpnorm = orc_norm2(Zr, n);
    double gap = 0.0;

    double t0 = now_seconds();
    OPV(x, Rr);
    for (long i = 0; i < n; i++) Rr[i] = b[i] + (-1.0) * Rr[i];     /* VecAYPX(Rr, -1, b) */
    memcpy(Rl, Rr, sizeof(double) * n);
    pc_apply(&pc, Rr, Zr, n, 0);
    pc_apply(&pc, Rl, Zl, n, 1);              /* B^T r~ */
    double dp = orc_norm2(Zr, n);
    if (trace && tlen < trace_cap) trace[tlen++] = dp;
    int its = 0;
    int reason = ksp_converged(&cv_, 0, dp);
    double beta, betaold = 0.0;
    int i = 0;
    if (!reason) {
        do {
            beta = orc_dot(Zr, Rl, n);
            if (!i) {
                if (beta == 0.0) {
                    reason = ORC_KSP_DIVERGED_BREAKDOWN_BICG;
                    break;
                }
                memcpy(Pr, Zr, sizeof(double) * n);
                memcpy(Pl, Zl, sizeof(double) * n);
            } else {
                const double bb = beta / betaold;
                for (long k = 0; k < n; k++) {
                    Pr[k] = Zr[k] + bb * Pr[k];
                    Pl[k] = Zl[k] + bb * Pl[k];
                }
            }
            betaold = beta;
            OPV(Pr, Zr);
            OPV(Pl, Zl);                      /* A^T p~: the FD12 operator is symmetric */
            const double dpi = orc_dot(Zr, Pl, n);
            const double al = beta / dpi;
            for (long k = 0; k < n; k++) {
                x[k] = x[k] + al * Pr[k];
                Rr[k] = Rr[k] + (-al) * Zr[k];
                Rl[k] = Rl[k] + (-al) * Zl[k];
            }
            pc_apply(&pc, Rr, Zr, n, 0);
            pc_apply(&pc, Rl, Zl, n, 1);
            dp = orc_norm2(Zr, n);
            for (long k = 0; k < n; k++) {
                const double d = fabs(Rl[k] - Rr[k]);
                if (d > gap) gap = d;
            }
            its = i + 1;
            if (trace && tlen < trace_cap) trace[tlen++] = dp;
            reason = ksp_converged(&cv_, i + 1, dp);
            if (reason) break;
            i++;
        } while (i < max_it);
        if (i >= max_it && !reason) reason = ORC_KSP_DIVERGED_ITS;
    }
    double t1 = now_seconds();
    if (shadow_gap) *shadow_gap = gap;
    ksp_finish(st, its, reason, x, b, Rr, nx, ny, nz, fdn, a, sigma, cv, coeff, tlen, t1 - t0);
    free(dinv); free(F); free(Rr); free(Rl); free(Zr); free(Zl); free(Pr); free(Pl);
    return 0;
}

/* KSPCG (PETSc 3.11 KSPSolve_CG, KSP_CG_SYMMETRIC, preconditioned norm),
 * the CG of PETSC_REAL/main.c:152-175.  r = b - A x, z = B r, dp = ||z||;
 * converged at 0?  beta = z.r; then per step i: its = i + 1; beta = 0 ->
 * CONVERGED_ATOL, beta * betaold < 0 -> DIVERGED_INDEFINITE_PC; p = z
 * (i = 0) or z + (beta/betaold) p; w = A p, dpi = p.w (0, or a sign change
 * of dpi -> DIVERGED_INDEFINITE_MAT); a = beta/dpi; x += a p; r -= a w;
 * z = B r; dp = ||z||; convergence test; beta = z.r. */
static int kspcg_impl(int nx, int ny, int nz, int fdn, double a, double sigma, const double *cv,
                      double *x, const double *b, int max_it, double rtol, int precond,
                      double *trace, int trace_cap, orc_stats *st)
{
    const long n = (long)nx * ny * nz;
    double coeff[32] = {0};
    orc_lap_coefficient(coeff, fdn);
    double diag = a * (3 * coeff[0]);
    if (sigma != 0.0) diag += sigma;
    double *dinv, *F;
    ksp_pc pc;
    if (pc_setup(&pc, precond, nx, ny, nz, fdn, a, sigma, cv, diag, &dinv, &F)) {
        free(dinv); free(F);
        return -2;
    }
    double *R_ = (double *)malloc(sizeof(double) * n), *Z = (double *)malloc(sizeof(double) * n);
    double *P = (double *)malloc(sizeof(double) * n), *W = (double *)malloc(sizeof(double) * n);
    int tlen = 0;
    ksp_conv cv_ = {0.0, 0.0, rtol, 0.0};
    pc_apply(&pc, b, Z, n, 0);
    cv_.bpnorm = orc_norm2(Z, n);

    double t0 = now_seconds();
    OPV(x, R_);
    for (long i = 0; i < n; i++) R_[i] = b[i] + (-1.0) * R_[i];     /* VecAYPX(R, -1, b) */
    pc_apply(&pc, R_, Z, n, 0);
    double dp = orc_norm2(Z, n);
    if (trace && tlen < trace_cap) trace[tlen++] = dp;
    int its = 0;
    int reason = ksp_converged(&cv_, 0, dp);
    if (!reason) {
        double beta = orc_dot(Z, R_, n), betaold = 0.0, dpi = 0.0, dpiold;
        int i = 0;
        do {
            its = i + 1;
            if (beta == 0.0) {
                reason = ORC_KSP_CONVERGED_ATOL;
                break;
            } else if (i > 0 && beta * betaold < 0.0) {
                reason = ORC_KSP_DIVERGED_INDEFINITE_PC;
                break;
            }
            if (!i) {
                memcpy(P, Z, sizeof(double) * n);
            } else {
                const double bb = beta / betaold;
                for (long k = 0; k < n; k++) P[k] = Z[k] + bb * P[k];  /* VecAYPX(P, b, Z) */
            }
            betaold = beta;
            dpiold = dpi;
            OPV(P, W);
            dpi = orc_dot(P, W, n);
            if (dpi == 0.0 || (i > 0 && dpi * dpiold <= 0.0)) {
                reason = ORC_KSP_DIVERGED_INDEFINITE_MAT;
                break;
            }
            const double al = beta / dpi;
            for (long k = 0; k < n; k++) {
                x[k] = x[k] + al * P[k];
                R_[k] = R_[k] + (-al) * W[k];
            }
            pc_apply(&pc, R_, Z, n, 0);
            dp = orc_norm2(Z, n);
            if (trace && tlen < trace_cap) trace[tlen++] = dp;
            reason = ksp_converged(&cv_, i + 1, dp);
            if (reason) break;
            beta = orc_dot(Z, R_, n);
            i++;
        } while (i < max_it);
        if (i >= max_it && !reason) reason = ORC_KSP_DIVERGED_ITS;
    }
    double t1 = now_seconds();
    ksp_finish(st, its, reason, x, b, W, nx, ny, nz, fdn, a, sigma, cv, coeff, tlen, t1 - t0);
    free(dinv); free(F); free(R_); free(Z); free(P); free(W);
    return 0;
}

int orc_kspcg(int nx, int ny, int nz, int fdn, double a, double sigma, const double *c,
              double *x, const double *b, int max_it, double rtol, int precond,
              double *trace, int trace_cap, orc_stats *st)
{
    return kspcg_impl(nx, ny, nz, fdn, a, sigma, c, x, b, max_it, rtol, precond, trace, trace_cap, st);
}

int orc_bicg(int nx, int ny, int nz, int fdn, double a, double sigma, const double *c,
             double *x, const double *b, int max_it, double rtol, int precond,
             double *trace, int trace_cap, orc_stats *st, double *shadow_gap)
{
    return bicg_impl(nx, ny, nz, fdn, a, sigma, c, x, b, max_it, rtol, precond, trace, trace_cap, st, shadow_gap);
}

/* ------------------------------------------------------------------ */
/* ILU(0) of the periodic FD12 star matrix: the preconditioner of the  */
/* PETSc variant's pc = 1 (PCBJACOBI with its default sub-PC, PCILU at */
/* 0 levels, PETSC_REAL/main.c:88-91; on one rank the block is the     */
/* whole matrix, assembled as in PETSC_REAL/system.c:243-273).  For    */
/* this symmetric matrix ICC(0) (PCICC of AAR_Real.c:58-60) defines    */
/* the same preconditioner in exact arithmetic; no pivot is            */
/* nonpositive on these operators, so PETSc's shifts never act.        */
/* Needs every grid side >= 13 (the 37 columns of a row distinct).     */
/*                                                                     */
/* Entry layout per row: slot 0 = 1/u_ii (stored inverted, as PETSc's */
/* factored AIJ keeps its diagonal), slot 1 + 12*axis + k for the      */
/* column at periodic offset d (k = d + 6 for d < 0, d + 5 for d > 0) */
/* along axis 0 = x, 1 = y, 2 = z: l_ij for columns below the row,    */
/* u_ij above.  Elimination (IKJ, PETSc's MatLUFactorNumeric_SeqAIJ    */
/* order): for the row's lower columns k in ascending index, l_ik =    */
/* a_ik * (1/u_kk), then a_ij -= l_ik u_kj for the columns j > k of    */
/* row k that are in row i's pattern.  On the star pattern those j lie */
/* on the axis shared by i and k.  Solves: forward y_i = r_i - sum     */
/* l_ik y_k (ascending k), backward z_i = (y_i - sum u_ij z_j)         */
/* (ascending j) * (1/u_ii).                                           */
/* ------------------------------------------------------------------ */
static inline int ilu_slot(int axis, int d) { return 1 + 12 * axis + (d < 0 ? d + 6 : d + 5); }

/* periodic offset of coordinate cj from c, in [-n/2, n/2] */
static inline int pdiff(int cj, int c, int n)
{
    int d = cj - c;
    if (d > n / 2) d -= n;
    if (d < -n / 2) d += n;
    return d;
}

int orc_ilu0_factor(int nx, int ny, int nz, int fdn, double a, double sigma, const double *cv, double *F)
{
    if (fdn != 6 || nx < 13 || ny < 13 || nz < 13) return -1;
    const long n = (long)nx * ny * nz;
    const int dims[3] = {nx, ny, nz};
    const long strides[3] = {1, nx, (long)nx * ny};
    double coeff[32] = {0};
    orc_lap_coefficient(coeff, fdn);
    double v[ILU_SLOTS];
    for (long i = 0; i < n; i++) {
        const int cr[3] = {(int)(i % nx), (int)((i / nx) % ny), (int)(i / ((long)nx * ny))};
        v[0] = a * (3 * coeff[0]);
        if (sigma != 0.0) v[0] += sigma;
        if (cv) v[0] = v[0] + cv[i];
        for (int ax = 0; ax < 3; ax++)
            for (int d = 1; d <= 6; d++) {
                v[ilu_slot(ax, -d)] = a * coeff[d];
                v[ilu_slot(ax, d)] = a * coeff[d];
            }
        for (int ax = 2; ax >= 0; ax--) {           /* z, y, x: ascending global index */
            const int c = cr[ax], na = dims[ax];
            for (int cp = 0; cp < c; cp++) {
                const int dk = pdiff(cp, c, na);
                if (dk < -6 || dk > 6) continue;
                const long kk = i + (long)(cp - c) * strides[ax];
                const double l = v[ilu_slot(ax, dk)] * F[kk * ILU_SLOTS];
                v[ilu_slot(ax, dk)] = l;
                /* row kk's upper columns on this axis, ascending */
                for (int cj = cp + 1; cj < na; cj++) {
                    const int dkj = pdiff(cj, cp, na);
                    if (dkj < -6 || dkj > 6) continue;
                    const int dij = pdiff(cj, c, na);
                    if (dij < -6 || dij > 6) continue;
                    const int si = dij == 0 ? 0 : ilu_slot(ax, dij);
                    v[si] = v[si] - l * F[kk * ILU_SLOTS + ilu_slot(ax, dkj)];
                }
            }
        }
        if (v[0] == 0.0) return -2;
        v[0] = 1.0 / v[0];
        memcpy(F + i * ILU_SLOTS, v, sizeof(v));
    }
    return 0;
}

/* lower (below = 1) or upper (below = 0) columns of row i along `ax`, in
   ascending coordinate order: returns their count, coordinates in cs[] and
   periodic offsets in ds[] */
static int axis_cols(int c, int na, int below, int *cs, int *ds)
{
    int m = 0;
    const int lo = below ? 0 : c + 1, hi = below ? c : na;
    for (int cj = lo; cj < hi; cj++) {
        const int d = pdiff(cj, c, na);
        if (d < -6 || d > 6) continue;
        cs[m] = cj;
        ds[m] = d;
        m++;
    }
    return m;
}

/* out = M^{-1} in (trans = 0) or M^{-T} in (trans = 1); in-place allowed */
void orc_ilu0_apply(int nx, int ny, int nz, const double *F, const double *in, double *out, int trans)
{
    const long n = (long)nx * ny * nz;
    const int dims[3] = {nx, ny, nz};
    const long strides[3] = {1, nx, (long)nx * ny};
    int cs[12], ds[12];
    if (out != in) memcpy(out, in, sizeof(double) * n);
    if (!trans) {
        /* L y = r: lower columns ascending (z, y, x groups) */
        for (long i = 0; i < n; i++) {
            const int cr[3] = {(int)(i % nx), (int)((i / nx) % ny), (int)(i / ((long)nx * ny))};
            double s = out[i];
            for (int ax = 2; ax >= 0; ax--) {
                const int m = axis_cols(cr[ax], dims[ax], 1, cs, ds);
                for (int q = 0; q < m; q++)
                    s = s - F[i * ILU_SLOTS + ilu_slot(ax, ds[q])] * out[i + (long)(cs[q] - cr[ax]) * strides[ax]];
            }
            out[i] = s;
        }
        /* U z = y: upper columns ascending (x, y, z groups), times 1/u_ii */
        for (long i = n - 1; i >= 0; i--) {
            const int cr[3] = {(int)(i % nx), (int)((i / nx) % ny), (int)(i / ((long)nx * ny))};
            double s = out[i];
            for (int ax = 0; ax < 3; ax++) {
                const int m = axis_cols(cr[ax], dims[ax], 0, cs, ds);
                for (int q = 0; q < m; q++)
                    s = s - F[i * ILU_SLOTS + ilu_slot(ax, ds[q])] * out[i + (long)(cs[q] - cr[ax]) * strides[ax]];
            }
            out[i] = s * F[i * ILU_SLOTS];
        }
    } else {
        /* U^T w = r: w_i = (r_i - sum_{j < i} u_ji w_j) / u_ii, j ascending */
        for (long i = 0; i < n; i++) {
            const int cr[3] = {(int)(i % nx), (int)((i / nx) % ny), (int)(i / ((long)nx * ny))};
            double s = out[i];
            for (int ax = 2; ax >= 0; ax--) {
                const int m = axis_cols(cr[ax], dims[ax], 1, cs, ds);
                for (int q = 0; q < m; q++) {
                    const long j = i + (long)(cs[q] - cr[ax]) * strides[ax];
                    s = s - F[j * ILU_SLOTS + ilu_slot(ax, -ds[q])] * out[j];
                }
            }
            out[i] = s * F[i * ILU_SLOTS];
        }
        /* L^T z = w: z_i = w_i - sum_{j > i} l_ji z_j, j ascending */
        for (long i = n - 1; i >= 0; i--) {
            const int cr[3] = {(int)(i % nx), (int)((i / nx) % ny), (int)(i / ((long)nx * ny))};
            double s = out[i];
            for (int ax = 0; ax < 3; ax++) {
                const int m = axis_cols(cr[ax], dims[ax], 0, cs, ds);
                for (int q = 0; q < m; q++) {
                    const long j = i + (long)(cs[q] - cr[ax]) * strides[ax];
                    s = s - F[j * ILU_SLOTS + ilu_slot(ax, -ds[q])] * out[j];
                }
            }
            out[i] = s;
        }
    }
}

/* ------------------------------------------------------------------ */
/* AAR of the PETSc variant, PETSC_REAL/AAR_Real.c:32-199, on one rank */
/* with its preconditioner choice (AAR_Real.c:57-66): precond 1 =      */
/* PCJACOBI (multiply by 1/diag), 2 = ILU(0) in place of PCICC (the    */
/* same preconditioner in exact arithmetic on this symmetric matrix;   */
/* no pivot shift acts), 0 = none.  Differences from MPI_REAL/AAR.c:   */
/* the Anderson step runs whenever iter % p == 0, iter = 1 included    */
/* (AAR_Real.c:116); the update is x += sum_i (-s_i)(beta DF_i + DX_i) */
/* one column at a time, then x += beta f (AAR_Real.c:132-137); the    */
/* least-squares solve is MKL's dgelsd (AAR_Real.c:223), absent here:  */
/* restated as orc_pinv_times_vec_gelsd, the minimum-norm solution     */
/* over singular values above eps * s_max (rcond = -1); parity with    */
/* dgelsd's own arithmetic is unpinned.                                */
/* ------------------------------------------------------------------ */
int orc_aar_petsc(int nx, int ny, int nz, int fdn, double a, double sigma, const double *cv,
                  double *x, const double *b, double omega, double beta, int m, int p, int max_iter,
                  double tol, int precond, double *trace, int trace_cap, orc_stats *st)
{
    if (m < 1 || m > 32 || p < 1) return -1;
    const long n = (long)nx * ny * nz;
    double coeff[32] = {0};
    orc_lap_coefficient(coeff, fdn);
    double diag = a * (3 * coeff[0]);
    if (sigma != 0.0) diag += sigma;
    double *dinv, *F;
    ksp_pc pc;
    if (pc_setup(&pc, precond, nx, ny, nz, fdn, a, sigma, cv, diag, &dinv, &F)) {
        free(dinv); free(F);
        return -2;
    }
    double *res = (double *)malloc(sizeof(double) * n), *f = (double *)malloc(sizeof(double) * n);
    double *x_old = (double *)malloc(sizeof(double) * n), *f_old = (double *)malloc(sizeof(double) * n);
    double *DX = (double *)calloc((size_t)m * n, sizeof(double)), *DF = (double *)calloc((size_t)m * n, sizeof(double));
    double *G = (double *)malloc(sizeof(double) * m * m), *g = (double *)malloc(sizeof(double) * m);
    double *s = (double *)malloc(sizeof(double) * m);
    int tlen = 0, status = 0;

    const double b_2norm = orc_norm2(b, n);
    tol *= b_2norm;
    OPV(x, res);
    for (long i = 0; i < n; i++) res[i] = b[i] + (-1.0) * res[i];      /* VecAYPX(res, -1, b) */
    double r_2norm = tol + 1;
    int iter = 1;
    double t0 = now_seconds();
    while (r_2norm > tol && iter <= max_iter) {
        pc_apply(&pc, res, f, n, 0);
        if (iter > 1) {
            const int k = (iter - 2) % m;
            for (long i = 0; i < n; i++) {
                DX[(long)k * n + i] = (-1.0) * x_old[i] + x[i];
                DF[(long)k * n + i] = (-1.0) * f_old[i] + f[i];
            }
        }
        memcpy(x_old, x, sizeof(double) * n);
        memcpy(f_old, f, sizeof(double) * n);
        if (iter % p) {
            for (long i = 0; i < n; i++) x[i] = x[i] + omega * f[i];
        } else {
            for (int i = 0; i < m; i++) {
                for (int j = 0; j <= i; j++) {
                    const double v = orc_dot(DF + (long)i * n, DF + (long)j * n, n);
                    G[m * i + j] = v;
                    G[i + m * j] = v;
                }
                g[i] = orc_dot(DF + (long)i * n, f, n);
            }
            if (orc_pinv_times_vec_gelsd(G, g, s, m, NULL, NULL)) status = 1;
            for (int q = 0; q < m; q++)
                for (long i = 0; i < n; i++)
                    x[i] = x[i] + (-s[q]) * (beta * DF[(long)q * n + i] + DX[(long)q * n + i]);
            for (long i = 0; i < n; i++) x[i] = x[i] + beta * f[i];
        }
        OPV(x, res);
        for (long i = 0; i < n; i++) res[i] = b[i] + (-1.0) * res[i];
        if (iter % p == 0) {
            r_2norm = orc_norm2(res, n);
            if (trace && tlen < trace_cap) trace[tlen++] = r_2norm;
        }
        iter++;
    }
    double t1 = now_seconds();
    finish_stats(st, iter - 1, r_2norm, b_2norm, tol, iter, max_iter, tlen, t1 - t0, status);
    free(dinv); free(F); free(res); free(f); free(x_old); free(f_old); free(DX); free(DF); free(G); free(g); free(s);
    return 0;
}

/* ------------------------------------------------------------------ */
/* PGR / PL2R of the PETSc variant, PETSC_REAL/PGR_Real.c:32-195 and   */
/* PL2R_Real.c:32-197 (l2 = 1), with the preconditioner choice of      */
/* PGR_Real.c:55-62 (precond as orc_aar_petsc).  The loop of           */
/* PGR_Real.c:98-161: f = B res; x_old = x, Ax_prev = Ax; k = (iter-2  */
/* + m) % m; x += omega f; DX_k = x - x_old; Ax = A x; res = b - Ax;    */
/* DF_k = Ax - Ax_prev; every p: s from the m x m system, x += sum     */
/* s_i DX_i, DX_k = x - x_old, Ax += sum s_i DF_i, DF_k = Ax - Ax_prev;*/
/* res = b - Ax, its norm checked every p.  The m x m system is        */
/* A[i][j] = DX_i . DF_j (PGR; PL2R: DF_i . DF_j), rhs DX_i . res (DF_i */
/* . res), solved with dgelsd's rcond = -1 cutoff (eps * s_max,       */
/* orc_pinv_times_vec_gelsd).                                          */
/* The variant's column-major dgelsd call receives the transpose of    */
/* PGR's matrix; for these symmetric operators DF = A DX makes the two */
/* equal in exact arithmetic.                                          */
/* ------------------------------------------------------------------ */
int orc_pgr_petsc(int nx, int ny, int nz, int fdn, double a, double sigma, const double *cv,
                  double *x, const double *b, double omega, int m, int p, int max_iter, double tol,
                  int precond, int l2, double *trace, int trace_cap, orc_stats *st)
{
    if (m < 1 || m > 32 || p < 1) return -1;
    const long n = (long)nx * ny * nz;
    double coeff[32] = {0};
    orc_lap_coefficient(coeff, fdn);
    double diag = a * (3 * coeff[0]);
    if (sigma != 0.0) diag += sigma;
    double *dinv, *F;
    ksp_pc pc;
    if (pc_setup(&pc, precond, nx, ny, nz, fdn, a, sigma, cv, diag, &dinv, &F)) {
        free(dinv); free(F);
        return -2;
    }
    double *res = (double *)malloc(sizeof(double) * n), *f = (double *)malloc(sizeof(double) * n);
    double *Ax = (double *)malloc(sizeof(double) * n), *Ax_prev = (double *)malloc(sizeof(double) * n);
    double *x_old = (double *)malloc(sizeof(double) * n);
    double *DX = (double *)calloc((size_t)m * n, sizeof(double)), *DF = (double *)calloc((size_t)m * n, sizeof(double));
    double *G = (double *)malloc(sizeof(double) * m * m), *g = (double *)malloc(sizeof(double) * m);
    double *s = (double *)malloc(sizeof(double) * m);
    int tlen = 0, status = 0;
    const double *L = l2 ? DF : DX;           /* the left factor of the normal equations */

    const double b_2norm = orc_norm2(b, n);
    tol *= b_2norm;
    OPV(x, Ax);
    for (long i = 0; i < n; i++) res[i] = (-1.0) * Ax[i] + b[i];        /* VecWAXPY(res, -1, Ax, b) */
    double r_2norm = tol + 1;
    int iter = 1;
    double t0 = now_seconds();
    while (r_2norm > tol && iter <= max_iter) {
        pc_apply(&pc, res, f, n, 0);
        memcpy(x_old, x, sizeof(double) * n);
        memcpy(Ax_prev, Ax, sizeof(double) * n);
        const int k = (iter - 2 + m) % m;
        for (long i = 0; i < n; i++) {
            x[i] = x[i] + omega * f[i];
            DX[(long)k * n + i] = (-1.0) * x_old[i] + x[i];
        }
        OPV(x, Ax);
        for (long i = 0; i < n; i++) {
            res[i] = (-1.0) * Ax[i] + b[i];
            DF[(long)k * n + i] = (-1.0) * Ax_prev[i] + Ax[i];
        }
        if (iter % p == 0) {
            for (int i = 0; i < m; i++) {
                for (int j = 0; j < m; j++) G[i * m + j] = orc_dot(L + (long)i * n, DF + (long)j * n, n);
                g[i] = orc_dot(L + (long)i * n, res, n);
            }
            if (orc_pinv_times_vec_gelsd(G, g, s, m, NULL, NULL)) status = 1;
            for (int q = 0; q < m; q++)
                for (long i = 0; i < n; i++) x[i] = x[i] + s[q] * DX[(long)q * n + i];
            for (long i = 0; i < n; i++) DX[(long)k * n + i] = (-1.0) * x_old[i] + x[i];
            for (int q = 0; q < m; q++)
                for (long i = 0; i < n; i++) Ax[i] = Ax[i] + s[q] * DF[(long)q * n + i];
            for (long i = 0; i < n; i++) DF[(long)k * n + i] = (-1.0) * Ax_prev[i] + Ax[i];
        }
        for (long i = 0; i < n; i++) res[i] = (-1.0) * Ax[i] + b[i];
        if (iter % p == 0) {
            r_2norm = orc_norm2(res, n);
            if (trace && tlen < trace_cap) trace[tlen++] = r_2norm;
        }
        iter++;
    }
    double t1 = now_seconds();
    finish_stats(st, iter - 1, r_2norm, b_2norm, tol, iter, max_iter, tlen, t1 - t0, status);
    free(dinv); free(F); free(res); free(f); free(Ax); free(Ax_prev); free(x_old); free(DX); free(DF);
    free(G); free(g); free(s);
    return 0;
}

/* ------------------------------------------------------------------ */
/* GMRES of the complex PETSc variant (PETSC_COMPLEX/main.c:80-118:    */
/* KSPGMRES on the Helmholtz operator, PCJACOBI, rtol 1e-6, max_it     */
/* 2000, nonzero guess), the algorithm of gmres_impl in complex        */
/* arithmetic as PETSc's complex build runs it: <x, y> = sum x conj(y) */
/* (VecMDot), Givens rotations with complex c and s,                   */
/* h <- conj(c) h + s h', h' <- c h' - s h.  Complex operations use    */
/* the explicit formulas below (no C99 complex type), so pointwise     */
/* results are reproducible elsewhere bit for bit.  Jacobi: 1/D with  */
/* D = (a*3*coeff[0] + c) + i eta, as (dr - i eta) / (dr^2 + eta^2).  */
/* Vectors are complex128 interleaved (re, im).  Parity with PETSc is */
/* unpinned (PETSc absent); pinned by the Krylov minimum and the FFT  */
/* solution (tests).                                                   */
/* ------------------------------------------------------------------ */
typedef struct { double re, im; } zc;
int orc_ilu0_factor_complex(int nx, int ny, int nz, int fdn, double a, const double *c, double c0, double eta,
                            double *Fd);
void orc_ilu0_apply_complex(int nx, int ny, int nz, const double *Fd, const double *ind, double *outd, int trans);
static inline zc zmul(zc a, zc b) { zc r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static inline zc zconj(zc a) { zc r = {a.re, -a.im}; return r; }
static inline zc zadd(zc a, zc b) { zc r = {a.re + b.re, a.im + b.im}; return r; }
static inline zc zsub(zc a, zc b) { zc r = {a.re - b.re, a.im - b.im}; return r; }
static inline zc zdiv(zc a, zc b)
{
    const double den = b.re * b.re + b.im * b.im;
    zc r = {(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
    return r;
}
static inline double zabs(zc a) { return sqrt(a.re * a.re + a.im * a.im); }

/* LGMRES(m, k) (PETSC_COMPLEX/main.c:120-147: KSPLGMRES, the same       */
/* tolerances and preconditioners) in the same loop, aug = k > 0: the     */
/* published algorithm of Baker, Jessup and Manteuffel (SIAM J. Matrix    */
/* Anal. Appl. 26, 2005) with PETSc's defaults, restart = m + k = 30 and  */
/* k = 2 error approximations, the space not held constant:               */
/*  * every cycle takes restart - aug Arnoldi steps, then one step per    */
/*    stored approximation z, newest first, with w = B A z (kept from    */
/*    when z was made) orthogonalised like an Arnoldi vector; every step  */
/*    counts as an iteration and is tested;                               */
/*  * the cycle's correction is sum_j y_j W_j (W_j the Arnoldi vector or  */
/*    the z of step j); it becomes the new z, and B A z = V H y by the     */
/*    unrotated Hessenberg H, with no operator application; both scaled  */
/*    by 1/||z||; the oldest z is replaced once k are stored.             */
/* aug = 0 is GMRES(restart), operation for operation.                   */
static int gmres_complex_impl(int nx, int ny, int nz, int fdn, double a, const double *c, double c0, double eta,
                              double *u, const double *rhs, int restart, int aug, int max_it, double rtol,
                              int precond, double *trace, int trace_cap, orc_stats *st)
{
    if (restart < 1 || restart > ORC_GMRES_MAXK) return -1;
    if (aug < 0 || aug > ORC_LGMRES_MAXAUG || aug >= restart) return -1;
    const long n = (long)nx * ny * nz;
    const int K = restart;
    double coeff[32] = {0};
    orc_lap_coefficient(coeff, fdn);
    const double dlap = a * (3 * coeff[0]);
    zc *x = (zc *)u;
    const zc *b = (const zc *)rhs;
    zc *dinv = (zc *)malloc(sizeof(zc) * n);
    zc *V = (zc *)malloc(sizeof(zc) * n * (K + 1));
    zc *t = (zc *)malloc(sizeof(zc) * n);
    zc *Fz = NULL;
    const int ilu = precond == 2;
    if (ilu) {
        Fz = (zc *)malloc(sizeof(zc) * n * ILU_SLOTS);
        if (orc_ilu0_factor_complex(nx, ny, nz, fdn, a, c, c0, eta, (double *)Fz)) {
            free(dinv); free(V); free(t); free(Fz);
            return -2;
        }
    }
    zc HH[(ORC_GMRES_MAXK + 2) * (ORC_GMRES_MAXK + 1)], HES[(ORC_GMRES_MAXK + 2) * (ORC_GMRES_MAXK + 1)];
    zc cc[ORC_GMRES_MAXK + 1], ss[ORC_GMRES_MAXK + 1], grs[ORC_GMRES_MAXK + 2], nrs[ORC_GMRES_MAXK + 1];
    zc h[ORC_GMRES_MAXK + 1], hw[ORC_GMRES_MAXK + 2];
    /* the error approximations z and B A z, slot order[s] (1 = newest) */
    zc *Z = aug ? (zc *)malloc(sizeof(zc) * n * aug) : NULL, *AZ = aug ? (zc *)malloc(sizeof(zc) * n * aug) : NULL;
    zc *corr = aug ? (zc *)malloc(sizeof(zc) * n) : NULL;
    int order[ORC_LGMRES_MAXAUG] = {0}, aug_ct = 0;
    const int ita = K - aug;       /* Arnoldi steps per cycle */
#define ZH(r, col) HH[(col) * (K + 2) + (r)]
#define ZS(r, col) HES[(col) * (K + 2) + (r)]
    for (long i = 0; i < n; i++) {
        if (precond) {
            const double dr = dlap + (c ? c[i] : c0);
            const double den = dr * dr + eta * eta;
            dinv[i].re = dr / den;
            dinv[i].im = -eta / den;
        } else {
            dinv[i].re = 1.0;
            dinv[i].im = 0.0;
        }
    }
    int tlen = 0;
    ksp_conv cv_ = {0.0, 0.0, rtol, 0.0};
    double s2 = 0.0;
    if (ilu) orc_ilu0_apply_complex(nx, ny, nz, (const double *)Fz, (const double *)b, (double *)t, 0);
    for (long i = 0; i < n; i++) {
        const zc z = ilu ? t[i] : (precond ? zmul(dinv[i], b[i]) : b[i]);
        s2 = s2 + (z.re * z.re + z.im * z.im);
    }
    cv_.bpnorm = sqrt(s2);
    int its = 0, reason = 0;
    double t0 = now_seconds();
    while (!reason) {
        zc *v0 = V;
        orc_matvec_complex(nx, ny, nz, fdn, coeff, a, c, c0, eta, (const double *)x, (double *)t);
        s2 = 0.0;
        if (ilu) {
            for (long i = 0; i < n; i++) v0[i] = zsub(b[i], t[i]);
            orc_ilu0_apply_complex(nx, ny, nz, (const double *)Fz, (const double *)v0, (double *)v0, 0);
            for (long i = 0; i < n; i++) s2 = s2 + (v0[i].re * v0[i].re + v0[i].im * v0[i].im);
        } else {
            for (long i = 0; i < n; i++) {
                const zc r = zsub(b[i], t[i]);
                v0[i] = precond ? zmul(dinv[i], r) : r;
                s2 = s2 + (v0[i].re * v0[i].re + v0[i].im * v0[i].im);
            }
        }
        double res = sqrt(s2);
        if (res != 0.0 && res != 1.0) {
            const double sc = 1.0 / res;
            for (long i = 0; i < n; i++) { v0[i].re = sc * v0[i].re; v0[i].im = sc * v0[i].im; }
        }
        grs[0].re = res;
        grs[0].im = 0.0;
        int it = 0;
        if (trace && tlen < trace_cap) trace[tlen++] = res;
        if (res == 0.0) { reason = ORC_KSP_CONVERGED_ATOL; break; }
        reason = ksp_converged(&cv_, its, res);
        const int itot = ita + aug_ct;
        while (!reason && it < itot && its < max_it) {
            zc *vn = V + (long)(it + 1) * n;
            if (it < ita) {
                orc_matvec_complex(nx, ny, nz, fdn, coeff, a, c, c0, eta, (const double *)(V + (long)it * n),
                                   (double *)t);
                if (ilu) orc_ilu0_apply_complex(nx, ny, nz, (const double *)Fz, (const double *)t, (double *)vn, 0);
                else
                    for (long i = 0; i < n; i++) vn[i] = precond ? zmul(dinv[i], t[i]) : t[i];
            } else {
                /* augmentation step: the approximation of order it - ita + 1 */
                int spot = 0;
                for (int q = 0; q < aug_ct; q++)
                    if (order[q] == it - ita + 1) spot = q;
                memcpy(vn, AZ + (long)spot * n, sizeof(zc) * n);
            }
            for (int j = 0; j <= it; j++) {
                zc sj = {0.0, 0.0};
                const zc *vj = V + (long)j * n;
                for (long i = 0; i < n; i++) sj = zadd(sj, zmul(vn[i], zconj(vj[i])));
                h[j] = sj;
            }
            s2 = 0.0;
            for (long i = 0; i < n; i++) {
                zc w = vn[i];
                for (int j = 0; j <= it; j++) {
                    const zc nh = {-h[j].re, -h[j].im};
                    w = zadd(w, zmul(nh, V[(long)j * n + i]));
                }
                vn[i] = w;
                s2 = s2 + (w.re * w.re + w.im * w.im);
            }
            for (int j = 0; j <= it; j++) ZH(j, it) = ZS(j, it) = h[j];
            const double tt = sqrt(s2);
            if (tt != 0.0 && tt != 1.0) {
                const double sc = 1.0 / tt;
                for (long i = 0; i < n; i++) { vn[i].re = sc * vn[i].re; vn[i].im = sc * vn[i].im; }
            }
            ZH(it + 1, it).re = tt;
            ZH(it + 1, it).im = 0.0;
            ZS(it + 1, it) = ZH(it + 1, it);
            double hapbnd = tt / zabs(grs[it]);
            if (hapbnd > ORC_GMRES_HAPTOL) hapbnd = ORC_GMRES_HAPTOL;
            const int hapend = tt < hapbnd;
            zc *hh = &ZH(0, it);
            for (int j = 1; j <= it; j++) {
                const zc t1 = hh[j - 1];
                hh[j - 1] = zadd(zmul(zconj(cc[j - 1]), t1), zmul(ss[j - 1], hh[j]));
                hh[j] = zsub(zmul(cc[j - 1], hh[j]), zmul(ss[j - 1], t1));
            }
            if (!hapend) {
                const zc q = zadd(zmul(zconj(hh[it]), hh[it]), zmul(zconj(hh[it + 1]), hh[it + 1]));
                const double r = sqrt(q.re);
                if (r == 0.0) {
                    reason = ORC_KSP_DIVERGED_NULL;
                } else {
                    const zc rr = {r, 0.0};
                    cc[it] = zdiv(hh[it], rr);
                    ss[it] = zdiv(hh[it + 1], rr);
                    const zc g = grs[it];
                    const zc ng = zmul(ss[it], g);
                    grs[it + 1].re = -ng.re;
                    grs[it + 1].im = -ng.im;
                    grs[it] = zmul(zconj(cc[it]), g);
                    hh[it] = zadd(zmul(zconj(cc[it]), hh[it]), zmul(ss[it], hh[it + 1]));
                    res = zabs(grs[it + 1]);
                }
            } else {
                res = 0.0;
            }
            it++;
            its++;
            if (trace && tlen < trace_cap) trace[tlen++] = res;
            if (!reason) reason = ksp_converged(&cv_, its, res);
            if (hapend) {
                if (!reason) reason = ORC_KSP_DIVERGED_BREAKDOWN;
                break;
            }
        }
        if (it > 0) {
            const int k = it - 1;
            nrs[k] = zdiv(grs[k], ZH(k, k));
            for (int ii = 1; ii <= k; ii++) {
                const int kk = k - ii;
                zc tt = grs[kk];
                for (int j = kk + 1; j <= k; j++) tt = zsub(tt, zmul(ZH(kk, j), nrs[j]));
                nrs[kk] = zdiv(tt, ZH(kk, kk));
            }
            if (!aug) {
                for (long i = 0; i < n; i++) {
                    zc sacc = {0.0, 0.0};
                    for (int j = 0; j <= k; j++) sacc = zadd(sacc, zmul(nrs[j], V[(long)j * n + i]));
                    x[i] = zadd(x[i], sacc);
                }
            } else {
                const zc *W[ORC_GMRES_MAXK];
                for (int j = 0; j <= k; j++) {
                    W[j] = V + (long)j * n;
                    if (j >= ita)
                        for (int q = 0; q < aug_ct; q++)
                            if (order[q] == j - ita + 1) W[j] = Z + (long)q * n;
                }
                double zn2 = 0.0;
                for (long i = 0; i < n; i++) {
                    zc sacc = {0.0, 0.0};
                    for (int j = 0; j <= k; j++) sacc = zadd(sacc, zmul(nrs[j], W[j][i]));
                    corr[i] = sacc;
                    x[i] = zadd(x[i], sacc);
                    zn2 = zn2 + (sacc.re * sacc.re + sacc.im * sacc.im);
                }
                /* B A z = V_{it+1} (H y), H unrotated */
                for (int r = 0; r <= it; r++) {
                    zc acc = {0.0, 0.0};
                    for (int j = 0; j <= k; j++)
                        if (r <= j + 1) acc = zadd(acc, zmul(ZS(r, j), nrs[j]));
                    hw[r] = acc;
                }
                int spot;
                if (aug_ct < aug) {
                    spot = aug_ct++;
                } else {
                    spot = 0;
                    for (int q = 0; q < aug; q++)
                        if (order[q] == aug) spot = q;
                }
                for (int q = 0; q < aug_ct; q++)
                    if (q != spot) order[q]++;
                order[spot] = 1;
                const double sc = 1.0 / sqrt(zn2);
                zc *zs = Z + (long)spot * n, *azs = AZ + (long)spot * n;
                for (long i = 0; i < n; i++) {
                    zc acc = {0.0, 0.0};
                    for (int r = 0; r <= it; r++) acc = zadd(acc, zmul(hw[r], V[(long)r * n + i]));
                    azs[i].re = sc * acc.re;
                    azs[i].im = sc * acc.im;
                    zs[i].re = sc * corr[i].re;
                    zs[i].im = sc * corr[i].im;
                }
            }
        }
        if (its >= max_it) {
            if (!reason) reason = ORC_KSP_DIVERGED_ITS;
            break;
        }
    }
#undef ZH
#undef ZS
    double t1 = now_seconds();
    if (st) {
        orc_matvec_complex(nx, ny, nz, fdn, coeff, a, c, c0, eta, (const double *)x, (double *)t);
        double rn = 0.0, bn = 0.0;
        for (long i = 0; i < n; i++) {
            const zc r = zsub(b[i], t[i]);
            rn = rn + (r.re * r.re + r.im * r.im);
            bn = bn + (b[i].re * b[i].re + b[i].im * b[i].im);
        }
        st->iterations = its;
        st->relres = sqrt(rn) / sqrt(bn);
        st->rhs_norm = sqrt(bn);
        st->converged = reason > 0;
        st->printed_converged = reason > 0;
        st->printed_warning = reason == ORC_KSP_DIVERGED_ITS;
        st->trace_len = tlen;
        st->seconds = t1 - t0;
        st->status = reason;
    }
    free(dinv); free(V); free(t); free(Fz); free(Z); free(AZ); free(corr);
    return 0;
}

int orc_gmres_complex(int nx, int ny, int nz, int fdn, double a, const double *c, double c0, double eta,
                      double *u, const double *rhs, int restart, int max_it, double rtol, int precond,
                      double *trace, int trace_cap, orc_stats *st)
{
    return gmres_complex_impl(nx, ny, nz, fdn, a, c, c0, eta, u, rhs, restart, 0, max_it, rtol, precond, trace,
                              trace_cap, st);
}

int orc_lgmres_complex(int nx, int ny, int nz, int fdn, double a, const double *c, double c0, double eta,
                       double *u, const double *rhs, int restart, int aug, int max_it, double rtol, int precond,
                       double *trace, int trace_cap, orc_stats *st)
{
    return gmres_complex_impl(nx, ny, nz, fdn, a, c, c0, eta, u, rhs, restart, aug, max_it, rtol, precond, trace,
                              trace_cap, st);
}

/* BiCG of the complex PETSc variant (PETSC_COMPLEX/main.c:150-175,
 * KSPBICG with PCJACOBI): the shadow sequence of KSPBICG's complex build,
 * z~ = B^H r~ and A^H p~ (VecConjugate around the transposed products);
 * the Helmholtz operator is complex symmetric, so A^H is the operator
 * with eta -> -eta and B^H = conj(1/D).  <x, y> = sum x conj(y). */
int orc_bicg_complex(int nx, int ny, int nz, int fdn, double a, const double *c, double c0, double eta,
                     double *u, const double *rhs, int max_it, double rtol, int precond,
                     double *trace, int trace_cap, orc_stats *st)
{
    const long n = (long)nx * ny * nz;
    double coeff[32] = {0};
    orc_lap_coefficient(coeff, fdn);
    const double dlap = a * (3 * coeff[0]);
    zc *x = (zc *)u;
    const zc *b = (const zc *)rhs;
    zc *dinv = (zc *)malloc(sizeof(zc) * n);
    zc *Rr = (zc *)malloc(sizeof(zc) * n), *Rl = (zc *)malloc(sizeof(zc) * n);
    zc *Zr = (zc *)malloc(sizeof(zc) * n), *Zl = (zc *)malloc(sizeof(zc) * n);
    zc *Pr = (zc *)malloc(sizeof(zc) * n), *Pl = (zc *)malloc(sizeof(zc) * n);
    zc *Fz = NULL;
    const int ilu = precond == 2;
    if (ilu) {
        Fz = (zc *)malloc(sizeof(zc) * n * ILU_SLOTS);
        if (orc_ilu0_factor_complex(nx, ny, nz, fdn, a, c, c0, eta, (double *)Fz)) {
            free(dinv); free(Rr); free(Rl); free(Zr); free(Zl); free(Pr); free(Pl); free(Fz);
            return -2;
        }
    }
    for (long i = 0; i < n; i++) {
        if (precond) {
            const double dr = dlap + (c ? c[i] : c0);
            const double den = dr * dr + eta * eta;
            dinv[i].re = dr / den;
            dinv[i].im = -eta / den;
        } else {
            dinv[i].re = 1.0;
            dinv[i].im = 0.0;
        }
    }
#define ZPC(i, v) (precond ? zmul(dinv[i], (v)) : (v))
#define ZPCH(i, v) (precond ? zmul(zconj(dinv[i]), (v)) : (v))
    int tlen = 0;
    ksp_conv cv_ = {0.0, 0.0, rtol, 0.0};
    double s2 = 0.0;
    if (ilu) orc_ilu0_apply_complex(nx, ny, nz, (const double *)Fz, (const double *)b, (double *)Zr, 0);
    for (long i = 0; i < n; i++) {
        const zc z = ilu ? Zr[i] : ZPC(i, b[i]);
        s2 = s2 + (z.re * z.re + z.im * z.im);
    }
    cv_.bpnorm = sqrt(s2);
    double t0 = now_seconds();
    orc_matvec_complex(nx, ny, nz, fdn, coeff, a, c, c0, eta, (const double *)x, (double *)Rr);
    s2 = 0.0;
    for (long i = 0; i < n; i++) {
        Rr[i].re = b[i].re + (-1.0) * Rr[i].re;
        Rr[i].im = b[i].im + (-1.0) * Rr[i].im;
        Rl[i] = Rr[i];
        if (!ilu) {
            Zr[i] = ZPC(i, Rr[i]);
            Zl[i] = ZPCH(i, Rl[i]);
        }
    }
    if (ilu) {
        orc_ilu0_apply_complex(nx, ny, nz, (const double *)Fz, (const double *)Rr, (double *)Zr, 0);
        orc_ilu0_apply_complex(nx, ny, nz, (const double *)Fz, (const double *)Rl, (double *)Zl, 2);
    }
    for (long i = 0; i < n; i++) s2 = s2 + (Zr[i].re * Zr[i].re + Zr[i].im * Zr[i].im);
    double dp = sqrt(s2);
    if (trace && tlen < trace_cap) trace[tlen++] = dp;
    int its = 0;
    int reason = ksp_converged(&cv_, 0, dp);
    zc beta, betaold = {0.0, 0.0};
    int i = 0;
    if (!reason) {
        do {
            zc sb = {0.0, 0.0};
            for (long k = 0; k < n; k++) sb = zadd(sb, zmul(Zr[k], zconj(Rl[k])));
            beta = sb;
            if (!i) {
                if (beta.re == 0.0 && beta.im == 0.0) {
                    reason = ORC_KSP_DIVERGED_BREAKDOWN_BICG;
                    break;
                }
                memcpy(Pr, Zr, sizeof(zc) * n);
                memcpy(Pl, Zl, sizeof(zc) * n);
            } else {
                const zc bb = zdiv(beta, betaold), bbc = zconj(bb);
                for (long k = 0; k < n; k++) {
                    Pr[k] = zadd(Zr[k], zmul(bb, Pr[k]));
                    Pl[k] = zadd(Zl[k], zmul(bbc, Pl[k]));
                }
            }
            betaold = beta;
            orc_matvec_complex(nx, ny, nz, fdn, coeff, a, c, c0, eta, (const double *)Pr, (double *)Zr);
            orc_matvec_complex(nx, ny, nz, fdn, coeff, a, c, c0, -eta, (const double *)Pl, (double *)Zl);
            zc dpi = {0.0, 0.0};
            for (long k = 0; k < n; k++) dpi = zadd(dpi, zmul(Zr[k], zconj(Pl[k])));
            const zc al = zdiv(beta, dpi), ma = {-al.re, -al.im}, mac = zconj(ma);
            s2 = 0.0;
            for (long k = 0; k < n; k++) {
                x[k] = zadd(x[k], zmul(al, Pr[k]));
                Rr[k] = zadd(Rr[k], zmul(ma, Zr[k]));
                Rl[k] = zadd(Rl[k], zmul(mac, Zl[k]));
                if (!ilu) {
                    Zr[k] = ZPC(k, Rr[k]);
                    Zl[k] = ZPCH(k, Rl[k]);
                }
            }
            if (ilu) {
                orc_ilu0_apply_complex(nx, ny, nz, (const double *)Fz, (const double *)Rr, (double *)Zr, 0);
                orc_ilu0_apply_complex(nx, ny, nz, (const double *)Fz, (const double *)Rl, (double *)Zl, 2);
            }
            for (long k = 0; k < n; k++) s2 = s2 + (Zr[k].re * Zr[k].re + Zr[k].im * Zr[k].im);
            dp = sqrt(s2);
            its = i + 1;
            if (trace && tlen < trace_cap) trace[tlen++] = dp;
            reason = ksp_converged(&cv_, i + 1, dp);
            if (reason) break;
            i++;
        } while (i < max_it);
        if (i >= max_it && !reason) reason = ORC_KSP_DIVERGED_ITS;
    }
#undef ZPC
#undef ZPCH
    double t1 = now_seconds();
    if (st) {
        orc_matvec_complex(nx, ny, nz, fdn, coeff, a, c, c0, eta, (const double *)x, (double *)Zr);
        double rn = 0.0, bn = 0.0;
        for (long k = 0; k < n; k++) {
            const zc r = zsub(b[k], Zr[k]);
            rn = rn + (r.re * r.re + r.im * r.im);
            bn = bn + (b[k].re * b[k].re + b[k].im * b[k].im);
        }
        st->iterations = its;
        st->relres = sqrt(rn) / sqrt(bn);
        st->rhs_norm = sqrt(bn);
        st->converged = reason > 0;
        st->printed_converged = reason > 0;
        st->printed_warning = reason == ORC_KSP_DIVERGED_ITS;
        st->trace_len = tlen;
        st->seconds = t1 - t0;
        st->status = reason;
    }
    free(dinv); free(Rr); free(Rl); free(Zr); free(Zl); free(Pr); free(Pl); free(Fz);
    return 0;
}

/* ILU(0) of the complex Helmholtz matrix (PETSC_COMPLEX pc = 1:
 * PCBJACOBI -> PCILU on one block): orc_ilu0_factor's elimination in
 * complex arithmetic (explicit formulas), diagonal (a*3*coeff[0] + c) +
 * i eta, off-diagonals a*coeff[d].  F: ORC_ILU_SLOTS complex entries per
 * row (interleaved), slot 0 = 1/u_ii.  apply: z = M^{-1} r (trans = 0),
 * M^{-T} r (trans = 1) or M^{-H} r = conj(M^{-T} conj(r)) (trans = 2). */
int orc_ilu0_factor_complex(int nx, int ny, int nz, int fdn, double a, const double *c, double c0, double eta,
                            double *Fd)
{
    if (fdn != 6 || nx < 13 || ny < 13 || nz < 13) return -1;
    const long n = (long)nx * ny * nz;
    const int dims[3] = {nx, ny, nz};
    const long strides[3] = {1, nx, (long)nx * ny};
    double coeff[32] = {0};
    orc_lap_coefficient(coeff, fdn);
    zc *F = (zc *)Fd;
    zc v[ILU_SLOTS];
    for (long i = 0; i < n; i++) {
        const int cr[3] = {(int)(i % nx), (int)((i / nx) % ny), (int)(i / ((long)nx * ny))};
        v[0].re = a * (3 * coeff[0]) + (c ? c[i] : c0);
        v[0].im = eta;
        for (int ax = 0; ax < 3; ax++)
            for (int d = 1; d <= 6; d++) {
                v[ilu_slot(ax, -d)].re = a * coeff[d];
                v[ilu_slot(ax, -d)].im = 0.0;
                v[ilu_slot(ax, d)].re = a * coeff[d];
                v[ilu_slot(ax, d)].im = 0.0;
            }
        for (int ax = 2; ax >= 0; ax--) {
            const int cc = cr[ax], na = dims[ax];
            for (int cp = 0; cp < cc; cp++) {
                const int dk = pdiff(cp, cc, na);
                if (dk < -6 || dk > 6) continue;
                const long kk = i + (long)(cp - cc) * strides[ax];
                const zc l = zmul(v[ilu_slot(ax, dk)], F[kk * ILU_SLOTS]);
                v[ilu_slot(ax, dk)] = l;
                for (int cj = cp + 1; cj < na; cj++) {
                    const int dkj = pdiff(cj, cp, na);
                    if (dkj < -6 || dkj > 6) continue;
                    const int dij = pdiff(cj, cc, na);
                    if (dij < -6 || dij > 6) continue;
                    const int si = dij == 0 ? 0 : ilu_slot(ax, dij);
                    v[si] = zsub(v[si], zmul(l, F[kk * ILU_SLOTS + ilu_slot(ax, dkj)]));
                }
            }
        }
        if (v[0].re == 0.0 && v[0].im == 0.0) return -2;
        const zc one = {1.0, 0.0};
        v[0] = zdiv(one, v[0]);
        memcpy(F + i * ILU_SLOTS, v, sizeof(v));
    }
    return 0;
}

void orc_ilu0_apply_complex(int nx, int ny, int nz, const double *Fd, const double *ind, double *outd, int trans)
{
    const long n = (long)nx * ny * nz;
    const int dims[3] = {nx, ny, nz};
    const long strides[3] = {1, nx, (long)nx * ny};
    const zc *F = (const zc *)Fd;
    const zc *in = (const zc *)ind;
    zc *out = (zc *)outd;
    int cs[12], ds[12];
    if (out != in) memcpy(out, in, sizeof(zc) * n);
    if (trans == 2)
        for (long i = 0; i < n; i++) out[i] = zconj(out[i]);
    if (!trans) {
        for (long i = 0; i < n; i++) {
            const int cr[3] = {(int)(i % nx), (int)((i / nx) % ny), (int)(i / ((long)nx * ny))};
            zc s = out[i];
            for (int ax = 2; ax >= 0; ax--) {
                const int m = axis_cols(cr[ax], dims[ax], 1, cs, ds);
                for (int q = 0; q < m; q++)
                    s = zsub(s, zmul(F[i * ILU_SLOTS + ilu_slot(ax, ds[q])],
                                     out[i + (long)(cs[q] - cr[ax]) * strides[ax]]));
            }
            out[i] = s;
        }
        for (long i = n - 1; i >= 0; i--) {
            const int cr[3] = {(int)(i % nx), (int)((i / nx) % ny), (int)(i / ((long)nx * ny))};
            zc s = out[i];
            for (int ax = 0; ax < 3; ax++) {
                const int m = axis_cols(cr[ax], dims[ax], 0, cs, ds);
                for (int q = 0; q < m; q++)
                    s = zsub(s, zmul(F[i * ILU_SLOTS + ilu_slot(ax, ds[q])],
                                     out[i + (long)(cs[q] - cr[ax]) * strides[ax]]));
            }
            out[i] = zmul(s, F[i * ILU_SLOTS]);
        }
    } else {
        for (long i = 0; i < n; i++) {
            const int cr[3] = {(int)(i % nx), (int)((i / nx) % ny), (int)(i / ((long)nx * ny))};
            zc s = out[i];
            for (int ax = 2; ax >= 0; ax--) {
                const int m = axis_cols(cr[ax], dims[ax], 1, cs, ds);
                for (int q = 0; q < m; q++) {
                    const long j = i + (long)(cs[q] - cr[ax]) * strides[ax];
                    s = zsub(s, zmul(F[j * ILU_SLOTS + ilu_slot(ax, -ds[q])], out[j]));
                }
            }
            out[i] = zmul(s, F[i * ILU_SLOTS]);
        }
        for (long i = n - 1; i >= 0; i--) {
            const int cr[3] = {(int)(i % nx), (int)((i / nx) % ny), (int)(i / ((long)nx * ny))};
            zc s = out[i];
            for (int ax = 0; ax < 3; ax++) {
                const int m = axis_cols(cr[ax], dims[ax], 0, cs, ds);
                for (int q = 0; q < m; q++) {
                    const long j = i + (long)(cs[q] - cr[ax]) * strides[ax];
                    s = zsub(s, zmul(F[j * ILU_SLOTS + ilu_slot(ax, -ds[q])], out[j]));
                }
            }
            out[i] = s;
        }
    }
    if (trans == 2)
        for (long i = 0; i < n; i++) out[i] = zconj(out[i]);
}
