// krylov.hip -- the solvers of the PETSc variants of the reference:
//  * the Krylov comparators of PETSC_REAL/main.c:86-150, GMRES and BiCG,
//    run the way PETSc's KSPGMRES and KSPBICG run them there (rtol, max_it
//    2000, nonzero initial guess), with PCJACOBI or the pc = 1 block-Jacobi
//    ILU(0);
//  * that ILU(0) (level-scheduled factorization and triangular solves);
//  * the variant's own AAR, PGR and PL2R (AAR_Real.c, PGR_Real.c,
//    PL2R_Real.c) with PCJACOBI or PCICC (as ILU(0));
//  * GMRES and BiCG of PETSC_COMPLEX/main.c on the Helmholtz operator, with
//    Jacobi or the complex ILU(0).
//
// PETSc itself is not part of the reference tree.  The control flow below
// follows its published algorithms and defaults:
//  * convergence (KSPConvergedDefault): ttol = max(rtol*||B b||, atol) and
//    rnorm0 = rnorm at the first test; converged when rnorm <= ttol,
//    diverged when rnorm >= dtol*rnorm0 (atol 1e-50, dtol 1e5);
//  * Jacobi (PCJACOBI): y = x * (1/diag), the reciprocal taken once;
//  * GMRES: restart k (30 by default), left preconditioning, the
//    preconditioned residual norm, classical Gram-Schmidt without
//    refinement, Givens rotations, happy-breakdown tolerance 1e-30, the
//    solution built by back-substitution at each cycle's end;
//  * BiCG: left-preconditioned BiCG with the shadow residual r~0 = r0.
//    The operators here are symmetric and the Jacobi scaling is diagonal,
//    so the shadow sequence (r~, z~, p~) equals the primal one value for
//    value (same operations on the same numbers); it is computed once.
//
// Data layout: the GMRES basis is restart+1 padded slots of one buffer; all
// vectors FP64 (complex128 interleaved for the Helmholtz solvers).  Each iteration is a few streaming kernels (the FD12 stencil
// for A v, one pass of dot products, one pass that orthogonalises and takes
// the norm, one scaling pass) and one-block control kernels that reduce the
// per-block partial sums in a fixed order and run the scalar recurrences
// (Hessenberg rotations, convergence test) on the device, so the host only
// enqueues and polls.
#include "aar_device.h"
#include "aar_solver.h"

#include <algorithm>
#include <cstring>

namespace aar {

#define HC(x)                                                                         \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess)                                                         \
            return fail(AAR_HIP_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int KMAX = AAR_HIP_GMRES_MAXK;   // largest restart
constexpr int KB = 256;                    // threads per block
constexpr int KP = KMAX + 1;               // partial sums per block (stride)

// PETSc's KSPConvergedReason codes
enum KReason : int {
    K_ITERATING = 0,
    K_CONVERGED_RTOL = 2,
    K_CONVERGED_ATOL = 3,
    K_DIVERGED_NULL = -2,
    K_DIVERGED_ITS = -3,
    K_DIVERGED_DTOL = -4,
    K_DIVERGED_BREAKDOWN = -5,
    K_DIVERGED_BREAKDOWN_BICG = -6,
    K_DIVERGED_INDEFINITE_PC = -8,
    K_DIVERGED_NANORINF = -9,
    K_DIVERGED_INDEFINITE_MAT = -10,
};
constexpr double K_ATOL = 1e-50, K_DTOL = 1e5, K_HAPTOL = 1e-30;

// the Jacobi scaling: 1/diag per point (1 without a preconditioner)
struct KDiag {
    const double *c;   // variable coefficient c(x) (diagonal diag + c) or null
    double diag, dinv;
    int precond;
};

__device__ __forceinline__ double kdinv(const KDiag &d, long i)
{
    if (!d.precond) return 1.0;
    return d.c ? 1.0 / (d.diag + d.c[i]) : d.dinv;
}

struct KCtrl {
    Ctrl c;        // c.done: the solve has a reason (stencil early exit, host polling)
    Ctrl gate;     // gate.done: no Arnoldi step pending (early exit of the step's kernels)
    int reason, its, it, open, noscale, nosol, nh, hcap;
    int max_it, restart, first, zero;   // zero: never set (a skip flag that never skips)
    double rtol, ttol, rnorm0, bpnorm, rhs_norm, scale, relres;
    double beta, betaold, bcoef, alpha;
    double dpi;                          // CG: p.A p of the last step
    double h[KMAX + 1];
    double HH[(KMAX + 2) * (KMAX + 1)];   // column-major, columns of KMAX + 2
    double cc[KMAX + 1], ss[KMAX + 1], grs[KMAX + 2], nrs[KMAX + 1];
};

__device__ __forceinline__ double &HHe(KCtrl *k, int r, int c) { return k->HH[c * (KMAX + 2) + r]; }

__device__ int k_converged(KCtrl *k, int n, double rnorm)
{
    if (n == 0) {
        k->ttol = fmax(k->rtol * k->bpnorm, K_ATOL);
        k->rnorm0 = rnorm;
    }
    if (isnan(rnorm) || isinf(rnorm)) return K_DIVERGED_NANORINF;
    if (rnorm <= k->ttol) return rnorm < K_ATOL ? K_CONVERGED_ATOL : K_CONVERGED_RTOL;
    if (rnorm >= K_DTOL * k->rnorm0) return K_DIVERGED_DTOL;
    return K_ITERATING;
}

__device__ void k_push(KCtrl *k, double *hist, double v)
{
    if (hist && k->nh < k->hcap) hist[k->nh] = v;
    k->nh++;
}

__device__ void k_stop(KCtrl *k)
{
    k->c.done = 1;
    k->c.final_iter = k->its;
    k->gate.done = 1;
}

// Block sums of NV per-thread accumulators (the first nv used) into
// part[blockIdx.x * KP + j], fixed order.
template <int NV>
__device__ void k_block_sums(const double (&acc)[NV], int nv, double *part)
{
    __shared__ double red[KB / 64][NV];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; j++) {
        if (j < nv) {
            const double v = wave_sum(acc[j]);
            if (lane == 0) red[wv][j] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x < nv) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < KB / 64; w++) s += red[w][threadIdx.x];
        part[(size_t)blockIdx.x * KP + threadIdx.x] = s;
    }
}

// One-block reduction of the nv partial sums of nb blocks, fixed order;
// out[j] valid for every thread after return.
__device__ void k_reduce(const double *part, int nb, int nv, double *out)
{
    __shared__ double red[KB / 64][KP];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j = 0; j < nv; j++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += KB) s += part[(size_t)b * KP + j];
        s = wave_sum(s);
        if (lane == 0) red[wv][j] = s;
    }
    __syncthreads();
    if (threadIdx.x < nv) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < KB / 64; w++) s += red[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// streaming kernels (grid-stride over the n points)
// ---------------------------------------------------------------------------
#define KLOOP(i) for (long i = (long)blockIdx.x * KB + threadIdx.x; i < n; i += (long)gridDim.x * KB)

__global__ __launch_bounds__(KB) void k_zero_kernel(unsigned *p, long words)
{
    for (long i = (long)blockIdx.x * KB + threadIdx.x; i < words; i += (long)gridDim.x * KB) p[i] = 0u;
}

// ||B b||^2 and ||b||^2
__global__ __launch_bounds__(KB) void k_bnorm_kernel(long n, const double *b, KDiag dg, double *part)
{
    double acc[2] = {0.0, 0.0};
    KLOOP(i)
    {
        const double bi = b[i];
        const double z = kdinv(dg, i) * bi;
        acc[0] += z * z;
        acc[1] += bi * bi;
    }
    k_block_sums<2>(acc, 2, part);
}

// ||b - t||^2 (t = A x): the true residual of the returned solution
__global__ __launch_bounds__(KB) void k_resnorm_kernel(long n, const double *b, const double *t, double *part)
{
    double acc[1] = {0.0};
    KLOOP(i)
    {
        const double r = b[i] - t[i];
        acc[0] += r * r;
    }
    k_block_sums<1>(acc, 1, part);
}

// GMRES cycle start: v0 = B (b - t), t = A x; ||v0||^2
__global__ __launch_bounds__(KB) void k_gm_resid_kernel(long n, const double *b, const double *t, double *v0,
                                                       KDiag dg, double *part, const int *skip)
{
    if (*skip) return;
    double acc[1] = {0.0};
    KLOOP(i)
    {
        const double v = kdinv(dg, i) * (b[i] - t[i]);
        v0[i] = v;
        acc[0] += v * v;
    }
    k_block_sums<1>(acc, 1, part);
}

// h_j = <B t, v_j>, j < nv (t = A v_{nv-1})
template <int NV>
__global__ __launch_bounds__(KB) void k_gm_mdot_kernel(long n, const double *t, const double *V, long ld, int nv,
                                                      KDiag dg, double *part, const int *skip)
{
    if (*skip) return;
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) acc[j] = 0.0;
    KLOOP(i)
    {
        const double w = kdinv(dg, i) * t[i];
#pragma unroll
        for (int j = 0; j < NV; j++)
            if (j < nv) acc[j] += w * V[(size_t)j * ld + i];
    }
    k_block_sums<NV>(acc, nv, part);
}

// v_nv = B t - sum_j h_j v_j (classical Gram-Schmidt), ||v_nv||^2
template <int NV>
__global__ __launch_bounds__(KB) void k_gm_orth_kernel(long n, const double *t, double *V, long ld, int nv,
                                                      KDiag dg, const KCtrl *k, double *part)
{
    if (k->gate.done) return;
    double nh[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) nh[j] = j < nv ? -k->h[j] : 0.0;
    double *out = V + (size_t)nv * ld;
    double acc[1] = {0.0};
    KLOOP(i)
    {
        double w = kdinv(dg, i) * t[i];
#pragma unroll
        for (int j = 0; j < NV; j++)
            if (j < nv) w = w + nh[j] * V[(size_t)j * ld + i];
        out[i] = w;
        acc[0] += w * w;
    }
    k_block_sums<1>(acc, 1, part);
}

// v *= scale (VecNormalize), unless the control says otherwise; an
// Arnoldi step's vector (gated) only when that step ran
__global__ __launch_bounds__(KB) void k_gm_scale_kernel(long n, double *v, const KCtrl *k, int gated)
{
    if (k->noscale || (gated && k->gate.done)) return;
    const double s = k->scale;
    KLOOP(i) v[i] = s * v[i];
}

// x += sum_j nrs_j v_j, j < it (the cycle's solution update)
template <int NV>
__global__ __launch_bounds__(KB) void k_gm_soln_kernel(long n, double *x, const double *V, long ld, const KCtrl *k)
{
    if (k->nosol) return;
    const int nv = k->it;
    double c[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) c[j] = j < nv ? k->nrs[j] : 0.0;
    KLOOP(i)
    {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < NV; j++)
            if (j < nv) s = s + c[j] * V[(size_t)j * ld + i];
        x[i] = x[i] + s;
    }
}

// BiCG start: r = b - t (t = A x), z = B r; ||z||^2, <z, r>
__global__ __launch_bounds__(KB) void k_bi_init_kernel(long n, const double *b, const double *t, double *r, double *z,
                                                      KDiag dg, double *part)
{
    double acc[2] = {0.0, 0.0};
    KLOOP(i)
    {
        const double ri = b[i] + (-1.0) * t[i];
        const double zi = kdinv(dg, i) * ri;
        r[i] = ri;
        z[i] = zi;
        acc[0] += zi * zi;
        acc[1] += zi * ri;
    }
    k_block_sums<2>(acc, 2, part);
}

// p = z (first iteration) or z + (beta/betaold) p
__global__ __launch_bounds__(KB) void k_bi_p_kernel(long n, const double *z, double *p, const KCtrl *k)
{
    if (k->c.done) return;
    if (k->first) {
        KLOOP(i) p[i] = z[i];
    } else {
        const double bc = k->bcoef;
        KLOOP(i) p[i] = z[i] + bc * p[i];
    }
}

// <q, p> (q = A p)
__global__ __launch_bounds__(KB) void k_bi_dot_kernel(long n, const double *q, const double *p, double *part,
                                                     const KCtrl *k)
{
    if (k->c.done) return;
    double acc[1] = {0.0};
    KLOOP(i) acc[0] += q[i] * p[i];
    k_block_sums<1>(acc, 1, part);
}

// x += a p; r -= a q; z = B r; ||z||^2, <z, r>
__global__ __launch_bounds__(KB) void k_bi_upd_kernel(long n, double *x, double *r, double *z, const double *p,
                                                     const double *q, KDiag dg, double *part, const KCtrl *k)
{
    if (k->c.done) return;
    const double al = k->alpha, ma = -k->alpha;
    double acc[2] = {0.0, 0.0};
    KLOOP(i)
    {
        x[i] = x[i] + al * p[i];
        const double ri = r[i] + ma * q[i];
        const double zi = kdinv(dg, i) * ri;
        r[i] = ri;
        z[i] = zi;
        acc[0] += zi * zi;
        acc[1] += zi * ri;
    }
    k_block_sums<2>(acc, 2, part);
}

// ||v||^2
__global__ __launch_bounds__(KB) void k_norm_kernel(long n, const double *v, double *part, const int *skip)
{
    if (*skip) return;
    double acc[1] = {0.0};
    KLOOP(i) acc[0] += v[i] * v[i];
    k_block_sums<1>(acc, 1, part);
}

// ||B b||^2 and ||b||^2 with B b given (zb)
__global__ __launch_bounds__(KB) void k_bnorm2_kernel(long n, const double *b, const double *zb, double *part)
{
    double acc[2] = {0.0, 0.0};
    KLOOP(i)
    {
        acc[0] += zb[i] * zb[i];
        acc[1] += b[i] * b[i];
    }
    k_block_sums<2>(acc, 2, part);
}

// BiCG with a non-diagonal preconditioner, whose transpose differs from
// it in rounding: the shadow sequence is carried explicitly.
// r = b - t (t = A x); r~ = r
__global__ __launch_bounds__(KB) void k_bi2_init_kernel(long n, const double *b, const double *t, double *r, double *rl)
{
    KLOOP(i)
    {
        const double ri = b[i] + (-1.0) * t[i];
        r[i] = ri;
        rl[i] = ri;
    }
}

// ||z||^2, <z, r~>
__global__ __launch_bounds__(KB) void k_bi2_norms_kernel(long n, const double *z, const double *rl, double *part,
                                                        const KCtrl *k)
{
    if (k->c.done) return;
    double acc[2] = {0.0, 0.0};
    KLOOP(i)
    {
        const double zi = z[i];
        acc[0] += zi * zi;
        acc[1] += zi * rl[i];
    }
    k_block_sums<2>(acc, 2, part);
}

// p = z, p~ = z~ (first) or p = z + b p, p~ = z~ + b p~
__global__ __launch_bounds__(KB) void k_bi2_p_kernel(long n, const double *z, const double *zl, double *p, double *pl,
                                                    const KCtrl *k)
{
    if (k->c.done) return;
    if (k->first) {
        KLOOP(i)
        {
            p[i] = z[i];
            pl[i] = zl[i];
        }
    } else {
        const double bc = k->bcoef;
        KLOOP(i)
        {
            p[i] = z[i] + bc * p[i];
            pl[i] = zl[i] + bc * pl[i];
        }
    }
}

// x += a p; r -= a q; r~ -= a q~
__global__ __launch_bounds__(KB) void k_bi2_upd_kernel(long n, double *x, double *r, double *rl, const double *p,
                                                      const double *q, const double *ql, const KCtrl *k)
{
    if (k->c.done) return;
    const double al = k->alpha, ma = -k->alpha;
    KLOOP(i)
    {
        x[i] = x[i] + al * p[i];
        r[i] = r[i] + ma * q[i];
        rl[i] = rl[i] + ma * ql[i];
    }
}

// --- AAR of the PETSc variant (AAR_Real.c:32-199) with a general B ---
// f = B r for the pointwise preconditioners (Jacobi or none)
__global__ __launch_bounds__(KB) void k_pcj_kernel(long n, const double *r, double *f, KDiag dg, const int *skip)
{
    if (*skip) return;
    KLOOP(i) f[i] = kdinv(dg, i) * r[i];
}

// AAR_Real.c:107-119: DX_k = x - x_old, DF_k = f - f_old (iter > 1);
// x_old = x, f_old = f; Richardson x += omega f
__global__ __launch_bounds__(KB) void k_aa_hist_kernel(long n, double *x, const double *f, double *x_old, double *f_old,
                                                      double *dx, double *df, int hist, int rich, double omega,
                                                      const KCtrl *k)
{
    if (k->c.done) return;
    KLOOP(i)
    {
        const double xi = x[i], fi = f[i];
        if (hist) {
            dx[i] = (-1.0) * x_old[i] + xi;
            df[i] = (-1.0) * f_old[i] + fi;
        }
        x_old[i] = xi;
        f_old[i] = fi;
        if (rich) x[i] = xi + omega * fi;
    }
}

// AAR_Real.c:128-137: x += sum_q (-s_q)(beta DF_q + DX_q), then x += beta f
template <int MM>
__global__ __launch_bounds__(KB) void k_aa_upd_kernel(long n, double *x, const double *f, const double *W, long ld,
                                                     int m, double beta, const KCtrl *k)
{
    if (k->c.done) return;
    double ns[MM];
#pragma unroll
    for (int q = 0; q < MM; q++) ns[q] = q < m ? -k->c.s[q] : 0.0;
    KLOOP(i)
    {
        double xi = x[i];
#pragma unroll
        for (int q = 0; q < MM; q++)
            if (q < m) xi = xi + ns[q] * (beta * W[(size_t)(1 + q) * ld + i] + W[(size_t)(1 + m + q) * ld + i]);
        x[i] = xi + beta * f[i];
    }
}

// res = b - t (VecAYPX(res, -1, b)); ||res||^2 when checked
__global__ __launch_bounds__(KB) void k_aa_res_kernel(long n, const double *b, double *t, int check, double *part,
                                                     const int *skip)
{
    if (*skip) return;
    double acc[1] = {0.0};
    KLOOP(i)
    {
        const double r = b[i] + (-1.0) * t[i];
        t[i] = r;
        acc[0] += r * r;
    }
    if (check) k_block_sums<1>(acc, 1, part);
}

__global__ void k_aa_init_kernel(KCtrl *k, double tol)
{
    k->c.tol_abs = tol * k->rhs_norm;      // AAR_Real.c:81
    k->c.relres = k->c.tol_abs + 1;        // r_2norm = tol + 1 (AAR_Real.c:91)
}

// the check of iteration `iter` (AAR_Real.c:149-150, 98)
__global__ __launch_bounds__(KB) void k_aa_check_kernel(KCtrl *k, const double *part, int nb, double *hist, int iter)
{
    if (k->c.done) return;
    __shared__ double s[KP];
    k_reduce(part, nb, 1, s);
    if (threadIdx.x != 0) return;
    const double r = sqrt(s[0]);
    k->c.relres = r;
    k->its = iter;
    k_push(k, hist, r);
    if (!(r > k->c.tol_abs)) {
        k->c.done = 1;
        k->c.final_iter = iter;
    }
}

// --- PGR / PL2R of the PETSc variant (PGR_Real.c, PL2R_Real.c) ---
// x_old = x, Ax_prev = Ax; x += omega f; DX_k = x - x_old
__global__ __launch_bounds__(KB) void k_pg_rich_kernel(long n, double *x, const double *f, double *x_old,
                                                      const double *Ax, double *Ax_prev, double *dx, double omega,
                                                      const KCtrl *k)
{
    if (k->c.done) return;
    KLOOP(i)
    {
        const double xo = x[i];
        x_old[i] = xo;
        Ax_prev[i] = Ax[i];
        const double xn = xo + omega * f[i];
        x[i] = xn;
        dx[i] = (-1.0) * xo + xn;
    }
}

// res = b - Ax; DF_k = Ax - Ax_prev; ||res||^2 when checked (no Galerkin step)
__global__ __launch_bounds__(KB) void k_pg_res_kernel(long n, const double *b, const double *Ax, const double *Ax_prev,
                                                     double *res, double *df, int check, double *part, const KCtrl *k)
{
    if (k->c.done) return;
    double acc[1] = {0.0};
    KLOOP(i)
    {
        const double ax = Ax[i];
        const double r = (-1.0) * ax + b[i];
        res[i] = r;
        df[i] = (-1.0) * Ax_prev[i] + ax;
        acc[0] += r * r;
    }
    if (check) k_block_sums<1>(acc, 1, part);
}

// the Galerkin / L2 step: x += sum s_q DX_q; DX_k = x - x_old; Ax += sum
// s_q DF_q; DF_k = Ax - Ax_prev; res = b - Ax with ||res||^2
template <int MM>
__global__ __launch_bounds__(KB) void k_pg_upd_kernel(long n, double *x, const double *x_old, double *Ax,
                                                     const double *Ax_prev, double *res, const double *b,
                                                     double *X, double *Y, long ld, int m, int kk, double *part,
                                                     const KCtrl *k)
{
    if (k->c.done) return;
    double s[MM];
#pragma unroll
    for (int q = 0; q < MM; q++) s[q] = q < m ? k->c.s[q] : 0.0;
    double acc[1] = {0.0};
    KLOOP(i)
    {
        double xi = x[i];
#pragma unroll
        for (int q = 0; q < MM; q++)
            if (q < m) xi = xi + s[q] * X[(size_t)(1 + q) * ld + i];
        x[i] = xi;
        X[(size_t)(1 + kk) * ld + i] = (-1.0) * x_old[i] + xi;
        double ax = Ax[i];
#pragma unroll
        for (int q = 0; q < MM; q++)
            if (q < m) ax = ax + s[q] * Y[(size_t)(1 + q) * ld + i];
        Ax[i] = ax;
        Y[(size_t)(1 + kk) * ld + i] = (-1.0) * Ax_prev[i] + ax;
        const double r = (-1.0) * ax + b[i];
        res[i] = r;
        acc[0] += r * r;
    }
    k_block_sums<1>(acc, 1, part);
}

// ---------------------------------------------------------------------------
// control kernels (one block; thread 0 runs the scalar recurrences)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(KB) void k_ctl_init_kernel(KCtrl *k, double rtol, int max_it, int restart, int hcap,
                                                       const double *part, int nb)
{
    __shared__ double s[KP];
    k_reduce(part, nb, 2, s);
    if (threadIdx.x != 0) return;
    k->rtol = rtol;
    k->max_it = max_it;
    k->restart = restart;
    k->hcap = hcap;
    k->bpnorm = sqrt(s[0]);
    k->rhs_norm = sqrt(s[1]);
    k->c.rhs_norm = k->rhs_norm;
    k->gate.done = 1;
    k->nosol = 1;
    k->noscale = 1;
}

// GMRES cycle start (KSPGMRESCycle's prologue): normalise v0, test
__global__ __launch_bounds__(KB) void k_gm_start_kernel(KCtrl *k, const double *part, int nb, double *hist)
{
    if (k->c.done) return;
    __shared__ double s[KP];
    k_reduce(part, nb, 1, s);
    if (threadIdx.x != 0) return;
    const double res = sqrt(s[0]);
    k->open = 1;
    k->it = 0;
    k->grs[0] = res;
    k->scale = 1.0 / res;
    k->noscale = (res == 0.0 || res == 1.0);
    k_push(k, hist, res);
    k->reason = res == 0.0 ? K_CONVERGED_ATOL : k_converged(k, k->its, res);
    const bool active = !k->reason && k->restart > 0 && k->its < k->max_it;
    k->gate.done = active ? 0 : 1;
}

__global__ __launch_bounds__(KB) void k_gm_dots_kernel(KCtrl *k, const double *part, int nb, int nv)
{
    if (k->gate.done) return;
    __shared__ double s[KP];
    k_reduce(part, nb, nv, s);
    if (threadIdx.x < nv) k->h[threadIdx.x] = s[threadIdx.x];
}

// after the orthogonalisation: the new Hessenberg column, its rotations
// (KSPGMRESUpdateHessenberg), the residual estimate and the test
__global__ __launch_bounds__(KB) void k_gm_arnoldi_kernel(KCtrl *k, const double *part, int nb, double *hist)
{
    if (k->gate.done) return;
    __shared__ double s[KP];
    k_reduce(part, nb, 1, s);
    if (threadIdx.x != 0) return;
    const int it = k->it;
    const double tt = sqrt(s[0]);
    for (int j = 0; j <= it; j++) HHe(k, j, it) = k->h[j];
    HHe(k, it + 1, it) = tt;
    k->scale = 1.0 / tt;
    double hapbnd = fabs(tt / k->grs[it]);
    if (hapbnd > K_HAPTOL) hapbnd = K_HAPTOL;
    const bool hapend = tt < hapbnd;
    double *hh = &HHe(k, 0, it);
    for (int j = 1; j <= it; j++) {
        const double t1 = hh[j - 1];
        hh[j - 1] = k->cc[j - 1] * t1 + k->ss[j - 1] * hh[j];
        hh[j] = k->cc[j - 1] * hh[j] - k->ss[j - 1] * t1;
    }
    double res = 0.0;
    int reason = K_ITERATING;
    if (!hapend) {
        const double r = sqrt(hh[it] * hh[it] + hh[it + 1] * hh[it + 1]);
        if (r == 0.0) {
            reason = K_DIVERGED_NULL;
        } else {
            k->cc[it] = hh[it] / r;
            k->ss[it] = hh[it + 1] / r;
            k->grs[it + 1] = -(k->ss[it] * k->grs[it]);
            k->grs[it] = k->cc[it] * k->grs[it];
            hh[it] = k->cc[it] * hh[it] + k->ss[it] * hh[it + 1];
            res = fabs(k->grs[it + 1]);
        }
    }
    k->it = it + 1;
    k->its++;
    k_push(k, hist, res);
    if (!reason) reason = k_converged(k, k->its, res);
    if (hapend && !reason) reason = K_DIVERGED_BREAKDOWN;
    k->reason = reason;
    const bool active = !reason && !hapend && k->it < k->restart && k->its < k->max_it;
    k->gate.done = active ? 0 : 1;
    k->noscale = (!active || tt == 0.0 || tt == 1.0);
}

// KSPGMRESBuildSoln: back-substitution for the cycle's coefficients
__global__ void k_gm_soln_ctl_kernel(KCtrl *k)
{
    if (!k->open || k->it == 0) {
        k->nosol = 1;
        return;
    }
    const int kk = k->it - 1;
    k->nrs[kk] = HHe(k, kk, kk) != 0.0 ? k->grs[kk] / HHe(k, kk, kk) : 0.0;
    for (int ii = 1; ii <= kk; ii++) {
        const int r = kk - ii;
        double tt = k->grs[r];
        for (int j = r + 1; j <= kk; j++) tt = tt - HHe(k, r, j) * k->nrs[j];
        k->nrs[r] = tt / HHe(k, r, r);
    }
    k->nosol = 0;
}

__global__ void k_gm_end_kernel(KCtrl *k)
{
    if (!k->open) return;
    if (k->its >= k->max_it && !k->reason) k->reason = K_DIVERGED_ITS;
    if (k->reason) k_stop(k);
    k->open = 0;
    k->nosol = 1;
    k->noscale = 1;
    k->gate.done = 1;
}

__global__ __launch_bounds__(KB) void k_bi_start_kernel(KCtrl *k, const double *part, int nb, double *hist)
{
    __shared__ double s[KP];
    k_reduce(part, nb, 2, s);
    if (threadIdx.x != 0) return;
    const double dp = sqrt(s[0]);
    k->beta = s[1];
    k->first = 1;
    k_push(k, hist, dp);
    int reason = k_converged(k, 0, dp);
    if (!reason && k->beta == 0.0) reason = K_DIVERGED_BREAKDOWN_BICG;
    k->reason = reason;
    if (reason) k_stop(k);
}

__global__ __launch_bounds__(KB) void k_bi_alpha_kernel(KCtrl *k, const double *part, int nb)
{
    if (k->c.done) return;
    __shared__ double s[KP];
    k_reduce(part, nb, 1, s);
    if (threadIdx.x != 0) return;
    k->alpha = k->beta / s[0];
}

__global__ __launch_bounds__(KB) void k_bi_step_kernel(KCtrl *k, const double *part, int nb, double *hist)
{
    if (k->c.done) return;
    __shared__ double s[KP];
    k_reduce(part, nb, 2, s);
    if (threadIdx.x != 0) return;
    const double dp = sqrt(s[0]);
    k->its++;
    k_push(k, hist, dp);
    int reason = k_converged(k, k->its, dp);
    if (!reason && k->its >= k->max_it) reason = K_DIVERGED_ITS;
    k->reason = reason;
    if (reason) {
        k_stop(k);
        return;
    }
    k->betaold = k->beta;
    k->beta = s[1];
    k->bcoef = k->beta / k->betaold;
    k->first = 0;
}

// ---------------------------------------------------------------------------
// CG of the PETSc variant (PETSC_REAL/main.c:152-175: KSPCG, PETSc's
// KSPSolve_CG with the preconditioned norm).  The vector kernels are
// BiCG's (with the shadow sequence equal to the primal one, BiCG's
// recurrences are CG's); these scalar steps add CG's exits: beta = 0
// (converged), beta * betaold < 0 (indefinite preconditioner) and a zero
// or sign-changing p.A p (indefinite matrix), with PETSc's its = i + 1.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(KB) void k_cg_start_kernel(KCtrl *k, const double *part, int nb, double *hist)
{
    __shared__ double s[KP];
    k_reduce(part, nb, 2, s);
    if (threadIdx.x != 0) return;
    const double dp = sqrt(s[0]);
    k->beta = s[1];
    k->first = 1;
    k->dpi = 0.0;
    k_push(k, hist, dp);
    int reason = k_converged(k, 0, dp);
    if (!reason && k->max_it > 0 && k->beta == 0.0) {
        k->its = 1;
        reason = K_CONVERGED_ATOL;
    }
    k->reason = reason;
    if (reason) k_stop(k);
}

__global__ __launch_bounds__(KB) void k_cg_alpha_kernel(KCtrl *k, const double *part, int nb)
{
    if (k->c.done) return;
    __shared__ double s[KP];
    k_reduce(part, nb, 1, s);
    if (threadIdx.x != 0) return;
    const double dpiold = k->dpi, dpi = s[0];
    k->dpi = dpi;
    if (dpi == 0.0 || (!k->first && dpi * dpiold <= 0.0)) {
        k->its++;
        k->reason = K_DIVERGED_INDEFINITE_MAT;
        k_stop(k);
        return;
    }
    k->alpha = k->beta / dpi;
}

__global__ __launch_bounds__(KB) void k_cg_step_kernel(KCtrl *k, const double *part, int nb, double *hist)
{
    if (k->c.done) return;
    __shared__ double s[KP];
    k_reduce(part, nb, 2, s);
    if (threadIdx.x != 0) return;
    const double dp = sqrt(s[0]);
    k->its++;
    k_push(k, hist, dp);
    int reason = k_converged(k, k->its, dp);
    if (!reason && k->its >= k->max_it) reason = K_DIVERGED_ITS;
    if (!reason) {
        // the checks at the top of PETSc's next step (its = i + 1 there)
        const double beta = s[1];
        if (beta == 0.0) {
            k->its++;
            reason = K_CONVERGED_ATOL;
        } else if (beta * k->beta < 0.0) {
            k->its++;
            reason = K_DIVERGED_INDEFINITE_PC;
        }
        k->betaold = k->beta;
        k->beta = beta;
        k->bcoef = k->beta / k->betaold;
        k->first = 0;
    }
    k->reason = reason;
    if (reason) k_stop(k);
}

// x += a p; r -= a q (CG with ILU(0): z = M^{-1} r follows)
__global__ __launch_bounds__(KB) void k_cg_upd_kernel(long n, double *x, double *r, const double *p, const double *q,
                                                     const KCtrl *k)
{
    if (k->c.done) return;
    const double al = k->alpha, ma = -k->alpha;
    KLOOP(i)
    {
        x[i] = x[i] + al * p[i];
        r[i] = r[i] + ma * q[i];
    }
}

__global__ __launch_bounds__(KB) void k_final_kernel(KCtrl *k, const double *part, int nb)
{
    __shared__ double s[KP];
    k_reduce(part, nb, 1, s);
    if (threadIdx.x != 0) return;
    k->relres = sqrt(s[0]);
    k->c.relres = k->relres;
}

// ---------------------------------------------------------------------------
// ILU(0) of the periodic FD12 star matrix (the PETSc variant's pc = 1:
// PCBJACOBI -> PCILU on one block, PETSC_REAL/main.c:88-91; PCICC of
// AAR_Real.c:58-60 is the same preconditioner in exact arithmetic on this
// symmetric matrix).  Every dependency of a row (x, y, z) in the
// elimination and in the forward solve has a smaller x + y + z (periodic
// wrap included), so each level x + y + z = L is one launch; the backward
// solves run the levels downwards.  Rows are eliminated in PETSc's IKJ
// order and solved in ascending column order, with no sums across rows:
// the factors and solves are bit-identical to the oracle's sequential ones.
//
// Factor layout: IFS doubles per row (37 used): [0] = 1/u_ii, [1 + 12*axis
// + k] the entry at periodic offset d along axis (k = d + 6 for d < 0,
// d + 5 for d > 0): l_ij below the row, u_ij above.
// ---------------------------------------------------------------------------
constexpr int IFS = 40;

__host__ __device__ constexpr int islot(int ax, int d) { return 1 + 12 * ax + (d < 0 ? d + 6 : d + 5); }

template <typename T>
struct IluArgsT {
    int nx, ny, nz;
    int level, x0, y0, wx, wy;   // the level and its (x, y) box
    double w[R + 1];             // a * coeff[o]
    double diag;                 // real part of the diagonal without c(x): a * 3 * coeff[0] (+ sigma, + c0)
    double eta;                  // its imaginary part (complex matrix)
    const double *c;             // + c(x) on the diagonal, or null
    T *F;
    const int *skip;             // nonzero: no work (solver stopped / step gated)
};
using IluArgs = IluArgsT<double>;

// scalar operations of the real and complex factorizations (complex:
// explicit formulas, as the oracle's)
__device__ __forceinline__ double s_mul(double a, double b) { return a * b; }
__device__ __forceinline__ double2 s_mul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double s_sub(double a, double b) { return a - b; }
__device__ __forceinline__ double2 s_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double s_inv(double a) { return 1.0 / a; }
__device__ __forceinline__ double2 s_inv(double2 b)
{
    const double den = b.x * b.x + b.y * b.y;
    return make_double2((1.0 * b.x + 0.0 * b.y) / den, (0.0 * b.x - 1.0 * b.y) / den);
}
template <typename T> __device__ __forceinline__ T s_make(double re, double im);
template <> __device__ __forceinline__ double s_make<double>(double re, double) { return re; }
template <> __device__ __forceinline__ double2 s_make<double2>(double re, double im) { return make_double2(re, im); }

// Threads per block of the per-level kernels.  A level's points lie on the
// plane x + y + z = L, so neighbouring lanes are 1 - nx*ny elements apart
// and every load of a wave touches 64 cache lines; one-wave blocks spread a
// level over as many CUs as it has waves instead of stacking four per CU.
constexpr int ILU_WG = 64;

// the point of this thread on the level, or false
template <typename T>
__device__ __forceinline__ bool ilu_point(const IluArgsT<T> &a, int &x, int &y, int &z)
{
    const int t = blockIdx.x * ILU_WG + threadIdx.x;
    if (t >= a.wx * a.wy) return false;
    x = a.x0 + t % a.wx;
    y = a.y0 + t / a.wx;
    z = a.level - x - y;
    return z >= 0 && z < a.nz;
}

// periodic offsets of a row's columns along one axis in ascending column
// order: q = 0..5 -> d = 1..6 (below the row only when wrapped), q = 6..11
// -> d = -6..-1 (below unless wrapped)
__device__ __forceinline__ constexpr int qoff(int q) { return q < 6 ? q + 1 : q - 12; }

template <typename T>
__device__ __forceinline__ void ilu_factor_point(const IluArgsT<T> &a, int x, int y, int z)
{
    const int cr[3] = {x, y, z}, dims[3] = {a.nx, a.ny, a.nz};
    const long strides[3] = {1, a.nx, (long)a.nx * a.ny};
    const long i = x + (long)y * a.nx + (long)z * a.nx * a.ny;
    T v[37];
    v[0] = s_make<T>(a.c ? a.diag + a.c[i] : a.diag, a.eta);
#pragma unroll
    for (int ax = 0; ax < 3; ax++)
#pragma unroll
        for (int d = 1; d <= R; d++) {
            v[islot(ax, -d)] = s_make<T>(a.w[d], 0.0);
            v[islot(ax, d)] = s_make<T>(a.w[d], 0.0);
        }
#pragma unroll
    for (int ax = 2; ax >= 0; ax--) {            // z, y, x: ascending column index
        const int c = cr[ax], n = dims[ax];
        const long st = strides[ax];
#pragma unroll
        for (int q = 0; q < 12; q++) {
            const int dk = qoff(q);
            const bool lower = dk > 0 ? (c + dk >= n) : (c + dk >= 0);
            if (!lower) continue;
            const int ck = dk > 0 ? c + dk - n : c + dk;
            const T *Fk = a.F + (size_t)(i + (long)(ck - c) * st) * IFS;
            const T l = s_mul(v[islot(ax, dk)], Fk[0]);
            v[islot(ax, dk)] = l;
            // row k's columns above k that are in row i's pattern: on this axis
#pragma unroll
            for (int dij = -R; dij <= R; dij++) {
                if (dij == dk) continue;
                int cj = c + dij;
                cj = cj < 0 ? cj + n : (cj >= n ? cj - n : cj);
                if (cj <= ck) continue;
                int dkj = dij - dk;
                if (dkj > R) dkj -= n;
                else if (dkj < -R) dkj += n;
                if (dkj < -R || dkj > R) continue;
                const int si = dij == 0 ? 0 : islot(ax, dij);
                v[si] = s_sub(v[si], s_mul(l, Fk[islot(ax, dkj)]));
            }
        }
    }
    T *Fi = a.F + (size_t)i * IFS;
    Fi[0] = s_inv(v[0]);
#pragma unroll
    for (int s = 1; s < 37; s++) Fi[s] = v[s];
}

template <typename T>
__global__ __launch_bounds__(ILU_WG) void ilu_factor_kernel(IluArgsT<T> a)
{
    int x, y, z;
    if (!ilu_point(a, x, y, z)) return;
    ilu_factor_point(a, x, y, z);
}

// One level of a triangular solve, in place on io (in -> io when in != io
// at the first solve of an apply).  lowerpass: the columns below the row
// (forward levels); trans: the transposed factor (entries of the other
// row); diagmul: multiply by 1/u_ii at the end.
template <typename T, bool LOWER, bool TRANS, bool DIAG>
__device__ __forceinline__ void ilu_solve_point(const IluArgsT<T> &a, const T *in, T *io, int x, int y, int z)
{
    const int cr[3] = {x, y, z}, dims[3] = {a.nx, a.ny, a.nz};
    const long strides[3] = {1, a.nx, (long)a.nx * a.ny};
    const long i = x + (long)y * a.nx + (long)z * a.nx * a.ny;
    const T *Fi = a.F + (size_t)i * IFS;
    T s = in[i];
#pragma unroll
    for (int t = 0; t < 3; t++) {
        const int ax = LOWER ? 2 - t : t;        // ascending column index
        const int c = cr[ax], n = dims[ax];
        const long st = strides[ax];
#pragma unroll
        for (int q = 0; q < 12; q++) {
            const int d = qoff(q);
            const bool below = d > 0 ? (c + d >= n) : (c + d >= 0);
            if (below != LOWER) continue;
            const int cj = d > 0 ? (below ? c + d - n : c + d) : (below ? c + d : c + d + n);
            const long j = i + (long)(cj - c) * st;
            const T f = TRANS ? a.F[(size_t)j * IFS + islot(ax, -d)] : Fi[islot(ax, d)];
            s = s_sub(s, s_mul(f, io[j]));
        }
    }
    io[i] = DIAG ? s_mul(s, Fi[0]) : s;
}

template <typename T, bool LOWER, bool TRANS, bool DIAG>
__global__ __launch_bounds__(ILU_WG) void ilu_solve_kernel(IluArgsT<T> a, const T *in, T *io)
{
    if (*a.skip) return;
    int x, y, z;
    if (!ilu_point(a, x, y, z)) return;
    ilu_solve_point<T, LOWER, TRANS, DIAG>(a, in, io, x, y, z);
}

// The levels L0, L0 + step, ..., L1 (each with at most ILU_SMALL_PTS points
// in its (x, y) box) in one workgroup, a barrier between levels: the same
// per-point work as ilu_solve_kernel without a launch per level.
constexpr int ILU_SMALL_WG = 1024, ILU_SMALL_PTS = 1024;

template <typename T>
__device__ __forceinline__ void ilu_level_box(const IluArgsT<T> &a, int L, int &x0, int &y0, int &wx, int &wy)
{
    x0 = max(0, L - (a.ny - 1) - (a.nz - 1));
    y0 = max(0, L - (a.nx - 1) - (a.nz - 1));
    wx = min(a.nx - 1, L) - x0 + 1;
    wy = min(a.ny - 1, L) - y0 + 1;
}

template <typename T, bool LOWER, bool TRANS, bool DIAG>
__global__ __launch_bounds__(ILU_SMALL_WG) void ilu_solve_small_kernel(IluArgsT<T> a, const T *in, T *io, int L0,
                                                                      int L1, int step)
{
    if (*a.skip) return;
    for (int L = L0;; L += step) {
        int x0, y0, wx, wy;
        ilu_level_box(a, L, x0, y0, wx, wy);
        for (int t = threadIdx.x; t < wx * wy; t += ILU_SMALL_WG) {
            const int x = x0 + t % wx, y = y0 + t / wx, z = L - x - y;
            if (z >= 0 && z < a.nz) ilu_solve_point<T, LOWER, TRANS, DIAG>(a, in, io, x, y, z);
        }
        if (L == L1) break;
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(ILU_SMALL_WG) void ilu_factor_small_kernel(IluArgsT<T> a, int L0, int L1)
{
    for (int L = L0;; L++) {
        int x0, y0, wx, wy;
        ilu_level_box(a, L, x0, y0, wx, wy);
        for (int t = threadIdx.x; t < wx * wy; t += ILU_SMALL_WG) {
            const int x = x0 + t % wx, y = y0 + t / wx, z = L - x - y;
            if (z >= 0 && z < a.nz) ilu_factor_point(a, x, y, z);
        }
        if (L == L1) break;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
namespace {

struct KrylovRun {
    KCtrl *k = nullptr;
    double *part = nullptr, *hist = nullptr;
    int nb = 0, hcap = 0;
    long n = 0;
    KDiag dg{};
    IluArgs ilu{};
};

int check_ilu_grid(const Op &op)
{
    if (op.nx < 13 || op.ny < 13 || op.nz < 13)
        return fail(AAR_HIP_ERR_ARG, "the ILU(0) preconditioner needs every grid side >= 13");
    return 0;
}

// the (x, y) box of level L: x + y + z = L inside the grid
template <typename T>
void ilu_box(IluArgsT<T> &a, int L)
{
    const int x0 = std::max(0, L - (a.ny - 1) - (a.nz - 1)), x1 = std::min(a.nx - 1, L);
    const int y0 = std::max(0, L - (a.nx - 1) - (a.nz - 1)), y1 = std::min(a.ny - 1, L);
    a.level = L;
    a.x0 = x0;
    a.y0 = y0;
    a.wx = x1 - x0 + 1;
    a.wy = y1 - y0 + 1;
}

IluArgs ilu_args(const Op &op, double *F)
{
    IluArgs a{};
    a.eta = 0.0;
    a.nx = op.nx;
    a.ny = op.ny;
    a.nz = op.nz;
    for (int o = 0; o <= R; o++) a.w[o] = op.w[o];
    a.diag = op.diag;
    a.c = op.c;
    a.F = F;
    return a;
}

template <typename T>
int ilu_nlev(const IluArgsT<T> &a) { return a.nx + a.ny + a.nz - 2; }

// a level whose (x, y) box has at most ILU_SMALL_PTS points
template <typename T>
bool ilu_small(IluArgsT<T> a, int L)
{
    ilu_box(a, L);
    return (long)a.wx * a.wy <= ILU_SMALL_PTS;
}

// Launch the levels from..to (step +1 or -1): runs of small levels go to
// `small(L0, L1)` in one launch, the others to `big(L)` one launch each.
template <typename T, typename Small, typename Big>
void ilu_levels(const IluArgsT<T> &a, int from, int to, int step, Small small, Big big)
{
    for (int L = from; L != to + step;) {
        if (ilu_small(a, L)) {
            int L1 = L;
            while (L1 != to && ilu_small(a, L1 + step)) L1 += step;
            small(L, L1);
            L = L1 + step;
        } else {
            big(L);
            L += step;
        }
    }
}

template <typename T>
hipError_t ilu_factor(IluArgsT<T> a, hipStream_t st)
{
    ilu_levels(
        a, 0, ilu_nlev(a) - 1, 1,
        [&](int L0, int L1) {
            hipLaunchKernelGGL((ilu_factor_small_kernel<T>), dim3(1), dim3(ILU_SMALL_WG), 0, st, a, L0, L1);
        },
        [&](int L) {
            IluArgsT<T> b = a;
            ilu_box(b, L);
            hipLaunchKernelGGL((ilu_factor_kernel<T>), dim3((b.wx * b.wy + ILU_WG - 1) / ILU_WG), dim3(ILU_WG), 0, st, b);
        });
    return hipGetLastError();
}

// out = M^{-1} in (trans: M^{-T} in); out may be in
template <typename T>
hipError_t ilu_apply(IluArgsT<T> a, const T *in, T *out, bool trans, const int *skip, hipStream_t st)
{
    a.skip = skip;
    const int nl = ilu_nlev(a);
    const dim3 sw(ILU_SMALL_WG);
    // forward: L (or U^T) over ascending levels, reading `in`
    ilu_levels(
        a, 0, nl - 1, 1,
        [&](int L0, int L1) {
            if (trans) hipLaunchKernelGGL((ilu_solve_small_kernel<T, true, true, true>), dim3(1), sw, 0, st, a, in, out,
                                          L0, L1, 1);
            else hipLaunchKernelGGL((ilu_solve_small_kernel<T, true, false, false>), dim3(1), sw, 0, st, a, in, out,
                                    L0, L1, 1);
        },
        [&](int L) {
            IluArgsT<T> b = a;
            ilu_box(b, L);
            const dim3 g((b.wx * b.wy + ILU_WG - 1) / ILU_WG), wg(ILU_WG);
            if (trans) hipLaunchKernelGGL((ilu_solve_kernel<T, true, true, true>), g, wg, 0, st, b, in, out);
            else hipLaunchKernelGGL((ilu_solve_kernel<T, true, false, false>), g, wg, 0, st, b, in, out);
        });
    // backward: U (or L^T) over descending levels, in place
    ilu_levels(
        a, nl - 1, 0, -1,
        [&](int L0, int L1) {
            if (trans) hipLaunchKernelGGL((ilu_solve_small_kernel<T, false, true, false>), dim3(1), sw, 0, st, a,
                                          (const T *)out, out, L0, L1, -1);
            else hipLaunchKernelGGL((ilu_solve_small_kernel<T, false, false, true>), dim3(1), sw, 0, st, a,
                                    (const T *)out, out, L0, L1, -1);
        },
        [&](int L) {
            IluArgsT<T> b = a;
            ilu_box(b, L);
            const dim3 g((b.wx * b.wy + ILU_WG - 1) / ILU_WG), wg(ILU_WG);
            if (trans) hipLaunchKernelGGL((ilu_solve_kernel<T, false, true, false>), g, wg, 0, st, b,
                                          (const T *)out, out);
            else hipLaunchKernelGGL((ilu_solve_kernel<T, false, false, true>), g, wg, 0, st, b, (const T *)out,
                                    out);
        });
    return hipGetLastError();
}

// algorithmic bytes of one ILU(0) apply: each row's 37 factor entries, the
// input and both passes' output
constexpr double ILU_APPLY_BYTES = (37 + 4) * 8.0;

int k_setup(const Op &op, int precond, double tol, int max_it, int restart, int hcap, const double *b,
            KrylovRun &R, double *F = nullptr, double *zb = nullptr)
{
    Context &C = ctx();
    R.n = op.n;
    R.nb = std::min(nblocks_for(op.n), 1024);
    R.hcap = hcap;
    HC(ensure(C.kctl, sizeof(KCtrl)));
    HC(ensure(C.hist0, sizeof(double) * (size_t)hcap));
    HC(ensure(C.part0, sizeof(double) * (size_t)R.nb * KP));
    R.k = (KCtrl *)C.kctl.p;
    R.part = (double *)C.part0.p;
    R.hist = (double *)C.hist0.p;
    R.dg.c = op.c;
    R.dg.diag = op.diag;
    R.dg.dinv = 1.0 / op.diag;
    R.dg.precond = precond == 1;     // the fused Jacobi scaling (ILU(0) is applied by its own sweeps)
    const long words = (long)(sizeof(KCtrl) / sizeof(unsigned));
    hipLaunchKernelGGL(k_zero_kernel, dim3(64), dim3(KB), 0, C.stream, (unsigned *)R.k, words);
    if (F) {
        // PCSetUp: factor, then B b for the tolerance
        R.ilu = ilu_args(op, F);
        HC(ilu_factor(R.ilu, C.stream));
        HC(ilu_apply(R.ilu, b, zb, false, &R.k->zero, C.stream));
        hipLaunchKernelGGL(k_bnorm2_kernel, dim3(R.nb), dim3(KB), 0, C.stream, R.n, b, (const double *)zb, R.part);
        C.bytes_acc += op.n * (37 * 8.0 + ILU_APPLY_BYTES);
    } else {
        hipLaunchKernelGGL(k_bnorm_kernel, dim3(R.nb), dim3(KB), 0, C.stream, R.n, b, R.dg, R.part);
    }
    hipLaunchKernelGGL(k_ctl_init_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, tol, max_it, restart, hcap,
                       (const double *)R.part, R.nb);
    HC(hipGetLastError());
    C.bytes_acc += 8.0 * op.n * (op.c ? 2 : 1);
    return 0;
}

// y = A x, skipped once *gate->done is set
int k_matvec(const Op &op, const double *x, double *y, Ctrl *gate)
{
    SArgs<double> a = base_sargs<double>(op);
    a.in0 = x;
    a.out0 = y;
    a.ctrl = gate;
    return run_stencil<double>(SMode::MATVEC, a, (op.c ? 3.0 : 2.0) * 8.0 * op.n);
}

int k_finish(const Op &op, const double *x, const double *b, double *t, KrylovRun &R, aar_hip_stats *st)
{
    Context &C = ctx();
    {
        // the true residual of the returned x (PETSC_REAL/main.c:112-118)
        SArgs<double> a = base_sargs<double>(op);
        a.in0 = x;
        a.out0 = t;
        if (int rc = run_stencil<double>(SMode::MATVEC, a, (op.c ? 3.0 : 2.0) * 8.0 * op.n)) return rc;
        hipLaunchKernelGGL(k_resnorm_kernel, dim3(R.nb), dim3(KB), 0, C.stream, R.n, b, (const double *)t, R.part);
        hipLaunchKernelGGL(k_final_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, (const double *)R.part, R.nb);
        HC(hipGetLastError());
    }
    HC(hipEventRecord(C.t1, C.stream));
    KCtrl hk;
    HC(hipMemcpyAsync(&hk, R.k, sizeof(KCtrl), hipMemcpyDeviceToHost, C.stream));
    HC(hipStreamSynchronize(C.stream));
    float ms = 0.f;
    HC(hipEventElapsedTime(&ms, C.t0, C.t1));
    prof_flush();
    const int hl = std::min(hk.nh, R.hcap);
    C.hist_host.assign(std::max(hl, 0), 0.0);
    if (hl > 0) HC(hipMemcpy(C.hist_host.data(), R.hist, sizeof(double) * hl, hipMemcpyDeviceToHost));
    aar_hip_stats s;
    std::memset(&s, 0, sizeof(s));
    s.iterations = hk.its;
    s.rhs_norm = hk.rhs_norm;
    s.relres = hk.rhs_norm > 0 ? hk.relres / hk.rhs_norm : hk.relres;
    s.converged = hk.reason > 0;
    s.printed_converged = hk.reason > 0;
    s.printed_warning = hk.reason == K_DIVERGED_ITS;
    s.status = hk.reason == K_DIVERGED_NANORINF ? AAR_HIP_ERR_NONFINITE : 0;
    s.hist_len = hl;
    s.solve_ms = ms;
    s.bytes = C.bytes_acc;
    s.reason = hk.reason;
    C.last = s;
    if (st) *st = s;
    if (hk.reason == K_DIVERGED_NANORINF) return fail(AAR_HIP_ERR_NONFINITE, "residual is not finite");
    return 0;
}

template <int NV>
hipError_t launch_gm_step(const Op &op, const double *t, double *V, long ld, int nv, KrylovRun &R)
{
    Context &C = ctx();
    hipLaunchKernelGGL((k_gm_mdot_kernel<NV>), dim3(R.nb), dim3(KB), 0, C.stream, R.n, t, (const double *)V, ld, nv,
                       R.dg, R.part, (const int *)&R.k->gate.done);
    hipLaunchKernelGGL(k_gm_dots_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, (const double *)R.part, R.nb, nv);
    hipLaunchKernelGGL((k_gm_orth_kernel<NV>), dim3(R.nb), dim3(KB), 0, C.stream, R.n, t, V, ld, nv, R.dg,
                       (const KCtrl *)R.k, R.part);
    return hipGetLastError();
}

template <int NV>
hipError_t launch_gm_soln(double *x, const double *V, long ld, KrylovRun &R)
{
    hipLaunchKernelGGL((k_gm_soln_kernel<NV>), dim3(R.nb), dim3(KB), 0, ctx().stream, R.n, x, V, ld,
                       (const KCtrl *)R.k);
    return hipGetLastError();
}

}  // namespace

int gmres_solve(const Op &op, double *x, const double *b, int restart, int max_it, double tol, int precond,
                aar_hip_stats *st)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!x || !b) return fail(AAR_HIP_ERR_ARG, "null vector");
    if (max_it < 0) return fail(AAR_HIP_ERR_ARG, "max_iter must be >= 0");
    if (restart < 1 || restart > KMAX) return fail(AAR_HIP_ERR_ARG, "restart must be in 1..32");
    if (precond < 0 || precond > 2) return fail(AAR_HIP_ERR_ARG, "precond must be 0 (none), 1 (Jacobi) or 2 (ILU(0))");
    const bool ilu = precond == 2;
    if (ilu && op.fdn != R)
        return fail(AAR_HIP_ERR_ARG, "ILU(0) is built for the full FD12 stencil (fdn = 6)");
    if (ilu)
        if (int rc = check_ilu_grid(op)) return rc;
    const long n = op.n;
    const long ld = ring_ld(n, sizeof(double));
    // basis v_0 .. v_restart, then t
    HC(ensure(C.kvec, sizeof(double) * (size_t)ld * (restart + 2)));
    double *V = (double *)C.kvec.p, *t = V + (size_t)ld * (restart + 1);
    double *F = nullptr;
    if (ilu) {
        HC(ensure(C.kilu, sizeof(double) * (size_t)n * IFS));
        F = (double *)C.kilu.p;
    }
    const int hcap = max_it + max_it / restart + 8;
    C.bytes_acc = 0.0;
    HC(hipEventRecord(C.t0, C.stream));
    KrylovRun R;
    if (int rc = k_setup(op, precond, tol, max_it, restart, hcap, b, R, F, t)) return rc;
    const double ib = ilu ? ILU_APPLY_BYTES * n : 0.0;
    const double s = 8.0 * n;
    const double cb = op.c ? s : 0.0;      // the coefficient field's reads, where 1/diag is per point
    bool stopped = false;
    for (int cyc = 0; cyc <= max_it && !stopped; cyc++) {
        // v0 = B (b - A x), ||v0||, test
        if (int rc = k_matvec(op, x, t, &R.k->c)) return rc;
        hipLaunchKernelGGL(k_gm_resid_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, b, (const double *)t, V, R.dg,
                           R.part, (const int *)&R.k->c.done);
        if (ilu) {
            HC(ilu_apply(R.ilu, V, V, false, &R.k->c.done, C.stream));
            hipLaunchKernelGGL(k_norm_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, (const double *)V, R.part,
                               (const int *)&R.k->c.done);
            C.bytes_acc += ib + s;
        }
        hipLaunchKernelGGL(k_gm_start_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, (const double *)R.part, R.nb,
                           R.hist);
        hipLaunchKernelGGL(k_gm_scale_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, V, (const KCtrl *)R.k, 0);
        HC(hipGetLastError());
        C.bytes_acc += 5 * s + cb;
        for (int it = 0; it < restart; it++) {
            const int nv = it + 1;
            if (int rc = k_matvec(op, V + (size_t)it * ld, t, &R.k->gate)) return rc;
            const double *w = t;
            if (ilu) {
                // w = M^{-1} A v_it, built in the next basis slot
                prof_begin("krylov_ilu_apply");
                HC(ilu_apply(R.ilu, t, V + (size_t)nv * ld, false, &R.k->gate.done, C.stream));
                prof_end(ib);
                C.bytes_acc += ib;
                w = V + (size_t)nv * ld;
            }
            prof_begin("krylov_gmres_orth");
            hipError_t e = nv <= 8 ? launch_gm_step<8>(op, w, V, ld, nv, R)
                         : nv <= 16 ? launch_gm_step<16>(op, w, V, ld, nv, R)
                                    : launch_gm_step<32>(op, w, V, ld, nv, R);
            prof_end((2.0 * nv + 3) * s + 2 * cb);
            if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("gmres step: ") + hipGetErrorString(e));
            hipLaunchKernelGGL(k_gm_arnoldi_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, (const double *)R.part, R.nb,
                               R.hist);
            hipLaunchKernelGGL(k_gm_scale_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, V + (size_t)nv * ld,
                               (const KCtrl *)R.k, 1);
            HC(hipGetLastError());
            C.bytes_acc += (2.0 * nv + 5) * s + 2 * cb;
        }
        hipLaunchKernelGGL(k_gm_soln_ctl_kernel, dim3(1), dim3(1), 0, C.stream, R.k);
        hipError_t e = restart <= 8 ? launch_gm_soln<8>(x, V, ld, R)
                     : restart <= 16 ? launch_gm_soln<16>(x, V, ld, R)
                                     : launch_gm_soln<32>(x, V, ld, R);
        if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("gmres solution: ") + hipGetErrorString(e));
        hipLaunchKernelGGL(k_gm_end_kernel, dim3(1), dim3(1), 0, C.stream, R.k);
        HC(hipGetLastError());
        C.bytes_acc += (restart + 2.0) * s;
        if (int rc = poll_record(&R.k->c, cyc)) return rc;
        if (poll_check(cyc - 2)) stopped = true;
    }
    return k_finish(op, x, b, t, R, st);
}

// z = M^{-1} r (trans: M^{-T} r) with the ILU(0) factors of op, factored here
int ilu0_apply(const Op &op, const double *r, double *z, int trans)
{
    if (op.fdn != R) return fail(AAR_HIP_ERR_ARG, "ILU(0) is built for the full FD12 stencil (fdn = 6)");
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!r || !z) return fail(AAR_HIP_ERR_ARG, "null vector");
    if (int rc = check_ilu_grid(op)) return rc;
    HC(ensure(C.kilu, sizeof(double) * (size_t)op.n * IFS));
    HC(ensure(C.kctl, sizeof(KCtrl)));
    KCtrl *k = (KCtrl *)C.kctl.p;
    hipLaunchKernelGGL(k_zero_kernel, dim3(64), dim3(KB), 0, C.stream, (unsigned *)k,
                       (long)(sizeof(KCtrl) / sizeof(unsigned)));
    IluArgs a = ilu_args(op, (double *)C.kilu.p);
    HC(ilu_factor(a, C.stream));
    HC(ilu_apply(a, r, z, trans != 0, &k->zero, C.stream));
    HC(hipStreamSynchronize(C.stream));
    return 0;
}

// AAR of the PETSc variant (PETSC_REAL/AAR_Real.c:32-199) with its
// preconditioner choice: Jacobi by the reciprocal diagonal (precond 1),
// ILU(0) standing in for PCICC (2; the same preconditioner in exact
// arithmetic on this symmetric matrix) or none (0).  A plain iteration on
// x: each step applies B to r = b - A x, updates the history columns,
// takes the Richardson or Anderson step and recomputes r; the Anderson
// least squares uses the Gram and one-wavefront SVD kernels of the
// MPI_REAL solver with dgelsd's rcond = -1 cutoff (singular values at or
// below eps * s_max dropped, LS_GELSD), as the variant's LAPACKE_dgelsd
// call does (PETSC_REAL/AAR_Real.c:223).
int aar_petsc_solve(const Op &op, double *x, const double *b, double omega, double beta, int m, int p,
                    int max_iter, double tol, int precond, aar_hip_stats *st)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!x || !b) return fail(AAR_HIP_ERR_ARG, "null vector");
    if (max_iter < 0) return fail(AAR_HIP_ERR_ARG, "max_iter must be >= 0");
    if (m < 1 || m > MAXM) return fail(AAR_HIP_ERR_ARG, "history m must be in 1..32");
    if (p < 1) return fail(AAR_HIP_ERR_ARG, "p must be >= 1");
    if (precond < 0 || precond > 2) return fail(AAR_HIP_ERR_ARG, "precond must be 0 (none), 1 (Jacobi) or 2 (ILU(0))");
    const bool ilu = precond == 2;
    if (ilu && op.fdn != R)
        return fail(AAR_HIP_ERR_ARG, "ILU(0) is built for the full FD12 stencil (fdn = 6)");
    if (ilu)
        if (int rc = check_ilu_grid(op)) return rc;
    const long n = op.n;
    const long ld = ring_ld(n, sizeof(double));
    // slots: 0 f, 1..m DF, m+1..2m DX, 2m+1 x_old, 2m+2 f_old, 2m+3 r
    HC(ensure(C.kvec, sizeof(double) * (size_t)ld * (2 * m + 4)));
    double *W = (double *)C.kvec.p;
    double *f = W, *x_old = W + (size_t)(2 * m + 1) * ld, *f_old = W + (size_t)(2 * m + 2) * ld;
    double *r = W + (size_t)(2 * m + 3) * ld;
    double *F = nullptr;
    if (ilu) {
        HC(ensure(C.kilu, sizeof(double) * (size_t)n * IFS));
        F = (double *)C.kilu.p;
    }
    const int gnb = nblocks_for(n);
    HC(ensure(C.part0, sizeof(double) * (size_t)gnb * GRAM_NA_MAX));
    const int hcap = max_iter / p + 8;
    C.bytes_acc = 0.0;
    HC(hipEventRecord(C.t0, C.stream));
    KrylovRun R;
    if (int rc = k_setup(op, precond, tol, max_iter, 0, hcap, b, R, F, r)) return rc;
    // the history columns start at zero (VecDuplicateVecs)
    hipLaunchKernelGGL(k_zero_kernel, dim3(1024), dim3(KB), 0, C.stream, (unsigned *)(W + ld),
                       (long)(2 * m * ld * 2));
    hipLaunchKernelGGL(k_aa_init_kernel, dim3(1), dim3(1), 0, C.stream, R.k, tol);
    const double s = 8.0 * n, ib = ilu ? ILU_APPLY_BYTES * n : 0.0;
    const double cb = op.c ? s : 0.0;
    const int *done = &R.k->c.done;
    if (int rc = k_matvec(op, x, r, nullptr)) return rc;
    hipLaunchKernelGGL(k_aa_res_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, b, r, 0, R.part, (const int *)&R.k->zero);
    HC(hipGetLastError());
    C.bytes_acc += 5 * s + cb;
    bool stopped = false;
    int cyc = 0;
    for (int iter = 1; iter <= max_iter && !stopped; iter++) {
        if (ilu) {
            prof_begin("krylov_ilu_apply");
            HC(ilu_apply(R.ilu, r, f, false, done, C.stream));
            prof_end(ib);
        } else {
            hipLaunchKernelGGL(k_pcj_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, (const double *)r, f, R.dg, done);
        }
        const int hk = iter > 1 ? (iter - 2) % m : 0;
        const bool rich = (iter % p) != 0;
        hipLaunchKernelGGL(k_aa_hist_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, x, (const double *)f, x_old, f_old,
                           W + (size_t)(1 + m + hk) * ld, W + (size_t)(1 + hk) * ld, iter > 1 ? 1 : 0, rich ? 1 : 0,
                           omega, (const KCtrl *)R.k);
        C.bytes_acc += (ilu ? ib : 2 * s + cb) + (iter > 1 ? 8 : 6) * s;
        if (!rich) {
            GramArgs<double> g;
            g.n = n;
            g.m = m;
            g.kind = 0;
            g.direct = 1;
            g.F = W;
            g.ld = ld;
            g.nvalid = m;
            for (int j = 0; j < m; j++) g.slot[j] = 1 + j;
            g.fslot = 0;
            g.nblocks = gnb;
            g.partials = R.part;
            g.counter = nullptr;
            g.ctrl = &R.k->c;
            HC(launch_gram<double>(g, C.stream));
            HC(launch_ls(&R.k->c, m, LS_GELSD, C.stream));   // dgelsd, rcond = -1
            if (m <= 8) hipLaunchKernelGGL((k_aa_upd_kernel<8>), dim3(R.nb), dim3(KB), 0, C.stream, n, x,
                                           (const double *)f, (const double *)W, ld, m, beta, (const KCtrl *)R.k);
            else hipLaunchKernelGGL((k_aa_upd_kernel<16>), dim3(R.nb), dim3(KB), 0, C.stream, n, x,
                                    (const double *)f, (const double *)W, ld, m, beta, (const KCtrl *)R.k);
            HC(hipGetLastError());
            C.bytes_acc += (m + 1) * s + (2.0 * m + 3) * s;
        }
        if (int rc = k_matvec(op, x, r, &R.k->c)) return rc;
        const bool check = !rich;
        hipLaunchKernelGGL(k_aa_res_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, b, r, check ? 1 : 0, R.part, done);
        if (check)
            hipLaunchKernelGGL(k_aa_check_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, (const double *)R.part, R.nb,
                               R.hist, iter);
        HC(hipGetLastError());
        C.bytes_acc += 3 * s;
        if (check) {
            if (int rc = poll_record(&R.k->c, cyc)) return rc;
            if (poll_check(cyc - 2)) stopped = true;
            cyc++;
        }
    }
    HC(hipEventRecord(C.t1, C.stream));
    KCtrl hk;
    HC(hipMemcpyAsync(&hk, R.k, sizeof(KCtrl), hipMemcpyDeviceToHost, C.stream));
    HC(hipStreamSynchronize(C.stream));
    float ms = 0.f;
    HC(hipEventElapsedTime(&ms, C.t0, C.t1));
    prof_flush();
    const int hl = std::min(hk.nh, R.hcap);
    C.hist_host.assign(std::max(hl, 0), 0.0);
    if (hl > 0) HC(hipMemcpy(C.hist_host.data(), R.hist, sizeof(double) * hl, hipMemcpyDeviceToHost));
    // AAR_Real.c:162-166: iter - 1 iterations (iter stops one past the last)
    const int iterations = hk.c.done ? hk.c.final_iter : max_iter;
    aar_hip_stats sres;
    std::memset(&sres, 0, sizeof(sres));
    sres.iterations = iterations;
    sres.rhs_norm = hk.rhs_norm;
    sres.relres = hk.rhs_norm > 0 ? hk.c.relres / hk.rhs_norm : hk.c.relres;
    sres.converged = hk.c.relres <= hk.c.tol_abs;
    sres.printed_converged = iterations + 1 < max_iter;
    sres.printed_warning = !(iterations + 1 < max_iter);
    sres.status = hk.c.status == 1 ? AAR_HIP_ERR_SVD : 0;
    sres.hist_len = hl;
    sres.solve_ms = ms;
    sres.bytes = C.bytes_acc;
    C.last = sres;
    if (st) *st = sres;
    if (hk.c.status == 1) return fail(AAR_HIP_ERR_SVD, "no convergence in 250 svd iterations");
    return 0;
}

// PGR (l2 = 0) / PL2R (l2 = 1) of the PETSc variant, PGR_Real.c:32-195 /
// PL2R_Real.c:32-197, preconditioner as aar_petsc_solve.  The normal
// equations come from the Gram kernel (kind 1: DX^T DF, kind 2: DF^T DF,
// with res = b - Ax formed in its loads) and the one-wavefront SVD with
// dgelsd's rcond = -1 cutoff (LS_GELSD; PGR_Real.c:216, PL2R_Real.c:220).
int pgr_petsc_solve(const Op &op, double *x, const double *b, double omega, int m, int p, int max_iter, double tol,
                    int precond, int l2, aar_hip_stats *st)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!x || !b) return fail(AAR_HIP_ERR_ARG, "null vector");
    if (max_iter < 0) return fail(AAR_HIP_ERR_ARG, "max_iter must be >= 0");
    if (m < 1 || m > MAXM) return fail(AAR_HIP_ERR_ARG, "history m must be in 1..32");
    if (p < 1) return fail(AAR_HIP_ERR_ARG, "p must be >= 1");
    if (precond < 0 || precond > 2) return fail(AAR_HIP_ERR_ARG, "precond must be 0 (none), 1 (Jacobi) or 2 (ILU(0))");
    const bool ilu = precond == 2;
    if (ilu && op.fdn != R)
        return fail(AAR_HIP_ERR_ARG, "ILU(0) is built for the full FD12 stencil (fdn = 6)");
    if (ilu)
        if (int rc = check_ilu_grid(op)) return rc;
    const long n = op.n;
    const long ld = ring_ld(n, sizeof(double));
    // Y: slot 0 Ax, 1..m DF;  X: slot 0 f, 1..m DX;  then x_old, Ax_prev, res
    HC(ensure(C.kvec, sizeof(double) * (size_t)ld * (2 * m + 5)));
    double *Y = (double *)C.kvec.p, *X = Y + (size_t)(m + 1) * ld;
    double *Ax = Y, *f = X;
    double *x_old = X + (size_t)(m + 1) * ld, *Ax_prev = x_old + ld, *res = Ax_prev + ld;
    double *F = nullptr;
    if (ilu) {
        HC(ensure(C.kilu, sizeof(double) * (size_t)n * IFS));
        F = (double *)C.kilu.p;
    }
    const int gnb = nblocks_for(n);
    HC(ensure(C.part0, sizeof(double) * (size_t)gnb * GRAM_NA_MAX));
    const int hcap = max_iter / p + 8;
    C.bytes_acc = 0.0;
    HC(hipEventRecord(C.t0, C.stream));
    KrylovRun R;
    if (int rc = k_setup(op, precond, tol, max_iter, 0, hcap, b, R, F, res)) return rc;
    // history columns start at zero
    hipLaunchKernelGGL(k_zero_kernel, dim3(1024), dim3(KB), 0, C.stream, (unsigned *)(Y + ld), (long)(m * ld * 2));
    hipLaunchKernelGGL(k_zero_kernel, dim3(1024), dim3(KB), 0, C.stream, (unsigned *)(X + ld), (long)(m * ld * 2));
    hipLaunchKernelGGL(k_aa_init_kernel, dim3(1), dim3(1), 0, C.stream, R.k, tol);
    const double s = 8.0 * n, ib = ilu ? ILU_APPLY_BYTES * n : 0.0;
    const double cb = op.c ? s : 0.0;
    const int *done = &R.k->c.done;
    if (int rc = k_matvec(op, x, Ax, nullptr)) return rc;
    HC(hipMemcpyAsync(res, Ax, sizeof(double) * n, hipMemcpyDeviceToDevice, C.stream));
    hipLaunchKernelGGL(k_aa_res_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, b, res, 0, R.part,
                       (const int *)&R.k->zero);
    HC(hipGetLastError());
    C.bytes_acc += 6 * s + cb;
    bool stopped = false;
    int cyc = 0;
    for (int iter = 1; iter <= max_iter && !stopped; iter++) {
        if (ilu) {
            prof_begin("krylov_ilu_apply");
            HC(ilu_apply(R.ilu, res, f, false, done, C.stream));
            prof_end(ib);
        } else {
            hipLaunchKernelGGL(k_pcj_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, (const double *)res, f, R.dg,
                               done);
        }
        const int kk = (iter - 2 + m) % m;
        const bool gal = (iter % p) == 0;
        hipLaunchKernelGGL(k_pg_rich_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, x, (const double *)f, x_old,
                           (const double *)Ax, Ax_prev, X + (size_t)(1 + kk) * ld, omega, (const KCtrl *)R.k);
        if (int rc = k_matvec(op, x, Ax, &R.k->c)) return rc;
        hipLaunchKernelGGL(k_pg_res_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, b, (const double *)Ax,
                           (const double *)Ax_prev, res, Y + (size_t)(1 + kk) * ld, 0, R.part, (const KCtrl *)R.k);
        C.bytes_acc += (ilu ? ib : 2 * s + cb) + 7 * s + 2 * s + cb + 5 * s;
        if (gal) {
            GramArgs<double> g;
            g.n = n;
            g.m = m;
            g.kind = l2 ? 2 : 1;
            g.direct = 1;
            g.F = X;
            g.Y = Y;
            g.b = b;
            g.ld = ld;
            g.nvalid = m;
            for (int j = 0; j < m; j++) g.slot[j] = 1 + j;
            g.fslot = 0;
            g.nblocks = gnb;
            g.partials = R.part;
            g.counter = nullptr;
            g.ctrl = &R.k->c;
            HC(launch_gram<double>(g, C.stream));
            HC(launch_ls(&R.k->c, m, LS_GELSD, C.stream));   // dgelsd, rcond = -1
            if (m <= 8)
                hipLaunchKernelGGL((k_pg_upd_kernel<8>), dim3(R.nb), dim3(KB), 0, C.stream, n, x,
                                   (const double *)x_old, Ax, (const double *)Ax_prev, res, b, X, Y, ld, m, kk, R.part,
                                   (const KCtrl *)R.k);
            else
                hipLaunchKernelGGL((k_pg_upd_kernel<16>), dim3(R.nb), dim3(KB), 0, C.stream, n, x,
                                   (const double *)x_old, Ax, (const double *)Ax_prev, res, b, X, Y, ld, m, kk, R.part,
                                   (const KCtrl *)R.k);
            hipLaunchKernelGGL(k_aa_check_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, (const double *)R.part, R.nb,
                               R.hist, iter);
            HC(hipGetLastError());
            C.bytes_acc += ((l2 ? 1.0 : 2.0) * m + 2) * s + (2.0 * m + 9) * s;
            if (int rc = poll_record(&R.k->c, cyc)) return rc;
            if (poll_check(cyc - 2)) stopped = true;
            cyc++;
        }
    }
    HC(hipEventRecord(C.t1, C.stream));
    KCtrl hk;
    HC(hipMemcpyAsync(&hk, R.k, sizeof(KCtrl), hipMemcpyDeviceToHost, C.stream));
    HC(hipStreamSynchronize(C.stream));
    float ms = 0.f;
    HC(hipEventElapsedTime(&ms, C.t0, C.t1));
    prof_flush();
    const int hl = std::min(hk.nh, R.hcap);
    C.hist_host.assign(std::max(hl, 0), 0.0);
    if (hl > 0) HC(hipMemcpy(C.hist_host.data(), R.hist, sizeof(double) * hl, hipMemcpyDeviceToHost));
    const int iterations = hk.c.done ? hk.c.final_iter : max_iter;
    aar_hip_stats sres;
    std::memset(&sres, 0, sizeof(sres));
    sres.iterations = iterations;
    sres.rhs_norm = hk.rhs_norm;
    sres.relres = hk.rhs_norm > 0 ? hk.c.relres / hk.rhs_norm : hk.c.relres;
    sres.converged = hk.c.relres <= hk.c.tol_abs;
    sres.printed_converged = iterations + 1 < max_iter;
    sres.printed_warning = !(iterations + 1 < max_iter);
    sres.status = hk.c.status == 1 ? AAR_HIP_ERR_SVD : 0;
    sres.hist_len = hl;
    sres.solve_ms = ms;
    sres.bytes = C.bytes_acc;
    C.last = sres;
    if (st) *st = sres;
    if (hk.c.status == 1) return fail(AAR_HIP_ERR_SVD, "no convergence in 250 svd iterations");
    return 0;
}

// BiCG with ILU(0): M^{-T} differs from M^{-1} in rounding, so the shadow
// sequence (r~, z~ = M^{-T} r~, p~, A^T p~ = A p~) is carried explicitly.
static int bicg_ilu_solve(const Op &op, double *x, const double *b, int max_it, double tol, aar_hip_stats *st)
{
    Context &C = ctx();
    if (int rc = check_ilu_grid(op)) return rc;
    const long n = op.n;
    const long ld = ring_ld(n, sizeof(double));
    HC(ensure(C.kvec, sizeof(double) * (size_t)ld * 8));
    HC(ensure(C.kilu, sizeof(double) * (size_t)n * IFS));
    double *r = (double *)C.kvec.p, *rl = r + ld, *z = r + 2 * ld, *zl = r + 3 * ld;
    double *p = r + 4 * ld, *pl = r + 5 * ld, *q = r + 6 * ld, *ql = r + 7 * ld;
    const int hcap = max_it + 8;
    C.bytes_acc = 0.0;
    HC(hipEventRecord(C.t0, C.stream));
    KrylovRun R;
    if (int rc = k_setup(op, 2, tol, max_it, 0, hcap, b, R, (double *)C.kilu.p, z)) return rc;
    const double s = 8.0 * n, ib = ILU_APPLY_BYTES * n;
    const double cb = op.c ? s : 0.0;
    if (int rc = k_matvec(op, x, q, nullptr)) return rc;
    hipLaunchKernelGGL(k_bi2_init_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, b, (const double *)q, r, rl);
    HC(ilu_apply(R.ilu, r, z, false, &R.k->zero, C.stream));
    HC(ilu_apply(R.ilu, rl, zl, true, &R.k->zero, C.stream));
    hipLaunchKernelGGL(k_bi2_norms_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, (const double *)z,
                       (const double *)rl, R.part, (const KCtrl *)R.k);
    hipLaunchKernelGGL(k_bi_start_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, (const double *)R.part, R.nb, R.hist);
    HC(hipGetLastError());
    C.bytes_acc += 6 * s + cb + 2 * ib;
    bool stopped = false;
    const int loops = std::max(max_it, 1);
    int cyc = 0;
    for (int i = 0; i < loops && !stopped; i++) {
        hipLaunchKernelGGL(k_bi2_p_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, (const double *)z,
                           (const double *)zl, p, pl, (const KCtrl *)R.k);
        if (int rc = k_matvec(op, p, q, &R.k->c)) return rc;
        if (int rc = k_matvec(op, pl, ql, &R.k->c)) return rc;     // A^T p~ (symmetric operator)
        hipLaunchKernelGGL(k_bi_dot_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, (const double *)q,
                           (const double *)pl, R.part, (const KCtrl *)R.k);
        hipLaunchKernelGGL(k_bi_alpha_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, (const double *)R.part, R.nb);
        hipLaunchKernelGGL(k_bi2_upd_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, x, r, rl, (const double *)p,
                           (const double *)q, (const double *)ql, (const KCtrl *)R.k);
        prof_begin("krylov_ilu_apply");
        HC(ilu_apply(R.ilu, r, z, false, &R.k->c.done, C.stream));
        HC(ilu_apply(R.ilu, rl, zl, true, &R.k->c.done, C.stream));
        prof_end(2 * ib);
        hipLaunchKernelGGL(k_bi2_norms_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, (const double *)z,
                           (const double *)rl, R.part, (const KCtrl *)R.k);
        hipLaunchKernelGGL(k_bi_step_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, (const double *)R.part, R.nb,
                           R.hist);
        HC(hipGetLastError());
        C.bytes_acc += 21 * s + 2 * cb + 2 * ib;
        if ((i + 1) % 8 == 0) {
            if (int rc = poll_record(&R.k->c, cyc)) return rc;
            if (poll_check(cyc - 2)) stopped = true;
            cyc++;
        }
    }
    return k_finish(op, x, b, q, R, st);
}

int bicg_solve(const Op &op, double *x, const double *b, int max_it, double tol, int precond, aar_hip_stats *st)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!x || !b) return fail(AAR_HIP_ERR_ARG, "null vector");
    if (max_it < 0) return fail(AAR_HIP_ERR_ARG, "max_iter must be >= 0");
    if (precond < 0 || precond > 2) return fail(AAR_HIP_ERR_ARG, "precond must be 0 (none), 1 (Jacobi) or 2 (ILU(0))");
    if (precond == 2) {
        if (op.fdn != R) return fail(AAR_HIP_ERR_ARG, "ILU(0) is built for the full FD12 stencil (fdn = 6)");
        return bicg_ilu_solve(op, x, b, max_it, tol, st);
    }
    const long n = op.n;
    const long ld = ring_ld(n, sizeof(double));
    HC(ensure(C.kvec, sizeof(double) * (size_t)ld * 4));
    double *r = (double *)C.kvec.p, *z = r + ld, *p = r + 2 * ld, *q = r + 3 * ld;
    const int hcap = max_it + 8;
    C.bytes_acc = 0.0;
    HC(hipEventRecord(C.t0, C.stream));
    KrylovRun R;
    if (int rc = k_setup(op, precond, tol, max_it, 0, hcap, b, R)) return rc;
    const double s = 8.0 * n;
    const double cb = op.c ? s : 0.0;
    if (int rc = k_matvec(op, x, q, nullptr)) return rc;
    hipLaunchKernelGGL(k_bi_init_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, b, (const double *)q, r, z, R.dg,
                       R.part);
    hipLaunchKernelGGL(k_bi_start_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, (const double *)R.part, R.nb, R.hist);
    HC(hipGetLastError());
    C.bytes_acc += 4 * s + cb;
    bool stopped = false;
    const int loops = std::max(max_it, 1);    // PETSc's do-while runs one step even at max_it = 0
    int cyc = 0;
    for (int i = 0; i < loops && !stopped; i++) {
        hipLaunchKernelGGL(k_bi_p_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, (const double *)z, p,
                           (const KCtrl *)R.k);
        if (int rc = k_matvec(op, p, q, &R.k->c)) return rc;
        prof_begin("krylov_bicg_update");
        hipLaunchKernelGGL(k_bi_dot_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, (const double *)q,
                           (const double *)p, R.part, (const KCtrl *)R.k);
        hipLaunchKernelGGL(k_bi_alpha_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, (const double *)R.part, R.nb);
        hipLaunchKernelGGL(k_bi_upd_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, x, r, z, (const double *)p,
                           (const double *)q, R.dg, R.part, (const KCtrl *)R.k);
        hipLaunchKernelGGL(k_bi_step_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, (const double *)R.part, R.nb,
                           R.hist);
        prof_end(9 * s + cb);
        HC(hipGetLastError());
        C.bytes_acc += 12 * s + cb;
        if ((i + 1) % 8 == 0) {
            if (int rc = poll_record(&R.k->c, cyc)) return rc;
            if (poll_check(cyc - 2)) stopped = true;
            cyc++;
        }
    }
    return k_finish(op, x, b, q, R, st);
}

int kspcg_solve(const Op &op, double *x, const double *b, int max_it, double tol, int precond, aar_hip_stats *st)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!x || !b) return fail(AAR_HIP_ERR_ARG, "null vector");
    if (max_it < 0) return fail(AAR_HIP_ERR_ARG, "max_iter must be >= 0");
    if (precond < 0 || precond > 2) return fail(AAR_HIP_ERR_ARG, "precond must be 0 (none), 1 (Jacobi) or 2 (ILU(0))");
    const bool ilu = precond == 2;
    if (ilu && op.fdn != R) return fail(AAR_HIP_ERR_ARG, "ILU(0) is built for the full FD12 stencil (fdn = 6)");
    if (ilu)
        if (int rc = check_ilu_grid(op)) return rc;
    const long n = op.n;
    const long ld = ring_ld(n, sizeof(double));
    HC(ensure(C.kvec, sizeof(double) * (size_t)ld * 4));
    if (ilu) HC(ensure(C.kilu, sizeof(double) * (size_t)n * IFS));
    double *r = (double *)C.kvec.p, *z = r + ld, *p = r + 2 * ld, *q = r + 3 * ld;
    const int hcap = max_it + 8;
    C.bytes_acc = 0.0;
    HC(hipEventRecord(C.t0, C.stream));
    KrylovRun R;
    if (ilu) {
        if (int rc = k_setup(op, 2, tol, max_it, 0, hcap, b, R, (double *)C.kilu.p, z)) return rc;
    } else {
        if (int rc = k_setup(op, precond, tol, max_it, 0, hcap, b, R)) return rc;
    }
    const double s = 8.0 * n, ib = ilu ? ILU_APPLY_BYTES * n : 0.0;
    const double cb = op.c ? s : 0.0;
    if (int rc = k_matvec(op, x, q, nullptr)) return rc;
    if (ilu) {
        hipLaunchKernelGGL(k_bi2_init_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, b, (const double *)q, r, p);
        HC(ilu_apply(R.ilu, r, z, false, &R.k->zero, C.stream));
        hipLaunchKernelGGL(k_bi2_norms_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, (const double *)z,
                           (const double *)r, R.part, (const KCtrl *)R.k);
    } else {
        hipLaunchKernelGGL(k_bi_init_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, b, (const double *)q, r, z, R.dg,
                           R.part);
    }
    hipLaunchKernelGGL(k_cg_start_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, (const double *)R.part, R.nb, R.hist);
    HC(hipGetLastError());
    C.bytes_acc += 4 * s + cb + ib;
    bool stopped = false;
    const int loops = std::max(max_it, 1);
    int cyc = 0;
    for (int i = 0; i < loops && !stopped; i++) {
        hipLaunchKernelGGL(k_bi_p_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, (const double *)z, p,
                           (const KCtrl *)R.k);
        if (int rc = k_matvec(op, p, q, &R.k->c)) return rc;
        hipLaunchKernelGGL(k_bi_dot_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, (const double *)q,
                           (const double *)p, R.part, (const KCtrl *)R.k);
        hipLaunchKernelGGL(k_cg_alpha_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, (const double *)R.part, R.nb);
        if (ilu) {
            hipLaunchKernelGGL(k_cg_upd_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, x, r, (const double *)p,
                               (const double *)q, (const KCtrl *)R.k);
            prof_begin("krylov_ilu_apply");
            HC(ilu_apply(R.ilu, r, z, false, &R.k->c.done, C.stream));
            prof_end(ib);
            hipLaunchKernelGGL(k_bi2_norms_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, (const double *)z,
                               (const double *)r, R.part, (const KCtrl *)R.k);
        } else {
            hipLaunchKernelGGL(k_bi_upd_kernel, dim3(R.nb), dim3(KB), 0, C.stream, n, x, r, z, (const double *)p,
                               (const double *)q, R.dg, R.part, (const KCtrl *)R.k);
        }
        hipLaunchKernelGGL(k_cg_step_kernel, dim3(1), dim3(KB), 0, C.stream, R.k, (const double *)R.part, R.nb,
                           R.hist);
        HC(hipGetLastError());
        C.bytes_acc += (ilu ? 13 * s + ib : 12 * s) + cb;
        if ((i + 1) % 8 == 0) {
            if (int rc = poll_record(&R.k->c, cyc)) return rc;
            if (poll_check(cyc - 2)) stopped = true;
            cyc++;
        }
    }
    return k_finish(op, x, b, q, R, st);
}

}  // namespace aar

// ===========================================================================
// GMRES of the complex PETSc variant (PETSC_COMPLEX/main.c:80-118): the
// algorithm above on the complex Helmholtz operator, as PETSc's complex
// build runs it: <x, y> = sum x conj(y), Givens rotations with complex
// c and s.  Vectors complex128 interleaved (double2).  Complex products,
// quotients and the Jacobi reciprocal (dr - i eta) / (dr^2 + eta^2), D =
// (a*3*coeff[0] + c) + i eta, use explicit formulas (no fused operations),
// identical to the oracle's, so pointwise results agree bit for bit.
// ===========================================================================
namespace aar {

constexpr int KPZ = 2 * (KMAX + 1);   // partial sums per block, complex dots

__device__ __forceinline__ double2 zmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 zdiv(double2 a, double2 b)
{
    const double den = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den);
}
__device__ __forceinline__ double zabs(double2 a) { return sqrt(a.x * a.x + a.y * a.y); }

struct ZDiag {
    const double *c;
    double dlap, c0, eta;
    int precond;
};

// B x for the complex Jacobi preconditioner (identity without one)
__device__ __forceinline__ double2 zpc(const ZDiag &d, long i, double2 x)
{
    if (!d.precond) return x;
    const double dr = d.dlap + (d.c ? d.c[i] : d.c0);
    const double den = dr * dr + d.eta * d.eta;
    return zmul(make_double2(dr / den, -d.eta / den), x);
}

struct KCtrlZ {
    Ctrl c;
    Ctrl gate;
    int reason, its, it, open, noscale, nosol, nh, hcap;
    int max_it, restart, zero, pad1;   // zero: never set (a skip flag that never skips)
    double rtol, ttol, rnorm0, bpnorm, rhs_norm, scale, relres;
    double2 h[KMAX + 1];
    double2 HH[(KMAX + 2) * (KMAX + 1)];
    double2 cc[KMAX + 1], ss[KMAX + 1], grs[KMAX + 2], nrs[KMAX + 1];
    double2 beta, betaold, bcoef, alpha;   // BiCG
    int first, pad2;
    // LGMRES: the unrotated Hessenberg, H y and 1/||z|| of the new error
    // approximation
    double2 HES[(KMAX + 2) * (KMAX + 1)];
    double2 hw[KMAX + 2];
    double ascale;
    int lg;          // LGMRES: the last vector of a cycle is normalised too (V H y uses it)
    int scale_nv;    // the basis slot the pending step's scale applies to
};

__device__ __forceinline__ double2 &ZHe(KCtrlZ *k, int r, int c) { return k->HH[c * (KMAX + 2) + r]; }
__device__ __forceinline__ double2 &ZSe(KCtrlZ *k, int r, int c) { return k->HES[c * (KMAX + 2) + r]; }

__device__ int kz_converged(KCtrlZ *k, int n, double rnorm)
{
    if (n == 0) {
        k->ttol = fmax(k->rtol * k->bpnorm, K_ATOL);
        k->rnorm0 = rnorm;
    }
    if (isnan(rnorm) || isinf(rnorm)) return K_DIVERGED_NANORINF;
    if (rnorm <= k->ttol) return rnorm < K_ATOL ? K_CONVERGED_ATOL : K_CONVERGED_RTOL;
    if (rnorm >= K_DTOL * k->rnorm0) return K_DIVERGED_DTOL;
    return K_ITERATING;
}

__device__ void kz_push(KCtrlZ *k, double *hist, double v)
{
    if (hist && k->nh < k->hcap) hist[k->nh] = v;
    k->nh++;
}

// block sums of NA accumulators (first na used) into part[blockIdx.x * KPZ + j]
template <int NA>
__device__ void kz_block_sums(const double (&acc)[NA], int na, double *part)
{
    __shared__ double red[KB / 64][NA];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NA; j++) {
        if (j < na) {
            const double v = wave_sum(acc[j]);
            if (lane == 0) red[wv][j] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x < na) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < KB / 64; w++) s += red[w][threadIdx.x];
        part[(size_t)blockIdx.x * KPZ + threadIdx.x] = s;
    }
}

__device__ void kz_reduce(const double *part, int nb, int na, double *out)
{
    __shared__ double red[KB / 64][KPZ];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j = 0; j < na; j++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += KB) s += part[(size_t)b * KPZ + j];
        s = wave_sum(s);
        if (lane == 0) red[wv][j] = s;
    }
    __syncthreads();
    if (threadIdx.x < na) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < KB / 64; w++) s += red[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

__global__ __launch_bounds__(KB) void kz_bnorm_kernel(long n, const double2 *b, ZDiag dg, double *part)
{
    double acc[2] = {0.0, 0.0};
    KLOOP(i)
    {
        const double2 bi = b[i];
        const double2 z = zpc(dg, i, bi);
        acc[0] += z.x * z.x + z.y * z.y;
        acc[1] += bi.x * bi.x + bi.y * bi.y;
    }
    kz_block_sums<2>(acc, 2, part);
}

__global__ __launch_bounds__(KB) void kz_resnorm_kernel(long n, const double2 *b, const double2 *t, double *part)
{
    double acc[1] = {0.0};
    KLOOP(i)
    {
        const double2 r = zsub(b[i], t[i]);
        acc[0] += r.x * r.x + r.y * r.y;
    }
    kz_block_sums<1>(acc, 1, part);
}

__global__ __launch_bounds__(KB) void kz_resid_kernel(long n, const double2 *b, const double2 *t, double2 *v0,
                                                     ZDiag dg, double *part, const int *skip)
{
    if (*skip) return;
    double acc[1] = {0.0};
    KLOOP(i)
    {
        const double2 v = zpc(dg, i, zsub(b[i], t[i]));
        v0[i] = v;
        acc[0] += v.x * v.x + v.y * v.y;
    }
    kz_block_sums<1>(acc, 1, part);
}

// h_j = <B t, v_j> = sum (B t) conj(v_j)
template <int NV>
__global__ __launch_bounds__(KB) void kz_mdot_kernel(long n, const double2 *t, const double2 *V, long ld, int nv,
                                                    ZDiag dg, double *part, const int *skip)
{
    if (*skip) return;
    double acc[2 * NV];
#pragma unroll
    for (int j = 0; j < 2 * NV; j++) acc[j] = 0.0;
    KLOOP(i)
    {
        const double2 w = zpc(dg, i, t[i]);
#pragma unroll
        for (int j = 0; j < NV; j++)
            if (j < nv) {
                const double2 p = zmul(w, zconj(V[(size_t)j * ld + i]));
                acc[2 * j] += p.x;
                acc[2 * j + 1] += p.y;
            }
    }
    kz_block_sums<2 * NV>(acc, 2 * nv, part);
}

template <int NV>
__global__ __launch_bounds__(KB) void kz_orth_kernel(long n, const double2 *t, double2 *V, long ld, int nv, ZDiag dg,
                                                    const KCtrlZ *k, double *part)
{
    if (k->gate.done) return;
    double2 nh[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) nh[j] = j < nv ? make_double2(-k->h[j].x, -k->h[j].y) : make_double2(0.0, 0.0);
    double2 *out = V + (size_t)nv * ld;
    double acc[1] = {0.0};
    KLOOP(i)
    {
        double2 w = zpc(dg, i, t[i]);
#pragma unroll
        for (int j = 0; j < NV; j++)
            if (j < nv) w = zadd(w, zmul(nh[j], V[(size_t)j * ld + i]));
        out[i] = w;
        acc[0] += w.x * w.x + w.y * w.y;
    }
    kz_block_sums<1>(acc, 1, part);
}

// v *= scale (VecNormalize) unless the control says otherwise; a step's
// vector (gated, basis slot nv) only when that step ran and produced it
__global__ __launch_bounds__(KB) void kz_scale_kernel(long n, double2 *v, const KCtrlZ *k, int gated, int nv)
{
    if (k->noscale || (gated && k->scale_nv != nv)) return;
    const double s = k->scale;
    KLOOP(i) v[i] = make_double2(s * v[i].x, s * v[i].y);
}

template <int NV>
__global__ __launch_bounds__(KB) void kz_soln_kernel(long n, double2 *x, const double2 *V, long ld, const KCtrlZ *k)
{
    if (k->nosol) return;
    const int nv = k->it;
    double2 c[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) c[j] = j < nv ? k->nrs[j] : make_double2(0.0, 0.0);
    KLOOP(i)
    {
        double2 s = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < NV; j++)
            if (j < nv) s = zadd(s, zmul(c[j], V[(size_t)j * ld + i]));
        x[i] = zadd(x[i], s);
    }
}

__global__ __launch_bounds__(KB) void kz_ctl_init_kernel(KCtrlZ *k, double rtol, int max_it, int restart, int hcap,
                                                        const double *part, int nb, int lg)
{
    __shared__ double s[KPZ];
    kz_reduce(part, nb, 2, s);
    if (threadIdx.x != 0) return;
    k->lg = lg;
    k->rtol = rtol;
    k->max_it = max_it;
    k->restart = restart;
    k->hcap = hcap;
    k->bpnorm = sqrt(s[0]);
    k->rhs_norm = sqrt(s[1]);
    k->gate.done = 1;
    k->nosol = 1;
    k->noscale = 1;
}

// itot: the steps of this cycle (GMRES: restart; LGMRES: the Arnoldi steps
// plus one per stored error approximation)
__global__ __launch_bounds__(KB) void kz_start_kernel(KCtrlZ *k, const double *part, int nb, double *hist, int itot)
{
    if (k->c.done) return;
    __shared__ double s[KPZ];
    kz_reduce(part, nb, 1, s);
    if (threadIdx.x != 0) return;
    const double res = sqrt(s[0]);
    k->restart = itot;
    k->scale_nv = 0;
    k->open = 1;
    k->it = 0;
    k->grs[0] = make_double2(res, 0.0);
    k->scale = 1.0 / res;
    k->noscale = (res == 0.0 || res == 1.0);
    kz_push(k, hist, res);
    k->reason = res == 0.0 ? K_CONVERGED_ATOL : kz_converged(k, k->its, res);
    const bool active = !k->reason && k->restart > 0 && k->its < k->max_it;
    k->gate.done = active ? 0 : 1;
}

__global__ __launch_bounds__(KB) void kz_dots_kernel(KCtrlZ *k, const double *part, int nb, int nv)
{
    if (k->gate.done) return;
    __shared__ double s[KPZ];
    kz_reduce(part, nb, 2 * nv, s);
    if (threadIdx.x < nv) k->h[threadIdx.x] = make_double2(s[2 * threadIdx.x], s[2 * threadIdx.x + 1]);
}

__global__ __launch_bounds__(KB) void kz_arnoldi_kernel(KCtrlZ *k, const double *part, int nb, double *hist)
{
    if (k->gate.done) return;
    __shared__ double s[KPZ];
    kz_reduce(part, nb, 1, s);
    if (threadIdx.x != 0) return;
    const int it = k->it;
    const double tt = sqrt(s[0]);
    for (int j = 0; j <= it; j++) ZHe(k, j, it) = ZSe(k, j, it) = k->h[j];
    ZHe(k, it + 1, it) = ZSe(k, it + 1, it) = make_double2(tt, 0.0);
    k->scale = 1.0 / tt;
    double hapbnd = tt / zabs(k->grs[it]);
    if (hapbnd > K_HAPTOL) hapbnd = K_HAPTOL;
    const bool hapend = tt < hapbnd;
    double2 *hh = &ZHe(k, 0, it);
    for (int j = 1; j <= it; j++) {
        const double2 t1 = hh[j - 1];
        hh[j - 1] = zadd(zmul(zconj(k->cc[j - 1]), t1), zmul(k->ss[j - 1], hh[j]));
        hh[j] = zsub(zmul(k->cc[j - 1], hh[j]), zmul(k->ss[j - 1], t1));
    }
    double res = 0.0;
    int reason = K_ITERATING;
    if (!hapend) {
        const double2 q = zadd(zmul(zconj(hh[it]), hh[it]), zmul(zconj(hh[it + 1]), hh[it + 1]));
        const double r = sqrt(q.x);
        if (r == 0.0) {
            reason = K_DIVERGED_NULL;
        } else {
            const double2 rr = make_double2(r, 0.0);
            k->cc[it] = zdiv(hh[it], rr);
            k->ss[it] = zdiv(hh[it + 1], rr);
            const double2 g = k->grs[it];
            const double2 ng = zmul(k->ss[it], g);
            k->grs[it + 1] = make_double2(-ng.x, -ng.y);
            k->grs[it] = zmul(zconj(k->cc[it]), g);
            hh[it] = zadd(zmul(zconj(k->cc[it]), hh[it]), zmul(k->ss[it], hh[it + 1]));
            res = zabs(k->grs[it + 1]);
        }
    }
    k->it = it + 1;
    k->its++;
    kz_push(k, hist, res);
    if (!reason) reason = kz_converged(k, k->its, res);
    if (hapend && !reason) reason = K_DIVERGED_BREAKDOWN;
    k->reason = reason;
    const bool active = !reason && !hapend && k->it < k->restart && k->its < k->max_it;
    k->gate.done = active ? 0 : 1;
    // GMRES never reads the last vector of a cycle; LGMRES's B A z does
    k->scale_nv = it + 1;
    k->noscale = ((!active && !k->lg) || tt == 0.0 || tt == 1.0);
}

__global__ void kz_soln_ctl_kernel(KCtrlZ *k)
{
    if (!k->open || k->it == 0) {
        k->nosol = 1;
        return;
    }
    const int kk = k->it - 1;
    k->nrs[kk] = zdiv(k->grs[kk], ZHe(k, kk, kk));
    for (int ii = 1; ii <= kk; ii++) {
        const int r = kk - ii;
        double2 tt = k->grs[r];
        for (int j = r + 1; j <= kk; j++) tt = zsub(tt, zmul(ZHe(k, r, j), k->nrs[j]));
        k->nrs[r] = zdiv(tt, ZHe(k, r, r));
    }
    k->nosol = 0;
}

// --- LGMRES (PETSC_COMPLEX/main.c:120-147): the cycle's correction and its
// error approximation.  za.z[q] is the stored approximation of order q + 1
// (newest first), the basis of step za.ita + q.
constexpr int ZAUG_MAX = AAR_HIP_LGMRES_MAXAUG;
struct ZAug {
    const double2 *z[ZAUG_MAX];
    int ita, nz;
};

// c = sum_j y_j W_j over the cycle's steps (W_j = v_j for the Arnoldi
// steps, z of the augmentation steps); x += c; znew = c; ||c||^2
template <int NV>
__global__ __launch_bounds__(KB) void kz_lsoln_kernel(long n, double2 *x, const double2 *V, long ld, ZAug za,
                                                     double2 *znew, const KCtrlZ *k, double *part)
{
    if (k->nosol) return;
    const int nv = k->it;
    const int na = nv < za.ita ? nv : za.ita;
    double2 c[NV], ca[ZAUG_MAX];
#pragma unroll
    for (int j = 0; j < NV; j++) c[j] = j < na ? k->nrs[j] : make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < ZAUG_MAX; q++)
        ca[q] = (q < za.nz && za.ita + q < nv) ? k->nrs[za.ita + q] : make_double2(0.0, 0.0);
    double acc[1] = {0.0};
    KLOOP(i)
    {
        double2 s = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < NV; j++)
            if (j < na) s = zadd(s, zmul(c[j], V[(size_t)j * ld + i]));
#pragma unroll
        for (int q = 0; q < ZAUG_MAX; q++)
            if (q < za.nz && za.ita + q < nv) s = zadd(s, zmul(ca[q], za.z[q][i]));
        znew[i] = s;
        x[i] = zadd(x[i], s);
        acc[0] += s.x * s.x + s.y * s.y;
    }
    kz_block_sums<1>(acc, 1, part);
}

// 1/||c|| and H y (H the unrotated Hessenberg): B A c = V_{it+1} H y
__global__ __launch_bounds__(KB) void kz_aug_ctl_kernel(KCtrlZ *k, const double *part, int nb)
{
    if (k->nosol) return;
    __shared__ double s[KPZ];
    kz_reduce(part, nb, 1, s);
    if (threadIdx.x != 0) return;
    k->ascale = 1.0 / sqrt(s[0]);
    const int it = k->it;
    for (int r = 0; r <= it; r++) {
        double2 acc = make_double2(0.0, 0.0);
        for (int j = 0; j < it; j++)
            if (r <= j + 1) acc = zadd(acc, zmul(ZSe(k, r, j), k->nrs[j]));
        k->hw[r] = acc;
    }
}

// the new approximation, normalised: az = (V H y)/||c||, z = c/||c||
template <int NV>
__global__ __launch_bounds__(KB) void kz_aug_kernel(long n, const double2 *V, long ld, double2 *z, double2 *az,
                                                   const KCtrlZ *k)
{
    if (k->nosol) return;
    const int nv = k->it;
    const double sc = k->ascale;
    double2 hw[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) hw[j] = j <= nv ? k->hw[j] : make_double2(0.0, 0.0);
    KLOOP(i)
    {
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < NV; j++)
            if (j <= nv) acc = zadd(acc, zmul(hw[j], V[(size_t)j * ld + i]));
        az[i] = make_double2(sc * acc.x, sc * acc.y);
        const double2 c = z[i];
        z[i] = make_double2(sc * c.x, sc * c.y);
    }
}

__global__ void kz_end_kernel(KCtrlZ *k)
{
    if (!k->open) return;
    if (k->its >= k->max_it && !k->reason) k->reason = K_DIVERGED_ITS;
    if (k->reason) {
        k->c.done = 1;
        k->c.final_iter = k->its;
        k->gate.done = 1;
    }
    k->open = 0;
    k->nosol = 1;
    k->noscale = 1;
    k->gate.done = 1;
}

__global__ __launch_bounds__(KB) void kz_final_kernel(KCtrlZ *k, const double *part, int nb)
{
    __shared__ double s[KPZ];
    kz_reduce(part, nb, 1, s);
    if (threadIdx.x != 0) return;
    k->relres = sqrt(s[0]);
}

// --- ILU(0) of the complex matrix (the complex variant's pc = 1) ---
__global__ __launch_bounds__(KB) void kz_norm_kernel(long n, const double2 *v, double *part, const int *skip)
{
    if (*skip) return;
    double acc[1] = {0.0};
    KLOOP(i) acc[0] += v[i].x * v[i].x + v[i].y * v[i].y;
    kz_block_sums<1>(acc, 1, part);
}

// ||B b||^2 and ||b||^2 with B b given
__global__ __launch_bounds__(KB) void kz_bnorm2_kernel(long n, const double2 *b, const double2 *zb, double *part)
{
    double acc[2] = {0.0, 0.0};
    KLOOP(i)
    {
        acc[0] += zb[i].x * zb[i].x + zb[i].y * zb[i].y;
        acc[1] += b[i].x * b[i].x + b[i].y * b[i].y;
    }
    kz_block_sums<2>(acc, 2, part);
}

__global__ __launch_bounds__(KB) void kz_conj_kernel(long n, double2 *v, const int *skip)
{
    if (*skip) return;
    KLOOP(i) v[i].y = -v[i].y;
}

// ||z||^2, <z, r~>
__global__ __launch_bounds__(KB) void kz_bi_norms_kernel(long n, const double2 *z, const double2 *rl, double *part,
                                                        const int *skip)
{
    if (*skip) return;
    double acc[3] = {0.0, 0.0, 0.0};
    KLOOP(i)
    {
        const double2 zi = z[i];
        const double2 pp = zmul(zi, zconj(rl[i]));
        acc[0] += zi.x * zi.x + zi.y * zi.y;
        acc[1] += pp.x;
        acc[2] += pp.y;
    }
    kz_block_sums<3>(acc, 3, part);
}

// --- BiCG of the complex variant: shadow sequence with A^H and B^H ---
__device__ __forceinline__ double2 zpch(const ZDiag &d, long i, double2 x)
{
    if (!d.precond) return x;
    const double dr = d.dlap + (d.c ? d.c[i] : d.c0);
    const double den = dr * dr + d.eta * d.eta;
    return zmul(zconj(make_double2(dr / den, -d.eta / den)), x);
}

// r = b - t, r~ = r, z = B r, z~ = B^H r~; ||z||^2, <z, r~>
__global__ __launch_bounds__(KB) void kz_bi_init_kernel(long n, const double2 *b, const double2 *t, double2 *r,
                                                       double2 *rl, double2 *z, double2 *zl, ZDiag dg, double *part)
{
    double acc[3] = {0.0, 0.0, 0.0};
    KLOOP(i)
    {
        const double2 ri = make_double2(b[i].x + (-1.0) * t[i].x, b[i].y + (-1.0) * t[i].y);
        const double2 zi = zpc(dg, i, ri), zli = zpch(dg, i, ri);
        r[i] = ri;
        rl[i] = ri;
        z[i] = zi;
        zl[i] = zli;
        const double2 p = zmul(zi, zconj(ri));
        acc[0] += zi.x * zi.x + zi.y * zi.y;
        acc[1] += p.x;
        acc[2] += p.y;
    }
    kz_block_sums<3>(acc, 3, part);
}

__global__ __launch_bounds__(KB) void kz_bi_p_kernel(long n, const double2 *z, const double2 *zl, double2 *p,
                                                    double2 *pl, const KCtrlZ *k)
{
    if (k->c.done) return;
    if (k->first) {
        KLOOP(i)
        {
            p[i] = z[i];
            pl[i] = zl[i];
        }
    } else {
        const double2 bc = k->bcoef, bcc = zconj(k->bcoef);
        KLOOP(i)
        {
            p[i] = zadd(z[i], zmul(bc, p[i]));
            pl[i] = zadd(zl[i], zmul(bcc, pl[i]));
        }
    }
}

// <q, p~> = sum q conj(p~)
__global__ __launch_bounds__(KB) void kz_bi_dot_kernel(long n, const double2 *q, const double2 *pl, double *part,
                                                      const KCtrlZ *k)
{
    if (k->c.done) return;
    double acc[2] = {0.0, 0.0};
    KLOOP(i)
    {
        const double2 v = zmul(q[i], zconj(pl[i]));
        acc[0] += v.x;
        acc[1] += v.y;
    }
    kz_block_sums<2>(acc, 2, part);
}

// x += a p; r -= a q; r~ -= conj(a) q~; z = B r; z~ = B^H r~; ||z||^2, <z, r~>
__global__ __launch_bounds__(KB) void kz_bi_upd_kernel(long n, double2 *x, double2 *r, double2 *rl, double2 *z,
                                                      double2 *zl, const double2 *p, const double2 *q,
                                                      const double2 *ql, ZDiag dg, double *part, const KCtrlZ *k)
{
    if (k->c.done) return;
    const double2 al = k->alpha, ma = make_double2(-k->alpha.x, -k->alpha.y), mac = zconj(ma);
    double acc[3] = {0.0, 0.0, 0.0};
    KLOOP(i)
    {
        x[i] = zadd(x[i], zmul(al, p[i]));
        const double2 ri = zadd(r[i], zmul(ma, q[i]));
        const double2 rli = zadd(rl[i], zmul(mac, ql[i]));
        const double2 zi = zpc(dg, i, ri), zli = zpch(dg, i, rli);
        r[i] = ri;
        rl[i] = rli;
        z[i] = zi;
        zl[i] = zli;
        const double2 pp = zmul(zi, zconj(rli));
        acc[0] += zi.x * zi.x + zi.y * zi.y;
        acc[1] += pp.x;
        acc[2] += pp.y;
    }
    kz_block_sums<3>(acc, 3, part);
}

__global__ __launch_bounds__(KB) void kz_bi_start_kernel(KCtrlZ *k, const double *part, int nb, double *hist)
{
    __shared__ double s[KPZ];
    kz_reduce(part, nb, 3, s);
    if (threadIdx.x != 0) return;
    const double dp = sqrt(s[0]);
    k->beta = make_double2(s[1], s[2]);
    k->first = 1;
    kz_push(k, hist, dp);
    int reason = kz_converged(k, 0, dp);
    if (!reason && k->beta.x == 0.0 && k->beta.y == 0.0) reason = K_DIVERGED_BREAKDOWN_BICG;
    k->reason = reason;
    if (reason) {
        k->c.done = 1;
        k->c.final_iter = k->its;
    }
}

__global__ __launch_bounds__(KB) void kz_bi_alpha_kernel(KCtrlZ *k, const double *part, int nb)
{
    if (k->c.done) return;
    __shared__ double s[KPZ];
    kz_reduce(part, nb, 2, s);
    if (threadIdx.x != 0) return;
    k->alpha = zdiv(k->beta, make_double2(s[0], s[1]));
}

__global__ __launch_bounds__(KB) void kz_bi_step_kernel(KCtrlZ *k, const double *part, int nb, double *hist)
{
    if (k->c.done) return;
    __shared__ double s[KPZ];
    kz_reduce(part, nb, 3, s);
    if (threadIdx.x != 0) return;
    const double dp = sqrt(s[0]);
    k->its++;
    kz_push(k, hist, dp);
    int reason = kz_converged(k, k->its, dp);
    if (!reason && k->its >= k->max_it) reason = K_DIVERGED_ITS;
    k->reason = reason;
    if (reason) {
        k->c.done = 1;
        k->c.final_iter = k->its;
        return;
    }
    k->betaold = k->beta;
    k->beta = make_double2(s[1], s[2]);
    k->bcoef = zdiv(k->beta, k->betaold);
    k->first = 0;
}

namespace {
template <int NV>
hipError_t launch_gz_step(const double2 *t, double2 *V, long ld, int nv, long n, int nb, ZDiag dg, KCtrlZ *k,
                          double *part)
{
    hipStream_t st = ctx().stream;
    hipLaunchKernelGGL((kz_mdot_kernel<NV>), dim3(nb), dim3(KB), 0, st, n, t, (const double2 *)V, ld, nv, dg, part,
                       (const int *)&k->gate.done);
    hipLaunchKernelGGL(kz_dots_kernel, dim3(1), dim3(KB), 0, st, k, (const double *)part, nb, nv);
    hipLaunchKernelGGL((kz_orth_kernel<NV>), dim3(nb), dim3(KB), 0, st, n, t, V, ld, nv, dg, (const KCtrlZ *)k, part);
    return hipGetLastError();
}
template <int NV>
hipError_t launch_gz_soln(double2 *x, const double2 *V, long ld, long n, int nb, KCtrlZ *k)
{
    hipLaunchKernelGGL((kz_soln_kernel<NV>), dim3(nb), dim3(KB), 0, ctx().stream, n, x, V, ld, (const KCtrlZ *)k);
    return hipGetLastError();
}
template <int NV>
hipError_t launch_gz_lsoln(double2 *x, const double2 *V, long ld, const ZAug &za, double2 *znew, long n, int nb,
                           KCtrlZ *k, double *part)
{
    hipLaunchKernelGGL((kz_lsoln_kernel<NV>), dim3(nb), dim3(KB), 0, ctx().stream, n, x, V, ld, za, znew,
                       (const KCtrlZ *)k, part);
    return hipGetLastError();
}
template <int NV>
hipError_t launch_gz_aug(const double2 *V, long ld, double2 *z, double2 *az, long n, int nb, KCtrlZ *k)
{
    hipLaunchKernelGGL((kz_aug_kernel<NV>), dim3(nb), dim3(KB), 0, ctx().stream, n, V, ld, z, az, (const KCtrlZ *)k);
    return hipGetLastError();
}
}  // namespace

static IluArgsT<double2> helm_ilu_args(const HOp &op, double2 *F)
{
    IluArgsT<double2> a{};
    a.nx = op.nx;
    a.ny = op.ny;
    a.nz = op.nz;
    for (int o = 0; o <= R; o++) a.w[o] = op.w[o];
    a.diag = op.c ? op.dlap : op.dlap + op.c0;   // (a*3*coeff[0] + c) + i eta
    a.eta = op.eta;
    a.c = op.c;
    a.F = F;
    return a;
}

// z = M^{-1} r (trans 0), M^{-T} r (1) or M^{-H} r (2) for the complex
// ILU(0) factors of op, factored here
int helm_ilu0_apply(const HOp &op, const double *r, double *z, int trans)
{
    if (op.fdn != R) return fail(AAR_HIP_ERR_ARG, "ILU(0) is built for the full FD12 stencil (fdn = 6)");
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!r || !z) return fail(AAR_HIP_ERR_ARG, "null vector");
    if (trans < 0 || trans > 2) return fail(AAR_HIP_ERR_ARG, "trans must be 0, 1 or 2");
    if (op.nx < 13 || op.ny < 13 || op.nz < 13)
        return fail(AAR_HIP_ERR_ARG, "the ILU(0) preconditioner needs every grid side >= 13");
    const long n = op.n;
    HC(ensure(C.kilu, sizeof(double2) * (size_t)n * IFS));
    HC(ensure(C.kctl, sizeof(KCtrlZ)));
    KCtrlZ *k = (KCtrlZ *)C.kctl.p;
    hipLaunchKernelGGL(k_zero_kernel, dim3(64), dim3(KB), 0, C.stream, (unsigned *)k,
                       (long)(sizeof(KCtrlZ) / sizeof(unsigned)));
    IluArgsT<double2> a = helm_ilu_args(op, (double2 *)C.kilu.p);
    HC(ilu_factor(a, C.stream));
    double2 *zz = (double2 *)z;
    const int nb = std::min(nblocks_for(n), 1024);
    HC(hipMemcpyAsync(zz, r, sizeof(double2) * n, hipMemcpyDeviceToDevice, C.stream));
    if (trans == 2) hipLaunchKernelGGL(kz_conj_kernel, dim3(nb), dim3(KB), 0, C.stream, n, zz, (const int *)&k->zero);
    HC(ilu_apply(a, (const double2 *)zz, zz, trans != 0, &k->zero, C.stream));
    if (trans == 2) hipLaunchKernelGGL(kz_conj_kernel, dim3(nb), dim3(KB), 0, C.stream, n, zz, (const int *)&k->zero);
    HC(hipGetLastError());
    HC(hipStreamSynchronize(C.stream));
    return 0;
}

// GMRES(restart) (aug = 0) and LGMRES(restart - aug, aug) of the complex
// variant.  LGMRES (PETSC_COMPLEX/main.c:120-147, KSPLGMRES; the algorithm
// of Baker, Jessup and Manteuffel with PETSc's defaults restart = 30,
// aug = 2, the space not held constant): each cycle takes restart - aug
// Arnoldi steps, then one step per stored error approximation z (newest
// first) whose vector is the stored B A z; every step is an iteration.
// The cycle's correction becomes the newest z, and B A z = V H y from the
// unrotated Hessenberg (no operator application), both scaled by 1/||z||.
// The approximations take aug + 1 slots: the new one is written to the free
// slot, and the oldest slot is freed once aug are stored.  The schedule
// (steps per cycle, slot of each order) depends only on the cycle count,
// so the host enqueues it ahead of the convergence flag.
static int helm_gmres_impl(const HOp &op, double *u, const double *b, int restart, int aug, int max_it, double tol,
                           int precond, aar_hip_stats *st)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!u || !b) return fail(AAR_HIP_ERR_ARG, "null vector");
    if (max_it < 0) return fail(AAR_HIP_ERR_ARG, "max_iter must be >= 0");
    if (restart < 1 || restart > KMAX) return fail(AAR_HIP_ERR_ARG, "restart must be in 1..32");
    if (aug < 0 || aug > ZAUG_MAX || aug >= restart)
        return fail(AAR_HIP_ERR_ARG, "aug must be in 0..min(8, restart - 1)");
    if (precond < 0 || precond > 2) return fail(AAR_HIP_ERR_ARG, "precond must be 0 (none), 1 (Jacobi) or 2 (ILU(0))");
    const bool ilu = precond == 2;
    if (ilu && op.fdn != R)
        return fail(AAR_HIP_ERR_ARG, "ILU(0) is built for the full FD12 stencil (fdn = 6)");
    if (ilu && (op.nx < 13 || op.ny < 13 || op.nz < 13))
        return fail(AAR_HIP_ERR_ARG, "the ILU(0) preconditioner needs every grid side >= 13");
    const long n = op.n;
    const long ld = ring_ld(n, sizeof(double2));
    // basis v_0 .. v_restart, t, then aug + 1 slots of z and of B A z
    const int nslot = aug ? aug + 1 : 0;
    HC(ensure(C.kvec, sizeof(double2) * (size_t)ld * (restart + 2 + 2 * nslot)));
    double2 *V = (double2 *)C.kvec.p, *t = V + (size_t)ld * (restart + 1);
    double2 *Zs = t + ld, *AZs = Zs + (size_t)ld * nslot;
    int ord[ZAUG_MAX + 1] = {0}, aug_ct = 0;   // ord[slot]: 1 = newest, 0 = free
    const int ita = restart - aug;             // Arnoldi steps per cycle
    const int nb = std::min(nblocks_for(n), 1024);
    const int hcap = max_it + max_it / ita + 8;
    HC(ensure(C.kctl, sizeof(KCtrlZ)));
    HC(ensure(C.hist0, sizeof(double) * (size_t)hcap));
    HC(ensure(C.part0, sizeof(double) * (size_t)nb * KPZ));
    KCtrlZ *k = (KCtrlZ *)C.kctl.p;
    double *part = (double *)C.part0.p, *hist = (double *)C.hist0.p;
    ZDiag dg;
    dg.c = op.c;
    dg.dlap = op.dlap;
    dg.c0 = op.c0;
    dg.eta = op.eta;
    dg.precond = precond == 1;
    ZDiag dg0 = dg;            // the augmentation steps' vectors are already B A z
    dg0.precond = 0;
    const double2 *bz = (const double2 *)b;
    double2 *x = (double2 *)u;
    C.bytes_acc = 0.0;
    HC(hipEventRecord(C.t0, C.stream));
    hipLaunchKernelGGL(k_zero_kernel, dim3(64), dim3(KB), 0, C.stream, (unsigned *)k,
                       (long)(sizeof(KCtrlZ) / sizeof(unsigned)));
    const double s = 16.0 * n, cb = op.c ? 8.0 * n : 0.0;
    IluArgsT<double2> za{};
    if (ilu) {
        HC(ensure(C.kilu, sizeof(double2) * (size_t)n * IFS));
        za = helm_ilu_args(op, (double2 *)C.kilu.p);
        HC(ilu_factor(za, C.stream));
        HC(ilu_apply(za, bz, t, false, &k->zero, C.stream));
        hipLaunchKernelGGL(kz_bnorm2_kernel, dim3(nb), dim3(KB), 0, C.stream, n, bz, (const double2 *)t, part);
        C.bytes_acc += 2 * ILU_APPLY_BYTES * n;
    } else {
        hipLaunchKernelGGL(kz_bnorm_kernel, dim3(nb), dim3(KB), 0, C.stream, n, bz, dg, part);
    }
    hipLaunchKernelGGL(kz_ctl_init_kernel, dim3(1), dim3(KB), 0, C.stream, k, tol, max_it, restart, hcap,
                       (const double *)part, nb, aug > 0 ? 1 : 0);
    HC(hipGetLastError());
    C.bytes_acc += s + cb;
    bool stopped = false;
    for (int cyc = 0; cyc <= max_it && !stopped; cyc++) {
        if (int rc = helm_matvec_async(op, u, (double *)t, &k->c)) return rc;
        hipLaunchKernelGGL(kz_resid_kernel, dim3(nb), dim3(KB), 0, C.stream, n, bz, (const double2 *)t, V, dg, part,
                           (const int *)&k->c.done);
        if (ilu) {
            HC(ilu_apply(za, (const double2 *)V, V, false, &k->c.done, C.stream));
            hipLaunchKernelGGL(kz_norm_kernel, dim3(nb), dim3(KB), 0, C.stream, n, (const double2 *)V, part,
                               (const int *)&k->c.done);
            C.bytes_acc += 2 * ILU_APPLY_BYTES * n;
        }
        // slot of the approximation of each order (1 = newest)
        int slot_of[ZAUG_MAX + 1] = {0};
        for (int q = 0; q < nslot; q++)
            if (ord[q]) slot_of[ord[q]] = q;
        const int itot = ita + aug_ct;
        hipLaunchKernelGGL(kz_start_kernel, dim3(1), dim3(KB), 0, C.stream, k, (const double *)part, nb, hist, itot);
        hipLaunchKernelGGL(kz_scale_kernel, dim3(nb), dim3(KB), 0, C.stream, n, V, (const KCtrlZ *)k, 0, 0);
        HC(hipGetLastError());
        C.bytes_acc += 5 * s + 2 * cb;
        for (int it = 0; it < itot; it++) {
            const int nv = it + 1;
            const double2 *w = t;
            const bool arnoldi = it < ita;
            if (arnoldi) {
                if (int rc = helm_matvec_async(op, (const double *)(V + (size_t)it * ld), (double *)t, &k->gate))
                    return rc;
                if (ilu) {
                    HC(ilu_apply(za, (const double2 *)t, V + (size_t)nv * ld, false, &k->gate.done, C.stream));
                    w = V + (size_t)nv * ld;
                    C.bytes_acc += 2 * ILU_APPLY_BYTES * n;
                }
            } else {
                w = AZs + (size_t)ld * slot_of[it - ita + 1];
            }
            const ZDiag &dw = (arnoldi && !ilu) ? dg : dg0;
            prof_begin("krylov_gmres_orth");
            hipError_t e = nv <= 8 ? launch_gz_step<8>(w, V, ld, nv, n, nb, dw, k, part)
                         : nv <= 16 ? launch_gz_step<16>(w, V, ld, nv, n, nb, dw, k, part)
                                    : launch_gz_step<32>(w, V, ld, nv, n, nb, dw, k, part);
            prof_end((2.0 * nv + 3) * s + 2 * cb);
            if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("gmres step: ") + hipGetErrorString(e));
            hipLaunchKernelGGL(kz_arnoldi_kernel, dim3(1), dim3(KB), 0, C.stream, k, (const double *)part, nb, hist);
            hipLaunchKernelGGL(kz_scale_kernel, dim3(nb), dim3(KB), 0, C.stream, n, V + (size_t)nv * ld,
                               (const KCtrlZ *)k, 1, nv);
            HC(hipGetLastError());
            C.bytes_acc += (2.0 * nv + 5) * s + 2 * cb;
        }
        hipLaunchKernelGGL(kz_soln_ctl_kernel, dim3(1), dim3(1), 0, C.stream, k);
        if (!aug) {
            hipError_t e = restart <= 8 ? launch_gz_soln<8>(x, V, ld, n, nb, k)
                         : restart <= 16 ? launch_gz_soln<16>(x, V, ld, n, nb, k)
                                         : launch_gz_soln<32>(x, V, ld, n, nb, k);
            if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("gmres solution: ") + hipGetErrorString(e));
            C.bytes_acc += (restart + 2.0) * s;
        } else {
            // the new approximation goes to the free slot
            int fs = 0;
            while (ord[fs]) fs++;
            double2 *zn = Zs + (size_t)ld * fs, *azn = AZs + (size_t)ld * fs;
            ZAug zq{};
            zq.ita = ita;
            zq.nz = aug_ct;
            for (int q = 0; q < aug_ct; q++) zq.z[q] = Zs + (size_t)ld * slot_of[q + 1];
            hipError_t e = ita <= 8 ? launch_gz_lsoln<8>(x, V, ld, zq, zn, n, nb, k, part)
                         : ita <= 16 ? launch_gz_lsoln<16>(x, V, ld, zq, zn, n, nb, k, part)
                                     : launch_gz_lsoln<32>(x, V, ld, zq, zn, n, nb, k, part);
            if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("lgmres solution: ") + hipGetErrorString(e));
            hipLaunchKernelGGL(kz_aug_ctl_kernel, dim3(1), dim3(KB), 0, C.stream, k, (const double *)part, nb);
            e = itot < 16 ? launch_gz_aug<16>((const double2 *)V, ld, zn, azn, n, nb, k)
                          : launch_gz_aug<KMAX + 1>((const double2 *)V, ld, zn, azn, n, nb, k);
            if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("lgmres approximation: ") + hipGetErrorString(e));
            C.bytes_acc += (itot + 3.0) * s + (itot + 4.0) * s;
            // the order of the stored approximations after this cycle
            if (aug_ct == aug) ord[slot_of[aug]] = 0;
            else aug_ct++;
            for (int q = 0; q < nslot; q++)
                if (ord[q]) ord[q]++;
            ord[fs] = 1;
        }
        hipLaunchKernelGGL(kz_end_kernel, dim3(1), dim3(1), 0, C.stream, k);
        HC(hipGetLastError());
        if (int rc = poll_record(&k->c, cyc)) return rc;
        if (poll_check(cyc - 2)) stopped = true;
    }
    // the true residual of the returned u (PETSC_COMPLEX/main.c:110-113)
    if (int rc = helm_matvec_async(op, u, (double *)t, nullptr)) return rc;
    hipLaunchKernelGGL(kz_resnorm_kernel, dim3(nb), dim3(KB), 0, C.stream, n, bz, (const double2 *)t, part);
    hipLaunchKernelGGL(kz_final_kernel, dim3(1), dim3(KB), 0, C.stream, k, (const double *)part, nb);
    HC(hipGetLastError());
    HC(hipEventRecord(C.t1, C.stream));
    KCtrlZ hk;
    HC(hipMemcpyAsync(&hk, k, sizeof(KCtrlZ), hipMemcpyDeviceToHost, C.stream));
    HC(hipStreamSynchronize(C.stream));
    float ms = 0.f;
    HC(hipEventElapsedTime(&ms, C.t0, C.t1));
    prof_flush();
    const int hl = std::min(hk.nh, hcap);
    C.hist_host.assign(std::max(hl, 0), 0.0);
    if (hl > 0) HC(hipMemcpy(C.hist_host.data(), hist, sizeof(double) * hl, hipMemcpyDeviceToHost));
    aar_hip_stats sres;
    std::memset(&sres, 0, sizeof(sres));
    sres.iterations = hk.its;
    sres.rhs_norm = hk.rhs_norm;
    sres.relres = hk.rhs_norm > 0 ? hk.relres / hk.rhs_norm : hk.relres;
    sres.converged = hk.reason > 0;
    sres.printed_converged = hk.reason > 0;
    sres.printed_warning = hk.reason == K_DIVERGED_ITS;
    sres.status = hk.reason == K_DIVERGED_NANORINF ? AAR_HIP_ERR_NONFINITE : 0;
    sres.hist_len = hl;
    sres.solve_ms = ms;
    sres.bytes = C.bytes_acc;
    sres.reason = hk.reason;
    C.last = sres;
    if (st) *st = sres;
    if (hk.reason == K_DIVERGED_NANORINF) return fail(AAR_HIP_ERR_NONFINITE, "residual is not finite");
    return 0;
}

int helm_gmres_solve(const HOp &op, double *u, const double *b, int restart, int max_it, double tol, int precond,
                     aar_hip_stats *st)
{
    return helm_gmres_impl(op, u, b, restart, 0, max_it, tol, precond, st);
}

int helm_lgmres_solve(const HOp &op, double *u, const double *b, int restart, int aug, int max_it, double tol,
                      int precond, aar_hip_stats *st)
{
    return helm_gmres_impl(op, u, b, restart, aug, max_it, tol, precond, st);
}

// BiCG of the complex PETSc variant (PETSC_COMPLEX/main.c:150-175)
int helm_bicg_solve(const HOp &op, double *u, const double *b, int max_it, double tol, int precond,
                    aar_hip_stats *st)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!u || !b) return fail(AAR_HIP_ERR_ARG, "null vector");
    if (max_it < 0) return fail(AAR_HIP_ERR_ARG, "max_iter must be >= 0");
    if (precond < 0 || precond > 2) return fail(AAR_HIP_ERR_ARG, "precond must be 0 (none), 1 (Jacobi) or 2 (ILU(0))");
    const bool ilu = precond == 2;
    if (ilu && op.fdn != R)
        return fail(AAR_HIP_ERR_ARG, "ILU(0) is built for the full FD12 stencil (fdn = 6)");
    if (ilu && (op.nx < 13 || op.ny < 13 || op.nz < 13))
        return fail(AAR_HIP_ERR_ARG, "the ILU(0) preconditioner needs every grid side >= 13");
    const long n = op.n;
    const long ld = ring_ld(n, sizeof(double2));
    HC(ensure(C.kvec, sizeof(double2) * (size_t)ld * 8));
    double2 *r = (double2 *)C.kvec.p, *rl = r + ld, *z = r + 2 * ld, *zl = r + 3 * ld;
    double2 *p = r + 4 * ld, *pl = r + 5 * ld, *q = r + 6 * ld, *ql = r + 7 * ld;
    const int nb = std::min(nblocks_for(n), 1024);
    const int hcap = max_it + 8;
    HC(ensure(C.kctl, sizeof(KCtrlZ)));
    HC(ensure(C.hist0, sizeof(double) * (size_t)hcap));
    HC(ensure(C.part0, sizeof(double) * (size_t)nb * KPZ));
    KCtrlZ *k = (KCtrlZ *)C.kctl.p;
    double *part = (double *)C.part0.p, *hist = (double *)C.hist0.p;
    ZDiag dg;
    dg.c = op.c;
    dg.dlap = op.dlap;
    dg.c0 = op.c0;
    dg.eta = op.eta;
    dg.precond = precond == 1;
    HOp oph = op;                 // A^H: eta -> -eta (the operator is complex symmetric)
    oph.eta = -op.eta;
    const double2 *bz = (const double2 *)b;
    double2 *x = (double2 *)u;
    C.bytes_acc = 0.0;
    HC(hipEventRecord(C.t0, C.stream));
    hipLaunchKernelGGL(k_zero_kernel, dim3(64), dim3(KB), 0, C.stream, (unsigned *)k,
                       (long)(sizeof(KCtrlZ) / sizeof(unsigned)));
    IluArgsT<double2> za{};
    // z = B r, z~ = B^H r~ = conj(M^{-T} conj(r~)) with ILU(0); the
    // pointwise kernels then leave z = r, z~ = r~ for these sweeps
    auto ilu_z = [&](const int *skip) -> int {
        HC(ilu_apply(za, (const double2 *)z, z, false, skip, C.stream));
        hipLaunchKernelGGL(kz_conj_kernel, dim3(nb), dim3(KB), 0, C.stream, n, zl, skip);
        HC(ilu_apply(za, (const double2 *)zl, zl, true, skip, C.stream));
        hipLaunchKernelGGL(kz_conj_kernel, dim3(nb), dim3(KB), 0, C.stream, n, zl, skip);
        hipLaunchKernelGGL(kz_bi_norms_kernel, dim3(nb), dim3(KB), 0, C.stream, n, (const double2 *)z,
                           (const double2 *)rl, part, skip);
        HC(hipGetLastError());
        C.bytes_acc += 4 * ILU_APPLY_BYTES * n;
        return 0;
    };
    if (ilu) {
        HC(ensure(C.kilu, sizeof(double2) * (size_t)n * IFS));
        za = helm_ilu_args(op, (double2 *)C.kilu.p);
        HC(ilu_factor(za, C.stream));
        HC(ilu_apply(za, bz, z, false, &k->zero, C.stream));
        hipLaunchKernelGGL(kz_bnorm2_kernel, dim3(nb), dim3(KB), 0, C.stream, n, bz, (const double2 *)z, part);
    } else {
        hipLaunchKernelGGL(kz_bnorm_kernel, dim3(nb), dim3(KB), 0, C.stream, n, bz, dg, part);
    }
    hipLaunchKernelGGL(kz_ctl_init_kernel, dim3(1), dim3(KB), 0, C.stream, k, tol, max_it, 0, hcap,
                       (const double *)part, nb, 0);
    HC(hipGetLastError());
    const double s = 16.0 * n, cb = op.c ? 8.0 * n : 0.0;
    if (int rc = helm_matvec_async(op, u, (double *)q, nullptr)) return rc;
    hipLaunchKernelGGL(kz_bi_init_kernel, dim3(nb), dim3(KB), 0, C.stream, n, bz, (const double2 *)q, r, rl, z, zl,
                       dg, part);
    if (ilu)
        if (int rc = ilu_z(&k->zero)) return rc;
    hipLaunchKernelGGL(kz_bi_start_kernel, dim3(1), dim3(KB), 0, C.stream, k, (const double *)part, nb, hist);
    HC(hipGetLastError());
    C.bytes_acc += 7 * s + 2 * cb;
    bool stopped = false;
    const int loops = std::max(max_it, 1);
    int cyc = 0;
    for (int i = 0; i < loops && !stopped; i++) {
        hipLaunchKernelGGL(kz_bi_p_kernel, dim3(nb), dim3(KB), 0, C.stream, n, (const double2 *)z,
                           (const double2 *)zl, p, pl, (const KCtrlZ *)k);
        if (int rc = helm_matvec_async(op, (const double *)p, (double *)q, &k->c)) return rc;
        if (int rc = helm_matvec_async(oph, (const double *)pl, (double *)ql, &k->c)) return rc;
        prof_begin("krylov_bicg_update");
        hipLaunchKernelGGL(kz_bi_dot_kernel, dim3(nb), dim3(KB), 0, C.stream, n, (const double2 *)q,
                           (const double2 *)pl, part, (const KCtrlZ *)k);
        hipLaunchKernelGGL(kz_bi_alpha_kernel, dim3(1), dim3(KB), 0, C.stream, k, (const double *)part, nb);
        hipLaunchKernelGGL(kz_bi_upd_kernel, dim3(nb), dim3(KB), 0, C.stream, n, x, r, rl, z, zl,
                           (const double2 *)p, (const double2 *)q, (const double2 *)ql, dg, part, (const KCtrlZ *)k);
        if (ilu)
            if (int rc = ilu_z(&k->c.done)) return rc;
        hipLaunchKernelGGL(kz_bi_step_kernel, dim3(1), dim3(KB), 0, C.stream, k, (const double *)part, nb, hist);
        prof_end(16 * s + 2 * cb);
        HC(hipGetLastError());
        C.bytes_acc += 22 * s + 4 * cb;
        if ((i + 1) % 8 == 0) {
            if (int rc = poll_record(&k->c, cyc)) return rc;
            if (poll_check(cyc - 2)) stopped = true;
            cyc++;
        }
    }
    if (int rc = helm_matvec_async(op, u, (double *)q, nullptr)) return rc;
    hipLaunchKernelGGL(kz_resnorm_kernel, dim3(nb), dim3(KB), 0, C.stream, n, bz, (const double2 *)q, part);
    hipLaunchKernelGGL(kz_final_kernel, dim3(1), dim3(KB), 0, C.stream, k, (const double *)part, nb);
    HC(hipGetLastError());
    HC(hipEventRecord(C.t1, C.stream));
    KCtrlZ hk;
    HC(hipMemcpyAsync(&hk, k, sizeof(KCtrlZ), hipMemcpyDeviceToHost, C.stream));
    HC(hipStreamSynchronize(C.stream));
    float ms = 0.f;
    HC(hipEventElapsedTime(&ms, C.t0, C.t1));
    prof_flush();
    const int hl = std::min(hk.nh, hcap);
    C.hist_host.assign(std::max(hl, 0), 0.0);
    if (hl > 0) HC(hipMemcpy(C.hist_host.data(), hist, sizeof(double) * hl, hipMemcpyDeviceToHost));
    aar_hip_stats sres;
    std::memset(&sres, 0, sizeof(sres));
    sres.iterations = hk.its;
    sres.rhs_norm = hk.rhs_norm;
    sres.relres = hk.rhs_norm > 0 ? hk.relres / hk.rhs_norm : hk.relres;
    sres.converged = hk.reason > 0;
    sres.printed_converged = hk.reason > 0;
    sres.printed_warning = hk.reason == K_DIVERGED_ITS;
    sres.status = hk.reason == K_DIVERGED_NANORINF ? AAR_HIP_ERR_NONFINITE : 0;
    sres.hist_len = hl;
    sres.solve_ms = ms;
    sres.bytes = C.bytes_acc;
    sres.reason = hk.reason;
    C.last = sres;
    if (st) *st = sres;
    if (hk.reason == K_DIVERGED_NANORINF) return fail(AAR_HIP_ERR_NONFINITE, "residual is not finite");
    return 0;
}

}  // namespace aar
