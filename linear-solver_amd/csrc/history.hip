// history.hip -- Anderson / Galerkin / L2 history kernels of libaar_hip.
//
// The reference builds its normal equations with m(m+1)/2 + m (AAR, PL2R)
// or m^2 + m (PGR) separate N-length dot passes over stored difference
// columns (MPI_REAL/AAR.c:192-210, PGR.c:188-204, PL2R.c:187-204) and then
// re-reads 2m columns for the update (AAR.c:225-231, PGR.c:117-121).
// Here the iterates x_k and residuals f_k themselves live in ring buffers
// of m+1 slots, difference columns are formed on the fly with the same
// subtraction the reference stores (DX = x - x_old, DF = f - f_old), and
// the whole normal-equation build is ONE pass over the m+1 slots with all
// products accumulated in FP64 registers.  The m x m solve runs on device.
#include "aar_common.h"
#include "aar_device.h"
#include "aar_kernels.h"

namespace aar {

template <typename T> __device__ inline T tsub(T a, T b) { return a - b; }

// ---------------------------------------------------------------------------
// normal equations, one pass
// ---------------------------------------------------------------------------
template <typename T, int MM, int KIND>
__global__ __launch_bounds__(BLOCK) void gram_kernel(const GramArgs<T> a)
{
    constexpr int NG = (KIND == 1) ? MM * MM : MM * (MM + 1) / 2;
    constexpr int NA = NG + MM;
    __shared__ double scratch[NA * (BLOCK / 64)];
    __shared__ double sums[NA];
    if (a.ctrl->done) return;

    const int m = a.m;
    const T *pa[MM], *pb[MM], *qa[MM], *qb[MM];
    bool ok[MM];
#pragma unroll
    for (int j = 0; j < MM; j++) {
        ok[j] = (j < m) && a.valid[j];
        const int sa = ok[j] ? a.sa[j] : 0, sb = ok[j] ? a.sb[j] : 0;
        pa[j] = a.F + (size_t)sa * a.ld;
        pb[j] = a.F + (size_t)sb * a.ld;
        if constexpr (KIND != 0) {
            qa[j] = a.Y + (size_t)sa * a.ld;
            qb[j] = a.Y + (size_t)sb * a.ld;
        }
    }
    const T *pf = (KIND == 0 ? a.F : a.Y) + (size_t)a.fslot * a.ld;

    double acc[NA];
#pragma unroll
    for (int k = 0; k < NA; k++) acc[k] = 0.0;

    const long chunk = (a.n + a.nblocks - 1) / a.nblocks;
    const long beg = (long)blockIdx.x * chunk;
    const long end = min(a.n, beg + chunk);
    for (long i = beg + threadIdx.x; i < end; i += BLOCK) {
        T fv;
        if constexpr (KIND == 0) fv = pf[i];
        else fv = tsub(a.b[i], pf[i]);                     // PGR.c:102-103: f = rhs - Ax
        double d[MM], e[MM];
#pragma unroll
        for (int j = 0; j < MM; j++) {
            d[j] = 0.0;
            e[j] = 0.0;
            if (ok[j]) {
                if constexpr (KIND == 0) {
                    d[j] = (double)tsub(pa[j][i], pb[j][i]);   // DF = f - f_old (AAR.c:100)
                } else if constexpr (KIND == 1) {
                    e[j] = (double)tsub(pa[j][i], pb[j][i]);   // DX = x - x_old (PGR.c:107)
                    d[j] = (double)tsub(qa[j][i], qb[j][i]);   // DF = Ax - Ax_old (PGR.c:108)
                } else {
                    d[j] = (double)tsub(qa[j][i], qb[j][i]);
                }
            }
        }
        const double fd = (double)fv;
        if constexpr (KIND == 1) {
#pragma unroll
            for (int j = 0; j < MM; j++) {
#pragma unroll
                for (int q = 0; q < MM; q++) acc[j * MM + q] += e[q] * d[j];   // DX_q . DF_j
                acc[NG + j] += e[j] * fd;
            }
        } else {
            int t = 0;
#pragma unroll
            for (int j = 0; j < MM; j++) {
#pragma unroll
                for (int q = 0; q <= j; q++) acc[t++] += d[q] * d[j];
                acc[NG + j] += d[j] * fd;
            }
        }
    }

    block_reduce_to<NA>(acc, a.partials + (size_t)blockIdx.x * NA, scratch);
    if (last_block_arrive(a.counter, gridDim.x)) {
        // fixed-order sum over blocks, one entry per thread group
        for (int k = threadIdx.x; k < NA; k += BLOCK) {
            double s = 0.0;
            for (int bidx = 0; bidx < a.nblocks; bidx++) s += a.partials[(size_t)bidx * NA + k];
            sums[k] = s;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            // reference allredvec layout (AAR.c:192-212, PGR.c:188-206)
            double *G = a.ctrl->G;
            if constexpr (KIND == 1) {
                for (int j = 0; j < m; j++)
                    for (int q = 0; q < m; q++) G[j * m + q] = sums[j * MM + q];
            } else {
                for (int j = 0; j < m; j++)
                    for (int q = 0; q <= j; q++) {
                        const double v = sums[j * (j + 1) / 2 + q];
                        G[j * m + q] = v;
                        G[q * m + j] = v;
                    }
            }
            for (int j = 0; j < m; j++) G[m * m + j] = sums[NG + j];
            *a.counter = 0u;
        }
    }
}

template <typename T>
hipError_t launch_gram(const GramArgs<T> &a, hipStream_t st)
{
    if (a.m < 1 || a.m > MAXM) return hipErrorInvalidValue;
    dim3 grid(a.nblocks), block(BLOCK);
    switch (a.kind) {
    case 0: hipLaunchKernelGGL((gram_kernel<T, MAXM, 0>), grid, block, 0, st, a); break;
    case 1: hipLaunchKernelGGL((gram_kernel<T, MAXM, 1>), grid, block, 0, st, a); break;
    case 2: hipLaunchKernelGGL((gram_kernel<T, MAXM, 2>), grid, block, 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t launch_gram<double>(const GramArgs<double> &, hipStream_t);
template hipError_t launch_gram<float>(const GramArgs<float> &, hipStream_t);

// ---------------------------------------------------------------------------
// m x m pseudo-inverse solve: Golub-Reinsch SVD as in MPI_REAL/tools.c,
// register-resident with compile-time indices (loops run to the bound M
// and are predicated on the runtime size n), executed by one lane.  The
// operation sequence is the reference's, so for the same input matrix the
// result is bit-identical.  U is never formed: the reference replaces it
// by V (tools.c:324-329) and no U-only step feeds w, the superdiagonal or
// V.  The left reflector is applied to columns 1..n-1 (tools.c:137).
// ---------------------------------------------------------------------------
__device__ inline double d_pythag(double a, double b)
{
    const double absa = fabs(a), absb = fabs(b);
    if (absa > absb) return absa * sqrt(1.0 + absb * absb / (absa * absa));
    return absb == 0.0 ? 0.0 : absb * sqrt(1.0 + absa * absa / (absb * absb));
}
__device__ inline double d_sign(double mag, double s) { return s >= 0.0 ? fabs(mag) : -fabs(mag); }

template <int M>
__device__ int gr_pinv(double (&a)[M][M], const int n, const double (&bv)[M], double (&xv)[M],
                       double (&w)[M], double (&v)[M][M])
{
    double rv1[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
        rv1[i] = 0.0;
        w[i] = 0.0;
#pragma unroll
        for (int j = 0; j < M; j++) v[i][j] = 0.0;
    }
    double g = 0.0, scale = 0.0, anorm = 0.0, s = 0.0, f = 0.0, h = 0.0;

#pragma unroll
    for (int i = 0; i < M; i++) {
        if (i < n) {
            const int l = i + 1;
            rv1[i] = scale * g;
            g = s = scale = 0.0;
#pragma unroll
            for (int k = i; k < M; k++)
                if (k < n) scale += fabs(a[k][i]);
            if (scale != 0.0) {
#pragma unroll
                for (int k = i; k < M; k++)
                    if (k < n) {
                        a[k][i] /= scale;
                        s += a[k][i] * a[k][i];
                    }
                f = a[i][i];
                g = -d_sign(sqrt(s), f);
                h = f * g - s;
                a[i][i] = f - g;
#pragma unroll
                for (int j = 1; j < M; j++) {
                    if (j < n) {
                        double t = 0.0;
#pragma unroll
                        for (int k = i; k < M; k++)
                            if (k < n) t += a[k][i] * a[k][j];
                        const double fac = t / h;
#pragma unroll
                        for (int k = i; k < M; k++)
                            if (k < n) a[k][j] += fac * a[k][i];
                    }
                }
#pragma unroll
                for (int k = i; k < M; k++)
                    if (k < n) a[k][i] *= scale;
            }
            w[i] = scale * g;
            g = s = scale = 0.0;
            if (i != n - 1) {
#pragma unroll
                for (int k = l; k < M; k++)
                    if (k < n) scale += fabs(a[i][k]);
                if (scale != 0.0) {
#pragma unroll
                    for (int k = l; k < M; k++)
                        if (k < n) {
                            a[i][k] /= scale;
                            s += a[i][k] * a[i][k];
                        }
                    f = a[i][l < M ? l : M - 1];
                    g = -d_sign(sqrt(s), f);
                    h = f * g - s;
                    a[i][l < M ? l : M - 1] = f - g;
#pragma unroll
                    for (int k = l; k < M; k++)
                        if (k < n) rv1[k] = a[i][k] / h;
#pragma unroll
                    for (int j = l; j < M; j++) {
                        if (j < n) {
                            double t = 0.0;
#pragma unroll
                            for (int k = l; k < M; k++)
                                if (k < n) t += a[j][k] * a[i][k];
#pragma unroll
                            for (int k = l; k < M; k++)
                                if (k < n) a[j][k] += t * rv1[k];
                        }
                    }
#pragma unroll
                    for (int k = l; k < M; k++)
                        if (k < n) a[i][k] *= scale;
                }
            }
            {
                const double t = fabs(w[i]) + fabs(rv1[i]);
                anorm = anorm > t ? anorm : t;       // tools.h:21 max()
            }
        }
    }

    // accumulation of right-hand transforms (for i < n-1: l = i+1, g = rv1[i+1])
#pragma unroll
    for (int i = M - 1; i >= 0; i--) {
        if (i < n) {
            if (i < n - 1) {
                const int l = i + 1;
                const double gg = rv1[l < M ? l : M - 1];
                if (gg != 0.0) {
                    const double ail = a[i][l < M ? l : M - 1];
#pragma unroll
                    for (int j = l; j < M; j++)
                        if (j < n) v[j][i] = (a[i][j] / ail) / gg;
#pragma unroll
                    for (int j = l; j < M; j++) {
                        if (j < n) {
                            double t = 0.0;
#pragma unroll
                            for (int k = l; k < M; k++)
                                if (k < n) t += a[i][k] * v[k][j];
#pragma unroll
                            for (int k = l; k < M; k++)
                                if (k < n) v[k][j] += t * v[k][i];
                        }
                    }
                }
#pragma unroll
                for (int j = l; j < M; j++)
                    if (j < n) {
                        v[i][j] = 0.0;
                        v[j][i] = 0.0;
                    }
            }
            v[i][i] = 1.0;
        }
    }

    // implicit-shift QR on the bidiagonal
    int status = 0;
#pragma unroll
    for (int k = M - 1; k >= 0; k--) {
        if (k < n && !status) {
            for (int its = 0; its <= 250; its++) {
                int flag = 1, l = 0;
                bool found = false;
#pragma unroll
                for (int ll = k; ll >= 0; ll--) {
                    if (!found) {
                        if ((double)(fabs(rv1[ll]) + anorm) == anorm) {
                            flag = 0;
                            l = ll;
                            found = true;
                        } else if (ll >= 1 && (double)(fabs(w[ll >= 1 ? ll - 1 : 0]) + anorm) == anorm) {
                            l = ll;
                            found = true;
                        }
                    }
                }
                if (flag) {
                    double c = 0.0;
                    s = 1.0;
                    bool stop = false;
#pragma unroll
                    for (int i = 0; i <= k; i++) {
                        if (i >= l && !stop) {
                            f = s * rv1[i];
                            rv1[i] = c * rv1[i];
                            if ((double)(fabs(f) + anorm) == anorm) {
                                stop = true;
                            } else {
                                g = w[i];
                                h = d_pythag(f, g);
                                w[i] = h;
                                h = 1.0 / h;
                                c = g * h;
                                s = -f * h;
                            }
                        }
                    }
                }
                double z = w[k];
                if (l == k) {
                    if (z < 0.0) {
                        w[k] = -z;
#pragma unroll
                        for (int j = 0; j < M; j++)
                            if (j < n) v[j][k] = -v[j][k];
                    }
                    break;
                }
                if (its == 250) {
                    status = 1;
                    break;
                }
                double x = 0.0;
#pragma unroll
                for (int ll = 0; ll <= k; ll++)
                    if (ll == l) x = w[ll];
                const int nm = k - 1;
                double y = w[nm >= 0 ? nm : 0];
                g = rv1[nm >= 0 ? nm : 0];
                h = rv1[k];
                f = ((y - z) * (y + z) + (g - h) * (g + h)) / (2.0 * h * y);
                g = d_pythag(f, 1.0);
                f = ((x - z) * (x + z) + h * ((y / (f + d_sign(g, f))) - h)) / x;
                double c = 1.0;
                s = 1.0;
#pragma unroll
                for (int j = 0; j < k; j++) {
                    if (j >= l) {
                        const int i = j + 1;
                        g = rv1[i];
                        y = w[i];
                        h = s * g;
                        g = c * g;
                        z = d_pythag(f, h);
                        rv1[j] = z;
                        c = f / z;
                        s = h / z;
                        f = x * c + g * s;
                        g = g * c - x * s;
                        h = y * s;
                        y *= c;
#pragma unroll
                        for (int jj = 0; jj < M; jj++) {
                            if (jj < n) {
                                const double vx = v[jj][j], vz = v[jj][i];
                                v[jj][j] = vx * c + vz * s;
                                v[jj][i] = vz * c - vx * s;
                            }
                        }
                        z = d_pythag(f, h);
                        w[j] = z;
                        if (z != 0.0) {
                            z = 1.0 / z;
                            c = f * z;
                            s = h * z;
                        }
                        f = c * g + s * y;
                        x = c * y - s * g;
                    }
                }
#pragma unroll
                for (int ll = 0; ll <= k; ll++)
                    if (ll == l) rv1[ll] = 0.0;
                rv1[k] = f;
                w[k] = x;
            }
        }
    }

    // truncation (tools.c:331-335) and x = V diag(1/w) V^T b (tools.c:76-94)
    double wmax = 0.0;
#pragma unroll
    for (int j = 0; j < M; j++)
        if (j < n && w[j] > wmax) wmax = w[j];
    const double wmin = n * wmax * (2.22044605e-16);
#pragma unroll
    for (int j = 0; j < M; j++)
        if (j < n && w[j] < wmin) w[j] = 0.0;
    double tmp[M];
#pragma unroll
    for (int j = 0; j < M; j++) {
        double t = 0.0;
        if (j < n && w[j] != 0.0) {
#pragma unroll
            for (int i = 0; i < M; i++)
                if (i < n) t += v[i][j] * bv[i];
            t /= w[j];
        }
        tmp[j] = t;
    }
#pragma unroll
    for (int j = 0; j < M; j++) {
        double t = 0.0;
#pragma unroll
        for (int jj = 0; jj < M; jj++)
            if (jj < n) t += v[j][jj] * tmp[jj];
        xv[j] = t;
    }
    return status;
}

template <int M>
__global__ __launch_bounds__(64) void ls_kernel(Ctrl *c, int n)
{
    if (c->done || threadIdx.x != 0) return;
    double a[M][M], b[M], x[M], w[M], v[M][M];
    // FtF[i][j] = allredvec[j*m + i], Ftf[j] = allredvec[m*m + j] (AAR.c:214-221)
#pragma unroll
    for (int i = 0; i < M; i++) {
#pragma unroll
        for (int j = 0; j < M; j++) a[i][j] = (i < n && j < n) ? c->G[j * n + i] : 0.0;
        b[i] = i < n ? c->G[n * n + i] : 0.0;
    }
    const int st = gr_pinv<M>(a, n, b, x, w, v);
#pragma unroll
    for (int j = 0; j < M; j++) {
        if (j < n) {
            c->s[j] = x[j];
            c->w[j] = w[j];
#pragma unroll
            for (int q = 0; q < M; q++)
                if (q < n) c->V[j * n + q] = v[j][q];
        }
    }
    if (st) {
        // tools.c:270 exits the process; here the solve stops with a status
        c->status = 1;
        c->done = 1;
    }
}

hipError_t launch_ls(Ctrl *ctrl, int n, int /*complex_embed*/, hipStream_t st)
{
    if (n < 1 || n > MAXM_LS) return hipErrorInvalidValue;
    if (n <= 8) hipLaunchKernelGGL((ls_kernel<8>), dim3(1), dim3(64), 0, st, ctrl, n);
    else hipLaunchKernelGGL((ls_kernel<16>), dim3(1), dim3(64), 0, st, ctrl, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Anderson update (AAR.c:225-231 then AAR.c:116-117)
// ---------------------------------------------------------------------------
template <typename T, int MM>
__global__ __launch_bounds__(BLOCK) void aar_update_kernel(const UpdArgs<T> a)
{
    if (a.ctrl->done) return;
    const T *xa[MM], *xb[MM], *fa[MM], *fb[MM];
    bool ok[MM];
    double sv[MM];
#pragma unroll
    for (int j = 0; j < MM; j++) {
        ok[j] = (j < a.m) && a.valid[j];
        const int sa = ok[j] ? a.sa[j] : 0, sb = ok[j] ? a.sb[j] : 0;
        xa[j] = a.X + (size_t)sa * a.ld;
        xb[j] = a.X + (size_t)sb * a.ld;
        fa[j] = a.F + (size_t)sa * a.ld;
        fb[j] = a.F + (size_t)sb * a.ld;
        sv[j] = j < a.m ? a.ctrl->s[j] : 0.0;
    }
    const T *xk = a.X + (size_t)a.xslot * a.ld;
    const T *fk = a.F + (size_t)a.fslot * a.ld;
    T *xo = const_cast<T *>(a.X) + (size_t)a.oslot * a.ld;
    const T beta = (T)a.beta;
    const long stride = (long)gridDim.x * BLOCK;
    for (long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < a.n; i += stride) {
        double am = 0.0;
#pragma unroll
        for (int j = 0; j < MM; j++) {
            if (j < a.m) {
                T dx = (T)0, df = (T)0;
                if (ok[j]) {
                    dx = xa[j][i] - xb[j][i];
                    df = fa[j][i] - fb[j][i];
                }
                am = am + (double)(dx + beta * df) * sv[j];
            }
        }
        const T xb0 = xk[i] + beta * fk[i];
        xo[i] = (T)((double)xb0 - am);
    }
}

template <typename T>
hipError_t launch_aar_update(const UpdArgs<T> &a, hipStream_t st)
{
    if (a.m < 1 || a.m > MAXM) return hipErrorInvalidValue;
    long nb = (a.n + BLOCK - 1) / BLOCK;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL((aar_update_kernel<T, MAXM>), dim3((unsigned)nb), dim3(BLOCK), 0, st, a);
    return hipGetLastError();
}
template hipError_t launch_aar_update<double>(const UpdArgs<double> &, hipStream_t);
template hipError_t launch_aar_update<float>(const UpdArgs<float> &, hipStream_t);

// ---------------------------------------------------------------------------
// PGR / PL2R projection update (PGR.c:117-133)
// ---------------------------------------------------------------------------
template <typename T, int MM>
__global__ __launch_bounds__(BLOCK) void pgr_update_kernel(const PUpdArgs<T> a)
{
    __shared__ double scratch[BLOCK / 64];
    __shared__ double sums[1];
    if (a.ctrl->done) return;
    const T *xa[MM], *xb[MM], *ya[MM], *yb[MM];
    bool ok[MM];
    T sv[MM];
#pragma unroll
    for (int j = 0; j < MM; j++) {
        ok[j] = (j < a.m) && a.valid[j];
        const int sa = ok[j] ? a.sa[j] : 0, sb = ok[j] ? a.sb[j] : 0;
        xa[j] = a.X + (size_t)sa * a.ld;
        xb[j] = a.X + (size_t)sb * a.ld;
        ya[j] = a.Y + (size_t)sa * a.ld;
        yb[j] = a.Y + (size_t)sb * a.ld;
        sv[j] = j < a.m ? (T)a.ctrl->s[j] : (T)0;
    }
    T *xc = a.X + (size_t)a.cur * a.ld;
    T *yc = a.Y + (size_t)a.cur * a.ld;
    double acc[1] = {0.0};
    const long chunk = (a.n + a.nblocks - 1) / a.nblocks;
    const long beg = (long)blockIdx.x * chunk, end = min(a.n, beg + chunk);
    for (long i = beg + threadIdx.x; i < end; i += BLOCK) {
        T dx[MM], dy[MM];
#pragma unroll
        for (int j = 0; j < MM; j++) {
            dx[j] = (T)0;
            dy[j] = (T)0;
            if (ok[j]) {
                dx[j] = xa[j][i] - xb[j][i];
                dy[j] = ya[j][i] - yb[j][i];
            }
        }
        T xv = xc[i], yv = yc[i];
#pragma unroll
        for (int j = 0; j < MM; j++) {
            if (j < a.m) {
                xv = xv + dx[j] * sv[j];
                yv = yv + dy[j] * sv[j];
            }
        }
        xc[i] = xv;
        yc[i] = yv;
        const T r = a.b[i] - yv;                 // PGR.c:129-130
        acc[0] += (double)r * (double)r;
    }
    block_reduce_to<1>(acc, a.partials + blockIdx.x, scratch);
    if (last_block_arrive(a.counter, gridDim.x)) {
        final_sum(a.partials, a.nblocks, 1, sums, scratch);
        if (threadIdx.x == 0) {
            finalize(a.fin, a.ctrl, sums, 1);
            *a.counter = 0u;
        }
    }
}

template <typename T>
hipError_t launch_pgr_update(const PUpdArgs<T> &a, hipStream_t st)
{
    if (a.m < 1 || a.m > MAXM) return hipErrorInvalidValue;
    hipLaunchKernelGGL((pgr_update_kernel<T, MAXM>), dim3(a.nblocks), dim3(BLOCK), 0, st, a);
    return hipGetLastError();
}
template hipError_t launch_pgr_update<double>(const PUpdArgs<double> &, hipStream_t);
template hipError_t launch_pgr_update<float>(const PUpdArgs<float> &, hipStream_t);

// ---------------------------------------------------------------------------
// dots and CG pointwise step
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(BLOCK) void dot_kernel(const DotArgs<T> a)
{
    __shared__ double scratch[BLOCK / 64];
    __shared__ double sums[1];
    if (a.ctrl->done) return;
    double acc[1] = {0.0};
    const long chunk = (a.n + a.nblocks - 1) / a.nblocks;
    const long beg = (long)blockIdx.x * chunk, end = min(a.n, beg + chunk);
    const T *w = a.w ? a.w : a.v;
    for (long i = beg + threadIdx.x; i < end; i += BLOCK) acc[0] += (double)a.v[i] * (double)w[i];
    block_reduce_to<1>(acc, a.partials + blockIdx.x, scratch);
    if (last_block_arrive(a.counter, gridDim.x)) {
        final_sum(a.partials, a.nblocks, 1, sums, scratch);
        if (threadIdx.x == 0) {
            finalize(a.fin, a.ctrl, sums, 1);
            *a.counter = 0u;
        }
    }
}

template <typename T>
hipError_t launch_dot(const DotArgs<T> &a, hipStream_t st)
{
    hipLaunchKernelGGL((dot_kernel<T>), dim3(a.nblocks), dim3(BLOCK), 0, st, a);
    return hipGetLastError();
}
template hipError_t launch_dot<double>(const DotArgs<double> &, hipStream_t);
template hipError_t launch_dot<float>(const DotArgs<float> &, hipStream_t);

__global__ __launch_bounds__(BLOCK) void cg_update_kernel(const CgUpdArgs a)
{
    __shared__ double scratch[2 * (BLOCK / 64)];
    __shared__ double sums[2];
    if (a.ctrl->done) return;
    const double alpha = a.ctrl->alpha;
    double acc[2] = {0.0, 0.0};
    const long chunk = (a.n + a.nblocks - 1) / a.nblocks;
    const long beg = (long)blockIdx.x * chunk, end = min(a.n, beg + chunk);
    for (long i = beg + threadIdx.x; i < end; i += BLOCK) {
        a.x[i] = a.x[i] + alpha * a.d[i];            // CG.c:106-107
        if (a.update_r) {
            const double r = a.r[i] - alpha * a.q[i];   // CG.c:121-124
            a.r[i] = r;
            double s = r;
            if (a.precond) s = s / a.diag;
            acc[0] += r * s;
            acc[1] += r * r;
        }
    }
    if (!a.update_r) return;   // restart: a stencil pass recomputes r and reduces
    block_reduce_to<2>(acc, a.partials + (size_t)blockIdx.x * 2, scratch);
    if (last_block_arrive(a.counter, gridDim.x)) {
        final_sum(a.partials, a.nblocks, 2, sums, scratch);
        if (threadIdx.x == 0) {
            finalize(a.fin, a.ctrl, sums, 2);
            *a.counter = 0u;
        }
    }
}

hipError_t launch_cg_update(const CgUpdArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(cg_update_kernel, dim3(a.nblocks), dim3(BLOCK), 0, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__global__ void copy_f64_kernel(double *dst, const double *src, long n, const Ctrl *c)
{
    if (c && c->done) return;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
template <typename T>
__global__ void fill_kernel(T *dst, T v, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = v;
}
__global__ void axpy_f64_f32_kernel(double *x, const float *e, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = x[i] + (double)e[i];
}
__global__ void div_f64_kernel(double *r, double d, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        r[i] /= d;
}
__global__ void zero_ctrl_kernel(Ctrl *c)
{
    double *p = reinterpret_cast<double *>(c);
    const int nw = (int)(sizeof(Ctrl) / sizeof(double));
    for (int i = threadIdx.x; i < nw; i += blockDim.x) p[i] = 0.0;
}

static unsigned grid_for(long n)
{
    long g = (n + BLOCK - 1) / BLOCK;
    return (unsigned)(g > 8192 ? 8192 : (g < 1 ? 1 : g));
}

hipError_t launch_copy_f64(double *dst, const double *src, long n, const Ctrl *ctrl, hipStream_t st)
{
    hipLaunchKernelGGL(copy_f64_kernel, dim3(grid_for(n)), dim3(BLOCK), 0, st, dst, src, n, ctrl);
    return hipGetLastError();
}
hipError_t launch_fill_f64(double *dst, double v, long n, hipStream_t st)
{
    hipLaunchKernelGGL((fill_kernel<double>), dim3(grid_for(n)), dim3(BLOCK), 0, st, dst, v, n);
    return hipGetLastError();
}
hipError_t launch_fill_f32(float *dst, float v, long n, hipStream_t st)
{
    hipLaunchKernelGGL((fill_kernel<float>), dim3(grid_for(n)), dim3(BLOCK), 0, st, dst, v, n);
    return hipGetLastError();
}
hipError_t launch_axpy_f64_f32(double *x, const float *e, long n, hipStream_t st)
{
    hipLaunchKernelGGL(axpy_f64_f32_kernel, dim3(grid_for(n)), dim3(BLOCK), 0, st, x, e, n);
    return hipGetLastError();
}
hipError_t launch_div_f64(double *r, double diag, long n, hipStream_t st)
{
    hipLaunchKernelGGL(div_f64_kernel, dim3(grid_for(n)), dim3(BLOCK), 0, st, r, diag, n);
    return hipGetLastError();
}
hipError_t launch_zero_ctrl(Ctrl *c, hipStream_t st)
{
    hipLaunchKernelGGL(zero_ctrl_kernel, dim3(1), dim3(256), 0, st, c);
    return hipGetLastError();
}

}  // namespace aar
