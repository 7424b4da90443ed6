// helmholtz.hip -- complex Helmholtz kernels (no MPI_REAL counterpart; the
// AAR semantics follow MPI_REAL/AAR.c with the Hermitian Gram of
// PETSC_COMPLEX/AAR_Complex.c:207-246).
//
// Operator, complex128 interleaved (double2 = re, im), periodic:
//   A u = a*Lap(u) + (c(x) + i*eta) u,
// the Laplacian applied to re and im separately in the order of
// MPI_REAL/system.c:195-201, then y_re = lap_re + (c u_re - eta u_im),
// y_im = lap_im + (c u_im + eta u_re).  Jacobi: D = (a*3*coeff[0] + c) + i*eta,
// g = f / D computed as ((f_re D_re + f_im D_im) + i (f_im D_re - f_re D_im)) / |D|^2.
// The adjoint operator A^H is the same with eta -> -eta (the Laplacian is
// real symmetric and c real).
//
// The stencil keeps the design of stencil.hip (z-march with a register
// queue, LDS plane tile, XCD-aware tiles) on 16-byte points; its plane tile
// is twice as large, so halos are loaded per plane (no halo ring) and a
// thread owns two rows.
#include "aar_common.h"
#include "aar_device.h"
#include "aar_kernels.h"

namespace aar {

namespace {

constexpr int HWY = 8;   // wavefronts per block (stacked in y)
constexpr int HPY = 2;   // rows per thread

__device__ inline int hmodp(int i, int n) { i %= n; return i < 0 ? i + n : i; }

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(__dadd_rn(a.x, b.x), __dadd_rn(a.y, b.y)); }
__device__ inline double2 cscale(double2 a, double s) { return make_double2(__dmul_rn(a.x, s), __dmul_rn(a.y, s)); }

// (c + i eta) u added to the Laplacian part, in the oracle's order
__device__ inline double2 helm_diag_add(double2 lap, double2 u, double ck, double eta)
{
    const double tr = __dsub_rn(__dmul_rn(ck, u.x), __dmul_rn(eta, u.y));
    const double ti = __dadd_rn(__dmul_rn(ck, u.y), __dmul_rn(eta, u.x));
    return make_double2(__dadd_rn(lap.x, tr), __dadd_rn(lap.y, ti));
}

// f / D with D = dr + i di
__device__ inline double2 cdiv_d(double2 f, double dr, double di)
{
    const double den = __dadd_rn(__dmul_rn(dr, dr), __dmul_rn(di, di));
    const double re = __dadd_rn(__dmul_rn(f.x, dr), __dmul_rn(f.y, di)) / den;
    const double im = __dsub_rn(__dmul_rn(f.y, dr), __dmul_rn(f.x, di)) / den;
    return make_double2(re, im);
}

}  // namespace

template <HMode MODE>
__global__ __launch_bounds__(64 * HWY) void hstencil_kernel(const HArgs a)
{
    constexpr int WY = HWY, PY = HPY;
    constexpr int NT = 64 * WY;
    constexpr int TY = WY * PY;
    constexpr int LH = TY + 2 * R;
    constexpr int LW = TX + 2 * R;
    constexpr int NYH = (2 * R + WY - 1) / WY;
    constexpr int NXH = (2 * R * TY + NT - 1) / NT;
    constexpr int QN = 2 * R + 1;
    constexpr bool NEED_B = MODE == HMode::TRUE_RES;

    __shared__ double2 lds[2][LH][LW];
    __shared__ double red_scratch[8 * WY];

    if (a.ctrl && a.ctrl->done) return;

    const int item = xcd_item(blockIdx.x, a.nitems);
    const int itx = item % a.ntx;
    const int rest = item / a.ntx;
    const int ity = rest % a.nty;
    const int izc = rest / a.nty;
    const int x0 = itx * TX, y0 = ity * TY;
    const int z0 = izc * a.lz;
    const int z1 = min(z0 + a.lz, a.nz);
    const int tid = threadIdx.x, tx = tid & 63, ty = tid >> 6;
    const int nx = a.nx, ny = a.ny, nz = a.nz, nxy = a.nxy;

    const int cx = hmodp(x0 + tx, nx);
    int c_off[PY];
    bool c_ok[PY];
#pragma unroll
    for (int p = 0; p < PY; p++) {
        const int y = y0 + ty * PY + p;
        c_off[p] = cx + hmodp(y, ny) * nx;
        c_ok[p] = (x0 + tx < nx) && (y < ny);
    }
    int yh_off[NYH], yh_row[NYH];
    bool yh_ok[NYH];
#pragma unroll
    for (int k = 0; k < NYH; k++) {
        const int hr0 = ty + WY * k;
        yh_ok[k] = hr0 < 2 * R;
        const int hr = yh_ok[k] ? hr0 : 0;
        const int gy = hr < R ? y0 - R + hr : y0 + TY + (hr - R);
        yh_off[k] = cx + hmodp(gy, ny) * nx;
        yh_row[k] = hr < R ? hr : TY + hr;
    }
    int xh_off[NXH], xh_row[NXH], xh_col[NXH];
    bool xh_ok[NXH];
#pragma unroll
    for (int k = 0; k < NXH; k++) {
        const int c = tid + NT * k;
        xh_ok[k] = c < 2 * R * TY;
        const int r = xh_ok[k] ? c / (2 * R) : 0, cc = xh_ok[k] ? c % (2 * R) : 0;
        const int gx = cc < R ? x0 - R + cc : x0 + TX + (cc - R);
        xh_off[k] = xh_ok[k] ? hmodp(gx, nx) + hmodp(y0 + r, ny) * nx : c_off[0];
        xh_row[k] = R + r;
        xh_col[k] = cc < R ? cc : TX + cc;
    }
    double wt[R + 1];
#pragma unroll
    for (int o = 0; o <= R; o++) wt[o] = a.w[o];
    const double eta = a.eta, om = a.omega, dlap = a.dlap;

    double2 q[PY][QN];
    struct Pre {
        double2 qn[PY];
        double2 hy[NYH];
        double2 hx[NXH];
        double2 bb[PY];
        double cc[PY];
    };
    Pre s0, s1;

#pragma unroll
    for (int s = 0; s < 2 * R; s++) {
        const int zo = hmodp(z0 - R + s, nz) * nxy;
#pragma unroll
        for (int p = 0; p < PY; p++) q[p][s] = a.in0[c_off[p] + zo];
    }
    auto prefetch = [&](int zp, Pre &st) {
        const int zq = hmodp(zp + R, nz) * nxy, zc = hmodp(zp, nz) * nxy;
#pragma unroll
        for (int p = 0; p < PY; p++) st.qn[p] = a.in0[c_off[p] + zq];
#pragma unroll
        for (int k = 0; k < NYH; k++) st.hy[k] = a.in0[yh_off[k] + zc];
#pragma unroll
        for (int k = 0; k < NXH; k++) st.hx[k] = a.in0[xh_off[k] + zc];
#pragma unroll
        for (int p = 0; p < PY; p++) {
            if constexpr (NEED_B) st.bb[p] = a.bvec[c_off[p] + zc];
            st.cc[p] = a.c ? a.c[c_off[p] + zc] : a.c0;
        }
    };

    double acc = 0.0;
    auto plane = [&](const int z, Pre &cur, Pre &nxt) {
        const int buf = (z - z0) & 1;
#pragma unroll
        for (int p = 0; p < PY; p++) {
            q[p][2 * R] = cur.qn[p];
            lds[buf][R + ty * PY + p][R + tx] = q[p][R];
        }
#pragma unroll
        for (int k = 0; k < NYH; k++)
            if (yh_ok[k]) lds[buf][yh_row[k]][R + tx] = cur.hy[k];
#pragma unroll
        for (int k = 0; k < NXH; k++)
            if (xh_ok[k]) lds[buf][xh_row[k]][xh_col[k]] = cur.hx[k];
        prefetch(z + 1, nxt);
        __syncthreads();

        const int zoff = z * nxy;
#pragma unroll
        for (int p = 0; p < PY; p++) {
            const int lr = R + ty * PY + p, lc = R + tx;
            const double2 u = q[p][R];
            auto ynb = [&](int d) -> double2 {
                const int pp = p + d;
                if (pp >= 0 && pp < PY) return q[pp < 0 ? 0 : (pp < PY ? pp : PY - 1)][R];
                return lds[buf][lr + d][lc];
            };
            // system.c:195-201 order on each component
            double2 y = cscale(cscale(u, 3.0), wt[0]);
#pragma unroll
            for (int o = 1; o <= R; o++) {
                double2 nb = cadd(lds[buf][lr][lc - o], lds[buf][lr][lc + o]);
                nb = cadd(nb, ynb(-o));
                nb = cadd(nb, ynb(o));
                nb = cadd(nb, q[p][R - o]);
                nb = cadd(nb, q[p][R + o]);
                y = cadd(y, cscale(nb, wt[o]));
            }
            const double ck = cur.cc[p];
            const double2 au = helm_diag_add(y, u, ck, eta);
            const bool ok = c_ok[p];
            const int g = c_off[p] + zoff;
            const double dr = __dadd_rn(dlap, ck);
            if constexpr (MODE == HMode::MATVEC) {
                if (ok) a.out0[g] = au;
            } else if constexpr (MODE == HMode::TRUE_RES) {
                const double2 f = make_double2(__dsub_rn(cur.bb[p].x, au.x), __dsub_rn(cur.bb[p].y, au.y));
                acc += ok ? (f.x * f.x + f.y * f.y) : 0.0;
                if (ok) a.out0[g] = cdiv_d(f, dr, eta);
            } else {
                // g' = g - omega (A g)/D; |D g'|^2 = ||b - A x'||^2 at the point
                const double2 t = cdiv_d(au, dr, eta);
                const double2 gn = make_double2(__dsub_rn(u.x, __dmul_rn(om, t.x)), __dsub_rn(u.y, __dmul_rn(om, t.y)));
                if (ok) a.out0[g] = gn;
                if (a.trace) {
                    const double den = dr * dr + eta * eta;
                    acc += ok ? den * (gn.x * gn.x + gn.y * gn.y) : 0.0;
                }
            }
        }
#pragma unroll
        for (int p = 0; p < PY; p++)
#pragma unroll
            for (int s = 0; s < 2 * R; s++) q[p][s] = q[p][s + 1];
    };

    prefetch(z0, s0);
    int z = z0;
    for (; z + 1 < z1; z += 2) {
        plane(z, s0, s1);
        plane(z + 1, s1, s0);
    }
    if (z < z1) plane(z, s0, s1);

    if (a.partials) {
        double accv[1] = {acc};
        block_reduce_to<1, NT>(accv, a.partials + (size_t)item, red_scratch);
    }
}

static void hgeometry(HArgs &a)
{
    a.nxy = a.nx * a.ny;
    a.ntx = (a.nx + TX - 1) / TX;
    a.nty = (a.ny + HWY * HPY - 1) / (HWY * HPY);
    if (a.lz <= 0) {
        const int cols = a.ntx * a.nty;
        int nzc = std::max(1, 256 / std::max(1, cols));
        int lz = (a.nz + nzc - 1) / nzc;
        a.lz = std::max(lz, std::min(a.nz, 16));
    }
    if (a.lz > a.nz) a.lz = a.nz;
    a.nzc = (a.nz + a.lz - 1) / a.lz;
    a.nitems = a.ntx * a.nty * a.nzc;
}

int hstencil_items(const HArgs &a0)
{
    HArgs a = a0;
    hgeometry(a);
    return a.nitems;
}

hipError_t launch_hstencil(HMode m, const HArgs &a0, hipStream_t st)
{
    HArgs a = a0;
    hgeometry(a);
    dim3 grid(a.nitems), block(64 * HWY);
    switch (m) {
    case HMode::MATVEC: hipLaunchKernelGGL((hstencil_kernel<HMode::MATVEC>), grid, block, 0, st, a); break;
    case HMode::TRUE_RES: hipLaunchKernelGGL((hstencil_kernel<HMode::TRUE_RES>), grid, block, 0, st, a); break;
    case HMode::REC: hipLaunchKernelGGL((hstencil_kernel<HMode::REC>), grid, block, 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Hermitian Gram of the complex g ring (PETSC_COMPLEX/AAR_Complex.c:213-236)
// ---------------------------------------------------------------------------
namespace {
__device__ inline int href_col(int it, int k, int m) { return ((it - 2 - k) % m + m) % m; }
}

template <int MM>
__global__ __launch_bounds__(BLOCK) void hgram_kernel(const HGramArgs a)
{
    constexpr int NG = MM * (MM + 1) / 2;
    constexpr int NA = 2 * (NG + MM);
    static_assert(NA <= HGRAM_NA_MAX, "partials buffers hold HGRAM_NA_MAX per block");
    constexpr int NS = MM + 1;
    __shared__ double scratch[NA * (BLOCK / 64)];
    if (a.ctrl->done) return;
    const int nv = a.nvalid;
    const double2 *pd[NS];
    bool use[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        use[k] = k <= nv;
        pd[k] = a.F + (size_t)(use[k] ? a.slot[k] : 0) * a.ld;
    }
    double acc[NA];
#pragma unroll
    for (int k = 0; k < NA; k++) acc[k] = 0.0;
    const long stride = (long)gridDim.x * BLOCK;
    for (long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < a.n; i += stride) {
        double2 v[NS];
#pragma unroll
        for (int k = 0; k < NS; k++)
            v[k] = use[k] ? ld_nt2<double2>(reinterpret_cast<const double *>(pd[k] + i)) : make_double2(0.0, 0.0);
        double dr[MM], di[MM];
#pragma unroll
        for (int k = 0; k < MM; k++) {
            dr[k] = k < nv ? v[k].x - v[k + 1].x : 0.0;     // DF = f - f_old
            di[k] = k < nv ? v[k].y - v[k + 1].y : 0.0;
        }
        const double fr = v[0].x, fi = v[0].y;
        int t = 0;
#pragma unroll
        for (int j = 0; j < MM; j++) {
#pragma unroll
            for (int q = 0; q <= j; q++) {
                // <DF_q, DF_j> = conj(DF_q) DF_j
                acc[2 * t] += dr[q] * dr[j] + di[q] * di[j];
                acc[2 * t + 1] += dr[q] * di[j] - di[q] * dr[j];
                t++;
            }
            acc[2 * NG + 2 * j] += dr[j] * fr + di[j] * fi;       // <DF_j, f>
            acc[2 * NG + 2 * j + 1] += dr[j] * fi - di[j] * fr;
        }
    }
    block_reduce_to<NA>(acc, a.partials + (size_t)blockIdx.x * NA, scratch);
}

// one block: fixed-order sum of the partials, then the real 2m x 2m layout
// read by ls_kernel (a[i][j] = G[j*n + i], rhs G[n*n + i], n = 2m)
template <int MM>
__global__ __launch_bounds__(BLOCK) void hgram_finish_kernel(const HGramArgs a)
{
    constexpr int NG = MM * (MM + 1) / 2;
    constexpr int NA = 2 * (NG + MM);
    __shared__ double scratch[BLOCK / 64];
    __shared__ double sums[NA];
    if (a.ctrl->done) return;
    final_sum(a.partials, a.nblocks, NA, sums, scratch);
    if (threadIdx.x != 0) return;
    const int m = a.m, n2 = 2 * m, nv = a.nvalid;
    double *G = a.ctrl->G;
    for (int k = 0; k < n2 * n2 + n2; k++) G[k] = 0.0;
    auto put = [&](int i, int j, double v) { G[j * n2 + i] = v; };   // E[i][j]
    for (int j = 0; j < nv; j++)
        for (int q = 0; q <= j; q++) {
            const int t = j * (j + 1) / 2 + q;
            const double re = sums[2 * t], im = sums[2 * t + 1];
            const int ci = href_col(a.it, q, m), cj = href_col(a.it, j, m);
            // Gr[ci][cj] = Gr[cj][ci] = re; Gi[ci][cj] = im, Gi[cj][ci] = -im;
            // E = [Gr -Gi; Gi Gr]
            put(ci, cj, re);
            put(cj, ci, re);
            put(m + ci, m + cj, re);
            put(m + cj, m + ci, re);
            put(ci, m + cj, -im);
            put(cj, m + ci, im);
            put(m + ci, cj, im);
            put(m + cj, ci, -im);
        }
    for (int j = 0; j < nv; j++) {
        const int cj = href_col(a.it, j, m);
        G[n2 * n2 + cj] = sums[2 * NG + 2 * j];
        G[n2 * n2 + m + cj] = sums[2 * NG + 2 * j + 1];
    }
}

hipError_t launch_hgram(const HGramArgs &a, hipStream_t st)
{
    if (a.m < 1 || a.m > MAXM_CPLX || 2 * a.m > MAXM_LS || a.nvalid < 0 || a.nvalid > a.m) return hipErrorInvalidValue;
    if (a.m <= 7) {
        hipLaunchKernelGGL((hgram_kernel<7>), dim3(a.nblocks), dim3(BLOCK), 0, st, a);
        hipLaunchKernelGGL((hgram_finish_kernel<7>), dim3(1), dim3(BLOCK), 0, st, a);
    } else if (a.m <= 8) {
        hipLaunchKernelGGL((hgram_kernel<8>), dim3(a.nblocks), dim3(BLOCK), 0, st, a);
        hipLaunchKernelGGL((hgram_finish_kernel<8>), dim3(1), dim3(BLOCK), 0, st, a);
    } else {
        // deeper histories (the reference's complex README run uses m = 13):
        // 304 accumulators, partly spilled; correct, at lower throughput
        hipLaunchKernelGGL((hgram_kernel<MAXM_CPLX>), dim3(a.nblocks), dim3(BLOCK), 0, st, a);
        hipLaunchKernelGGL((hgram_finish_kernel<MAXM_CPLX>), dim3(1), dim3(BLOCK), 0, st, a);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// complex anchor update (see HAncArgs / AncArgs).  ONE (histories up to 8):
// one element per thread and one block per BLOCK elements, x read and
// written non-temporally, as the FP64 real update.  The deeper form keeps a
// grid stride over at most 4096 blocks (its registers leave too few waves
// for one element per thread, as measured for the real update).
// ---------------------------------------------------------------------------
template <int NGM, int NCM, bool ONE>
__global__ __launch_bounds__(BLOCK) void hanchor_update_kernel(const HAncArgs a)
{
    __shared__ double s_cg[2 * NGMAX], s_cc[2 * NCMAX];
    if (a.ctrl && a.ctrl->done) return;
    if (threadIdx.x == 0) {
        for (int i = 0; i < NGMAX; i++) {
            s_cg[2 * i] = (a.with_beta && i == a.ng - 1) ? a.beta : 0.0;
            s_cg[2 * i + 1] = 0.0;
        }
        for (int c = 0; c < 2 * NCMAX; c++) s_cc[c] = 0.0;
        for (int t = 0; t < a.ncol; t++) {
            const double gr = a.ctrl->s[a.col_ref[t]], gi = a.ctrl->s[a.m + a.col_ref[t]];
            const int j = a.col_j[t], cj = a.col_c[t];
            s_cg[2 * j] -= a.beta * gr;
            s_cg[2 * j + 1] -= a.beta * gi;
            const double f = (cj < 0) ? (a.beta - a.omega) : a.beta;
            s_cg[2 * (j - 1)] += f * gr;
            s_cg[2 * (j - 1) + 1] += f * gi;
            if (cj >= 0) {
                s_cc[2 * cj] -= gr;
                s_cc[2 * cj + 1] -= gi;
            }
        }
    }
    __syncthreads();
    const int ng = a.ng, nc = a.nc;
    double2 *co = (a.cout >= 0) ? a.C + (size_t)a.cout * a.ld : nullptr;
    const long stride = ONE ? a.n : (long)gridDim.x * BLOCK;
    for (long e = (long)blockIdx.x * BLOCK + threadIdx.x; e < a.n; e += stride) {
        double sxr = 0.0, sxi = 0.0, scr = 0.0, sci = 0.0;
        // every stream's load issued before the sums (static slot indices)
        double2 gv[NGM], cv[NCM];
#pragma unroll
        for (int i = 0; i < NGM; i++)
            if (i < ng) gv[i] = ld_nt2<double2>(reinterpret_cast<const double *>(a.G + (size_t)a.gslot[i] * a.ld + e));
#pragma unroll
        for (int c = 0; c < NCM; c++)
            if (c < nc) cv[c] = ld_nt2<double2>(reinterpret_cast<const double *>(a.C + (size_t)a.cslot[c] * a.ld + e));
#pragma unroll
        for (int i = 0; i < NGM; i++)
            if (i < ng) {
                const double2 g = gv[i];
                if (i >= a.xlo && i < a.xhi) {
                    sxr += g.x;
                    sxi += g.y;
                }
                const double cr = s_cg[2 * i], ci = s_cg[2 * i + 1];
                scr += cr * g.x - ci * g.y;
                sci += cr * g.y + ci * g.x;
            }
#pragma unroll
        for (int c = 0; c < NCM; c++)
            if (c < nc) {
                const double2 v = cv[c];
                const double cr = s_cc[2 * c], ci = s_cc[2 * c + 1];
                scr += cr * v.x - ci * v.y;
                sci += cr * v.y + ci * v.x;
            }
        if (ONE) {
            const double2 xv = ld_nt2<double2>(reinterpret_cast<const double *>(a.x + e));
            st_nt2(reinterpret_cast<double *>(a.x + e),
                   make_double2(xv.x + (a.omega * sxr + scr), xv.y + (a.omega * sxi + sci)));
        } else {
            const double2 xv = a.x[e];
            a.x[e] = make_double2(xv.x + (a.omega * sxr + scr), xv.y + (a.omega * sxi + sci));
        }
        if (co) st_nt2(reinterpret_cast<double *>(co + e), make_double2(scr, sci));
    }
}

hipError_t launch_hanchor_update(const HAncArgs &a, hipStream_t st)
{
    constexpr int HNG = MAXM_CPLX + 2, HNC = MAXM_CPLX + 1;   // the complex solver's m <= 16
    if (a.ng < 1 || a.ng > HNG || a.nc < 0 || a.nc > HNC || a.ncol < 0 || a.ncol > MAXM_CPLX)
        return hipErrorInvalidValue;
    if (a.xlo < 0 || a.xhi > a.ng || !a.G || !a.x) return hipErrorInvalidValue;
    if ((a.nc > 0 || a.cout >= 0) && !a.C) return hipErrorInvalidValue;
    if (a.ncol > 0 && !a.ctrl) return hipErrorInvalidValue;
    for (int t = 0; t < a.ncol; t++)
        if (a.col_j[t] < 1 || a.col_j[t] >= a.ng || a.col_c[t] >= a.nc || a.col_ref[t] < 0 ||
            a.col_ref[t] >= a.m)
            return hipErrorInvalidValue;
    long nb = (a.n + BLOCK - 1) / BLOCK;
    if (nb < 1) nb = 1;
    if (nb > 0x7fffffffL) return hipErrorInvalidValue;
    const long nstride = std::min(nb, 4096L);
    // histories up to 8 (an (8, 1) form with half the registers measured 4%
    // slower at 256^3), and the deeper ones
    if (a.ng <= MAXM_FAST + 2 && a.nc <= MAXM_FAST + 1)
        hipLaunchKernelGGL((hanchor_update_kernel<MAXM_FAST + 2, MAXM_FAST + 1, true>), dim3((unsigned)nb), dim3(BLOCK),
                           0, st, a);
    else
        hipLaunchKernelGGL((hanchor_update_kernel<HNG, HNC, false>), dim3((unsigned)nstride), dim3(BLOCK), 0, st, a);
    return hipGetLastError();
}

// dL/dc_k = -Re(conj(lambda_k) u_k)
__global__ __launch_bounds__(BLOCK) void helm_grad_c_kernel(long n, const double2 *lam, const double2 *u, double *dc)
{
    const long stride = (long)gridDim.x * BLOCK;
    for (long e = (long)blockIdx.x * BLOCK + threadIdx.x; e < n; e += stride) {
        const double2 l = lam[e], v = u[e];
        dc[e] = -(l.x * v.x + l.y * v.y);
    }
}

hipError_t launch_helm_grad_c(long n, const double2 *lambda, const double2 *u, double *dc, hipStream_t st)
{
    if (n < 1 || !lambda || !u || !dc) return hipErrorInvalidValue;
    long nb = (n + BLOCK - 1) / BLOCK;
    nb = nb < 1 ? 1 : (nb > 8192 ? 8192 : nb);
    hipLaunchKernelGGL(helm_grad_c_kernel, dim3((unsigned)nb), dim3(BLOCK), 0, st, n, lambda, u, dc);
    return hipGetLastError();
}

}  // namespace aar
