// stencil_body.h -- the FD12 stencil sweep shared by the single-operator
// and the batched kernels (see stencil.hip for the design notes).
#pragma once

#include "aar_common.h"
#include "aar_device.h"
#include "aar_kernels.h"

namespace aar {

template <typename T> struct Arith;
template <> struct Arith<double> {
    using W = double;
    __device__ static inline double add(double a, double b) { return __dadd_rn(a, b); }
    __device__ static inline double sub(double a, double b) { return __dsub_rn(a, b); }
    __device__ static inline double mul(double a, double b) { return __dmul_rn(a, b); }
    __device__ static inline double div(double a, double b) { return a / b; }
    __device__ static inline double sq(double a) { return a * a; }
};
template <> struct Arith<float> {
    using W = float;
    __device__ static inline float add(float a, float b) { return __fadd_rn(a, b); }
    __device__ static inline float sub(float a, float b) { return __fsub_rn(a, b); }
    __device__ static inline float mul(float a, float b) { return __fmul_rn(a, b); }
    __device__ static inline float div(float a, float b) { return __fdiv_rn(a, b); }
    __device__ static inline double sq(float a) { return (double)a * (double)a; }
};

template <SMode M> struct ModeInfo;
template <> struct ModeInfo<SMode::MATVEC> { static constexpr int NF = 1, NE = 0, NACC = 1; };
template <> struct ModeInfo<SMode::MATVEC_RES> { static constexpr int NF = 1, NE = 1, NACC = 1; };
template <> struct ModeInfo<SMode::AAR_TRUE> { static constexpr int NF = 1, NE = 1, NACC = 1; };
template <> struct ModeInfo<SMode::AAR_REC> { static constexpr int NF = 1, NE = 0, NACC = 1; };
template <> struct ModeInfo<SMode::RESID> { static constexpr int NF = 1, NE = 1, NACC = 1; };
template <> struct ModeInfo<SMode::PGR_RICH> { static constexpr int NF = 3, NE = 1, NACC = 1; };
template <> struct ModeInfo<SMode::CG_Q> { static constexpr int NF = 2, NE = 0, NACC = 1; };
template <> struct ModeInfo<SMode::CG_RESID> { static constexpr int NF = 1, NE = 1, NACC = 2; };

__device__ inline int modp(int i, int n) { i %= n; return i < 0 ? i + n : i; }

// base[eoff] with a uniform (per-plane) base pointer and a 32-bit in-plane
// offset: the access compiles to the scalar-base + vector-offset form, with
// no per-access 64-bit address arithmetic (in-plane byte offsets are below
// 2^32 for any plane the host accepts)
template <typename T> __device__ __forceinline__ T ld_at(const T *base, unsigned eoff)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + (size_t)(eoff * (unsigned)sizeof(T)));
}
template <typename T> __device__ __forceinline__ void st_at(T *base, unsigned eoff, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (size_t)(eoff * (unsigned)sizeof(T))) = v;
}
// Plane-relative access: the form above when SA, else one signed index
// (the sweeps with two or three streams keep fewer live registers with the
// plain index; measured per mode)
template <bool SA, typename T> __device__ __forceinline__ T ld_pl(const T *base, size_t zo, unsigned off)
{
    if constexpr (SA) return ld_at(base + zo, off);
    else return base[(int)off + (int)zo];
}
template <bool SA, typename T> __device__ __forceinline__ void st_pl(T *base, size_t zo, unsigned off, T v)
{
    if constexpr (SA) st_at(base + zo, off, v);
    else base[(int)off + (int)zo] = v;
}

// The stencil sweep of one work item (a 64 x TY column over a z-chunk) of
// the operator and fields in `a`; shared by the single-operator kernel
// below and the batched kernel of batch.hip.
// ADDR: 1 = plane pointer + 32-bit offset accesses, 0 = one signed index,
// -1 = the per-mode choice below.  VC: the variable-coefficient operator
// a*Lap + c(x), with a pointwise field a.vf read in the epilogue.
template <typename T, int WY, int PY, SMode MODE, int ADDR = -1, bool VC = false>
__device__ __forceinline__ void stencil_body(const SArgs<T> &a, const int item)
{
    static_assert(!VC || MODE != SMode::RESID, "no variable-coefficient RESID sweep");
    using A = Arith<T>;
    constexpr int NT = 64 * WY;                       // threads: WY wavefronts stacked in y
    constexpr int NF = ModeInfo<MODE>::NF;
    // VC sweeps whose pre-stencil step divides by the pointwise diagonal
    // (PGR's Richardson update, CG's preconditioned direction) load c(x) as
    // one more field at every queue and halo point
    constexpr bool VCF = VC && (MODE == SMode::PGR_RICH || MODE == SMode::CG_Q);
    constexpr int NFX = NF + (VCF ? 1 : 0);
    constexpr int NE = ModeInfo<MODE>::NE;
    constexpr int NACC = ModeInfo<MODE>::NACC;
    constexpr int TY = WY * PY;
    constexpr int LH = TY + 2 * R;
    constexpr int LW = TX + 2 * R;
    constexpr int NYH = (2 * R + WY - 1) / WY;        // y-halo rows per thread row
    constexpr int NXH = (2 * R * TY + NT - 1) / NT;   // x-halo cells per thread
    constexpr int QN = 2 * R + 1;
    // unconditional stores and an exact memory wait before the plane loop:
    // measured faster for these sweeps, slower for the others (which keep
    // fewer live registers with masked stores)
    constexpr bool USTORE = MODE == SMode::MATVEC || MODE == SMode::MATVEC_RES || MODE == SMode::AAR_TRUE ||
                            MODE == SMode::RESID || MODE == SMode::CG_RESID;
    constexpr bool SA = ADDR >= 0 ? ADDR == 1
                                  : MODE == SMode::AAR_REC ||
                                        (sizeof(T) == 8 && (MODE == SMode::MATVEC || (MODE == SMode::CG_Q && !VC)));
    // prefetch distance (planes): 2 for the FP64 sweeps without epilogue
    // inputs, whose one block per CU (LDS-limited) leaves VGPRs to spare; 1
    // where a third register set would cost a block per CU (FP32), spill
    // (two or three fields) or measured slower (epilogue inputs)
    constexpr int PD = (sizeof(T) == 8 && NF == 1 && NE == 0 && !VC) ? 2 : 1;

    constexpr int HR = R + 1;                         // halo ring depth (planes)
    constexpr int NH = NYH + NXH;                     // halo cells per thread
    __shared__ T lds[2][LH][LW];
    // Halo cells are fetched R planes ahead, in the same step as the
    // neighbouring blocks fetch those lines as their own queue planes, so
    // that they are served from L2; each thread parks its cells here until
    // their plane is computed (slots private to the thread: no barrier).
    __shared__ T hring[HR][NH][NT];
    __shared__ double red_scratch[8 * WY];
    __shared__ double red_out[8];

    const int itx = item % a.ntx;
    const int rest = item / a.ntx;
    const int ity = rest % a.nty;
    const int izc = rest / a.nty;
    const int x0 = itx * TX, y0 = ity * TY;
    const int z0 = izc * a.lz;
    const int z1 = min(z0 + a.lz, a.nz);
    const int tid = threadIdx.x, tx = tid & 63, ty = tid >> 6;
    const int nx = a.nx, ny = a.ny, nz = a.nz, nxy = a.nxy;

    // in-plane offsets of this thread's points, halo cells and LDS slots
    const int cx = modp(x0 + tx, nx);
    unsigned c_off[PY];
    bool c_ok[PY];
#pragma unroll
    for (int p = 0; p < PY; p++) {
        const int y = y0 + ty * PY + p;
        c_off[p] = (unsigned)(cx + modp(y, ny) * nx);
        c_ok[p] = (x0 + tx < nx) && (y < ny);
    }
    unsigned yh_off[NYH];
    int yh_row[NYH];
    bool yh_ok[NYH];
#pragma unroll
    for (int k = 0; k < NYH; k++) {
        const int hr0 = ty + WY * k;                       // 0 .. 2R-1 when valid
        yh_ok[k] = hr0 < 2 * R;
        const int hr = yh_ok[k] ? hr0 : 0;
        const int gy = hr < R ? y0 - R + hr : y0 + TY + (hr - R);
        yh_off[k] = (unsigned)(cx + modp(gy, ny) * nx);
        yh_row[k] = hr < R ? hr : TY + hr;
    }
    unsigned xh_off[NXH];
    int xh_row[NXH], xh_col[NXH];
    bool xh_ok[NXH];
#pragma unroll
    for (int k = 0; k < NXH; k++) {
        const int c = tid + NT * k;
        xh_ok[k] = c < 2 * R * TY;
        const int r = xh_ok[k] ? c / (2 * R) : 0, cc = xh_ok[k] ? c % (2 * R) : 0;
        const int gx = cc < R ? x0 - R + cc : x0 + TX + (cc - R);
        // cells past the halo load a valid (unused) address so that every
        // load is unconditional; only their LDS write is predicated
        xh_off[k] = xh_ok[k] ? (unsigned)(modp(gx, nx) + modp(y0 + r, ny) * nx) : c_off[0];
        xh_row[k] = R + r;
        xh_col[k] = cc < R ? cc : TX + cc;
    }

    const T wt0 = (T)a.w[0];
    T wt[R + 1];
#pragma unroll
    for (int o = 0; o <= R; o++) wt[o] = (T)a.w[o];
    const T sig = (T)a.sigma, dg = (T)a.diag, om = (T)a.omega;
    [[maybe_unused]] const T w3 = (T)(3.0 * a.w[0]);
    T cf[R + 1];
#pragma unroll
    for (int o = 0; o <= R; o++) cf[o] = (T)a.cf[o];
    const double trace_scale = a.precond ? a.diag : 1.0;
    // FP64 sweeps: norm terms of points past the grid edge are multiplied by
    // 0 (exact: x * 1 = x), so the accumulation needs no select per point.
    // The FP32 sweeps (four waves per SIMD at <= 128 VGPRs) keep the select:
    // the factors would take them past 128.
    constexpr bool MF = sizeof(T) == 8;
    double okf[PY], okts[PY];
#pragma unroll
    for (int p = 0; p < PY; p++) {
        okf[p] = c_ok[p] ? 1.0 : 0.0;
        okts[p] = c_ok[p] ? trace_scale : 0.0;
    }
    auto nmask = [&](int p, double v) -> double {
        if constexpr (MF) return v * okf[p];
        else return c_ok[p] ? v : 0.0;
    };
    const T cgb = (MODE == SMode::CG_Q && a.cg_beta) ? (T)(*a.cg_beta) : (T)0;

    // fields at in-plane offset `off` of the plane starting at element zo
    auto load_raw = [&](size_t zo, unsigned off, T (&v)[NFX]) {
        v[0] = ld_pl<SA>(a.in0, zo, off);
        if constexpr (NF > 1) v[1] = ld_pl<SA>(a.in1, zo, off);
        if constexpr (NF > 2) v[2] = ld_pl<SA>(a.in2, zo, off);
        if constexpr (VCF) v[NF] = ld_pl<SA>(a.vf, zo, off);
    };
    // the Jacobi divisor at a point: diag, or diag + c(x) (VCF)
    auto dv = [&](const T (&v)[NFX]) -> T {
        if constexpr (VCF) return A::add(dg, v[NFX - 1]);
        else return dg;
    };
    // stencil operand u at a point from its loaded fields
    auto make_u = [&](const T (&v)[NFX]) -> T {
        if constexpr (MODE == SMode::PGR_RICH) {
            // PGR.c:92-98: f = (rhs - Ax) / diag; x_new = x + omega*f
            T f = A::sub(v[1], v[2]);
            if (a.precond) f = A::div(f, dv(v));
            return A::add(v[0], A::mul(om, f));
        } else if constexpr (MODE == SMode::CG_Q) {
            // CG.c:128-150: s = r/diag; d = s + beta*d  (first pass: d as given)
            if (a.cg_first) return v[1];
            T s = v[0];
            if (a.precond) s = A::div(s, dv(v));
            return A::add(s, A::mul(cgb, v[1]));
        } else {
            return v[0];
        }
    };

    T q[PY][QN];
    // prefetch register set: halo + epilogue inputs of one plane and the
    // queue plane R ahead of it.  The loads of plane z+PD are issued while
    // plane z is computed, so PD planes of loads are in flight per block;
    // PD+1 sets rotate (the plane loop is unrolled by PD+1) so that no
    // pending load is copied between registers.
    struct Pre {
        T qn[PY][NFX];
        T hy[NYH][NFX];
        T hx[NXH][NFX];
        T en[PY][NE > 0 ? NE : 1];
        T vv[PY];                                     // VC: a.vf at the plane
    };
    Pre s0, s1;

    // prologue: queue planes z0-R .. z0+R-1
#pragma unroll
    for (int s = 0; s < 2 * R; s++) {
        const size_t zo = (size_t)modp(z0 - R + s, nz) * nxy;
#pragma unroll
        for (int p = 0; p < PY; p++) {
            T v[NFX];
            load_raw(zo, c_off[p], v);
            q[p][s] = make_u(v);
        }
    }
    // halo cells of planes z0 .. z0+R-1 into the ring
#pragma unroll
    for (int s = 0; s < R; s++) {
        const size_t zo = (size_t)modp(z0 + s, nz) * nxy;
#pragma unroll
        for (int k = 0; k < NYH; k++) {
            T v[NFX];
            load_raw(zo, yh_off[k], v);
            hring[s][k][tid] = make_u(v);
        }
#pragma unroll
        for (int k = 0; k < NXH; k++) {
            T v[NFX];
            load_raw(zo, xh_off[k], v);
            hring[s][NYH + k][tid] = make_u(v);
        }
    }
    // queue plane and halo cells of plane zp+R, epilogue inputs of plane zp
    // (indices wrap, so the prefetch past the last plane is a harmless
    // unused load)
    auto prefetch = [&](int zp, Pre &st) {
        const size_t zq = (size_t)modp(zp + R, nz) * nxy, zc = (size_t)modp(zp, nz) * nxy;
#pragma unroll
        for (int p = 0; p < PY; p++) load_raw(zq, c_off[p], st.qn[p]);
#pragma unroll
        for (int k = 0; k < NYH; k++) load_raw(zq, yh_off[k], st.hy[k]);
#pragma unroll
        for (int k = 0; k < NXH; k++) load_raw(zq, xh_off[k], st.hx[k]);
        if constexpr (NE > 0) {
#pragma unroll
            for (int p = 0; p < PY; p++) st.en[p][0] = ld_pl<SA>(a.bvec, zc, c_off[p]);
        }
        if constexpr (VC) {
#pragma unroll
            for (int p = 0; p < PY; p++) st.vv[p] = ld_pl<SA>(a.vf, zc, c_off[p]);
        }
    };

    double acc[NACC];
#pragma unroll
    for (int k = 0; k < NACC; k++) acc[k] = 0.0;

    auto plane = [&](const int z, Pre &cur, Pre &nxt) {
        const int buf = (z - z0) & 1;
        // queue top <- plane z+R; LDS <- plane z (centre + halo)
#pragma unroll
        for (int p = 0; p < PY; p++) {
            q[p][2 * R] = make_u(cur.qn[p]);
            lds[buf][R + ty * PY + p][R + tx] = q[p][R];
        }
        // halo of plane z from the ring; park the cells of plane z+R
        const int rs = (z - z0) % HR, ws = (z - z0 + R) % HR;
#pragma unroll
        for (int k = 0; k < NYH; k++)
            if (yh_ok[k]) lds[buf][yh_row[k]][R + tx] = hring[rs][k][tid];
#pragma unroll
        for (int k = 0; k < NXH; k++)
            if (xh_ok[k]) lds[buf][xh_row[k]][xh_col[k]] = hring[rs][NYH + k][tid];
#pragma unroll
        for (int k = 0; k < NYH; k++) hring[ws][k][tid] = make_u(cur.hy[k]);
#pragma unroll
        for (int k = 0; k < NXH; k++) hring[ws][NYH + k][tid] = make_u(cur.hx[k]);
        prefetch(z + PD, nxt);
        __syncthreads();

            const size_t zoff = (size_t)z * nxy;
#pragma unroll
            for (int p = 0; p < PY; p++) {
                const int lr = R + ty * PY + p, lc = R + tx;
                const T u = q[p][R];
                // this point's row of the LDS tile from column lc - R: the
                // x-neighbours lc -/+ o sit at offsets R -/+ o of one base
                const T *xrow = &lds[buf][lr][tx];
                // y-neighbour at row offset d: rows of this thread come from
                // its own registers, the others from the LDS tile (p and d are
                // compile-time constants after unrolling, so this folds)
                auto ynb = [&](int d) -> T {
                    const int pp = p + d;
                    if (pp >= 0 && pp < PY) return q[pp < 0 ? 0 : (pp < PY ? pp : PY - 1)][R];
                    return lds[buf][lr + d][lc];
                };
                // (FP32: indexed from lc, which keeps it within 128 VGPRs)
                auto xl = [&](int o) -> T {
                    if constexpr (MF) return xrow[R - o];
                    else return lds[buf][lr][lc - o];
                };
                auto xr = [&](int o) -> T {
                    if constexpr (MF) return xrow[R + o];
                    else return lds[buf][lr][lc + o];
                };
                // system.c:195-201 order: centre, then o = 1..6 with
                // (x-, x+, y-, y+, z-, z+) summed left to right
                T y;
                if constexpr (MODE == SMode::AAR_REC && VC) {
                    // variable coefficient, D = diag + c(x): A g = D g +
                    // sum_o w_o S_o, so with the Jacobi step g' = g -
                    // (omega/D) A g = (1 - omega) g - h sum_o w_o S_o, h =
                    // omega/D = a.vf; without a preconditioner a.vf = c and
                    // g' = (1 - omega (3 w0 + c)) g - omega sum_o w_o S_o
                    T ns = (T)0;
#pragma unroll
                    for (int o = 1; o <= R; o++) {
                        T nb = A::add(xl(o), xr(o));
                        nb = A::add(nb, ynb(-o));
                        nb = A::add(nb, ynb(o));
                        nb = A::add(nb, q[p][R - o]);
                        nb = A::add(nb, q[p][R + o]);
                        ns = fma(nb, wt[o], ns);
                    }
                    const T hv = cur.vv[p];
                    if (a.precond) y = fma(-hv, ns, A::mul(u, cf[0]));
                    else y = fma(-om, ns, A::mul(u, A::sub((T)1, A::mul(om, A::add(w3, hv)))));
                } else if constexpr (MODE == SMode::AAR_REC) {
                    // the Richardson step folded into the weights: g' = cf0 g +
                    // sum_o cf_o S_o, cf0 = 1 - c (3 w0 + sigma), cf_o = -c w_o,
                    // c = omega/diag (no reference counterpart; the reference
                    // recomputes b - A x, see the solver)
                    y = A::mul(u, cf[0]);
#pragma unroll
                    for (int o = 1; o <= R; o++) {
                        T nb = A::add(xl(o), xr(o));
                        nb = A::add(nb, ynb(-o));
                        nb = A::add(nb, ynb(o));
                        nb = A::add(nb, q[p][R - o]);
                        nb = A::add(nb, q[p][R + o]);
                        y = fma(nb, cf[o], y);
                    }
                } else if constexpr (sizeof(T) == 8) {
                    y = A::mul(A::mul(u, (T)3), wt0);
#pragma unroll
                    for (int o = 1; o <= R; o++) {
                        T nb = A::add(xl(o), xr(o));
                        nb = A::add(nb, ynb(-o));
                        nb = A::add(nb, ynb(o));
                        nb = A::add(nb, q[p][R - o]);
                        nb = A::add(nb, q[p][R + o]);
                        y = A::add(y, A::mul(nb, wt[o]));
                    }
                } else {
                    // FP32 (no reference counterpart): difference form
                    // sum_o w_o * sum_nb (x_nb - x_c), i.e. the centre weight taken
                    // as -2 sum_o w_o.  Row sums are exactly zero and each rounding
                    // is relative to a local difference, so the smooth modes, whose
                    // eigenvalues are tiny, keep their accuracy; the reference form
                    // would shift them by ~eps32 * |w_0| and stall FP32 iterations.
                    y = (T)0;
#pragma unroll
                    for (int o = 1; o <= R; o++) {
                        T nb = A::add(A::sub(xl(o), u), A::sub(xr(o), u));
                        nb = A::add(nb, A::sub(ynb(-o), u));
                        nb = A::add(nb, A::sub(ynb(o), u));
                        nb = A::add(nb, A::sub(q[p][R - o], u));
                        nb = A::add(nb, A::sub(q[p][R + o], u));
                        y = A::add(y, A::mul(nb, wt[o]));
                    }
                }
                if (MODE != SMode::AAR_REC && a.sigma != 0.0) y = A::add(y, A::mul(sig, u));
                // y + c(x) x, the order of the screened term above
                if constexpr (VC && MODE != SMode::AAR_REC) y = A::add(y, A::mul(cur.vv[p], u));
                // Points past the grid edge (partial tiles) are computed on
                // wrapped inputs.  Such a lane reads exactly what the lane
                // owning the wrapped point reads, through the same operations,
                // so it may store the same bits to the same address.  With
                // USTORE the stores are unconditional (no branch per point, so
                // the memory-counter waits before the next plane stay exact);
                // otherwise they are masked.  Norm terms are always masked.
                const bool ok = c_ok[p];
                const unsigned g = c_off[p];
                // outputs are next read a whole sweep later (beyond the
                // caches at the sizes where it matters): non-temporal stores
                // in the FP64 sweeps (2% faster for the recurrence sweep at
                // 512^3); the FP32 sweeps keep the plane-relative store, whose
                // 32-bit offsets hold them within 128 VGPRs
                auto put = [&](auto *base, auto v) {
                    if (USTORE || ok) {
                        if constexpr (MF) __builtin_nontemporal_store(v, base + zoff + g);
                        else st_pl<SA>(base, zoff, g, v);
                    }
                };
                if constexpr (MODE == SMode::MATVEC) {
                    put(a.out0, y);
                } else if constexpr (MODE == SMode::MATVEC_RES) {
                    // y = A x and ||b - A x||^2 (PGR.c:79-81 start-up)
                    put(a.out0, y);
                    acc[0] += nmask(p, A::sq(A::sub(cur.en[p][0], y)));
                } else if constexpr (MODE == SMode::AAR_TRUE) {
                    // AAR.c:80-82/93-94/119-123: f = (rhs - A x)/diag, stored in
                    // T or (mixed precision) rounded to FP32; ||rhs - A x||^2
                    const T r = A::sub(cur.en[p][0], y);
                    acc[0] += nmask(p, A::sq(r));
                    T f = r;
                    if constexpr (VC) {
                        if (a.precond) f = A::div(f, A::add(dg, cur.vv[p]));
                    } else {
                        if (a.precond) f = A::div(f, dg);
                    }
                    if (a.out32) put(a.out32, (float)f);
                    else put(a.out0, f);
                } else if constexpr (MODE == SMode::AAR_REC) {
                    // y = g_{k+1} = (rhs - A(x_k + omega g_k))/diag; scale^2 *
                    // ||g_{k+1}||^2 = ||rhs - A x_{k+1}||^2 for the history
                    put(a.out0, y);
                    if (a.trace) {
                        // ||rhs - A x||: D g' pointwise (VC: D = omega/h, by
                        // the hardware reciprocal -- the trace is a record,
                        // convergence is decided on the AAR_TRUE sweep)
                        double sv;
                        if constexpr (VC)
                            sv = (double)y * (a.precond ? a.omega * __builtin_amdgcn_rcp((double)cur.vv[p]) : 1.0);
                        else if constexpr (MF)
                            sv = (double)y * okts[p];
                        else
                            sv = (double)y * trace_scale;
                        if constexpr (VC || !MF) sv = nmask(p, sv);
                        acc[0] = fma(sv, sv, acc[0]);
                    }
                } else if constexpr (MODE == SMode::RESID) {
                    const T r = A::sub(cur.en[p][0], y);
                    acc[0] += nmask(p, A::sq(r));
                    if (a.out32) put(a.out32, (float)r);
                    if (a.out0) put(a.out0, r);
                } else if constexpr (MODE == SMode::PGR_RICH) {
                    // PGR.c:95-103: x_old=x; x += omega*f; Ax = A x; f = rhs - Ax
                    put(a.out0, u);
                    put(a.out1, y);
                    if (a.trace) acc[0] += ok ? A::sq(A::sub(cur.en[p][0], y)) : 0.0;
                } else if constexpr (MODE == SMode::CG_Q) {
                    // CG.c:94-100: q = A d; alpha partial d.q
                    put(a.out0, u);
                    put(a.out1, y);
                    acc[0] += ok ? (double)u * (double)y : 0.0;
                } else if constexpr (MODE == SMode::CG_RESID) {
                    // CG.c:65-84 / 112-119: r = b - A x; s = r/diag; r.s, r.r
                    const T r = A::sub(cur.en[p][0], y);
                    T s = r;
                    if constexpr (VC) {
                        if (a.precond) s = A::div(s, A::add(dg, cur.vv[p]));
                    } else {
                        if (a.precond) s = A::div(s, dg);
                    }
                    put(a.out0, r);
                    if (a.out1) put(a.out1, s);
                    acc[0] += ok ? (double)r * (double)s : 0.0;
                    acc[1] += ok ? A::sq(r) : 0.0;
                }
            }
        // shift the queue one plane
#pragma unroll
        for (int p = 0; p < PY; p++)
#pragma unroll
            for (int s = 0; s < 2 * R; s++) q[p][s] = q[p][s + 1];
    };

    // named sets (not an array) keep every set in registers
    [[maybe_unused]] Pre s2;
    int z = z0;
    if constexpr (PD == 1) {
        prefetch(z0, s0);
        if constexpr (USTORE) __builtin_amdgcn_s_waitcnt(0);
        for (; z + 1 < z1; z += 2) {
            plane(z, s0, s1);
            plane(z + 1, s1, s0);
        }
        if (z < z1) plane(z, s0, s1);
    } else {
        static_assert(PD == 2, "prefetch distance 1 or 2");
        prefetch(z0, s0);
        prefetch(z0 + 1, s1);
        // the prologue's loads complete here, so the loop's memory waits are
        // set by its own steady state rather than by the prologue's order
        if constexpr (USTORE) __builtin_amdgcn_s_waitcnt(0);
        for (; z + 2 < z1; z += 3) {
            plane(z, s0, s2);
            plane(z + 1, s1, s0);
            plane(z + 2, s2, s1);
        }
        if (z < z1) plane(z, s0, s2);
        if (z + 1 < z1) plane(z + 1, s1, s0);
    }

    if (a.partials) {
        block_reduce_to<NACC, NT>(acc, a.partials + (size_t)item * NACC, red_scratch);
        if (!a.counter) return;                 // two-pass: finish_kernel follows
        if (last_block_arrive(a.counter, gridDim.x)) {
            final_sum<NT>(a.partials, a.nitems, NACC, red_out, red_scratch);
            if (tid == 0) {
                finalize(a.fin, a.ctrl, red_out, NACC);
                *a.counter = 0u;
            }
        }
    }
}

template <typename T, int WY, int PY, SMode MODE, bool VC = false>
__global__ __launch_bounds__(64 * WY) void stencil_kernel(const SArgs<T> a)
{
    if (a.ctrl && a.ctrl->done) return;
    stencil_body<T, WY, PY, MODE, -1, VC>(a, xcd_item(blockIdx.x, a.nitems));
}

}  // namespace aar
