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
#include <type_traits>

#include "aar_common.h"
#include "aar_device.h"
#include "aar_kernels.h"

namespace aar {

template <typename T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };


// reference column index of the difference column of age k (AAR.c:97,
// PGR.c:105); identity in DIRECT mode
__device__ inline int ref_col(int it, int k, int m, int direct)
{
    return direct ? k : ((it - 2 - k) % m + m) % m;
}

// reference allredvec layout (AAR.c:192-212, PGR.c:188-206): [j*m + i] =
// col_i . col_j, [m*m + j] = col_j . f; missing columns are zero
template <typename T, int MM, int KIND>
__device__ inline void gram_layout(const GramArgs<T> &a, const double *sums)
{
    constexpr int NG = (KIND == 1) ? MM * MM : MM * (MM + 1) / 2;
    const int m = a.m, nv = a.nvalid, dir = a.direct;
    double *G = a.ctrl->G;
    for (int k = 0; k < m * m + m; k++) G[k] = 0.0;
    if constexpr (KIND == 1) {
        for (int j = 0; j < nv; j++)
            for (int q = 0; q < nv; q++)
                G[ref_col(a.it, j, m, dir) * m + ref_col(a.it, q, m, dir)] = sums[j * MM + q];
    } else {
        for (int j = 0; j < nv; j++)
            for (int q = 0; q <= j; q++) {
                const double v = sums[j * (j + 1) / 2 + q];
                const int cj = ref_col(a.it, j, m, dir), cq = ref_col(a.it, q, m, dir);
                G[cj * m + cq] = v;
                G[cq * m + cj] = v;
            }
    }
    for (int j = 0; j < nv; j++) G[m * m + ref_col(a.it, j, m, dir)] = sums[NG + j];
}

// ---------------------------------------------------------------------------
// normal equations, one pass over the m+1 ring slots
// ---------------------------------------------------------------------------
template <typename T, int MM, int KIND>
__global__ __launch_bounds__(BLOCK) void gram_kernel(const GramArgs<T> a)
{
    constexpr int NG = (KIND == 1) ? MM * MM : MM * (MM + 1) / 2;
    constexpr int NA = NG + MM;
    static_assert(NA <= GRAM_NA_MAX, "partials buffers hold GRAM_NA_MAX per block");
    constexpr int NS = MM + 1;
    __shared__ double scratch[NA * (BLOCK / 64)];
    __shared__ double sums[NA];
    if (a.ctrl->done) return;

    const int m = a.m, nv = a.nvalid, dir = a.direct;
    const T *D = (KIND == 0) ? a.F : a.Y;     // ring of the DF columns
    const T *pd[NS], *px[NS];
    bool use[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        use[k] = dir ? (k < nv) : (k <= nv);
        const size_t off = (size_t)(use[k] ? a.slot[k] : 0) * a.ld;
        pd[k] = D + off;
        px[k] = a.F + off;
    }
    const T *pf = (KIND == 0 ? a.F : a.Y) + (size_t)a.fslot * a.ld;

    double acc[NA];
#pragma unroll
    for (int k = 0; k < NA; k++) acc[k] = 0.0;

    // one point: DF (and DX) columns by age, then all products in FP64
    auto point = [&](T fv, const T (&vd)[NS], const T (&vx)[NS]) {
        double d[MM], e[MM];
#pragma unroll
        for (int k = 0; k < MM; k++) {
            d[k] = 0.0;
            e[k] = 0.0;
            if (k < nv) {
                d[k] = dir ? (double)vd[k] : (double)(vd[k] - vd[k + 1]);   // DF = f - f_old
                if constexpr (KIND == 1)
                    e[k] = dir ? (double)vx[k] : (double)(vx[k] - vx[k + 1]);   // DX = x - x_old
            }
        }
        const double fd = (double)fv;
        if constexpr (KIND == 1) {
#pragma unroll
            for (int j = 0; j < MM; j++) {
#pragma unroll
                for (int q = 0; q < MM; q++) acc[j * MM + q] += e[q] * d[j];   // DX_q . DF_j
                acc[NG + j] += e[j] * fd;                                      // DX_j . f
            }
        } else {
            int t = 0;
#pragma unroll
            for (int j = 0; j < MM; j++) {
#pragma unroll
                for (int q = 0; q <= j; q++) acc[t++] += d[q] * d[j];          // DF_q . DF_j
                acc[NG + j] += d[j] * fd;                                      // DF_j . f
            }
        }
    };

    // A fixed point-to-thread assignment, so the reduction order is fixed:
    // block b walks the chunks b, b + nblocks, ... of CHUNK_V 16-byte
    // vectors, each contiguous (measured 13% faster at 512^3 than a grid
    // stride of one vector per thread: every CU streams contiguous memory).
    // VW consecutive points per 16-byte load when N is a multiple of VW
    // (every ring slot is 256-byte aligned).  The newest slot is both the
    // age-0 column end and f (the residual ring of AAR; the A x ring of
    // PGR/PL2R, f = b - A x): one load.
    constexpr int VW = 16 / sizeof(T);
    using VT = T __attribute__((ext_vector_type(VW)));
    const bool f_is_d0 = !dir && a.fslot == a.slot[0];
    const long n = a.n;
    const long stride = (long)gridDim.x * BLOCK;
    const bool vec_ok = (n % VW) == 0 && (KIND == 0 || ((uintptr_t)a.b % 16) == 0);
    if (vec_ok) {
        const long nq = n / VW;
        for (long iq = chunk_first(); iq < nq; iq = chunk_next(iq)) {
            const long i = VW * iq;
            VT vd[NS], vx[NS];
#pragma unroll
            for (int k = 0; k < NS; k++) {
                vd[k] = (VT)(T)0;
                vx[k] = (VT)(T)0;
                if (use[k]) {
                    // streamed once per launch: not kept in the caches (2-8%
                    // faster at 512^3 than cached loads)
                    vd[k] = __builtin_nontemporal_load(reinterpret_cast<const VT *>(pd[k] + i));
                    if constexpr (KIND == 1) vx[k] = __builtin_nontemporal_load(reinterpret_cast<const VT *>(px[k] + i));
                }
            }
            VT fv = f_is_d0 ? vd[0] : *reinterpret_cast<const VT *>(pf + i);
            if constexpr (KIND != 0) fv = *reinterpret_cast<const VT *>(a.b + i) - fv;   // PGR.c:102-103
#pragma unroll
            for (int e = 0; e < VW; e++) {
                T sd[NS], sx[NS];
#pragma unroll
                for (int k = 0; k < NS; k++) {
                    sd[k] = vd[k][e];
                    sx[k] = vx[k][e];
                }
                point(fv[e], sd, sx);
            }
        }
    } else {
        for (long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) {
            T vd[NS], vx[NS];
#pragma unroll
            for (int k = 0; k < NS; k++) {
                vd[k] = use[k] ? pd[k][i] : (T)0;
                vx[k] = (KIND == 1 && use[k]) ? px[k][i] : (T)0;
            }
            T f = pf[i];
            if constexpr (KIND != 0) f = a.b[i] - f;
            point(f, vd, vx);
        }
    }

    // value-major partials: each value's per-block sums are contiguous
    block_reduce_to_vm<NA>(acc, a.partials + blockIdx.x, a.nblocks, scratch);
    if (!a.counter) return;                     // two-pass: gram_finish_kernel follows
    if (last_block_arrive(a.counter, gridDim.x)) {
        final_sum_vm(a.partials, a.nblocks, NA, sums);
        if (threadIdx.x == 0) {
            gram_layout<T, MM, KIND>(a, sums);
            *a.counter = 0u;
        }
    }
}

// second pass of the normal-equation build: one block sums the per-block
// partials in a fixed order (one wavefront per value, contiguous reads) and
// lays the result out for the m x m solve
template <typename T, int MM, int KIND>
__global__ __launch_bounds__(BLOCK) void gram_finish_kernel(const GramArgs<T> a)
{
    constexpr int NG = (KIND == 1) ? MM * MM : MM * (MM + 1) / 2;
    constexpr int NA = NG + MM;
    __shared__ double sums[NA];
    if (a.ctrl->done) return;
    final_sum_vm(a.partials, a.nblocks, NA, sums);
    if (threadIdx.x == 0) gram_layout<T, MM, KIND>(a, sums);
}

// ---------------------------------------------------------------------------
// Normal equations beyond m = 16 (560 or 1056 products: more than one pass
// of four wavefronts holds in registers).  The products are cut into NSL
// slices of consecutive products; each of a block's four wavefronts owns one
// slice and walks the block's chunks (the later wavefronts' loads come from
// the caches).  One launch (pass) covers four slices, so HBM reads each
// needed ring slot once per pass, and no accumulator leaves the registers.
// ---------------------------------------------------------------------------
template <int MM, int KIND>
struct GramIdx {
    static constexpr int NG = (KIND == 1) ? MM * MM : MM * (MM + 1) / 2;
    static constexpr int NA = NG + MM;
    // product t: column pair (j, q) of G, or (j, -1) for col_j . f
    static constexpr int j_of(int t)
    {
        if (t >= NG) return t - NG;
        if (KIND == 1) return t / MM;
        int j = 0;
        while ((j + 1) * (j + 2) / 2 <= t) j++;
        return j;
    }
    static constexpr int q_of(int t)
    {
        if (t >= NG) return -1;
        if (KIND == 1) return t % MM;
        const int j = j_of(t);
        return t - j * (j + 1) / 2;
    }
};

template <typename T, int MM, int KIND, int W, int NSL>
__device__ __forceinline__ void gram_split_part(const GramArgs<T> &a)
{
    using Ix = GramIdx<MM, KIND>;
    constexpr int NA = Ix::NA;
    constexpr int PER = (NA + NSL - 1) / NSL;
    constexpr int T0 = W * PER < NA ? W * PER : NA;
    constexpr int T1 = (W + 1) * PER < NA ? (W + 1) * PER : NA;
    if constexpr (T1 <= T0) return;     // a slice past the last product
    constexpr int NT = T1 > T0 ? T1 - T0 : 1;
    constexpr int NS = MM + 1;
    const int m = a.m, nv = a.nvalid, dir = a.direct;
    (void)m;
    const T *D = (KIND == 0) ? a.F : a.Y;
    const T *pd[NS], *px[NS];
    bool use[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        use[k] = dir ? (k < nv) : (k <= nv);
        const size_t off = (size_t)(use[k] ? a.slot[k] : 0) * a.ld;
        pd[k] = D + off;
        px[k] = a.F + off;
    }
    const T *pf = (KIND == 0 ? a.F : a.Y) + (size_t)a.fslot * a.ld;
    const bool f_is_d0 = !dir && a.fslot == a.slot[0];
    double acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = 0.0;

    // this wavefront's products at one point
    auto point = [&](T fv, const T (&vd)[NS], const T (&vx)[NS]) {
        double d[MM], e[MM];
#pragma unroll
        for (int k = 0; k < MM; k++) {
            d[k] = 0.0;
            e[k] = 0.0;
            if (k < nv) {
                d[k] = dir ? (double)vd[k] : (double)(vd[k] - vd[k + 1]);        // DF = f - f_old
                if constexpr (KIND == 1)
                    e[k] = dir ? (double)vx[k] : (double)(vx[k] - vx[k + 1]);    // DX = x - x_old
            }
        }
        const double fd = (double)fv;
#pragma unroll
        for (int t = T0; t < T1; t++) {
            constexpr int dummy = 0;
            (void)dummy;
            const int j = Ix::j_of(t), q = Ix::q_of(t);
            const double u = (q < 0) ? fd : (KIND == 1 ? e[q < 0 ? 0 : q] : d[q < 0 ? 0 : q]);
            const double v = (q < 0 && KIND == 1) ? e[j] : d[j];
            acc[t - T0] += u * v;
        }
    };

    const int lane = threadIdx.x & 63;
    constexpr int VW = 16 / sizeof(T);
    using VT = T __attribute__((ext_vector_type(VW)));
    const long n = a.n;
    const bool vec_ok = (n % VW) == 0 && (KIND == 0 || ((uintptr_t)a.b % 16) == 0);
    const long nq = vec_ok ? n / VW : n;
    for (long c0 = (long)blockIdx.x * CHUNK_V; c0 < nq; c0 += (long)gridDim.x * CHUNK_V) {
        const long c1 = c0 + CHUNK_V < nq ? c0 + CHUNK_V : nq;
        if (vec_ok) {
            for (long iq = c0 + lane; iq < c1; iq += 64) {
                const long i = VW * iq;
                VT vd[NS], vx[NS];
#pragma unroll
                for (int k = 0; k < NS; k++) {
                    vd[k] = (VT)(T)0;
                    vx[k] = (VT)(T)0;
                    if (use[k]) {
                        vd[k] = *reinterpret_cast<const VT *>(pd[k] + i);
                        if constexpr (KIND == 1) vx[k] = *reinterpret_cast<const VT *>(px[k] + i);
                    }
                }
                VT fv = f_is_d0 ? vd[0] : *reinterpret_cast<const VT *>(pf + i);
                if constexpr (KIND != 0) fv = *reinterpret_cast<const VT *>(a.b + i) - fv;   // PGR.c:102-103
#pragma unroll
                for (int e = 0; e < VW; e++) {
                    T sd[NS], sx[NS];
#pragma unroll
                    for (int k = 0; k < NS; k++) {
                        sd[k] = vd[k][e];
                        sx[k] = vx[k][e];
                    }
                    point(fv[e], sd, sx);
                }
            }
        } else {
            for (long i = c0 + lane; i < c1; i += 64) {
                T vd[NS], vx[NS];
#pragma unroll
                for (int k = 0; k < NS; k++) {
                    vd[k] = use[k] ? pd[k][i] : (T)0;
                    vx[k] = (KIND == 1 && use[k]) ? px[k][i] : (T)0;
                }
                T f = pf[i];
                if constexpr (KIND != 0) f = a.b[i] - f;
                point(f, vd, vx);
            }
        }
    }
    // this wavefront owns products T0..T1-1: one value-major partial each
#pragma unroll
    for (int t = T0; t < T1; t++) {
        const double v = wave_sum(acc[t - T0]);
        if (lane == 0) a.partials[(size_t)t * a.nblocks + blockIdx.x] = v;
    }
}

// pass PASS of NPASS: wavefront w owns slice 4 PASS + w of 4 NPASS
template <typename T, int MM, int KIND, int PASS = 0, int NPASS = 1>
__global__ __launch_bounds__(BLOCK) void gram_split_kernel(const GramArgs<T> a)
{
    static_assert(BLOCK == 256, "four wavefronts share the products");
    constexpr int NA = GramIdx<MM, KIND>::NA;
    constexpr int NSL = 4 * NPASS;
    static_assert(NA <= GRAM_NA_MAX, "partials buffers hold GRAM_NA_MAX per block");
    __shared__ double sums[NA];
    if (a.ctrl->done) return;
    switch (threadIdx.x >> 6) {
    case 0: gram_split_part<T, MM, KIND, 4 * PASS + 0, NSL>(a); break;
    case 1: gram_split_part<T, MM, KIND, 4 * PASS + 1, NSL>(a); break;
    case 2: gram_split_part<T, MM, KIND, 4 * PASS + 2, NSL>(a); break;
    default: gram_split_part<T, MM, KIND, 4 * PASS + 3, NSL>(a); break;
    }
    // two-pass (gram_finish_kernel follows), always so for several passes
    if (NPASS > 1 || !a.counter) return;
    if (last_block_arrive(a.counter, gridDim.x)) {
        final_sum_vm(a.partials, a.nblocks, NA, sums);
        if (threadIdx.x == 0) {
            gram_layout<T, MM, KIND>(a, sums);
            *a.counter = 0u;
        }
    }
}

// ---------------------------------------------------------------------------
// LDS-staged normal equations for the deep histories (8 < m <= 16).  The
// block loads each chunk of BLOCK points once (one point per thread,
// coalesced, non-temporal), forms the difference columns DF (and DX for
// PGR) and f there, and parks them in LDS; each of the four wavefronts then
// accumulates its quarter of the products over the chunk from LDS.  The
// next chunk's loads are issued before the current chunk's products, so
// HBM streams while they run, and HBM reads each needed ring slot once
// (the wavefront split of gram_split_kernel reads every slot once per
// wavefront, through L2).
// ---------------------------------------------------------------------------
template <typename T, int MM, int KIND>
struct GramLds {
    static constexpr int NC = (KIND == 1) ? 2 * MM + 1 : MM + 1;   // parked columns: DF, (DX), f
    static constexpr int LP = BLOCK;                             // points per chunk
};

template <typename T, int MM, int KIND, int W>
__device__ __forceinline__ void gram_lds_part(const GramArgs<T> &a, double (*sv)[BLOCK])
{
    using Ix = GramIdx<MM, KIND>;
    constexpr int NA = Ix::NA;
    constexpr int NS = MM + 1;
    constexpr int NC = GramLds<T, MM, KIND>::NC;
    constexpr int LP = GramLds<T, MM, KIND>::LP;
    constexpr int PER = (NA + BLOCK / 64 - 1) / (BLOCK / 64);
    constexpr int T0 = W * PER < NA ? W * PER : NA;
    constexpr int T1 = (W + 1) * PER < NA ? (W + 1) * PER : NA;
    constexpr int NT = T1 > T0 ? T1 - T0 : 1;
    const int nv = a.nvalid, dir = a.direct;
    const T *D = (KIND == 0) ? a.F : a.Y;
    const T *pd[NS], *px[NS];
    bool use[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        use[k] = dir ? (k < nv) : (k <= nv);
        const size_t off = (size_t)(use[k] ? a.slot[k] : 0) * a.ld;
        pd[k] = D + off;
        px[k] = a.F + off;
    }
    const T *pf = (KIND == 0 ? a.F : a.Y) + (size_t)a.fslot * a.ld;
    const bool f_is_d0 = !dir && a.fslot == a.slot[0];
    const long n = a.n;
    const long nch = (n + LP - 1) / LP;
    const int lane = threadIdx.x & 63;

    T rd[NS], rx[NS], rf = (T)0, rb = (T)0;
    auto load = [&](long c) {
        const long i = c * LP + threadIdx.x;
        const bool in = i < n;
#pragma unroll
        for (int k = 0; k < NS; k++) {
            rd[k] = (in && use[k]) ? __builtin_nontemporal_load(pd[k] + i) : (T)0;
            if constexpr (KIND == 1) rx[k] = (in && use[k]) ? __builtin_nontemporal_load(px[k] + i) : (T)0;
        }
        rf = in ? (f_is_d0 ? rd[0] : pf[i]) : (T)0;
        if constexpr (KIND != 0) rb = in ? a.b[i] : (T)0;
    };
    // the chunk's columns: DF = f - f_old (DX = x - x_old) by age, and f
    // (PGR.c:102-103: b - Ax), formed in T as the other Gram kernels do
    auto park = [&]() {
        const int t = threadIdx.x;
#pragma unroll
        for (int k = 0; k < MM; k++) {
            sv[k][t] = k < nv ? (dir ? (double)rd[k] : (double)(rd[k] - rd[k + 1])) : 0.0;
            if constexpr (KIND == 1) sv[MM + k][t] = k < nv ? (dir ? (double)rx[k] : (double)(rx[k] - rx[k + 1])) : 0.0;
        }
        if constexpr (KIND != 0) sv[NC - 1][t] = (double)(rb - rf);
        else sv[NC - 1][t] = (double)rf;
    };

    double acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = 0.0;
    long c = blockIdx.x;
    if (c < nch) load(c);
    for (; c < nch; c += gridDim.x) {
        park();
        __syncthreads();
        if (c + gridDim.x < nch) load(c + gridDim.x);
        if constexpr (T1 > T0) {
            for (int pp = lane; pp < LP; pp += 64) {
#pragma unroll
                for (int t = T0; t < T1; t++) {
                    const int j = Ix::j_of(t), q = Ix::q_of(t);
                    const double u = (q < 0) ? sv[NC - 1][pp] : (KIND == 1 ? sv[MM + (q < 0 ? 0 : q)][pp]
                                                                           : sv[q < 0 ? 0 : q][pp]);
                    const double v = (q < 0 && KIND == 1) ? sv[MM + j][pp] : sv[j][pp];
                    acc[t - T0] += u * v;
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = T0; t < T1; t++) {
        const double v = wave_sum(acc[t - T0]);
        if (lane == 0) a.partials[(size_t)t * a.nblocks + blockIdx.x] = v;
    }
}

template <typename T, int MM, int KIND>
__global__ __launch_bounds__(BLOCK) void gram_lds_kernel(const GramArgs<T> a)
{
    static_assert(BLOCK == 256, "four wavefronts share the products");
    constexpr int NA = GramIdx<MM, KIND>::NA;
    static_assert(NA <= GRAM_NA_MAX, "partials buffers hold GRAM_NA_MAX per block");
    __shared__ double sv[GramLds<T, MM, KIND>::NC][BLOCK];
    __shared__ double sums[NA];
    if (a.ctrl->done) return;
    switch (threadIdx.x >> 6) {
    case 0: gram_lds_part<T, MM, KIND, 0>(a, sv); break;
    case 1: gram_lds_part<T, MM, KIND, 1>(a, sv); break;
    case 2: gram_lds_part<T, MM, KIND, 2>(a, sv); break;
    default: gram_lds_part<T, MM, KIND, 3>(a, sv); break;
    }
    if (!a.counter) return;                     // two-pass: gram_finish_kernel follows
    if (last_block_arrive(a.counter, gridDim.x)) {
        final_sum_vm(a.partials, a.nblocks, NA, sums);
        if (threadIdx.x == 0) {
            gram_layout<T, MM, KIND>(a, sums);
            *a.counter = 0u;
        }
    }
}

template <typename T, int MM, int KIND, int NPASS, int PASS = 0>
void launch_split_passes(const GramArgs<T> &a, dim3 grid, dim3 block, hipStream_t st)
{
    hipLaunchKernelGGL((gram_split_kernel<T, MM, KIND, PASS, NPASS>), grid, block, 0, st, a);
    if constexpr (PASS + 1 < NPASS) launch_split_passes<T, MM, KIND, NPASS, PASS + 1>(a, grid, block, st);
}

template <typename T>
hipError_t launch_gram(const GramArgs<T> &a, hipStream_t st)
{
    if (a.m < 1 || a.m > MAXM || a.nvalid < 0 || a.nvalid > a.m) return hipErrorInvalidValue;
    dim3 grid(a.nblocks), block(BLOCK);
    // the reference's default history m = 7 gets its own instantiation: 35
    // accumulators and 8 loads per point instead of MAXM_FAST's 44 and 9
    if (a.kind == 0 && a.m == 7) {
        hipLaunchKernelGGL((gram_kernel<T, 7, 0>), grid, block, 0, st, a);
        if (!a.counter) hipLaunchKernelGGL((gram_finish_kernel<T, 7, 0>), dim3(1), block, 0, st, a);
        return hipGetLastError();
    }
    if (a.kind < 0 || a.kind > 2) return hipErrorInvalidValue;
    // m <= 8: every accumulator in registers
    auto go = [&](auto mm) {
        constexpr int M = decltype(mm)::value;
        switch (a.kind) {
        case 0: hipLaunchKernelGGL((gram_kernel<T, M, 0>), grid, block, 0, st, a); break;
        case 1: hipLaunchKernelGGL((gram_kernel<T, M, 1>), grid, block, 0, st, a); break;
        default: hipLaunchKernelGGL((gram_kernel<T, M, 2>), grid, block, 0, st, a); break;
        }
        if (!a.counter) {
            switch (a.kind) {
            case 0: hipLaunchKernelGGL((gram_finish_kernel<T, M, 0>), dim3(1), block, 0, st, a); break;
            case 1: hipLaunchKernelGGL((gram_finish_kernel<T, M, 1>), dim3(1), block, 0, st, a); break;
            default: hipLaunchKernelGGL((gram_finish_kernel<T, M, 2>), dim3(1), block, 0, st, a); break;
            }
        }
    };
    auto go_lds = [&](auto mm) {
        constexpr int M = decltype(mm)::value;
        switch (a.kind) {
        case 0: hipLaunchKernelGGL((gram_lds_kernel<T, M, 0>), grid, block, 0, st, a); break;
        case 1: hipLaunchKernelGGL((gram_lds_kernel<T, M, 1>), grid, block, 0, st, a); break;
        default: hipLaunchKernelGGL((gram_lds_kernel<T, M, 2>), grid, block, 0, st, a); break;
        }
        if (!a.counter) {
            switch (a.kind) {
            case 0: hipLaunchKernelGGL((gram_finish_kernel<T, M, 0>), dim3(1), block, 0, st, a); break;
            case 1: hipLaunchKernelGGL((gram_finish_kernel<T, M, 1>), dim3(1), block, 0, st, a); break;
            default: hipLaunchKernelGGL((gram_finish_kernel<T, M, 2>), dim3(1), block, 0, st, a); break;
            }
        }
    };
    // 8 < m <= 16: the LDS-staged form, with an exact instantiation for the
    // READMEs' m = 13.  At 512^3, m = 13: AAR 2.63 ms against 2.91 ms with
    // every product in one thread's registers; PGR's square Gram 5.7 ms
    // against 10.9 ms split over wavefronts reading through L2
    if (a.m <= MAXM_FAST) go(std::integral_constant<int, MAXM_FAST>());
    else if (a.m <= MAXM_MID) {
        if (a.m == 13) go_lds(std::integral_constant<int, 13>());
        else go_lds(std::integral_constant<int, MAXM_MID>());
    } else {
        // 16 < m <= 32: 560 (AAR, PL2R) or 1056 (PGR) products, four and
        // eight passes of four slices, then the fixed-order finish
        switch (a.kind) {
        case 0: launch_split_passes<T, MAXM, 0, 4>(a, grid, block, st); break;
        case 1: launch_split_passes<T, MAXM, 1, 8>(a, grid, block, st); break;
        default: launch_split_passes<T, MAXM, 2, 4>(a, grid, block, st); break;
        }
        switch (a.kind) {
        case 0: hipLaunchKernelGGL((gram_finish_kernel<T, MAXM, 0>), dim3(1), block, 0, st, a); break;
        case 1: hipLaunchKernelGGL((gram_finish_kernel<T, MAXM, 1>), dim3(1), block, 0, st, a); break;
        default: hipLaunchKernelGGL((gram_finish_kernel<T, MAXM, 2>), dim3(1), block, 0, st, a); break;
        }
    }
    return hipGetLastError();
}

template hipError_t launch_gram<double>(const GramArgs<double> &, hipStream_t);
template hipError_t launch_gram<float>(const GramArgs<float> &, hipStream_t);

// ---------------------------------------------------------------------------
// m x m pseudo-inverse solve: the Golub-Reinsch SVD of MPI_REAL/tools.c
// (tools.c:113-352) and x = V diag(1/w) V^T b (tools.c:52-105).  U is
// never formed: the reference replaces it by V (tools.c:324-329) and no
// U-only step feeds w, the superdiagonal or V.
// ---------------------------------------------------------------------------
__device__ __attribute__((always_inline)) inline double d_pythag(double a, double b)
{
    const double absa = fabs(a), absb = fabs(b);
    if (absa > absb) return absa * sqrt(1.0 + absb * absb / (absa * absa));
    return absb == 0.0 ? 0.0 : absb * sqrt(1.0 + absa * absa / (absb * absb));
}
__device__ inline double d_sign(double mag, double s) { return s >= 0.0 ? fabs(mag) : -fabs(mag); }

// ---------------------------------------------------------------------------
// The pinv is executed by one wavefront.  The operation sequence is the
// reference's (bit-identical results); loops whose iterations are
// independent are spread over the lanes -- lane = column (left reflector,
// right-transform accumulation, the final products) or row (right
// reflector; the QR phase, where each V row is rotated on its own and
// stays in the lane's registers).  Scalars (w, rv1, the reflector and
// rotation coefficients) are computed by every lane alike.  The left
// reflector of tools.c:137 also runs over columns j < i and j = i; column
// i itself is updated before the columns j > i read it, as in the
// sequential loop.  textbook = 1 (the complex solver, no MPI_REAL
// counterpart) runs the textbook step over columns i+1..n-1 only: on some
// embedded Hermitian systems the tools.c order does not converge.
// ---------------------------------------------------------------------------
template <int M, int k>
__device__ __attribute__((always_inline)) inline void wgr_qr_step(double (&w)[M], double (&rv1)[M], double (&vr)[M],
                                                                   const double anorm, const int n, int &status)
{
    double g = 0.0, s = 0.0, f = 0.0, h = 0.0;
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
                    vr[k] = -vr[k];
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
                    {
                        const double vx = vr[j], vz = vr[i];
                        vr[j] = vx * c + vz * s;
                        vr[i] = vz * c - vx * s;
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

template <int M, int k>
struct WGrQr {
    __device__ __attribute__((always_inline)) static inline void run(double (&w)[M], double (&rv1)[M], double (&vr)[M],
                                                                      const double anorm, const int n, int &status)
    {
        wgr_qr_step<M, k>(w, rv1, vr, anorm, n, status);
        WGrQr<M, k - 1>::run(w, rv1, vr, anorm, n, status);
    }
};
template <int M>
struct WGrQr<M, -1> {
    __device__ static inline void run(double (&)[M], double (&)[M], double (&)[M], const double, const int, int &) {}
};

template <int M>
__device__ __forceinline__ void ls_body(Ctrl *c, const int n, const int flags)
{
    __shared__ double A[M][M];
    __shared__ double V[M][M];
    __shared__ double T[M];
    if (c->done) return;
    const int lane = threadIdx.x;
    // FtF[i][j] = allredvec[j*m + i], Ftf[j] = allredvec[m*m + j] (AAR.c:214-221)
    for (int e = lane; e < M * M; e += 64) {
        const int i = e / M, j = e % M;
        A[i][j] = (i < n && j < n) ? c->G[j * n + i] : 0.0;
        V[i][j] = 0.0;
    }
    __syncthreads();

    double w[M], rv1[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
        w[i] = 0.0;
        rv1[i] = 0.0;
    }
    double g = 0.0, scale = 0.0, anorm = 0.0, s = 0.0, f = 0.0, h = 0.0;

    // Householder bidiagonalisation (tools.c:122-176)
#pragma unroll
    for (int i = 0; i < M; i++) {
        if (i < n) {
            const int l = i + 1;
            rv1[i] = scale * g;
            g = s = scale = 0.0;
            for (int k = i; k < n; k++) scale += fabs(A[k][i]);
            if (scale != 0.0) {
                __syncthreads();
                if (lane >= i && lane < n) A[lane][i] /= scale;
                __syncthreads();
                for (int k = i; k < n; k++) s += A[k][i] * A[k][i];
                f = A[i][i];
                g = -d_sign(sqrt(s), f);
                h = f * g - s;
                __syncthreads();
                if (lane == 0) A[i][i] = f - g;
                __syncthreads();
                // columns j = 1 .. n-1 in the order the sequential loop sees
                // column i: j < i, then j = i, then j > i
                auto colupd = [&](int j) {
                    double t = 0.0;
                    for (int k = i; k < n; k++) t += A[k][i] * A[k][j];
                    const double fac = t / h;
                    for (int k = i; k < n; k++) A[k][j] += fac * A[k][i];
                };
                if (!(flags & LS_TEXTBOOK)) {
                    if (lane >= 1 && lane < i && lane < n) colupd(lane);
                    __syncthreads();
                    if (i >= 1 && lane == i) colupd(i);
                    __syncthreads();
                }
                if (lane > i && lane >= 1 && lane < n) colupd(lane);
                __syncthreads();
                if (lane >= i && lane < n) A[lane][i] *= scale;
                __syncthreads();
            }
            w[i] = scale * g;
            g = s = scale = 0.0;
            if (i != n - 1) {
                for (int k = l; k < n; k++) scale += fabs(A[i][k]);
                if (scale != 0.0) {
                    __syncthreads();
                    if (lane >= l && lane < n) A[i][lane] /= scale;
                    __syncthreads();
                    for (int k = l; k < n; k++) s += A[i][k] * A[i][k];
                    f = A[i][l < M ? l : M - 1];
                    g = -d_sign(sqrt(s), f);
                    h = f * g - s;
                    __syncthreads();
                    if (lane == 0) A[i][l < M ? l : M - 1] = f - g;
                    __syncthreads();
                    // rv1[k] = a[i][k] / h, kept by every lane
#pragma unroll
                    for (int k = 0; k < M; k++)
                        if (k >= l && k < n) rv1[k] = A[i][k] / h;
                    if (lane >= l && lane < n) {
                        double t = 0.0;
                        for (int k = l; k < n; k++) t += A[lane][k] * A[i][k];
#pragma unroll
                        for (int k = 0; k < M; k++)
                            if (k >= l && k < n) A[lane][k] += t * rv1[k];
                    }
                    __syncthreads();
                    if (lane >= l && lane < n) A[i][lane] *= scale;
                    __syncthreads();
                }
            }
            {
                const double t = fabs(w[i]) + fabs(rv1[i]);
                anorm = anorm > t ? anorm : t;       // tools.h:21 max()
            }
        }
    }

    // accumulation of right-hand transforms (tools.c:178-196)
#pragma unroll
    for (int i = M - 1; i >= 0; i--) {
        if (i < n) {
            if (i < n - 1) {
                const int l = i + 1;
                const double gg = rv1[l < M ? l : M - 1];
                if (gg != 0.0) {
                    const double ail = A[i][l < M ? l : M - 1];
                    if (lane >= l && lane < n) V[lane][i] = (A[i][lane] / ail) / gg;
                    __syncthreads();
                    if (lane >= l && lane < n) {
                        double t = 0.0;
                        for (int k = l; k < n; k++) t += A[i][k] * V[k][lane];
                        for (int k = l; k < n; k++) V[k][lane] += t * V[k][i];
                    }
                    __syncthreads();
                }
                if (lane >= l && lane < n) {
                    V[i][lane] = 0.0;
                    V[lane][i] = 0.0;
                }
                __syncthreads();
            }
            if (lane == 0) V[i][i] = 1.0;
            __syncthreads();
        }
    }

    // implicit-shift QR on the bidiagonal (tools.c:254-350): lane = V row
    double vr[M];
#pragma unroll
    for (int j = 0; j < M; j++) vr[j] = (lane < n) ? V[lane < M ? lane : 0][j] : 0.0;
    int status = 0;
    WGrQr<M, M - 1>::run(w, rv1, vr, anorm, n, status);
    __syncthreads();
    if (lane < n) {
#pragma unroll
        for (int j = 0; j < M; j++) V[lane][j] = vr[j];
    }
    __syncthreads();

    // truncation (tools.c:331-335: w < n w_max 2.22044605e-16; with
    // LS_GELSD, LAPACK dgelsd's rcond = -1 rule of the PETSc variant:
    // w <= 2^-53 w_max) and x = V diag(1/w) V^T b (tools.c:76-94)
    double wmax = 0.0;
#pragma unroll
    for (int j = 0; j < M; j++)
        if (j < n && w[j] > wmax) wmax = w[j];
    if (flags & LS_GELSD) {
        const double tol = 1.1102230246251565e-16 * wmax;
#pragma unroll
        for (int j = 0; j < M; j++)
            if (j < n && w[j] <= tol) w[j] = 0.0;
    } else {
        const double wmin = n * wmax * (2.22044605e-16);
#pragma unroll
        for (int j = 0; j < M; j++)
            if (j < n && w[j] < wmin) w[j] = 0.0;
    }
    if (lane < n) {
        double wl = 0.0;
#pragma unroll
        for (int j = 0; j < M; j++)
            if (j == lane) wl = w[j];
        double t = 0.0;
        if (wl != 0.0) {
            for (int i = 0; i < n; i++) t += V[i][lane] * c->G[n * n + i];
            t /= wl;
        }
        T[lane] = t;
    }
    __syncthreads();
    if (lane < n) {
        double t = 0.0;
        for (int jj = 0; jj < n; jj++) t += V[lane][jj] * T[jj];
        c->s[lane] = t;
        double wl = 0.0;
#pragma unroll
        for (int j = 0; j < M; j++)
            if (j == lane) wl = w[j];
        c->w[lane] = wl;
        for (int q = 0; q < n; q++) c->V[lane * n + q] = V[lane][q];
    }
    if (status && lane == 0) {
        // tools.c:270 exits the process; here the solve stops with a status
        c->status = 1;
        c->done = 1;
    }
}

template <int M>
__global__ __launch_bounds__(64) void ls_kernel(Ctrl *c, int n, int flags)
{
    ls_body<M>(c, n, flags);
}

// one wavefront per solve of a batch
template <int M>
__global__ __launch_bounds__(64) void ls_batch_kernel(Ctrl *const *ctrls, int n)
{
    ls_body<M>(ctrls[blockIdx.x], n, 0);
}

hipError_t launch_ls_batch(Ctrl *const *ctrls, int count, int n, hipStream_t st)
{
    if (n < 1 || n > MAXM_LS || count < 1) return hipErrorInvalidValue;
    if (n > MAXM_BATCH) return hipErrorInvalidValue;
    if (n <= 8) hipLaunchKernelGGL((ls_batch_kernel<8>), dim3(count), dim3(64), 0, st, ctrls, n);
    else hipLaunchKernelGGL((ls_batch_kernel<16>), dim3(count), dim3(64), 0, st, ctrls, n);
    return hipGetLastError();
}

hipError_t launch_ls(Ctrl *ctrl, int n, int flags, hipStream_t st)
{
    if (n < 1 || n > MAXM_LS) return hipErrorInvalidValue;
    if (n <= 8) hipLaunchKernelGGL((ls_kernel<8>), dim3(1), dim3(64), 0, st, ctrl, n, flags);
    else if (n <= 16) hipLaunchKernelGGL((ls_kernel<16>), dim3(1), dim3(64), 0, st, ctrl, n, flags);
    else hipLaunchKernelGGL((ls_kernel<32>), dim3(1), dim3(64), 0, st, ctrl, n, flags);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Anderson update (AAR.c:225-231 then AAR.c:116-117):
//   x_new = (x_k + beta f_k) - sum_k (DX_k + beta DF_k) s_col(k)
// The m-term sum runs in age order (the reference sums in column order;
// the two differ by summation order only).
// ---------------------------------------------------------------------------
template <typename T, int MM>
__global__ __launch_bounds__(BLOCK) void aar_update_kernel(const UpdArgs<T> a)
{
    constexpr int NS = MM + 1;
    if (a.ctrl->done) return;
    const int nv = a.nvalid, dir = a.direct;
    const T *px[NS], *pf[NS];
    bool use[NS];
    double sv[MM];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        use[k] = dir ? (k < nv) : (k <= nv);
        const size_t off = (size_t)(use[k] ? a.slot[k] : 0) * a.ld;
        px[k] = a.X + off;
        pf[k] = a.F + off;
    }
#pragma unroll
    for (int k = 0; k < MM; k++) sv[k] = k < nv ? a.ctrl->s[ref_col(a.it, k, a.m, dir)] : 0.0;
    const T *xk = dir ? a.X + (size_t)a.xslot * a.ld : px[0];
    const T *fk = dir ? a.F + (size_t)a.fslot * a.ld : pf[0];
    T *xo = const_cast<T *>(a.X) + (size_t)a.oslot * a.ld;
    const T beta = (T)a.beta;

    auto point = [&](T xb, T fb, const T (&vx)[NS], const T (&vf)[NS]) -> T {
        double am = 0.0;
#pragma unroll
        for (int k = 0; k < MM; k++) {
            if (k < nv) {
                const T dx = dir ? vx[k] : vx[k] - vx[k + 1];
                const T df = dir ? vf[k] : vf[k] - vf[k + 1];
                am = am + (double)(dx + beta * df) * sv[k];
            }
        }
        const T x0 = xb + beta * fb;
        return (T)((double)x0 - am);
    };

    const long n = a.n;
    const long stride = (long)gridDim.x * BLOCK;
    if ((n & 1) == 0) {
        using V2 = typename Vec2<T>::type;
        const long np = n >> 1;
        for (long ip = chunk_first(); ip < np; ip = chunk_next(ip)) {
            const long i = 2 * ip;
            T vx0[NS], vx1[NS], vf0[NS], vf1[NS];
#pragma unroll
            for (int k = 0; k < NS; k++) {
                vx0[k] = vx1[k] = vf0[k] = vf1[k] = (T)0;
                if (use[k]) {
                    const V2 u = *reinterpret_cast<const V2 *>(px[k] + i);
                    const V2 w = *reinterpret_cast<const V2 *>(pf[k] + i);
                    vx0[k] = u.x;
                    vx1[k] = u.y;
                    vf0[k] = w.x;
                    vf1[k] = w.y;
                }
            }
            V2 xb, fb;
            if (dir) {
                xb = *reinterpret_cast<const V2 *>(xk + i);
                fb = *reinterpret_cast<const V2 *>(fk + i);
            } else {
                xb.x = vx0[0];
                xb.y = vx1[0];
                fb.x = vf0[0];
                fb.y = vf1[0];
            }
            V2 o;
            o.x = point(xb.x, fb.x, vx0, vf0);
            o.y = point(xb.y, fb.y, vx1, vf1);
            *reinterpret_cast<V2 *>(xo + i) = o;
        }
    } else {
        for (long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) {
            T vx[NS], vf[NS];
#pragma unroll
            for (int k = 0; k < NS; k++) {
                vx[k] = use[k] ? px[k][i] : (T)0;
                vf[k] = use[k] ? pf[k][i] : (T)0;
            }
            xo[i] = point(xk[i], fk[i], vx, vf);
        }
    }
}

template <typename T>
hipError_t launch_aar_update(const UpdArgs<T> &a, hipStream_t st)
{
    if (a.m < 1 || a.m > MAXM || a.nvalid < 0 || a.nvalid > a.m) return hipErrorInvalidValue;
    long nb = (a.n / 2 + BLOCK - 1) / BLOCK;
    nb = nb < 1 ? 1 : (nb > 4096 ? 4096 : nb);
    if (a.m <= MAXM_FAST) hipLaunchKernelGGL((aar_update_kernel<T, MAXM_FAST>), dim3((unsigned)nb), dim3(BLOCK), 0, st, a);
    else if (a.m <= MAXM_MID) hipLaunchKernelGGL((aar_update_kernel<T, MAXM_MID>), dim3((unsigned)nb), dim3(BLOCK), 0, st, a);
    else hipLaunchKernelGGL((aar_update_kernel<T, MAXM>), dim3((unsigned)nb), dim3(BLOCK), 0, st, a);
    return hipGetLastError();
}
template hipError_t launch_aar_update<double>(const UpdArgs<double> &, hipStream_t);
template hipError_t launch_aar_update<float>(const UpdArgs<float> &, hipStream_t);

// ---------------------------------------------------------------------------
// Anderson step of the residual-recurrence AAR (AAR.c:109-117 with the
// iterates held as anchor + omega * g increments; see AncArgs).  The term
// coefficients are computed by every thread from uniform data (kernel
// arguments, ctrl->s) with static indices only, so they stay in registers
// without spills; NGM/NCM bound the term counts at compile time.
// ---------------------------------------------------------------------------
// Walks.  The arithmetic of each point is the same in every walk.
// * ONE (FP64 history, at most 8 residual terms and one correction): one
//   pair per thread and one block per NT pairs, x read and written
//   non-temporally (it is next read a whole cycle later); at 512^3 with
//   m = 7, p = 6 this took the update from 2.46 to 2.17 ms.
// * QUAD (FP32 history, FP64 x, the same term counts; the host checked
//   n % 4 == 0 and a 16-byte aligned x): four points per thread, 16-byte
//   history loads, one block per NT quads, every stream non-temporal
//   (1.35 -> 1.29 ms).  One pair per thread there halves the bytes a block
//   keeps in flight: 2.40 ms.
// * Otherwise (the wider forms, and FP32 history off those conditions):
//   chunks of CHUNK_V pairs, block b taking chunks b, b + grid, ..., over at
//   most 4096 blocks.  At 100-256 VGPRs one pair per thread measured 7.6
//   against 3.6 ms (m = 13, p = 11).
// * Odd n, or x off a pair: a grid stride of single points.
template <typename Tx, typename Tg, int NGM, int NCM, int NT, bool ONE, bool QUAD = false>
__global__ __launch_bounds__(NT) void anchor_update_kernel(const AncArgs<Tx, Tg> a)
{
    if (a.ctrl && a.ctrl->done) return;
    const int ng = a.ng, nc = a.nc;
    double cg[NGM], cc[NCM > 0 ? NCM : 1];
#pragma unroll
    for (int i = 0; i < NGM; i++) cg[i] = (a.with_beta && i == ng - 1) ? a.beta : 0.0;
#pragma unroll
    for (int c = 0; c < (NCM > 0 ? NCM : 1); c++) cc[c] = 0.0;
    for (int t = 0; t < a.ncol; t++) {
        const double gam = a.ctrl->s[a.col_ref[t]];
        const int j = a.col_j[t], cj = a.col_c[t];
#pragma unroll
        for (int i = 0; i < NGM; i++) {
            // dg_t = g_j - g_{j-1}; dx_t = omega g_{j-1} (Richardson step)
            if (i == j) cg[i] -= a.beta * gam;
            if (i == j - 1) cg[i] += (cj < 0) ? (a.beta - a.omega) * gam : a.beta * gam;
        }
#pragma unroll
        for (int c = 0; c < NCM; c++)
            if (c == cj) cc[c] -= gam;       // dx_t = an earlier Anderson correction
    }
    const int xlo = a.xlo, xhi = a.xhi;
    const double om = a.omega;
    Tg *co = (a.cout >= 0) ? a.C + (size_t)a.cout * a.ld : nullptr;

    const long n = a.n;
    if constexpr (QUAD) {
        static_assert(sizeof(Tg) == 4 && sizeof(Tx) == 8, "four-point form: FP32 history, FP64 x");
        typedef float F4 __attribute__((ext_vector_type(4)));
        typedef double D2 __attribute__((ext_vector_type(2)));
        const long iq = (long)blockIdx.x * NT + threadIdx.x;
        if (iq >= (n >> 2)) return;
        const long e = 4 * iq;
        F4 v[NGM];
#pragma unroll
        for (int i = 0; i < NGM; i++)
            if (i < ng) v[i] = __builtin_nontemporal_load(reinterpret_cast<const F4 *>(a.G + (size_t)a.gslot[i] * a.ld + e));
        F4 w[NCM > 0 ? NCM : 1];
#pragma unroll
        for (int c = 0; c < NCM; c++)
            if (c < nc) w[c] = __builtin_nontemporal_load(reinterpret_cast<const F4 *>(a.C + (size_t)a.cslot[c] * a.ld + e));
        const D2 x01 = __builtin_nontemporal_load(reinterpret_cast<const D2 *>(a.x + e));
        const D2 x23 = __builtin_nontemporal_load(reinterpret_cast<const D2 *>(a.x + e + 2));
        double sx[4] = {0.0, 0.0, 0.0, 0.0}, sc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int i = 0; i < NGM; i++)
            if (i < ng) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const double g = (double)v[i][q];
                    if (i >= xlo && i < xhi) sx[q] += g;
                    sc[q] += cg[i] * g;
                }
            }
#pragma unroll
        for (int c = 0; c < NCM; c++)
            if (c < nc) {
#pragma unroll
                for (int q = 0; q < 4; q++) sc[q] += cc[c] * (double)w[c][q];
            }
        D2 o01, o23;
        o01.x = x01.x + (om * sx[0] + sc[0]);
        o01.y = x01.y + (om * sx[1] + sc[1]);
        o23.x = x23.x + (om * sx[2] + sc[2]);
        o23.y = x23.y + (om * sx[3] + sc[3]);
        __builtin_nontemporal_store(o01, reinterpret_cast<D2 *>(a.x + e));
        __builtin_nontemporal_store(o23, reinterpret_cast<D2 *>(a.x + e + 2));
        if (co) {
            F4 c4;
            c4.x = (float)sc[0];
            c4.y = (float)sc[1];
            c4.z = (float)sc[2];
            c4.w = (float)sc[3];
            __builtin_nontemporal_store(c4, reinterpret_cast<F4 *>(co + e));
        }
        return;
    }
    const long stride = (long)gridDim.x * NT;
    // pairs when n is even and x is aligned to a pair (ring slots always are)
    if ((n & 1) == 0 && ((uintptr_t)a.x % (2 * sizeof(Tx))) == 0) {
        using VX = typename Vec2<Tx>::type;
        using VG = typename Vec2<Tg>::type;
        const long np = n >> 1;
        for (long ip = ONE ? (long)blockIdx.x * NT + threadIdx.x : chunk_first(); ip < np;
             ip = ONE ? np : chunk_next(ip)) {
            const long e = 2 * ip;
            double sx0 = 0.0, sx1 = 0.0, sc0 = 0.0, sc1 = 0.0;
            VG v[NGM];
#pragma unroll
            for (int i = 0; i < NGM; i++)
                if (i < ng) v[i] = ld_nt2<VG>(a.G + (size_t)a.gslot[i] * a.ld + e);
            VG w[NCM > 0 ? NCM : 1];
#pragma unroll
            for (int c = 0; c < NCM; c++)
                if (c < nc) w[c] = ld_nt2<VG>(a.C + (size_t)a.cslot[c] * a.ld + e);
            const VX xv = ONE ? ld_nt2<VX>(a.x + e) : *reinterpret_cast<const VX *>(a.x + e);
#pragma unroll
            for (int i = 0; i < NGM; i++)
                if (i < ng) {
                    const double g0 = (double)v[i].x, g1 = (double)v[i].y;
                    if (i >= xlo && i < xhi) {
                        sx0 += g0;
                        sx1 += g1;
                    }
                    sc0 += cg[i] * g0;
                    sc1 += cg[i] * g1;
                }
#pragma unroll
            for (int c = 0; c < NCM; c++)
                if (c < nc) {
                    sc0 += cc[c] * (double)w[c].x;
                    sc1 += cc[c] * (double)w[c].y;
                }
            VX xo;
            xo.x = (Tx)((double)xv.x + (om * sx0 + sc0));
            xo.y = (Tx)((double)xv.y + (om * sx1 + sc1));
            if (ONE) st_nt2(a.x + e, xo);
            else *reinterpret_cast<VX *>(a.x + e) = xo;
            if (co) {
                VG c2;
                c2.x = (Tg)sc0;
                c2.y = (Tg)sc1;
                st_nt2(co + e, c2);
            }
        }
    } else {
        for (long e = (long)blockIdx.x * NT + threadIdx.x; e < n; e += stride) {
            double sx = 0.0, sc = 0.0;
#pragma unroll
            for (int i = 0; i < NGM; i++)
                if (i < ng) {
                    const double g = (double)a.G[(size_t)a.gslot[i] * a.ld + e];
                    if (i >= xlo && i < xhi) sx += g;
                    sc += cg[i] * g;
                }
#pragma unroll
            for (int c = 0; c < NCM; c++)
                if (c < nc) sc += cc[c] * (double)a.C[(size_t)a.cslot[c] * a.ld + e];
            a.x[e] = (Tx)((double)a.x[e] + (om * sx + sc));
            if (co) co[e] = (Tg)sc;
        }
    }
}

template <typename Tx, typename Tg>
hipError_t launch_anchor_update(const AncArgs<Tx, Tg> &a, hipStream_t st)
{
    if (a.ng < 1 || a.ng > NGMAX || a.nc < 0 || a.nc > NCMAX || a.ncol < 0 || a.ncol > MAXM)
        return hipErrorInvalidValue;
    if (a.xlo < 0 || a.xhi > a.ng || !a.G || !a.x) return hipErrorInvalidValue;
    if ((a.nc > 0 || a.cout >= 0) && !a.C) return hipErrorInvalidValue;
    if (a.ncol > 0 && !a.ctrl) return hipErrorInvalidValue;
    for (int t = 0; t < a.ncol; t++)
        if (a.col_j[t] < 1 || a.col_j[t] >= a.ng || a.col_c[t] >= a.nc || a.col_ref[t] < 0 ||
            a.col_ref[t] >= MAXM)
            return hipErrorInvalidValue;
    // FP64 history: one block per 1024 pairs where the terms are at most 8
    // residuals and one correction (m = 7, p = 6 among them).  FP32 history with FP64
    // x: four points per thread in that common form when n % 4 == 0 and x
    // is 16-byte aligned; otherwise a chunk walk over at most 4096 blocks.
    // The scalar fallback (odd n, x off a pair) walks a grid stride over
    // the same grid.
    constexpr bool ONE = sizeof(Tg) == 8;
    const long np = (a.n + 1) / 2;
    auto blocks = [np](long nt) {
        if (!ONE) {
            const long nb = (np + CHUNK_V - 1) / CHUNK_V;
            return (unsigned)(nb < 1 ? 1 : (nb > 4096 ? 4096 : nb));
        }
        return (unsigned)((np + nt - 1) / nt);
    };
    if (np > 0x7fffffffL * 256L) return hipErrorInvalidValue;
    constexpr int NT8 = ONE ? 1024 : BLOCK;
    if constexpr (sizeof(Tg) == 4 && sizeof(Tx) == 8) {
        // FP32 history, FP64 x: four points per thread where n and x allow
        // (ring slots are 256-byte aligned)
        if (a.ng <= 8 && a.nc <= 1 && a.n % 4 == 0 && (uintptr_t)a.x % 16 == 0) {
            const long nq = a.n / 4;
            hipLaunchKernelGGL((anchor_update_kernel<Tx, Tg, 8, 1, 1024, false, true>),
                               dim3((unsigned)((nq + 1023) / 1024)), dim3(1024), 0, st, a);
            return hipGetLastError();
        }
    }
    // the wider forms keep the chunk walk: at 100-256 VGPRs (2-3 waves per
    // SIMD) one pair per thread keeps too few bytes in flight (m = 13,
    // p = 11 at 512^3: 7.6 ms against 2.0 ms for the walk)
    const long nwalk = std::max(1L, std::min(4096L, (np + CHUNK_V - 1) / CHUNK_V));
    if (a.ng <= 8 && a.nc <= 1)
        hipLaunchKernelGGL((anchor_update_kernel<Tx, Tg, 8, 1, NT8, ONE>), dim3(blocks(NT8)), dim3(NT8), 0, st, a);
    else if (a.ng <= MAXM_FAST + 2 && a.nc <= MAXM_FAST + 1)
        hipLaunchKernelGGL((anchor_update_kernel<Tx, Tg, MAXM_FAST + 2, MAXM_FAST + 1, BLOCK, false>),
                           dim3((unsigned)nwalk), dim3(BLOCK), 0, st, a);
    else if (a.ng <= MAXM_MID + 2 && a.nc <= MAXM_MID + 1)
        hipLaunchKernelGGL((anchor_update_kernel<Tx, Tg, MAXM_MID + 2, MAXM_MID + 1, BLOCK, false>),
                           dim3((unsigned)nwalk), dim3(BLOCK), 0, st, a);
    else
        hipLaunchKernelGGL((anchor_update_kernel<Tx, Tg, NGMAX, NCMAX, BLOCK, false>), dim3((unsigned)nwalk),
                           dim3(BLOCK), 0, st, a);
    return hipGetLastError();
}
template hipError_t launch_anchor_update<double, double>(const AncArgs<double, double> &, hipStream_t);
template hipError_t launch_anchor_update<double, float>(const AncArgs<double, float> &, hipStream_t);
template hipError_t launch_anchor_update<float, float>(const AncArgs<float, float> &, hipStream_t);

// ---------------------------------------------------------------------------
// PGR / PL2R projection update (PGR.c:117-133): x += sum DX s; Ax += sum DF s
// in place at the age-0 slot; ||b - Ax||^2 for the convergence check.
// Every stream is read and written non-temporally (each is next touched a
// whole sweep later).  ONE (m <= 8): one pair per thread, one block and one
// partial per NT pairs; at 512^3 with m = 7 this measured 3.71 ms against
// 4.11 ms for the chunk walk over 2048 blocks with plain accesses.  Deeper
// histories keep the chunk walk (BLOCK threads, many more registers).
// ---------------------------------------------------------------------------
constexpr int PGR_ONE_NT = 1024;

int pgr_update_blocks(long n, int m, int nb)
{
    if (m > MAXM_FAST) return nb;
    const long g = ((n + 1) / 2 + PGR_ONE_NT - 1) / PGR_ONE_NT;
    return (int)std::max(1L, g);
}

template <typename T, int MM, int NT, bool ONE>
__global__ __launch_bounds__(NT) void pgr_update_kernel(const PUpdArgs<T> a)
{
    constexpr int NS = MM + 1;
    __shared__ double scratch[NT / 64];
    __shared__ double sums[1];
    if (a.ctrl->done) return;
    const int nv = a.nvalid;
    const T *px[NS], *py[NS];
    bool use[NS];
    T sv[MM];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        use[k] = k <= nv;
        const size_t off = (size_t)(use[k] ? a.slot[k] : 0) * a.ld;
        px[k] = a.X + off;
        py[k] = a.Y + off;
    }
#pragma unroll
    for (int k = 0; k < MM; k++) sv[k] = k < nv ? (T)a.ctrl->s[ref_col(a.it, k, a.m, 0)] : (T)0;
    T *xc = a.X + (size_t)a.slot[0] * a.ld;
    T *yc = a.Y + (size_t)a.slot[0] * a.ld;
    double acc[1] = {0.0};

    auto point = [&](const T (&vx)[NS], const T (&vy)[NS], T &xv, T &yv) {
        xv = vx[0];
        yv = vy[0];
#pragma unroll
        for (int k = 0; k < MM; k++) {
            if (k < nv) {
                xv = xv + (vx[k] - vx[k + 1]) * sv[k];
                yv = yv + (vy[k] - vy[k + 1]) * sv[k];
            }
        }
    };

    const long n = a.n;
    const long stride = (long)gridDim.x * NT;
    if ((n & 1) == 0) {
        using V2 = typename Vec2<T>::type;
        const long np = n >> 1;
        for (long ip = ONE ? (long)blockIdx.x * NT + threadIdx.x : chunk_first(); ip < np;
             ip = ONE ? np : chunk_next(ip)) {
            const long i = 2 * ip;
            T vx0[NS], vx1[NS], vy0[NS], vy1[NS];
#pragma unroll
            for (int k = 0; k < NS; k++) {
                vx0[k] = vx1[k] = vy0[k] = vy1[k] = (T)0;
                if (use[k]) {
                    const V2 u = ld_nt2<V2>(px[k] + i);
                    const V2 w = ld_nt2<V2>(py[k] + i);
                    vx0[k] = u.x;
                    vx1[k] = u.y;
                    vy0[k] = w.x;
                    vy1[k] = w.y;
                }
            }
            V2 xo, yo;
            point(vx0, vy0, xo.x, yo.x);
            point(vx1, vy1, xo.y, yo.y);
            st_nt2(xc + i, xo);
            st_nt2(yc + i, yo);
            const V2 bv = ld_nt2<V2>(a.b + i);
            const T r0 = bv.x - yo.x, r1 = bv.y - yo.y;            // PGR.c:129-130
            acc[0] += (double)r0 * (double)r0;
            acc[0] += (double)r1 * (double)r1;
        }
    } else {
        for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += stride) {
            T vx[NS], vy[NS];
#pragma unroll
            for (int k = 0; k < NS; k++) {
                vx[k] = use[k] ? px[k][i] : (T)0;
                vy[k] = use[k] ? py[k][i] : (T)0;
            }
            T xv, yv;
            point(vx, vy, xv, yv);
            xc[i] = xv;
            yc[i] = yv;
            const T r = a.b[i] - yv;
            acc[0] += (double)r * (double)r;
        }
    }
    block_reduce_to<1, NT>(acc, a.partials + blockIdx.x, scratch);
    if (!a.counter) return;                     // two-pass: finish_kernel follows
    if (last_block_arrive(a.counter, gridDim.x)) {
        final_sum<NT>(a.partials, a.nblocks, 1, sums, scratch);
        if (threadIdx.x == 0) {
            finalize(a.fin, a.ctrl, sums, 1);
            *a.counter = 0u;
        }
    }
}

template <typename T>
hipError_t launch_pgr_update(const PUpdArgs<T> &a, hipStream_t st)
{
    if (a.m < 1 || a.m > MAXM || a.nvalid < 0 || a.nvalid > a.m) return hipErrorInvalidValue;
    // the grid is the caller's partial count: pgr_update_blocks()
    if (a.m <= MAXM_FAST && a.nblocks != pgr_update_blocks(a.n, a.m, a.nblocks)) return hipErrorInvalidValue;
    if (a.m <= MAXM_FAST)
        hipLaunchKernelGGL((pgr_update_kernel<T, MAXM_FAST, PGR_ONE_NT, true>), dim3(a.nblocks), dim3(PGR_ONE_NT), 0,
                           st, a);
    else if (a.m <= MAXM_MID)
        hipLaunchKernelGGL((pgr_update_kernel<T, MAXM_MID, BLOCK, false>), dim3(a.nblocks), dim3(BLOCK), 0, st, a);
    else hipLaunchKernelGGL((pgr_update_kernel<T, MAXM, BLOCK, false>), dim3(a.nblocks), dim3(BLOCK), 0, st, a);
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
    const T *w = a.w ? a.w : a.v;
    const long stride = (long)gridDim.x * BLOCK;
    for (long i = chunk_first(); i < a.n; i = chunk_next(i))
        acc[0] += (double)a.v[i] * (double)w[i];
    block_reduce_to<1>(acc, a.partials + blockIdx.x, scratch);
    if (!a.counter) return;                     // two-pass: finish_kernel follows
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

__device__ __forceinline__ void cg_point(const CgUpdArgs &a, double alpha, long i, double (&acc)[2])
{
    a.x[i] = a.x[i] + alpha * a.d[i];                // CG.c:106-107
    if (a.update_r) {
        const double r = a.r[i] - alpha * a.q[i];   // CG.c:121-124
        a.r[i] = r;
        double s = r;
        if (a.precond) s = s / (a.c ? a.diag + a.c[i] : a.diag);
        acc[0] += r * s;
        acc[1] += r * r;
    }
}

// One pair per thread and one block (and one pair of partials) per
// CG_ONE_NT pairs, every stream non-temporal; the element arithmetic is the
// scalar form's.  At 512^3 the chunk walk with 8-byte plain accesses took
// 1.24 ms.  Odd n, or a vector off a 16-byte boundary, walks a grid stride
// element by element over the same grid.
constexpr int CG_ONE_NT = 1024;

int cg_update_blocks(long n)
{
    const long g = ((n + 1) / 2 + CG_ONE_NT - 1) / CG_ONE_NT;
    return (int)std::max(1L, g);
}

__global__ __launch_bounds__(CG_ONE_NT) void cg_update_kernel(const CgUpdArgs a)
{
    constexpr int NT = CG_ONE_NT;
    __shared__ double scratch[2 * (NT / 64)];
    __shared__ double sums[2];
    if (a.ctrl->done) return;
    const double alpha = a.ctrl->alpha;
    double acc[2] = {0.0, 0.0};
    const long n = a.n;
    const uintptr_t al = (uintptr_t)a.x | (uintptr_t)a.d | (uintptr_t)(a.update_r ? a.r : a.x) |
                         (uintptr_t)(a.update_r ? a.q : a.x) | (uintptr_t)(a.c ? a.c : a.x);
    if ((n & 1) == 0 && al % 16 == 0) {
        using V2 = typename Vec2<double>::type;
        const long ip = (long)blockIdx.x * NT + threadIdx.x;
        if (ip < (n >> 1)) {
            const long i = 2 * ip;
            V2 xv = ld_nt2<V2>(a.x + i);
            const V2 dv = ld_nt2<V2>(a.d + i);
            xv.x = xv.x + alpha * dv.x;
            xv.y = xv.y + alpha * dv.y;
            st_nt2(a.x + i, xv);
            if (a.update_r) {
                const V2 rv = ld_nt2<V2>(a.r + i), qv = ld_nt2<V2>(a.q + i);
                V2 ro;
                ro.x = rv.x - alpha * qv.x;
                ro.y = rv.y - alpha * qv.y;
                st_nt2(a.r + i, ro);
                double s0 = ro.x, s1 = ro.y;
                if (a.precond) {
                    if (a.c) {
                        const V2 cv = ld_nt2<V2>(a.c + i);
                        s0 = s0 / (a.diag + cv.x);
                        s1 = s1 / (a.diag + cv.y);
                    } else {
                        s0 = s0 / a.diag;
                        s1 = s1 / a.diag;
                    }
                }
                acc[0] += ro.x * s0;
                acc[1] += ro.x * ro.x;
                acc[0] += ro.y * s1;
                acc[1] += ro.y * ro.y;
            }
        }
    } else {
        for (long i = (long)blockIdx.x * NT + threadIdx.x; i < n; i += (long)gridDim.x * NT) cg_point(a, alpha, i, acc);
    }
    if (!a.update_r) return;   // restart: a stencil pass recomputes r and reduces
    block_reduce_to<2, NT>(acc, a.partials + (size_t)blockIdx.x * 2, scratch);
    if (!a.counter) return;                     // two-pass: finish_kernel follows
    if (last_block_arrive(a.counter, gridDim.x)) {
        final_sum<NT>(a.partials, a.nblocks, 2, sums, scratch);
        if (threadIdx.x == 0) {
            finalize(a.fin, a.ctrl, sums, 2);
            *a.counter = 0u;
        }
    }
}

hipError_t launch_cg_update(const CgUpdArgs &a, hipStream_t st)
{
    // the grid is the caller's partial count: cg_update_blocks()
    if (a.nblocks != cg_update_blocks(a.n)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cg_update_kernel, dim3(a.nblocks), dim3(CG_ONE_NT), 0, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// second pass of a reduction: one block sums per-block partials in a fixed
// order and runs the finaliser (no device-scope fence or atomic needed: the
// kernel boundary orders it after the partials)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void finish_kernel(const double *partials, int nitems, int nacc, FinArgs fin,
                                                      Ctrl *ctrl)
{
    __shared__ double scratch[BLOCK / 64];
    __shared__ double sums[8];
    if (ctrl->done) return;
    final_sum(partials, nitems, nacc, sums, scratch);
    if (threadIdx.x == 0) finalize(fin, ctrl, sums, nacc);
}

hipError_t launch_finish(const double *partials, int nitems, int nacc, const FinArgs &fin, Ctrl *ctrl,
                         hipStream_t st)
{
    if (nacc < 1 || nacc > 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(BLOCK), 0, st, partials, nitems, nacc, fin, ctrl);
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
