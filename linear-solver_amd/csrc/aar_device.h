// aar_device.h -- device helpers shared by the kernels of libaar_hip.
//
// Reductions are deterministic: every block reduces its own values in a
// fixed butterfly/tree order into a per-work-item slot, and the last block
// to finish sums the slots in a fixed order.  Results therefore do not
// depend on block scheduling, and repeated solves are bit-reproducible.
#pragma once

#include "aar_common.h"

namespace aar {

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Reduce NACC per-thread accumulators over the block; thread k < NACC
// writes dst[k].  `scratch` must hold NACC * (NT/64) doubles.
template <int NACC, int NT = BLOCK>
__device__ inline void block_reduce_to(const double (&acc)[NACC], double *dst, double *scratch)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    constexpr int NW = NT / 64;
#pragma unroll
    for (int k = 0; k < NACC; k++) {
        double v = wave_sum(acc[k]);
        if (lane == 0) scratch[k * NW + wv] = v;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < NACC; k += NT) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NW; w++) s += scratch[k * NW + w];
        dst[k] = s;
    }
}

// Chunk-cyclic walk of the streaming kernels: the index space is cut into
// chunks of CHUNK_V elements; block b takes chunks b, b + gridDim.x, ...
// and walks each one in BLOCK-wide steps, so every block streams contiguous
// memory (a grid stride of one element per thread measured 10-13% slower
// at 512^3 for the many-stream history sweeps).  The element-to-thread
// assignment is fixed, so reductions over it keep a fixed order.
constexpr long CHUNK_V = 2048;
static_assert(CHUNK_V % BLOCK == 0, "a chunk is a whole number of block steps");
__device__ __forceinline__ long chunk_first()
{
    return (long)blockIdx.x * CHUNK_V + threadIdx.x;
}
__device__ __forceinline__ long chunk_next(long i)
{
    return ((i + BLOCK) % CHUNK_V) >= BLOCK ? i + BLOCK : i + BLOCK + (long)(gridDim.x - 1) * CHUNK_V;
}

// Value-major form of block_reduce_to: thread k < NACC writes the block's
// sum of accumulator k to dst[k * stride] (dst = partials + block index,
// stride = number of blocks), so that each value's partials are contiguous.
template <int NACC, int NT = BLOCK>
__device__ inline void block_reduce_to_vm(const double (&acc)[NACC], double *dst, long stride, double *scratch)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    constexpr int NW = NT / 64;
#pragma unroll
    for (int k = 0; k < NACC; k++) {
        double v = wave_sum(acc[k]);
        if (lane == 0) scratch[k * NW + wv] = v;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < NACC; k += NT) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NW; w++) s += scratch[k * NW + w];
        dst[(size_t)k * stride] = s;
    }
}

// Sum value-major partials (partials[k * nitems + i]) over i in a fixed
// order: wave w takes the values k = w, w + NW, ...; its lanes read the
// value's partials contiguously (lane l: i = l, l + 64, ... in order) and
// the wave sums the lanes with a fixed butterfly.  out[k] (shared) is valid
// for every thread after the closing barrier.
template <int NT = BLOCK>
__device__ inline void final_sum_vm(const double *partials, int nitems, int nacc, double *out)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    constexpr int NW = NT / 64;
    for (int k = wv; k < nacc; k += NW) {
        const double *p = partials + (size_t)k * nitems;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int i = lane;
        for (; i + 192 < nitems; i += 256) {
            s0 += p[i];
            s1 += p[i + 64];
            s2 += p[i + 128];
            s3 += p[i + 192];
        }
        for (; i < nitems; i += 64) s0 += p[i];
        const double s = wave_sum((s0 + s1) + (s2 + s3));
        if (lane == 0) out[k] = s;
    }
    __syncthreads();
}

// Signal completion; returns true in exactly one block (the last one),
// after which all other blocks' global writes are visible to it.
__device__ inline bool last_block_arrive(unsigned *counter, unsigned total)
{
    __shared__ int am_last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned prev = atomicAdd(counter, 1u);
        am_last = (prev == total - 1);
    }
    __syncthreads();
    if (am_last) __threadfence();
    return am_last != 0;
}

// Sum partials[item * nacc + k] over items in a fixed order; result in
// out[k] (shared or global), valid for thread 0 after return.
template <int NT = BLOCK>
__device__ inline void final_sum(const double *partials, int nitems, int nacc, double *out, double *scratch)
{
    for (int k = 0; k < nacc; k++) {
        double s = 0.0;
        for (int i = threadIdx.x; i < nitems; i += NT) s += partials[(size_t)i * nacc + k];
        s = wave_sum(s);
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        if (lane == 0) scratch[wv] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < NT / 64; w++) t += scratch[w];
            out[k] = t;
        }
        __syncthreads();
    }
}

// Finalisers (thread 0 of the last block).
__device__ inline void finalize(const FinArgs &f, Ctrl *c, const double *sums, int nacc)
{
    switch (f.fin) {
    case FIN_SUMS:
        for (int k = 0; k < nacc && k < 8; k++) c->red[k] = sums[k];
        break;
    case FIN_NORM_B: {
        const double nb = sqrt(sums[0]);
        c->rhs_norm = nb;
        c->tol_abs = f.tol * nb;          // AAR.c:79 (tol *= rhs_norm)
        // AAR.c:76 starts relres at tol+1 (unscaled); the loop does not run
        // when that is already within the scaled tolerance.
        if (f.check && !(f.tol + 1.0 > c->tol_abs)) {
            c->done = 1;
            c->final_iter = 0;
            c->relres = f.tol + 1.0;
        }
        break;
    }
    case FIN_AAR_RES:
    case FIN_PGR_RES: {
        const double r = sqrt(sums[0]);
        if (f.hist && f.it - 1 >= 0 && f.it - 1 < f.hist_cap) f.hist[f.it - 1] = r;
        if (f.check) {
            c->relres = r;
            if (!(r == r)) {            // NaN residual: stop, flag it
                c->status = 2;
                c->done = 1;
                c->final_iter = f.it - 1;
            } else if (r <= c->tol_abs) {
                c->done = 1;
                c->final_iter = f.it - 1;
            }
        }
        break;
    }
    case FIN_CG_ALPHA:
        c->alpha = c->delta_new / sums[0];   // CG.c:100-104
        break;
    case FIN_CG_BETA: {
        c->delta_old = c->delta_new;         // CG.c:131
        c->delta_new = sums[0];              // CG.c:134
        const double rr = sqrt(sums[1]);     // CG.c:137
        c->relres = rr;
        c->beta = c->delta_new / c->delta_old;  // CG.c:148
        if (f.hist && f.it >= 0 && f.it < f.hist_cap) f.hist[f.it] = rr;
        if (!(rr == rr)) {
            c->status = 2;
            c->done = 1;
            c->final_iter = f.it;
        } else if (!(rr > c->tol_abs)) {
            c->done = 1;
            c->final_iter = f.it;
        }
        break;
    }
    default:
        break;
    }
}

// Non-temporal pair access (two consecutive T) for the history streams: a
// ring slot or correction is read or written once per launch and next
// touched a whole Anderson cycle later, so its lines are not kept in the
// caches (measured 4.5% faster for the anchor update at 512^3)
template <typename T> struct ExtV2;
template <> struct ExtV2<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct ExtV2<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <typename V, typename T>
__device__ __forceinline__ V ld_nt2(const T *p)
{
    using E = typename ExtV2<T>::type;
    const E t = __builtin_nontemporal_load(reinterpret_cast<const E *>(p));
    V v;
    v.x = t.x;
    v.y = t.y;
    return v;
}
template <typename V, typename T>
__device__ __forceinline__ void st_nt2(T *p, V v)
{
    using E = typename ExtV2<T>::type;
    E t;
    t.x = v.x;
    t.y = v.y;
    __builtin_nontemporal_store(t, reinterpret_cast<E *>(p));
}

}  // namespace aar
