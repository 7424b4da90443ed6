// stream.hip -- the attainable HBM bandwidth of the device, measured by
// plain streaming kernels (read-only, write-only, copy) with 16-byte
// accesses, each block streaming one contiguous chunk.  Each kind is timed in
// three launch shapes and the best is reported: it is the ceiling against
// which the solver kernels' achieved bandwidth is read, beside the 8 TB/s
// specification.  On MI355X the shape moves copy between 5.0 and 5.9 TB/s
// and write between 4.4 and 6.1 TB/s; zero and random data stream alike.
#include <hip/hip_runtime.h>

#include <string>

#include "aar_solver.h"

namespace aar {

typedef double d2v __attribute__((ext_vector_type(2)));

// target of the read kernel's never-taken store
__device__ double stream_sink;

// KIND 0: read src (the sum is stored only if it equals a sentinel, which keeps
// every load live); 1: write dst; 2: copy.  Block b streams the contiguous
// chunk [b*chunk, (b+1)*chunk) of 16-byte elements, U accesses per thread
// issued before any is used; loads are non-temporal, stores non-temporal
// when NTS.
template <int KIND, int NT, int U, bool NTS>
__global__ __launch_bounds__(NT) void stream_kernel(const d2v *__restrict__ src, d2v *__restrict__ dst, long n16,
                                                    long chunk)
{
    const long c0 = (long)blockIdx.x * chunk;
    const long c1 = c0 + chunk < n16 ? c0 + chunk : n16;
    d2v acc = {0.0, 0.0};
    const d2v fill = {1.0, 2.0};
    for (long i = c0 + threadIdx.x; i < c1; i += NT * U) {
        d2v v[U];
        if constexpr (KIND != 1) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                const long k = i + u * NT;
                if (k < c1) v[u] = __builtin_nontemporal_load(src + k);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long k = i + u * NT;
            if (k < c1) {
                if constexpr (KIND == 0) acc += v[u];
                if constexpr (KIND != 0) {
                    const d2v w = KIND == 1 ? fill : v[u];
                    if constexpr (NTS) __builtin_nontemporal_store(w, dst + k);
                    else dst[k] = w;
                }
            }
        }
    }
    if constexpr (KIND == 0)
        if (acc.x + acc.y == -1.0) stream_sink = acc.x;
}

// the three shapes timed per kind: (threads, accesses per thread and step,
// chunk in 16-byte elements, non-temporal stores)
struct StreamShape {
    int nt, u;
    long chunk;
    bool nts;
};
constexpr StreamShape STREAM_SHAPES[3] = {{256, 4, 4096, true}, {256, 2, 512, true}, {1024, 4, 16384, false}};

template <int KIND>
hipError_t stream_launch(int shape, const d2v *src, d2v *dst, long n16, hipStream_t s)
{
    const StreamShape &h = STREAM_SHAPES[shape];
    const long nblk = (n16 + h.chunk - 1) / h.chunk;
    const dim3 g((unsigned)nblk);
    if (shape == 0) hipLaunchKernelGGL((stream_kernel<KIND, 256, 4, true>), g, dim3(256), 0, s, src, dst, n16, h.chunk);
    else if (shape == 1) hipLaunchKernelGGL((stream_kernel<KIND, 256, 2, true>), g, dim3(256), 0, s, src, dst, n16, h.chunk);
    else hipLaunchKernelGGL((stream_kernel<KIND, 1024, 4, false>), g, dim3(1024), 0, s, src, dst, n16, h.chunk);
    return hipGetLastError();
}

}  // namespace aar

using namespace aar;

extern "C" int aar_hip_stream_bandwidth(int kind, void *dst, const void *src, long bytes, int reps, double *gbps)
{
    if (int rc = context_init()) return rc;
    if (kind < 0 || kind > 2) return fail(AAR_HIP_ERR_ARG, "stream: kind must be 0 (read), 1 (write) or 2 (copy)");
    if (bytes <= 0 || bytes % 16 != 0) return fail(AAR_HIP_ERR_ARG, "stream: bytes must be a positive multiple of 16");
    if (reps < 1 || !gbps) return fail(AAR_HIP_ERR_ARG, "stream: reps >= 1 and a result pointer are required");
    if ((kind != 1 && !src) || (kind != 0 && !dst))
        return fail(AAR_HIP_ERR_ARG, "stream: null buffer");
    if (((size_t)src | (size_t)dst) % 16 != 0) return fail(AAR_HIP_ERR_ARG, "stream: buffers must be 16-byte aligned");
    const long n16 = bytes / 16;
    if ((n16 + STREAM_SHAPES[1].chunk - 1) / STREAM_SHAPES[1].chunk > 0x7fffffffL)
        return fail(AAR_HIP_ERR_ARG, "stream: buffer too large");
    hipStream_t s = ctx().stream;
    const d2v *sp = static_cast<const d2v *>(src);
    d2v *dp = static_cast<d2v *>(dst);
    auto launch = [&](int shape) {
        if (kind == 0) return stream_launch<0>(shape, sp, dp, n16, s);
        if (kind == 1) return stream_launch<1>(shape, sp, dp, n16, s);
        return stream_launch<2>(shape, sp, dp, n16, s);
    };
    const double moved = (double)bytes * (kind == 2 ? 2.0 : 1.0) * reps;
    double best = 0.0;
    for (int shape = 0; shape < 3; shape++) {
        hipError_t e = launch(shape);                      // warm-up
        if (e == hipSuccess) e = hipEventRecord(ctx().t0, s);
        for (int r = 0; r < reps && e == hipSuccess; r++) e = launch(shape);
        if (e == hipSuccess) e = hipEventRecord(ctx().t1, s);
        if (e == hipSuccess) e = hipEventSynchronize(ctx().t1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx().t0, ctx().t1);
        if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("stream: ") + hipGetErrorString(e));
        const double gb = moved / ((double)ms * 1e-3) / 1e9;
        if (gb > best) best = gb;
    }
    *gbps = best;
    return 0;
}
