// stream.hip -- the attainable HBM bandwidth of the device, measured by
// plain streaming kernels (read-only, write-only, copy) with 16-byte
// non-temporal accesses, each block streaming one contiguous chunk.  It is the ceiling against which the solver
// kernels' achieved bandwidth is read, beside the 8 TB/s specification.
#include <hip/hip_runtime.h>

#include <string>

#include "aar_solver.h"

namespace aar {

typedef double d2v __attribute__((ext_vector_type(2)));

constexpr int STREAM_NT = 256;    // threads per block
constexpr int STREAM_U = 4;       // independent 16-byte accesses in flight per thread and step

// target of the read kernel's never-taken store
__device__ double stream_sink;

// KIND 0: read src (the sum is stored only if it equals a sentinel, which keeps
// every load live); 1: write dst; 2: copy.  Block b streams the contiguous
// chunk [b*STREAM_CHUNK, (b+1)*STREAM_CHUNK) of 16-byte elements, STREAM_U
// accesses per thread issued before any is used (contiguous chunks per block
// measured 9-15% faster than a grid stride on MI355X).
constexpr long STREAM_CHUNK = 4096;
template <int KIND>
__global__ __launch_bounds__(STREAM_NT) void stream_kernel(const d2v *__restrict__ src, d2v *__restrict__ dst,
                                                           long n16)
{
    const long c0 = (long)blockIdx.x * STREAM_CHUNK;
    const long c1 = c0 + STREAM_CHUNK < n16 ? c0 + STREAM_CHUNK : n16;
    d2v acc = {0.0, 0.0};
    const d2v fill = {1.0, 2.0};
    for (long i = c0 + threadIdx.x; i < c1; i += STREAM_NT * STREAM_U) {
        d2v v[STREAM_U];
        if constexpr (KIND != 1) {
#pragma unroll
            for (int u = 0; u < STREAM_U; u++) {
                const long k = i + u * STREAM_NT;
                if (k < c1) v[u] = __builtin_nontemporal_load(src + k);
            }
        }
#pragma unroll
        for (int u = 0; u < STREAM_U; u++) {
            const long k = i + u * STREAM_NT;
            if (k < c1) {
                if constexpr (KIND == 0) acc += v[u];
                if constexpr (KIND == 1) __builtin_nontemporal_store(fill, dst + k);
                if constexpr (KIND == 2) __builtin_nontemporal_store(v[u], dst + k);
            }
        }
    }
    if constexpr (KIND == 0)
        if (acc.x + acc.y == -1.0) stream_sink = acc.x;
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
    const long nblk = (n16 + STREAM_CHUNK - 1) / STREAM_CHUNK;
    if (nblk > 0x7fffffffL) return fail(AAR_HIP_ERR_ARG, "stream: buffer too large");
    const int blocks = (int)nblk;
    hipError_t e = hipSuccess;
    hipStream_t s = ctx().stream;
    auto launch = [&]() {
        const d2v *sp = static_cast<const d2v *>(src);
        d2v *dp = static_cast<d2v *>(dst);
        if (kind == 0) hipLaunchKernelGGL(stream_kernel<0>, dim3(blocks), dim3(STREAM_NT), 0, s, sp, dp, n16);
        else if (kind == 1) hipLaunchKernelGGL(stream_kernel<1>, dim3(blocks), dim3(STREAM_NT), 0, s, sp, dp, n16);
        else hipLaunchKernelGGL(stream_kernel<2>, dim3(blocks), dim3(STREAM_NT), 0, s, sp, dp, n16);
        return hipGetLastError();
    };
    e = launch();                                          // warm-up
    if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("stream: ") + hipGetErrorString(e));
    e = hipEventRecord(ctx().t0, s);
    for (int r = 0; r < reps && e == hipSuccess; r++) e = launch();
    if (e == hipSuccess) e = hipEventRecord(ctx().t1, s);
    if (e == hipSuccess) e = hipEventSynchronize(ctx().t1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx().t0, ctx().t1);
    if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("stream: ") + hipGetErrorString(e));
    const double moved = (double)bytes * (kind == 2 ? 2.0 : 1.0) * reps;
    *gbps = moved / ((double)ms * 1e-3) / 1e9;
    return 0;
}
