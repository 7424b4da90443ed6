// stream.hip -- the attainable HBM bandwidth of the device, measured by
// plain streaming kernels (read-only, write-only, copy) with 16-byte
// non-temporal accesses.  It is the ceiling against which the solver
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
// every load live); 1: write dst; 2: copy.
// Grid-stride over n16 16-byte elements, STREAM_U accesses per step issued
// before any is used.
template <int KIND>
__global__ __launch_bounds__(STREAM_NT) void stream_kernel(const d2v *__restrict__ src, d2v *__restrict__ dst,
                                                           long n16)
{
    const long stride = (long)gridDim.x * STREAM_NT;
    long i = (long)blockIdx.x * STREAM_NT + threadIdx.x;
    d2v acc = {0.0, 0.0};
    const d2v fill = {1.0, 2.0};
    for (; i + (STREAM_U - 1) * stride < n16; i += STREAM_U * stride) {
        d2v v[STREAM_U];
        if constexpr (KIND != 1) {
#pragma unroll
            for (int u = 0; u < STREAM_U; u++) v[u] = __builtin_nontemporal_load(src + i + u * stride);
        }
#pragma unroll
        for (int u = 0; u < STREAM_U; u++) {
            if constexpr (KIND == 0) acc += v[u];
            if constexpr (KIND == 1) __builtin_nontemporal_store(fill, dst + i + u * stride);
            if constexpr (KIND == 2) __builtin_nontemporal_store(v[u], dst + i + u * stride);
        }
    }
    for (; i < n16; i += stride) {
        if constexpr (KIND == 0) acc += __builtin_nontemporal_load(src + i);
        if constexpr (KIND == 1) __builtin_nontemporal_store(fill, dst + i);
        if constexpr (KIND == 2) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
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
    hipDeviceProp_t prop;
    int devid = 0;
    hipError_t e = hipGetDevice(&devid);
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, devid);
    if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("stream: ") + hipGetErrorString(e));
    // eight 4-wave blocks per CU, fewer for a small buffer
    const long want = (long)prop.multiProcessorCount * 8;
    const int blocks = (int)std::max(1L, std::min(want, (n16 + STREAM_NT - 1) / STREAM_NT));
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
