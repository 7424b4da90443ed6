// noalloc.hip -- the m x m pseudo-inverse solve (tools.c:52-105 semantics,
// the one-wavefront Golub-Reinsch SVD of the solvers) on device pointers,
// for callers that cannot allocate device memory: code under HIP graph
// capture and tight serving loops.  Its work area is `scratch`, one static
// device buffer per GPU; nothing is allocated, nothing is copied to or
// from the host and the host does not wait.
#include <hip/hip_runtime.h>

#include <string>

#include "aar_solver.h"

namespace aar {

// The work area: one LS control block (the normal-equation matrix in the
// reference's allredvec layout, the solution, singular values, V, status).
// 18 KB per GPU, held until the process ends; every call on a GPU shares it
// and leaves its contents behind, so calls that overlap on one GPU (on two
// streams) race.
constexpr int SCRATCH_SLOTS = 1;
static __device__ __align__(256) Ctrl scratch[SCRATCH_SLOTS];

namespace {

constexpr int MAX_DEVICES = 64;
Ctrl *g_scratch[MAX_DEVICES] = {};   // device address of `scratch` per GPU (host code never uses the symbol)

// Clears the whole control block (done, status and every array the solve
// reads), then stages A (row-major, A[i*m + j]) and b in the allredvec
// layout G[j*m + i] = A[i][j], G[m*m + j] = b[j].
__global__ __launch_bounds__(256) void pinv_stage_kernel(Ctrl *c, const double *__restrict__ A,
                                                         const double *__restrict__ b, int m)
{
    unsigned long long *w = reinterpret_cast<unsigned long long *>(c);
    constexpr int NW = (int)(sizeof(Ctrl) / sizeof(unsigned long long));
    static_assert(sizeof(Ctrl) % sizeof(unsigned long long) == 0, "Ctrl is a whole number of words");
    for (int i = threadIdx.x; i < NW; i += blockDim.x) w[i] = 0ull;
    __syncthreads();
    for (int k = threadIdx.x; k < m * m; k += blockDim.x) {
        const int i = k / m, j = k % m;
        c->G[j * m + i] = A[k];
    }
    for (int j = threadIdx.x; j < m; j += blockDim.x) c->G[m * m + j] = b[j];
}

__global__ __launch_bounds__(256) void pinv_unstage_kernel(const Ctrl *c, double *__restrict__ x,
                                                           double *__restrict__ w, double *__restrict__ V,
                                                           int *__restrict__ status, int m)
{
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
        x[j] = c->s[j];
        if (w) w[j] = c->w[j];
    }
    if (V)
        for (int k = threadIdx.x; k < m * m; k += blockDim.x) V[k] = c->V[k];
    if (status && threadIdx.x == 0) *status = c->status;
}

}  // namespace
}  // namespace aar

using namespace aar;

extern "C" int aar_hip_pinv_noalloc(const double *A, const double *b, double *x, int m, double *w, double *V,
                                    int *status)
{
    if (!A || !b || !x) return fail(AAR_HIP_ERR_ARG, "pinv_noalloc: null pointer");
    if (m < 1 || m > MAXM_LS) return fail(AAR_HIP_ERR_ARG, "pinv_noalloc: m must be in 1..32");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("pinv_noalloc: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(AAR_HIP_ERR_ARG, "pinv_noalloc: device index out of range");
    if (!g_scratch[dev]) {
        // once per GPU, with that GPU current: the address of its copy of `scratch`
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(AAR_HIP_ERR_HIP, std::string("pinv_noalloc: scratch address: ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<Ctrl *>(p);
    }
    Ctrl *c = g_scratch[dev];
    // the library stream once the library is set up (aar_hip_set_stream),
    // else the null stream; setting the library up here would allocate
    Context &C = ctx();
    hipStream_t s = C.init ? C.stream : nullptr;
    hipLaunchKernelGGL(pinv_stage_kernel, dim3(1), dim3(256), 0, s, c, A, b, m);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_ls(c, m, 0, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(pinv_unstage_kernel, dim3(1), dim3(256), 0, s, c, x, w, V, status, m);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("pinv_noalloc: ") + hipGetErrorString(e));
    return 0;
}
