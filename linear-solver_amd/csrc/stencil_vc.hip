// stencil_vc.hip -- the variable-coefficient sweeps: A = a*Lap_FD12 + c(x)
// with a pointwise Jacobi diagonal D(x) = a*(3*coeff[0]) + c(x).
//
// No MPI_REAL counterpart: the reference's operator slot takes any
// Lap_Vec_mult callback (README:25), and this is the device-side operator
// behind libaar_hip's Lap_Vec_mult_vc / Precondition_vc slots.  The sweeps
// are the constant-coefficient ones (stencil_body.h) with one more
// pointwise input in the epilogue:
//   MATVEC    y = a*Lap(x) + c x          (the order of the screened term)
//   AAR_TRUE  r = b - A x; g = r / (diag + c); ||r||^2
//   AAR_REC   g' = (1 - omega) g - h sum_o w_o S_o,  h = omega / D
// and, for PGR/PL2R and CG, MATVEC_RES, CG_RESID (s = r / (diag + c)) and
// PGR_RICH, CG_Q, whose pre-stencil step divides by diag + c at every
// queue and halo point (c loaded there as one more field)
// Kept in their own translation unit so the constant-coefficient kernels
// compile exactly as before.
#include "stencil_body.h"

#include <algorithm>

namespace aar {

template <typename T, SMode M>
hipError_t launch_stencil_vc_t(SArgs<T> a, int wy, int py, hipStream_t st)
{
    stencil_geometry(a, wy, py);
    dim3 grid(a.nitems), block(64 * wy);
    switch (wy * 10 + py) {
    case 42: hipLaunchKernelGGL((stencil_kernel<T, 4, 2, M, true>), grid, block, 0, st, a); break;
    case 44: hipLaunchKernelGGL((stencil_kernel<T, 4, 4, M, true>), grid, block, 0, st, a); break;
    case 82: hipLaunchKernelGGL((stencil_kernel<T, 8, 2, M, true>), grid, block, 0, st, a); break;
    case 84: hipLaunchKernelGGL((stencil_kernel<T, 8, 4, M, true>), grid, block, 0, st, a); break;
    case 48: hipLaunchKernelGGL((stencil_kernel<T, 4, 8, M, true>), grid, block, 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_stencil_vc(SMode m, const SArgs<T> &a, int wy, int py, hipStream_t st)
{
    if (!a.vf || a.sigma != 0.0) return hipErrorInvalidValue;
    if constexpr (sizeof(T) == 4) {
        // FP32: the recurrence sweep of the mixed-precision solver only
        if (m != SMode::AAR_REC) return hipErrorInvalidValue;
        return launch_stencil_vc_t<T, SMode::AAR_REC>(a, wy, py, st);
    } else {
        switch (m) {
        case SMode::MATVEC: return launch_stencil_vc_t<T, SMode::MATVEC>(a, wy, py, st);
        case SMode::AAR_TRUE: return launch_stencil_vc_t<T, SMode::AAR_TRUE>(a, wy, py, st);
        case SMode::AAR_REC: return launch_stencil_vc_t<T, SMode::AAR_REC>(a, wy, py, st);
        case SMode::MATVEC_RES: return launch_stencil_vc_t<T, SMode::MATVEC_RES>(a, wy, py, st);
        case SMode::PGR_RICH: return launch_stencil_vc_t<T, SMode::PGR_RICH>(a, wy, py, st);
        case SMode::CG_Q: return launch_stencil_vc_t<T, SMode::CG_Q>(a, wy, py, st);
        case SMode::CG_RESID: return launch_stencil_vc_t<T, SMode::CG_RESID>(a, wy, py, st);
        default: return hipErrorInvalidValue;
        }
    }
}

template hipError_t launch_stencil_vc<double>(SMode, const SArgs<double> &, int, int, hipStream_t);
template hipError_t launch_stencil_vc<float>(SMode, const SArgs<float> &, int, int, hipStream_t);

// h(x) = omega / (diag + c(x)) with a Jacobi preconditioner, else c(x),
// rounded to Tg: the pointwise input of the recurrence sweep
template <typename Tg>
__global__ __launch_bounds__(256) void vc_field_kernel(long n, const double *c, double diag, double omega,
                                                       int precond, Tg *h)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        h[i] = precond ? (Tg)(omega / (diag + c[i])) : (Tg)c[i];
}

template <typename Tg>
hipError_t launch_vc_field(long n, const double *c, double diag, double omega, int precond, Tg *h,
                           hipStream_t st)
{
    if (n < 1 || !c || !h) return hipErrorInvalidValue;
    const long nb = std::min<long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL((vc_field_kernel<Tg>), dim3((unsigned)nb), dim3(256), 0, st, n, c, diag, omega, precond, h);
    return hipGetLastError();
}
template hipError_t launch_vc_field<double>(long, const double *, double, double, int, double *, hipStream_t);
template hipError_t launch_vc_field<float>(long, const double *, double, double, int, float *, hipStream_t);

// r /= (diag + c(x)): the pointwise Jacobi step (system.c:222-227 with a
// diagonal per point)
__global__ __launch_bounds__(256) void div_vc_kernel(double *r, double diag, const double *c, long n)
{
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) r[i] = r[i] / (diag + c[i]);
}

hipError_t launch_div_vc(double *r, double diag, const double *c, long n, hipStream_t st)
{
    if (n < 1 || !r || !c) return hipErrorInvalidValue;
    const long nb = std::min<long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(div_vc_kernel, dim3((unsigned)nb), dim3(256), 0, st, r, diag, c, n);
    return hipGetLastError();
}

}  // namespace aar
