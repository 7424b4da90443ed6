// stencil.hip -- the FD12 (radius-6, 37-point) periodic star stencil and
// the kernels that fuse it with the solver sweeps around it.
//
// Reference operator: MPI_REAL/system.c:189-205 (1-rank branch of
// Lap_Vec_mult): y = (x*3)*w0; for o = 1..6:
//   y += (x[i-o] + x[i+o] + x[j-o] + x[j+o] + x[k-o] + x[k+o]) * w_o,
// w_o = a*coeff_lap[o].  The same operation order is used here with
// contraction off, so a stencil result is bit-identical to the reference.
//
// Design (gfx950): a block of 256 threads = 4 wavefronts owns a 64 x TY
// column of the grid (64 = one wavefront across x, TY = 4*PY rows, PY
// rows per thread) over a z-chunk of `lz` planes and marches it in z.
// * z-neighbours come from a per-thread register queue of 13 planes, so
//   every plane of the column is read from HBM once per chunk;
// * x/y-neighbours come from a double-buffered LDS tile of the current
//   plane with a 6-wide halo (corners are not needed by a star stencil),
//   one barrier per plane; halo cells are fetched R planes ahead, together
//   with the queue plane, and parked in a per-thread LDS ring;
// * block ids are mapped XCD-aware onto tiles so that y-adjacent tiles
//   (whose halos overlap) run on the same XCD and share its L2.
// The epilogue of each mode fuses the pointwise work of the solver step
// that consumes A*x, so that step costs one pass over the grid.
#include "stencil_body.h"

namespace aar {

// ---------------------------------------------------------------------------
// host-side launch
// ---------------------------------------------------------------------------

template <typename T>
void stencil_geometry(SArgs<T> &a, int wy, int py)
{
    const int ty = wy * py;
    a.nxy = a.nx * a.ny;
    a.ntx = (a.nx + TX - 1) / TX;
    a.nty = (a.ny + ty - 1) / ty;
    if (a.lz <= 0) a.lz = a.nz;
    if (a.lz > a.nz) a.lz = a.nz;
    a.nzc = (a.nz + a.lz - 1) / a.lz;
    a.nitems = a.ntx * a.nty * a.nzc;
}

template <typename T, SMode M>
hipError_t launch_stencil_t(SArgs<T> a, int wy, int py, hipStream_t st)
{
    stencil_geometry(a, wy, py);
    dim3 grid(a.nitems), block(64 * wy);
    const int key = wy * 10 + py;
    switch (key) {
    case 42: hipLaunchKernelGGL((stencil_kernel<T, 4, 2, M>), grid, block, 0, st, a); break;
    case 44: hipLaunchKernelGGL((stencil_kernel<T, 4, 4, M>), grid, block, 0, st, a); break;
    case 82: hipLaunchKernelGGL((stencil_kernel<T, 8, 2, M>), grid, block, 0, st, a); break;
    case 84: hipLaunchKernelGGL((stencil_kernel<T, 8, 4, M>), grid, block, 0, st, a); break;
    case 48: hipLaunchKernelGGL((stencil_kernel<T, 4, 8, M>), grid, block, 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_stencil(SMode m, const SArgs<T> &a, int wy, int py, hipStream_t st)
{
    if (a.vf) return launch_stencil_vc<T>(m, a, wy, py, st);
    switch (m) {
    case SMode::MATVEC: return launch_stencil_t<T, SMode::MATVEC>(a, wy, py, st);
    case SMode::AAR_TRUE: return launch_stencil_t<T, SMode::AAR_TRUE>(a, wy, py, st);
    case SMode::AAR_REC: return launch_stencil_t<T, SMode::AAR_REC>(a, wy, py, st);
    case SMode::RESID: return launch_stencil_t<T, SMode::RESID>(a, wy, py, st);
    case SMode::PGR_RICH: return launch_stencil_t<T, SMode::PGR_RICH>(a, wy, py, st);
    case SMode::CG_Q: return launch_stencil_t<T, SMode::CG_Q>(a, wy, py, st);
    case SMode::CG_RESID: return launch_stencil_t<T, SMode::CG_RESID>(a, wy, py, st);
    case SMode::MATVEC_RES: return launch_stencil_t<T, SMode::MATVEC_RES>(a, wy, py, st);
    }
    return hipErrorInvalidValue;
}

template <typename T>
int stencil_items(const SArgs<T> &a0, int wy, int py)
{
    SArgs<T> a = a0;
    stencil_geometry(a, wy, py);
    return a.nitems;
}

template hipError_t launch_stencil<double>(SMode, const SArgs<double> &, int, int, hipStream_t);
template hipError_t launch_stencil<float>(SMode, const SArgs<float> &, int, int, hipStream_t);
template int stencil_items<double>(const SArgs<double> &, int, int);
template void stencil_geometry<double>(SArgs<double> &, int, int);
template void stencil_geometry<float>(SArgs<float> &, int, int);
template int stencil_items<float>(const SArgs<float> &, int, int);

}  // namespace aar
