// aar_common.h -- shared definitions of libaar_hip (HIP, gfx950).
//
// Layout everywhere: x fastest, idx = i + j*nx + k*nx*ny, periodic in all
// three directions (MPI_REAL/system.c:82, system.c:189-205).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace aar {

constexpr int R = 6;                 // stencil radius: FD12 (FDn = 6, MPI_REAL/main.c:35)
constexpr int TX = 64;               // tile width in x: one wavefront
constexpr int BLOCK = 256;           // threads per block of the streaming (history/dot) kernels
constexpr int MAXM = 32;             // largest Anderson history m of the real solvers (m x m LS)
constexpr int MAXM_MID = 16;         // histories up to this run the single-pass Gram and the middle update forms
constexpr int MAXM_FAST = 8;         // histories up to this use the register-resident Gram instantiations
constexpr int MAXM_CPLX = 16;        // complex solver: m <= 16 (its LS is the 2m x 2m real embedding)
constexpr int MAXM_BATCH = 16;       // batched solves: m <= 16
constexpr int MAXM_LS = 32;          // largest LS system (complex path embeds m -> 2m)
// per-block partial sums of the largest Gram instantiations, which every
// buffer handed to a Gram kernel must hold per block: PGR's full DX^T DF
// (m^2 + m) at m = MAXM, and the complex Gram (2 (m(m+1)/2 + m)) at MAXM_CPLX
constexpr int GRAM_NA_MAX = MAXM * MAXM + MAXM;
constexpr int HGRAM_NA_MAX = 2 * (MAXM_CPLX * (MAXM_CPLX + 1) / 2 + MAXM_CPLX);
constexpr int NXCD = 8;              // MI355X: 8 XCDs, blocks dispatched round-robin

// Per-solve control block in device memory.  Kernels read `done` on entry
// and return at once when it is set, so work enqueued past convergence is
// a no-op.
struct Ctrl {
    int done;          // 1: stop (converged or zero-iteration exit)
    int final_iter;    // iteration count reported when done was set
    int status;        // 0 ok, 1 SVD did not converge, 2 non-finite residual
    int pad0;
    double rhs_norm;   // ||b||
    double tol_abs;    // tol * ||b||
    double relres;     // last checked ||b - A x|| (absolute)
    double alpha, beta, delta_new, delta_old;   // CG scalars
    double red[64];    // reduction results (FIN_SUMS)
    double s[2 * MAXM_LS];                    // LS solution (Anderson / projection coefficients)
    double G[MAXM_LS * MAXM_LS + MAXM_LS];    // last normal-equation matrix, reference allredvec layout
    double w[MAXM_LS];                        // its truncated singular values
    double V[MAXM_LS * MAXM_LS];              // its right singular vectors, row-major
};

// finalisers run by the last block of a reducing kernel (thread 0)
enum Fin : int {
    FIN_NONE = 0,
    FIN_SUMS = 1,       // ctrl->red[k] = sum_k
    FIN_AAR_RES = 2,    // hist[it-1] = sqrt(sum_0); convergence check if `check`
    FIN_NORM_B = 3,     // ctrl->rhs_norm = sqrt(sum_0), tol_abs = tol * rhs_norm
    FIN_PGR_RES = 4,    // like FIN_AAR_RES for the post-projection residual of PGR/PL2R
    FIN_CG_ALPHA = 5,   // alpha = delta_new / sum_0
    FIN_CG_BETA = 6,    // delta_old = delta_new; delta_new = sum_0; relres = sqrt(sum_1); beta; check
};

struct FinArgs {
    int fin;
    int it;            // iteration number the residual belongs to
    int check;         // convergence check enabled for this residual (FIN_NORM_B: apply the
                       // zero-iteration rule of AAR.c:76-79)
    int max_iter;
    double tol;        // relative tolerance (FIN_NORM_B)
    double *hist;      // residual history (may be null)
    int hist_cap;
};

// complex128 as a plain pair (interleaved re, im)
struct cdouble {
    double re, im;
};

__host__ __device__ inline int wrapi(int i, int n)
{
    // valid for -n <= i < 2n (R <= n)
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

// XCD-aware bijection of block ids onto work items: blocks land on XCD
// (b % 8); each XCD receives a contiguous run of item ids so that
// neighbouring tiles share that XCD's L2.
__device__ inline int xcd_item(int b, int nitems)
{
    const int q = nitems / NXCD, r = nitems % NXCD;
    const int d = b % NXCD, k = b / NXCD;
    return d * q + (d < r ? d : r) + k;
}

}  // namespace aar
