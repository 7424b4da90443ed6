// aar_kernels.h -- kernel argument blocks and launchers of libaar_hip.
#pragma once

#include "aar_common.h"

namespace aar {

enum class SMode : int {
    MATVEC = 0,    // y = a*Lap(x) + sigma*x                      (system.c:189-205)
    AAR_RICH = 1,  // r = b - A x; f = r/diag; x_new = x + omega f (AAR.c:93-94,134-139)
    RESID = 2,     // r = b - A x (+ optional fp32 copy), ||r||^2
    PGR_RICH = 3,  // u = x + omega (b - Ax)/diag; Au = A u        (PGR.c:92-103)
    CG_Q = 4,      // d = s + beta d (s = r/diag); q = A d; d.q   (CG.c:94-100,148-150)
    CG_RESID = 5,  // r = b - A x; r.s, r.r                        (CG.c:65-84,112-137)
};

template <typename T>
struct SArgs {
    int nx = 0, ny = 0, nz = 0, nxy = 0;
    int lz = 0;                       // z-chunk length (planes per block)
    int ntx = 0, nty = 0, nzc = 0, nitems = 0;
    double w[R + 1] = {};             // a * coeff_lap[o]
    double sigma = 0.0;               // screened term
    double diag = 1.0;                // Jacobi diagonal a*(3*coeff[0]) + sigma
    double omega = 0.0;
    int precond = 1;
    int write_x = 0;
    int trace = 0;
    int cg_first = 0;
    const T *in0 = nullptr, *in1 = nullptr, *in2 = nullptr;   // stencil operand fields
    const T *bvec = nullptr;          // right-hand side for the epilogue
    T *out0 = nullptr, *out1 = nullptr;
    float *out32 = nullptr;
    const double *cg_beta = nullptr;
    double *partials = nullptr;       // [nitems * NACC]
    unsigned *counter = nullptr;
    Ctrl *ctrl = nullptr;
    FinArgs fin = {};
};

template <typename T>
hipError_t launch_stencil(SMode m, const SArgs<T> &a, int py, hipStream_t st);
template <typename T>
int stencil_items(const SArgs<T> &a, int py);

// ---------------------------------------------------------------------------
// history sweeps
// ---------------------------------------------------------------------------

// One Anderson / projection normal-equation build over ring buffers.
//  kind 0 (AAR, AAR.c:192-212):    DF_j = F[a_j] - F[b_j]; G = DF^T DF (sym), g = DF^T f
//  kind 1 (PGR, PGR.c:188-206):    DX_j = X[a_j]-X[b_j], DF_j = Y[a_j]-Y[b_j]; G = DX^T DF, g = DX^T f
//  kind 2 (PL2R, PL2R.c:187-206):  DF_j = Y[a_j]-Y[b_j]; G = DF^T DF (sym), g = DF^T f
// For kinds 1/2 f = b - Y[fslot] (unpreconditioned residual).
// Columns j with valid[j] == 0 are zero (never written yet, AAR.c:64-65).
template <typename T>
struct GramArgs {
    long n = 0;
    int m = 0;
    int kind = 0;
    const T *F = nullptr;     // ring base: slot s at F + s*ld (kind 0: f; kinds 1/2: X)
    const T *Y = nullptr;     // ring base of A x (kinds 1/2)
    const T *b = nullptr;     // rhs (kinds 1/2)
    long ld = 0;
    int fslot = 0;            // slot of f (kind 0) or of the current Ax (kinds 1/2)
    int sa[MAXM] = {}, sb[MAXM] = {}, valid[MAXM] = {};
    int nblocks = 0;
    double *partials = nullptr;
    unsigned *counter = nullptr;
    Ctrl *ctrl = nullptr;
};

template <typename T>
hipError_t launch_gram(const GramArgs<T> &a, hipStream_t st);

// m x m pseudo-inverse solve on device, one wavefront: s = V W+ V^T g with
// the Golub-Reinsch SVD of tools.c:52-352 (bit-faithful).  Reads
// ctrl->G (allredvec layout), writes ctrl->s, ctrl->w.
hipError_t launch_ls(Ctrl *ctrl, int m, int complex_embed, hipStream_t st);

// AAR.c:225-231 + 116-117:  x_new = (x_k + beta f_k) - sum_j (DX_j + beta DF_j) s_j
template <typename T>
struct UpdArgs {
    long n = 0;
    int m = 0;
    double beta = 0.0;
    const T *X = nullptr, *F = nullptr;
    long ld = 0;
    int xslot = 0, fslot = 0, oslot = 0;
    int sa[MAXM] = {}, sb[MAXM] = {}, valid[MAXM] = {};
    const Ctrl *ctrl = nullptr;
};
template <typename T>
hipError_t launch_aar_update(const UpdArgs<T> &a, hipStream_t st);

// PGR.c:117-130: x += sum_i DX_i s_i; Ax += sum_i DF_i s_i (in place at
// slot `cur`), ||b - Ax||^2 -> FIN_PGR_RES.
template <typename T>
struct PUpdArgs {
    long n = 0;
    int m = 0;
    T *X = nullptr, *Y = nullptr;
    const T *b = nullptr;
    long ld = 0;
    int cur = 0;
    int sa[MAXM] = {}, sb[MAXM] = {}, valid[MAXM] = {};
    int nblocks = 0;
    double *partials = nullptr;
    unsigned *counter = nullptr;
    Ctrl *ctrl = nullptr;
    FinArgs fin = {};
};
template <typename T>
hipError_t launch_pgr_update(const PUpdArgs<T> &a, hipStream_t st);

// ||v||^2 (or v.w) with a finaliser, accumulated in FP64
template <typename T>
struct DotArgs {
    long n = 0;
    const T *v = nullptr, *w = nullptr;
    int nblocks = 0;
    double *partials = nullptr;
    unsigned *counter = nullptr;
    Ctrl *ctrl = nullptr;
    FinArgs fin = {};
};
template <typename T>
hipError_t launch_dot(const DotArgs<T> &a, hipStream_t st);

// CG.c:106-126: x += alpha d; r -= alpha q; partials r.(r/diag), r.r
struct CgUpdArgs {
    long n = 0;
    double *x = nullptr, *r = nullptr;
    const double *d = nullptr, *q = nullptr;
    double diag = 1.0;
    int precond = 1;
    int update_r = 1;          // 0 on restart iterations (r recomputed by a stencil pass)
    int nblocks = 0;
    double *partials = nullptr;
    unsigned *counter = nullptr;
    Ctrl *ctrl = nullptr;
    FinArgs fin = {};
};
hipError_t launch_cg_update(const CgUpdArgs &a, hipStream_t st);

// small pointwise helpers
hipError_t launch_copy_f64(double *dst, const double *src, long n, const Ctrl *ctrl, hipStream_t st);
hipError_t launch_fill_f64(double *dst, double v, long n, hipStream_t st);
hipError_t launch_fill_f32(float *dst, float v, long n, hipStream_t st);
// x64 += e32 (mixed-precision correction)
hipError_t launch_axpy_f64_f32(double *x, const float *e, long n, hipStream_t st);
hipError_t launch_div_f64(double *r, double diag, long n, hipStream_t st);
hipError_t launch_zero_ctrl(Ctrl *c, hipStream_t st);

}  // namespace aar
