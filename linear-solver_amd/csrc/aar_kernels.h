// aar_kernels.h -- kernel argument blocks and launchers of libaar_hip.
#pragma once

#include "aar_common.h"

namespace aar {

enum class SMode : int {
    MATVEC = 0,    // y = a*Lap(x) + sigma*x                      (system.c:189-205)
    AAR_TRUE = 1,  // r = b - A x; g = r/diag (T or FP32 out); ||r||^2 (AAR.c:80-82,119-123)
    RESID = 2,     // r = b - A x (+ optional fp32 copy), ||r||^2
    PGR_RICH = 3,  // u = x + omega (b - Ax)/diag; Au = A u        (PGR.c:92-103)
    CG_Q = 4,      // d = s + beta d (s = r/diag); q = A d; d.q   (CG.c:94-100,148-150)
    CG_RESID = 5,  // r = b - A x; r.s, r.r                        (CG.c:65-84,112-137)
    MATVEC_RES = 6,  // y = A x and ||b - A x||^2                  (PGR.c:79-81)
    AAR_REC = 7,   // g' = g - (omega/diag) A g: the residual of x + omega g (AAR.c:134-139)
};

template <typename T>
struct SArgs {
    int nx = 0, ny = 0, nz = 0, nxy = 0;
    int lz = 0;                       // z-chunk length (planes per block)
    int ntx = 0, nty = 0, nzc = 0, nitems = 0;
    double w[R + 1] = {};             // a * coeff_lap[o]
    double cf[R + 1] = {};            // AAR_REC: g' = cf[0] g + sum_o cf[o] S_o (S_o: the 6 neighbours at o)
    double sigma = 0.0;               // screened term
    double diag = 1.0;                // Jacobi diagonal a*(3*coeff[0]) + sigma
    double omega = 0.0;
    int precond = 1;
    int trace = 0;
    int cg_first = 0;
    const T *in0 = nullptr, *in1 = nullptr, *in2 = nullptr;   // stencil operand fields
    const T *bvec = nullptr;          // right-hand side for the epilogue
    // variable-coefficient operator A = a*Lap + c(x) (the VC sweeps only):
    // MATVEC/AAR_TRUE read c(x); AAR_REC reads h(x) = omega/(diag + c(x))
    // with a Jacobi preconditioner, else c(x)
    const T *vf = nullptr;
    T *out0 = nullptr, *out1 = nullptr;
    float *out32 = nullptr;
    const double *cg_beta = nullptr;
    double *partials = nullptr;       // [nitems * NACC]
    unsigned *counter = nullptr;
    Ctrl *ctrl = nullptr;
    FinArgs fin = {};
};

// wy: wavefronts per block stacked in y (4 or 8); py: rows per thread (2 or 4)
template <typename T>
hipError_t launch_stencil(SMode m, const SArgs<T> &a, int wy, int py, hipStream_t st);
template <typename T>
int stencil_items(const SArgs<T> &a, int wy, int py);
// the variable-coefficient sweeps (MATVEC, AAR_TRUE, AAR_REC with a.vf set)
template <typename T>
hipError_t launch_stencil_vc(SMode m, const SArgs<T> &a, int wy, int py, hipStream_t st);
// h = omega/(diag + c) (precond) or c, rounded to Tg
template <typename Tg>
hipError_t launch_vc_field(long n, const double *c, double diag, double omega, int precond, Tg *h,
                           hipStream_t st);
// r /= (diag + c), pointwise
hipError_t launch_div_vc(double *r, double diag, const double *c, long n, hipStream_t st);
// fills the tile geometry of `a` (nxy, ntx, nty, nzc, nitems; lz clamped)
template <typename T>
void stencil_geometry(SArgs<T> &a, int wy, int py);

// ---------------------------------------------------------------------------
// history sweeps
// ---------------------------------------------------------------------------

// History columns are indexed by AGE: the ring slot of age k holds the
// iterate (or residual) k steps before the newest, and the difference
// column of age k is slot(k) - slot(k+1) -- the reference's
// DX = x - x_old, DF = f - f_old (AAR.c:96-102, PGR.c:105-109).  Columns
// of age < nvalid exist (a prefix; the rest are the reference's never
// written zero columns, AAR.c:64-65).  The reference stores column of age
// k at index (it - 2 - k) mod m (AAR.c:97; PGR.c:105 with it = t), which
// is how results are laid out for the m x m solve.  In DIRECT mode (the
// stand-alone Anderson/Galerkin/L2 entry points) slot k holds column k
// itself and the layout is the identity.
template <typename T>
struct GramArgs {
    long n = 0;
    int m = 0;
    int kind = 0;             // 0 AAR (DF^T DF), 1 PGR (DX^T DF), 2 PL2R (DF^T DF of A x)
    int direct = 0;
    const T *F = nullptr;     // kind 0: residual ring; kinds 1/2: iterate ring
    const T *Y = nullptr;     // kinds 1/2: ring of A x
    const T *b = nullptr;     // kinds 1/2: rhs, f = b - Y[fslot]
    long ld = 0;
    int slot[MAXM + 1] = {};  // ring slot of age k
    int fslot = 0;            // slot of f (kind 0) or of A x (kinds 1/2)
    int nvalid = 0;
    int it = 0;
    int nblocks = 0;
    double *partials = nullptr;
    unsigned *counter = nullptr;
    Ctrl *ctrl = nullptr;
};

template <typename T>
hipError_t launch_gram(const GramArgs<T> &a, hipStream_t st);

// m x m pseudo-inverse solve on device, one wavefront: s = V W+ V^T g with
// the Golub-Reinsch SVD of tools.c:52-352 (bit-faithful).  Reads ctrl->G
// (allredvec layout), writes ctrl->s, ctrl->w.  flags: LS_TEXTBOOK -- the
// textbook left reflector (the complex solver's 2m x 2m real embedding);
// LS_GELSD -- LAPACK dgelsd's rcond = -1 cutoff of the PETSc variant
// (s_j <= 2^-53 s_max dropped) instead of tools.c's (s_j < m s_max eps).
constexpr int LS_TEXTBOOK = 1, LS_GELSD = 2;
hipError_t launch_ls(Ctrl *ctrl, int m, int flags, hipStream_t st);
// the same for `count` independent solves, ctrls a device array of pointers
hipError_t launch_ls_batch(Ctrl *const *ctrls, int count, int m, hipStream_t st);

// AAR.c:225-231 + 116-117:  x_new = (x_k + beta f_k) - sum_k (DX_k + beta DF_k) s_col(k)
template <typename T>
struct UpdArgs {
    long n = 0;
    int m = 0;
    int direct = 0;
    double beta = 0.0;
    const T *X = nullptr, *F = nullptr;
    long ld = 0;
    int slot[MAXM + 1] = {};
    int nvalid = 0;
    int it = 0;
    int xslot = 0, fslot = 0, oslot = 0;
    const Ctrl *ctrl = nullptr;
};
template <typename T>
hipError_t launch_aar_update(const UpdArgs<T> &a, hipStream_t st);

// Anderson step of the residual-recurrence AAR (solver.hip).  The iterate
// is held as an anchor x (updated in place) plus Richardson increments
// omega*g_j of the stored preconditioned residuals g_j; a term i of the
// update is ring slot gslot[i] (g_{jlo+i}), a correction term c is
// cslot[c] of the correction ring (x_{q+1} - x_q of an earlier Anderson
// step q).  With gamma = ctrl->s (reference column order) and difference
// column t, j = it - t: dg = g_j - g_{j-1}, dx = omega g_{j-1} or a
// correction,
//   corr  = beta g_it - sum_t gamma_t (dx_t + beta dg_t)   (g_it: term ng-1)
//   x_new = x + omega sum_{xlo <= i < xhi} g_i + corr,   written back to x;
//   C[cout] = corr when cout >= 0.
// With ncol = 0 and with_beta = 0 the same kernel folds x += omega sum g_i.
constexpr int NGMAX = MAXM + 2;
constexpr int NCMAX = MAXM + 1;
template <typename Tx, typename Tg>
struct AncArgs {
    long n = 0;
    Tx *x = nullptr;
    const Tg *G = nullptr;
    Tg *C = nullptr;
    long ld = 0;
    int ng = 0, nc = 0, ncol = 0;
    int gslot[NGMAX] = {};
    int xlo = 0, xhi = 0;     // terms of the Richardson increments since the anchor
    int with_beta = 0;        // beta g_it term (Anderson step)
    int cslot[NCMAX] = {};
    int col_ref[MAXM] = {};   // reference column of difference column t
    int col_j[MAXM] = {};     // term of g_j
    int col_c[MAXM] = {};     // correction of dx_t, or -1 for omega g_{j-1} (term col_j - 1)
    double omega = 0.0, beta = 0.0;
    int cout = -1;
    const Ctrl *ctrl = nullptr;
};
template <typename Tx, typename Tg>
hipError_t launch_anchor_update(const AncArgs<Tx, Tg> &a, hipStream_t st);

// PGR.c:117-130: x += sum DX s; Ax += sum DF s (in place at slot `cur`,
// the age-0 slot), ||b - Ax||^2 -> FIN_PGR_RES.
template <typename T>
struct PUpdArgs {
    long n = 0;
    int m = 0;
    T *X = nullptr, *Y = nullptr;
    const T *b = nullptr;
    long ld = 0;
    int slot[MAXM + 1] = {};
    int nvalid = 0;
    int it = 0;
    int nblocks = 0;
    double *partials = nullptr;
    unsigned *counter = nullptr;
    Ctrl *ctrl = nullptr;
    FinArgs fin = {};
};
template <typename T>
hipError_t launch_pgr_update(const PUpdArgs<T> &a, hipStream_t st);
// grid (= partial count) of the PGR update on n points with history m, given
// the streaming kernels' block count nb
int pgr_update_blocks(long n, int m, int nb);

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
    const double *c = nullptr;  // variable coefficient: divisor diag + c(x)
    int precond = 1;
    int update_r = 1;          // 0 on restart iterations (r recomputed by a stencil pass)
    int nblocks = 0;
    double *partials = nullptr;
    unsigned *counter = nullptr;
    Ctrl *ctrl = nullptr;
    FinArgs fin = {};
};
hipError_t launch_cg_update(const CgUpdArgs &a, hipStream_t st);
// grid (= number of partial pairs) of the CG update on n points
int cg_update_blocks(long n);

// ---------------------------------------------------------------------------
// complex Helmholtz (no MPI_REAL counterpart): A u = a*Lap(u) + (c(x) + i eta) u,
// complex128 interleaved (double2 = re, im), Jacobi D = (a*3*coeff[0] + c) + i eta
// ---------------------------------------------------------------------------
enum class HMode : int {
    MATVEC = 0,    // y = A u
    TRUE_RES = 1,  // f = b - A u; g = f / D; ||f||^2
    REC = 2,       // g' = g - omega (A g) / D; ||D g'||^2 for the history
};

struct HArgs {
    int nx = 0, ny = 0, nz = 0, nxy = 0;
    int lz = 0;
    int ntx = 0, nty = 0, nzc = 0, nitems = 0;
    double w[R + 1] = {};
    double dlap = 0.0;                // a * 3 * coeff[0]
    const double *c = nullptr;        // c(x) (N) or null: c0 everywhere
    double c0 = 0.0, eta = 0.0, omega = 0.0;
    int trace = 0;
    const double2 *in0 = nullptr, *bvec = nullptr;
    double2 *out0 = nullptr;
    double *partials = nullptr;       // [nitems]
    Ctrl *ctrl = nullptr;
    FinArgs fin = {};
};
hipError_t launch_hstencil(HMode m, const HArgs &a, hipStream_t st);
int hstencil_items(const HArgs &a);

// Hermitian Gram of the complex g ring (ages as in GramArgs): G_ij = <DF_i, DF_j>,
// g_i = <DF_i, f> with <u, v> = sum conj(u) v, laid out as the real 2m x 2m
// system [Gr -Gi; Gi Gr] [s_re; s_im] = [g_re; g_im] for the m x m solve.
struct HGramArgs {
    long n = 0;
    int m = 0;
    const double2 *F = nullptr;
    long ld = 0;
    int slot[MAXM + 1] = {};
    int nvalid = 0;
    int it = 0;
    int nblocks = 0;
    double *partials = nullptr;
    Ctrl *ctrl = nullptr;
};
hipError_t launch_hgram(const HGramArgs &a, hipStream_t st);

// complex anchor update: AncArgs with complex gamma = s[t] + i s[m + t]
struct HAncArgs {
    long n = 0;
    double2 *x = nullptr;
    const double2 *G = nullptr;
    double2 *C = nullptr;
    long ld = 0;
    int m = 0;
    int ng = 0, nc = 0, ncol = 0;
    int gslot[NGMAX] = {};
    int xlo = 0, xhi = 0;
    int with_beta = 0;
    int cslot[NCMAX] = {};
    int col_ref[MAXM] = {};
    int col_j[MAXM] = {};
    int col_c[MAXM] = {};
    double omega = 0.0, beta = 0.0;
    int cout = -1;
    const Ctrl *ctrl = nullptr;
};
hipError_t launch_hanchor_update(const HAncArgs &a, hipStream_t st);

// dL/dc = -Re(conj(lambda) u) for a diagonal coefficient field c
hipError_t launch_helm_grad_c(long n, const double2 *lambda, const double2 *u, double *dc, hipStream_t st);

// Reducing kernels run the finaliser in their last block when given a
// counter, or only write per-block partials when counter == nullptr; then
// launch_finish (or, for the Gram, launch_gram itself) runs a one-block
// second pass.
hipError_t launch_finish(const double *partials, int nitems, int nacc, const FinArgs &fin, Ctrl *ctrl,
                         hipStream_t st);

// small pointwise helpers
hipError_t launch_copy_f64(double *dst, const double *src, long n, const Ctrl *ctrl, hipStream_t st);
hipError_t launch_fill_f64(double *dst, double v, long n, hipStream_t st);
hipError_t launch_fill_f32(float *dst, float v, long n, hipStream_t st);
// x64 += e32 (mixed-precision correction)
hipError_t launch_axpy_f64_f32(double *x, const float *e, long n, hipStream_t st);
hipError_t launch_div_f64(double *r, double diag, long n, hipStream_t st);
hipError_t launch_zero_ctrl(Ctrl *c, hipStream_t st);

}  // namespace aar
