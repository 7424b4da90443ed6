// solver.hip -- host drivers of the device-resident solvers (libaar_hip).
//
// Storage (per solve, cached across calls):
//   X ring: m+1 slots of N -- iterates x_k in slot k % (m+1)
//   F ring: m+1 slots of N -- AAR: preconditioned residuals f_k;
//                             PGR/PL2R: A x_k
//   b, and a control block (Ctrl) + residual history in device memory.
// A difference column of the reference (DX[c] = x - x_old, DF[c] = f -
// f_old, AAR.c:96-102, PGR.c:105-109) is the difference of two ring slots
// and is formed on the fly, so a Richardson step writes 2 vectors instead
// of the reference's 4 (x, f, DX[c], DF[c]) plus 2 copies.
//
// Control: the host enqueues iterations following the reference's static
// schedule (Anderson / projection every p-th iteration); convergence is
// decided on the GPU by the last block of the reducing kernel, which sets
// ctrl->done; every later kernel returns at once.  The host polls `done`
// with a lag of two Anderson cycles, so the GPU never waits for the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "aar_common.h"
#include "aar_kernels.h"
#include "aar_solver.h"

namespace aar {

static Context g_ctx;
Context &ctx() { return g_ctx; }

hipError_t ensure(DevBuf &b, size_t bytes)
{
    if (b.p && b.bytes >= bytes) return hipSuccess;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
        if (e != hipSuccess) return e;
    }
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e == hipSuccess) b.bytes = bytes;
    return e;
}

int fail(int code, const std::string &msg)
{
    ctx().err = msg;
    return code;
}

#define HC(x)                                                                         \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess)                                                         \
            return fail(AAR_HIP_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

int context_init()
{
    Context &C = ctx();
    if (C.init) return 0;
    // a blocking stream: ordered with work on the legacy default stream,
    // which is where framework callers (torch's default stream) put inputs
    HC(hipStreamCreateWithFlags(&C.own, hipStreamDefault));
    if (!C.stream_set) C.stream = C.own;
    HC(hipMalloc((void **)&C.ctrl, 2 * sizeof(Ctrl)));
    HC(hipHostMalloc((void **)&C.ctrl_h, 2 * sizeof(Ctrl), hipHostMallocDefault));
    HC(hipHostMalloc((void **)&C.poll_h, NPOLL * sizeof(int), hipHostMallocDefault));
    HC(hipMalloc((void **)&C.counters, 64 * sizeof(unsigned)));
    HC(hipMemset(C.counters, 0, 64 * sizeof(unsigned)));
    for (int i = 0; i < NPOLL; i++) HC(hipEventCreateWithFlags(&C.poll_ev[i], hipEventDisableTiming));
    HC(hipEventCreate(&C.t0));
    HC(hipEventCreate(&C.t1));
    C.init = true;
    return 0;
}

void lap_coefficient(double *c, int fdn)
{
    // system.c:480-498: c0 = sum_p -2/p^2; c_p = 2 (-1)^(p+1) (prod ratio) / p^2
    double acc = 0.0;
    for (int p = 1; p <= fdn; p++) acc += -(2.0 / (p * p));
    c[0] = acc;
    for (int p = 1; p <= fdn; p++) {
        double num = 1.0, den = 1.0;
        for (int i = fdn - p + 1; i <= fdn; i++) num *= i;
        for (int i = fdn + 1; i <= fdn + p; i++) den *= i;
        const double ratio = num / den;
        const double sgn = (p % 2 == 1) ? 1.0 : -1.0;
        c[p] = 2.0 * sgn * ratio / (p * p);
    }
}

// The 1D weights c[0..R] of an operator of order 2*fdn: the caller's
// (coeff[0..fdn], as MPI_REAL/system.h's coeff_lap holds them), the
// registered user weights (fdn = AAR_HIP_FDN_USER), else Lap_coefficient's
// (system.c:480-498).  Weights beyond fdn are 0: the radius-6 sweeps then
// add exact zeros for the absent offsets.
int stencil_weights(int fdn_in, const double *coeff, double (&c)[R + 1], int *fdn_out)
{
    int fdn = fdn_in;
    const double *cw = coeff;
    if (fdn == AAR_HIP_FDN_USER) {
        if (ctx().user_fdn < 1) return fail(AAR_HIP_ERR_ARG, "no stencil weights registered (aar_hip_set_stencil)");
        fdn = ctx().user_fdn;
        cw = ctx().user_w;
    }
    if (fdn < 1 || fdn > R)
        return fail(AAR_HIP_ERR_ARG, "fdn must be 1..6 (stencil radius at most 6) or AAR_HIP_FDN_USER");
    for (int o = 0; o <= R; o++) c[o] = 0.0;
    if (cw) for (int o = 0; o <= fdn; o++) c[o] = cw[o];
    else lap_coefficient(c, fdn);
    *fdn_out = fdn;
    return 0;
}

int make_op(const aar_hip_grid *g, const double *coeff, Op &op)
{
    if (!g) return fail(AAR_HIP_ERR_ARG, "null grid");
    if (g->nx < 1 || g->ny < 1 || g->nz < 1) return fail(AAR_HIP_ERR_ARG, "grid sizes must be positive");
    if ((double)g->nx * g->ny * g->nz > 2147483647.0) return fail(AAR_HIP_ERR_ARG, "grid too large");
    double c[R + 1];
    if (int rc = stencil_weights(g->fdn, coeff, c, &op.fdn)) return rc;
    op.nx = g->nx;
    op.ny = g->ny;
    op.nz = g->nz;
    op.n = (long)g->nx * g->ny * g->nz;
    for (int o = 0; o <= R; o++) op.w[o] = g->a * c[o];   // system.c:102-104
    op.sigma = g->sigma;
    op.diag = g->a * (3 * c[0]);                           // AAR.c:94
    if (g->sigma != 0.0) op.diag += g->sigma;
    return 0;
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
int auto_lz(const Op &op, int wy, int py, size_t es)
{
    const int ntx = (op.nx + TX - 1) / TX, nty = (op.ny + wy * py - 1) / (wy * py);
    const int cols = ntx * nty;
    // long z-marches amortise the 2R-plane warm-up: about one block per CU
    // for FP64 (its plane tile and halo ring take most of the LDS, so only
    // one block fits per CU) and two for FP32 (two fit; measured faster)
    const int target = es == 4 ? 512 : 256;
    int nzc = std::max(1, target / std::max(1, cols));
    int lz = (op.nz + nzc - 1) / nzc;
    lz = std::max(lz, std::min(op.nz, 16));
    return lz;
}

// Launch shape for a grid: the one fixed by aar_hip_set_tuning, else a
// tuned one (aar_hip_autotune, or tuned here on first use with the
// "autotune" option), else the defaults with the automatic z-chunk.
Shape stencil_shape(int nx, int ny, int nz, size_t es)
{
    Context &C = ctx();
    if (!C.tuning_set) {
        const std::array<int, 4> key = {nx, ny, nz, (int)es};
        auto it = C.tuned.find(key);
        if (it == C.tuned.end() && C.autotune && autotune_stencil(nx, ny, nz, es, false, nullptr) == 0)
            it = C.tuned.find(key);
        if (it != C.tuned.end()) return it->second;
    }
    Op op;
    op.nx = nx;
    op.ny = ny;
    op.nz = nz;
    Shape sh;
    sh.wy = C.wy;
    sh.py = C.py;
    sh.lz = C.lz > 0 ? C.lz : auto_lz(op, C.wy, C.py, es);
    return sh;
}

template <typename T>
SArgs<T> base_sargs(const Op &op)
{
    SArgs<T> a;
    a.nx = op.nx;
    a.ny = op.ny;
    a.nz = op.nz;
    a.lz = stencil_shape(op.nx, op.ny, op.nz, sizeof(T)).lz;
    for (int o = 0; o <= R; o++) a.w[o] = op.w[o];
    a.sigma = op.sigma;
    a.diag = op.diag;
    // the FP64 sweeps read c(x) itself; the FP32 recurrence gets its own field
    if constexpr (sizeof(T) == 8) a.vf = op.c;
    return a;
}

// Slot stride of the ring buffers: N rounded up to 256 B plus the padding
// option "ring_pad" (aar_hip_set_option), so that equal offsets in different slots
// do not fall on the same HBM channel when N*sizeof(T) is a large power of
// two (512^3 FP64 slots are exactly 2^30 B apart).
long ring_ld(long n, size_t es)
{
    const long q = 256 / (long)es;
    return (n + q - 1) / q * q + ctx().ring_pad;
}

int nblocks_for(long n)
{
    const long cap = ctx().hist_blocks > 0 ? ctx().hist_blocks : 2048;
    long nb = (n + 4095) / 4096;
    return (int)std::max(1L, std::min(nb, cap));
}

// Grid of the register Gram (m <= 8): two blocks per CU unless the caller
// set "hist_blocks".  At 512^3, m = 7 the FP64 Gram took 1.22 ms on 512
// blocks against 1.28 ms on 2048 (FP32 history 0.71 against 0.75 ms; 256
// blocks 1.39 ms).  Between: 384 blocks 1.22 ms (FP32 history 0.82 ms),
// 768 blocks 1.25 ms (0.74 ms), 1024 blocks 1.26 ms (0.83 ms).  The deeper
// Grams keep the streaming kernels' grid.
int gram_blocks(long n, int m)
{
    const int nb = nblocks_for(n);
    if (m > MAXM_FAST || ctx().hist_blocks > 0) return nb;
    return std::min(nb, 512);
}

void prof_begin(const char *name)
{
    Context &C = ctx();
    if (!C.prof) return;
    PendingEv ev;
    ev.name = name;
    if (C.evpool.size() >= 2) {
        ev.a = C.evpool.back();
        C.evpool.pop_back();
        ev.b = C.evpool.back();
        C.evpool.pop_back();
    } else {
        (void)hipEventCreate(&ev.a);
        (void)hipEventCreate(&ev.b);
    }
    (void)hipEventRecord(ev.a, C.stream);
    C.pending.push_back(ev);
}

void prof_end(double bytes)
{
    Context &C = ctx();
    if (!C.prof || C.pending.empty()) return;
    PendingEv &ev = C.pending.back();
    ev.bytes = bytes;
    (void)hipEventRecord(ev.b, C.stream);
}

void prof_flush()
{
    Context &C = ctx();
    for (auto &ev : C.pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev.a, ev.b) == hipSuccess) {
            ProfEntry &e = C.profile[ev.name];
            e.count++;
            e.ms += ms;
            e.bytes += ev.bytes;
        }
        C.evpool.push_back(ev.a);
        C.evpool.push_back(ev.b);
    }
    C.pending.clear();
}

static const char *smode_name(SMode m)
{
    switch (m) {
    case SMode::MATVEC: return "stencil_matvec";
    case SMode::AAR_TRUE: return "stencil_aar_true";
    case SMode::AAR_REC: return "stencil_aar_rec";
    case SMode::RESID: return "stencil_resid";
    case SMode::PGR_RICH: return "stencil_pgr_rich";
    case SMode::CG_Q: return "stencil_cg_q";
    case SMode::CG_RESID: return "stencil_cg_resid";
    case SMode::MATVEC_RES: return "stencil_matvec_res";
    }
    return "stencil";
}

// Reductions either finish in the last block of the reducing kernel
// (counter + device-scope fence) or, in two-pass mode, in a one-block
// second kernel; the choice is a measured tuning option.
template <typename T>
int run_stencil(SMode mode, SArgs<T> a, double bytes)
{
    Context &C = ctx();
    const bool two = C.two_pass && a.partials;
    if (two) a.counter = nullptr;
    const Shape sh = stencil_shape(a.nx, a.ny, a.nz, sizeof(T));
    prof_begin(smode_name(mode));
    hipError_t e = launch_stencil<T>(mode, a, sh.wy, sh.py, C.stream);
    prof_end(bytes);
    if (two && e == hipSuccess)
        e = launch_finish(a.partials, stencil_items<T>(a, sh.wy, sh.py), mode == SMode::CG_RESID ? 2 : 1, a.fin,
                          a.ctrl, C.stream);
    if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("stencil launch: ") + hipGetErrorString(e));
    C.bytes_acc += bytes;
    return 0;
}

template <typename T>
hipError_t run_dot(DotArgs<T> d)
{
    Context &C = ctx();
    const bool two = C.two_pass;
    if (two) d.counter = nullptr;
    hipError_t e = launch_dot<T>(d, C.stream);
    if (two && e == hipSuccess) e = launch_finish(d.partials, d.nblocks, 1, d.fin, d.ctrl, C.stream);
    return e;
}

template <typename T>
static hipError_t run_gram(GramArgs<T> g)
{
    if (ctx().two_pass) g.counter = nullptr;
    return launch_gram<T>(g, ctx().stream);
}

static hipError_t run_pgr_update(PUpdArgs<double> u)
{
    Context &C = ctx();
    const bool two = C.two_pass;
    if (two) u.counter = nullptr;
    hipError_t e = launch_pgr_update<double>(u, C.stream);
    if (two && e == hipSuccess) e = launch_finish(u.partials, u.nblocks, 1, u.fin, u.ctrl, C.stream);
    return e;
}

static hipError_t run_cg_update(CgUpdArgs u)
{
    Context &C = ctx();
    const bool two = C.two_pass && u.update_r;
    if (two) u.counter = nullptr;
    hipError_t e = launch_cg_update(u, C.stream);
    if (two && e == hipSuccess) e = launch_finish(u.partials, u.nblocks, 2, u.fin, u.ctrl, C.stream);
    return e;
}

template <typename T>
int stencil_partials_needed(const Op &op)
{
    SArgs<T> a = base_sargs<T>(op);
    const Shape sh = stencil_shape(op.nx, op.ny, op.nz, sizeof(T));
    return stencil_items<T>(a, sh.wy, sh.py) * 2;
}

// ---------------------------------------------------------------------------
// Autotuner: times the recurrence sweep (the dominant kernel of AAR) on the
// grid over the launch shapes -- wavefronts per block x rows per thread
// (the tile height; tiles are one wavefront = 64 points wide in x) and the
// z-slab depth -- and keeps the fastest.  batch_shape restricts the search
// to the slab depth of the batched 8 x 4 launch.
// ---------------------------------------------------------------------------
template <typename T>
static int autotune_t(int nx, int ny, int nz, bool batch_shape, Shape *out)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    aar_hip_grid g = {nx, ny, nz, R, -1.0, 0.0};
    Op op;
    if (int rc = make_op(&g, nullptr, op)) return rc;
    T *in = nullptr, *o = nullptr;
    HC(hipMalloc((void **)&in, sizeof(T) * op.n));
    if (hipMalloc((void **)&o, sizeof(T) * op.n) != hipSuccess) {
        (void)hipFree(in);
        return fail(AAR_HIP_ERR_HIP, "autotune: out of device memory");
    }
    if (sizeof(T) == 8) (void)launch_fill_f64((double *)in, 1.0, op.n, C.stream);
    else (void)launch_fill_f32((float *)in, 1.0f, op.n, C.stream);
    const int shapes[5][2] = {{8, 4}, {4, 8}, {8, 2}, {4, 4}, {4, 2}};
    const int nshape = batch_shape ? 1 : 5;
    Shape best;
    best.us = 1e300;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    int rc = 0;
    for (int si = 0; si < nshape && !rc; si++) {
        const int wy = shapes[si][0], py = shapes[si][1];
        std::vector<int> lzs = {auto_lz(op, wy, py, sizeof(T)), nz, (nz + 1) / 2, (nz + 3) / 4, (nz + 7) / 8};
        std::sort(lzs.begin(), lzs.end());
        lzs.erase(std::unique(lzs.begin(), lzs.end()), lzs.end());
        for (int lz : lzs) {
            if (lz < std::min(nz, 16)) continue;
            SArgs<T> a;
            a.nx = nx;
            a.ny = ny;
            a.nz = nz;
            a.lz = lz;
            for (int q = 0; q <= R; q++) {
                a.w[q] = op.w[q];
                a.cf[q] = q == 0 ? 0.5 : -1e-3 * op.w[q];
            }
            a.in0 = in;
            a.out0 = o;
            hipError_t e = launch_stencil<T>(SMode::AAR_REC, a, wy, py, C.stream);
            const int reps = 4;
            e = e == hipSuccess ? hipEventRecord(e0, C.stream) : e;
            for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_stencil<T>(SMode::AAR_REC, a, wy, py, C.stream);
            e = e == hipSuccess ? hipEventRecord(e1, C.stream) : e;
            e = e == hipSuccess ? hipEventSynchronize(e1) : e;
            float ms = 0.f;
            e = e == hipSuccess ? hipEventElapsedTime(&ms, e0, e1) : e;
            if (e != hipSuccess) {
                rc = fail(AAR_HIP_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
                break;
            }
            const double us = 1e3 * ms / reps;
            if (us < best.us) {
                best.us = us;
                best.lz = lz;
                best.wy = wy;
                best.py = py;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(C.stream);
    (void)hipFree(in);
    (void)hipFree(o);
    if (rc) return rc;
    const std::array<int, 4> key = {nx, ny, nz, (int)sizeof(T)};
    (batch_shape ? C.tuned_batch : C.tuned)[key] = best;
    if (out) *out = best;
    return 0;
}

int autotune_stencil(int nx, int ny, int nz, size_t es, bool batch_shape, Shape *out)
{
    return es == 8 ? autotune_t<double>(nx, ny, nz, batch_shape, out) : autotune_t<float>(nx, ny, nz, batch_shape, out);
}

// poll ring: record ctrl->done after a cycle; true once done is observed
int poll_record(Ctrl *ctrl, int cyc)
{
    Context &C = ctx();
    const int s = cyc % NPOLL;
    HC(hipMemcpyAsync(&C.poll_h[s], &ctrl->done, sizeof(int), hipMemcpyDeviceToHost, C.stream));
    HC(hipEventRecord(C.poll_ev[s], C.stream));
    return 0;
}

bool poll_check(int cyc)
{
    Context &C = ctx();
    if (cyc < 0) return false;
    const int s = cyc % NPOLL;
    if (hipEventSynchronize(C.poll_ev[s]) != hipSuccess) return true;
    return C.poll_h[s] != 0;
}

static int finish_ctrl(Ctrl *ctrl, Ctrl *host)
{
    Context &C = ctx();
    HC(hipMemcpyAsync(host, ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, C.stream));
    HC(hipStreamSynchronize(C.stream));
    return 0;
}

static void fill_stats(aar_hip_stats *st, const Ctrl &c, int iterations, int loop_iter, int max_iter,
                       double ms, double bytes, int hist_len)
{
    aar_hip_stats s;
    std::memset(&s, 0, sizeof(s));
    s.iterations = iterations;
    s.rhs_norm = c.rhs_norm;
    s.relres = c.rhs_norm > 0 ? c.relres / c.rhs_norm : c.relres;
    s.converged = c.relres <= c.tol_abs;
    s.printed_converged = (loop_iter < max_iter && c.relres <= c.tol_abs);
    s.printed_warning = (loop_iter >= max_iter);
    s.status = c.status == 1 ? AAR_HIP_ERR_SVD : (c.status == 2 ? AAR_HIP_ERR_NONFINITE : 0);
    s.hist_len = hist_len;
    s.solve_ms = ms;
    s.bytes = bytes;
    ctx().last = s;
    if (st) *st = s;
}

static int store_history(const double *hist_dev, int len)
{
    Context &C = ctx();
    C.hist_host.assign(std::max(len, 0), 0.0);
    if (len > 0) HC(hipMemcpy(C.hist_host.data(), hist_dev, sizeof(double) * len, hipMemcpyDeviceToHost));
    return 0;
}

// ---------------------------------------------------------------------------
// AAR (AAR.c:37-177) by residual recurrence.
//
// With the constant Jacobi diagonal d (AAR.c:94), a Richardson step
// x_{k+1} = x_k + omega g_k (AAR.c:134-135) changes the preconditioned
// residual g = (b - A x)/d by g_{k+1} = g_k - (omega/d) A g_k.  Between two
// Anderson steps the solver therefore advances g alone (AAR_REC: read g_k,
// write g_{k+1}) and keeps x as an anchor: x_k = x_a + omega sum_{a<=j<k} g_j,
// where x_a is the iterate after the last Anderson step (the caller's x
// buffer itself).  At an Anderson iteration the Gram of the g ring gives
// gamma exactly as before (DF columns are g_j - g_{j-1}), DX columns are
// omega g_{j-1} or, for the step after an earlier Anderson update, the
// stored correction x_{q+1} - x_q, and one streaming pass writes the new
// anchor and its correction.  The true residual b - A x of the new anchor
// (AAR_TRUE) restarts the recurrence and is the residual the reference
// checks (AAR.c:119-123), so convergence is decided on b - A x exactly
// where the reference decides it.  Per Anderson cycle (m = 7, p = 6) this
// moves 33 vectors instead of 49.
//
// Tx is the type of x and b; Tg of the g ring and corrections.  <double,
// double> is FP64 AAR, <float, float> FP32 AAR, <double, float> the mixed
// solver: FP64 true residual, anchor and normal equations, FP32 sweeps.
// ---------------------------------------------------------------------------
template <typename Tx, typename Tg>
int aar_rec(const Op &op, Tx *x, const Tx *b, double omega, double beta, int m, int p, int max_iter,
            double tol, int precond, int ci, aar_hip_stats *st, int keep_history, int ref_start)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!x || !b) return fail(AAR_HIP_ERR_ARG, "null vector");
    if (m < 1 || m > MAXM) return fail(AAR_HIP_ERR_ARG, "history m must be in 1..32");
    if (p < 1) return fail(AAR_HIP_ERR_ARG, "p must be >= 1");
    if (max_iter < 0) return fail(AAR_HIP_ERR_ARG, "max_iter must be >= 0");
    if (op.c && sizeof(Tx) != 8)
        return fail(AAR_HIP_ERR_ARG, "the variable-coefficient operator needs FP64 x and b (FP64 or mixed AAR)");
    const long n = op.n;
    // g ring: g_k in slot k % S.  Live at any time: the m+1 residuals of the
    // next Anderson window and the one being written.
    const int S = m + 2;
    // corrections of the Anderson steps inside a window [k-m, k-1]
    const int NC = (p <= m) ? m / p + 1 : 0;
    constexpr bool f64g = sizeof(Tg) == 8;
    DevBuf &RG = f64g ? C.ring64f : C.ring32f;
    DevBuf &RC = f64g ? C.ring64x : C.ring32x;
    const long ld = ring_ld(n, sizeof(Tg));
    HC(ensure(RG, sizeof(Tg) * ld * S));
    if (NC > 0) HC(ensure(RC, sizeof(Tg) * ld * NC));
    DevBuf &HB = ci == 0 ? C.hist0 : C.hist1;
    HC(ensure(HB, sizeof(double) * (size_t)(max_iter + 4)));
    const int nb = nblocks_for(n);
    const size_t need = std::max<size_t>(
        std::max((size_t)stencil_partials_needed<Tg>(op), (size_t)stencil_partials_needed<Tx>(op)), (size_t)nb * GRAM_NA_MAX);
    DevBuf &PB = ci == 0 ? C.part0 : C.part1;
    HC(ensure(PB, sizeof(double) * need));
    Tg *G = (Tg *)RG.p;
    Tg *CR = NC > 0 ? (Tg *)RC.p : nullptr;
    double *hist = (double *)HB.p, *part = (double *)PB.p;
    Ctrl *ctrl = C.ctrl + ci;
    unsigned *cnt = C.counters + 8 * ci;
    auto gs = [&](long k) { return (int)(k % S); };
    auto gptr = [&](long k) { return G + (size_t)gs(k) * ld; };
    auto is_and = [&](long k) { return k > 1 && k % p == 0; };          // AAR.c:109
    auto next_and = [&](long k) { long q = (k / p + 1) * p; return q > 1 ? q : q + p; };
    const double sx = (double)sizeof(Tx) * n, sg = (double)sizeof(Tg) * n;
    C.bytes_acc = 0.0;
    // variable coefficient: the recurrence reads h = omega/(diag + c(x))
    // (Jacobi) or c(x) itself (no preconditioner), rounded to Tg
    const Tg *vch = nullptr;
    if (op.c) {
        HC(ensure(C.vch, sizeof(Tg) * (size_t)n));
        vch = (const Tg *)C.vch.p;
    }
    const double sv_true = op.c ? sx : 0.0, sv_rec = op.c ? sg : 0.0;

    // true residual of x into g_k (AAR.c:80-82 / 119-123)
    auto true_res = [&](long k, int check) -> int {
        SArgs<Tx> a = base_sargs<Tx>(op);
        a.omega = omega;
        a.precond = precond;
        a.in0 = x;
        a.bvec = b;
        if constexpr (sizeof(Tx) == sizeof(Tg)) a.out0 = (Tx *)gptr(k);
        else a.out32 = (float *)gptr(k);
        a.partials = part;
        a.counter = cnt + 1;
        a.ctrl = ctrl;
        a.fin.fin = FIN_AAR_RES;
        a.fin.it = (int)k;
        a.fin.check = check;
        a.fin.hist = hist;
        a.fin.hist_cap = max_iter + 4;
        return run_stencil<Tx>(SMode::AAR_TRUE, a, 2 * sx + sg + sv_true);
    };

    HC(launch_zero_ctrl(ctrl, C.stream));
    HC(hipEventRecord(C.t0, C.stream));
    if (vch) {
        prof_begin("vc_field");
        HC(launch_vc_field<Tg>(n, op.c, op.diag, omega, precond, (Tg *)C.vch.p, C.stream));
        prof_end(sx + sg);
        C.bytes_acc += sx + sg;
    }
    // ||b|| and the scaled tolerance (AAR.c:78-79)
    {
        DotArgs<Tx> d;
        d.n = n;
        d.v = b;
        d.nblocks = nb;
        d.partials = part;
        d.counter = cnt + 0;
        d.ctrl = ctrl;
        d.fin.fin = FIN_NORM_B;
        d.fin.tol = tol;
        d.fin.check = ref_start;      // the reference's zero-iteration rule (AAR.c:76-79)
        prof_begin("dot");
        HC(run_dot<Tx>(d));
        prof_end(sx);
    }
    C.bytes_acc += sx;
    if (int rc = true_res(1, 0)) return rc;

    long first = 1;         // x = x_a + omega sum_{first <= j < k} g_j
    int cyc = 0;
    long it = 1;
    bool stopped = false;
    for (; it <= max_iter && !stopped; it++) {
        if (!is_and(it)) {
            // Richardson (AAR.c:131-139) on the residual
            SArgs<Tg> a = base_sargs<Tg>(op);
            a.omega = omega;
            a.precond = precond;
            if (vch) {
                a.vf = vch;
                a.cf[0] = 1.0 - omega;            // (1 - omega) g - h sum_o w_o S_o
            } else {
                const double c = precond ? omega / op.diag : omega;
                a.cf[0] = 1.0 - c * (3.0 * op.w[0] + op.sigma);
                for (int o = 1; o <= R; o++) a.cf[o] = -c * op.w[o];
            }
            a.in0 = gptr(it);
            a.out0 = gptr(it + 1);
            a.trace = 1;
            a.partials = part;
            a.counter = cnt + 1;
            a.ctrl = ctrl;
            a.fin.fin = FIN_AAR_RES;
            a.fin.it = (int)it + 1;
            a.fin.check = 0;
            a.fin.hist = hist;
            a.fin.hist_cap = max_iter + 4;
            if (int rc = run_stencil<Tg>(SMode::AAR_REC, a, 2 * sg + sv_rec)) return rc;
            // a residual that no Anderson window will read again is folded
            // into the anchor before its slot is reused (only when p > m+1)
            if (it < next_and(it) - m) {
                AncArgs<Tx, Tg> u;
                u.n = n;
                u.x = x;
                u.G = G;
                u.ld = ld;
                u.ng = 1;
                u.gslot[0] = gs(it);
                u.xlo = 0;
                u.xhi = 1;
                u.omega = omega;
                u.ctrl = ctrl;
                prof_begin("fold");
                HC((launch_anchor_update<Tx, Tg>(u, C.stream)));
                prof_end(2 * sx + sg);
                C.bytes_acc += 2 * sx + sg;
                first = it + 1;
            }
            continue;
        }
        // Anderson step (AAR.c:109-123)
        const int nv = (int)std::min<long>(m, it - 1);
        {
            GramArgs<Tg> g;
            for (int k = 0; k <= m; k++) g.slot[k] = it - k >= 1 ? gs(it - k) : 0;
            g.n = n;
            g.m = m;
            g.nvalid = nv;
            g.it = (int)it;
            g.kind = 0;
            g.F = G;
            g.ld = ld;
            g.fslot = gs(it);
            g.nblocks = gram_blocks(n, m);
            g.partials = part;
            g.counter = cnt + 2;
            g.ctrl = ctrl;
            prof_begin("gram");
            HC(run_gram<Tg>(g));
            prof_end(sg * (nv + 1));
            C.bytes_acc += sg * (nv + 1);
        }
        prof_begin("ls");
        HC(launch_ls(ctrl, m, 0, C.stream));
        prof_end(0);
        {
            AncArgs<Tx, Tg> u;
            const long jlo = std::min(first, it - nv);
            u.n = n;
            u.x = x;
            u.G = G;
            u.C = CR;
            u.ld = ld;
            u.ng = (int)(it - jlo + 1);
            for (int i = 0; i < u.ng; i++) u.gslot[i] = gs(jlo + i);
            u.xlo = (int)(first - jlo);
            u.xhi = (int)(it - jlo);
            u.with_beta = 1;
            u.ncol = nv;
            u.nc = 0;
            for (int t = 0; t < nv; t++) {
                const long j = it - t;
                u.col_ref[t] = (int)((j - 2) % m);                    // AAR.c:97
                u.col_j[t] = (int)(j - jlo);
                u.col_c[t] = -1;
                if (is_and(j - 1)) {
                    u.cslot[u.nc] = (int)(((j - 1) / p) % NC);
                    u.col_c[t] = u.nc++;
                }
            }
            u.cout = NC > 0 ? (int)((it / p) % NC) : -1;
            u.omega = omega;
            u.beta = beta;
            u.ctrl = ctrl;
            const double by = 2 * sx + sg * (u.ng + u.nc + (u.cout >= 0 ? 1 : 0));
            prof_begin("anchor_update");
            HC((launch_anchor_update<Tx, Tg>(u, C.stream)));
            prof_end(by);
            C.bytes_acc += by;
        }
        first = it + 1;
        if (int rc = true_res(it + 1, 1)) return rc;
        if (int rc = poll_record(ctrl, cyc)) return rc;
        if (poll_check(cyc - 2)) stopped = true;
        cyc++;
    }
    if (!stopped && first <= max_iter) {
        // ended on Richardson steps: x_{max_iter+1} = x_a + omega sum g_j
        AncArgs<Tx, Tg> u;
        u.n = n;
        u.x = x;
        u.G = G;
        u.ld = ld;
        u.ng = (int)(max_iter - first + 1);
        for (int i = 0; i < u.ng; i++) u.gslot[i] = gs(first + i);
        u.xlo = 0;
        u.xhi = u.ng;
        u.omega = omega;
        u.ctrl = ctrl;
        prof_begin("fold");
        HC((launch_anchor_update<Tx, Tg>(u, C.stream)));
        prof_end(2 * sx + sg * u.ng);
        C.bytes_acc += 2 * sx + sg * u.ng;
    }
    HC(hipEventRecord(C.t1, C.stream));
    Ctrl &hc = C.ctrl_h[ci];
    if (int rc = finish_ctrl(ctrl, &hc)) return rc;
    float ms = 0.f;
    HC(hipEventElapsedTime(&ms, C.t0, C.t1));
    prof_flush();
    int iterations, loop_iter;
    if (hc.done) {
        iterations = hc.final_iter;
        loop_iter = hc.final_iter + 1;
    } else {
        iterations = max_iter;
        loop_iter = max_iter + 1;
    }
    const int hl = std::min(iterations + 1, max_iter + 4);
    if (keep_history)
        if (int rc = store_history(hist, hl)) return rc;
    fill_stats(st, hc, iterations, loop_iter, max_iter, ms, C.bytes_acc, hl);
    if (hc.status == 1) return fail(AAR_HIP_ERR_SVD, "no convergence in 250 svd iterations");
    if (hc.status == 2) return fail(AAR_HIP_ERR_NONFINITE, "residual is not finite");
    return 0;
}

template <typename T>
int aar_solve(const Op &op, T *x, const T *b, double omega, double beta, int m, int p, int max_iter,
              double tol, int precond, int ci, aar_hip_stats *st, int keep_history, int ref_start)
{
    return aar_rec<T, T>(op, x, b, omega, beta, m, p, max_iter, tol, precond, ci, st, keep_history, ref_start);
}

int aar_mixed(const Op &op, double *x, const double *b, double omega, double beta, int m, int p,
              int max_iter, double tol, int precond, aar_hip_stats *st)
{
    return aar_rec<double, float>(op, x, b, omega, beta, m, p, max_iter, tol, precond, 0, st, 1, 1);
}

template int aar_solve<double>(const Op &, double *, const double *, double, double, int, int, int, double,
                               int, int, aar_hip_stats *, int, int);
template SArgs<double> base_sargs<double>(const Op &);
template SArgs<float> base_sargs<float>(const Op &);
template hipError_t run_dot<double>(DotArgs<double>);
template int run_stencil<double>(SMode, SArgs<double>, double);
template int aar_solve<float>(const Op &, float *, const float *, double, double, int, int, int, double,
                              int, int, aar_hip_stats *, int, int);

// ---------------------------------------------------------------------------
// PGR / PL2R (PGR.c:35-218, PL2R.c:35-218), FP64
// ---------------------------------------------------------------------------
int pgr_solve(const Op &op, double *x, const double *b, double omega, int m, int p, int max_iter,
              double tol, int precond, int l2, aar_hip_stats *st)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!x || !b) return fail(AAR_HIP_ERR_ARG, "null vector");
    if (m < 1 || m > MAXM) return fail(AAR_HIP_ERR_ARG, "history m must be in 1..32");
    if (p < 1) return fail(AAR_HIP_ERR_ARG, "p must be >= 1");
    if (max_iter < 0) return fail(AAR_HIP_ERR_ARG, "max_iter must be >= 0");
    const long n = op.n;
    const int ns = m + 1;
    const long ld = ring_ld(n, sizeof(double));
    HC(ensure(C.ring64x, sizeof(double) * ld * ns));
    HC(ensure(C.ring64f, sizeof(double) * ld * ns));
    HC(ensure(C.hist0, sizeof(double) * (size_t)(max_iter + 4)));
    const int nb = nblocks_for(n);
    const size_t need = std::max<size_t>((size_t)stencil_partials_needed<double>(op), (size_t)nb * GRAM_NA_MAX);
    HC(ensure(C.part0, sizeof(double) * need));
    double *X = (double *)C.ring64x.p, *Y = (double *)C.ring64f.p;
    double *hist = (double *)C.hist0.p, *part = (double *)C.part0.p;
    Ctrl *ctrl = C.ctrl;
    unsigned *cnt = C.counters;
    auto slot = [&](long t) { return (int)(t % ns); };
    const double s_b = 8.0 * n;
    const double vcb = op.c ? s_b : 0.0;    // the coefficient field's reads
    C.bytes_acc = 0.0;

    HC(launch_zero_ctrl(ctrl, C.stream));
    HC(hipEventRecord(C.t0, C.stream));
    HC(hipMemcpyAsync(X + (size_t)slot(1) * ld, x, sizeof(double) * n, hipMemcpyDeviceToDevice, C.stream));
    {
        DotArgs<double> d;
        d.n = n;
        d.v = b;
        d.nblocks = nb;
        d.partials = part;
        d.counter = cnt + 0;
        d.ctrl = ctrl;
        d.fin.fin = FIN_NORM_B;
        d.fin.tol = tol;
        d.fin.check = 1;
        HC(run_dot<double>(d));
    }
    // Ax_1 and ||b - A x_1|| (PGR.c:79-81)
    {
        SArgs<double> a = base_sargs<double>(op);
        a.in0 = X + (size_t)slot(1) * ld;
        a.bvec = b;
        a.out0 = Y + (size_t)slot(1) * ld;
        a.partials = part;
        a.counter = cnt + 1;
        a.ctrl = ctrl;
        a.fin.fin = FIN_AAR_RES;
        a.fin.it = 1;
        a.fin.check = 0;
        a.fin.hist = hist;
        a.fin.hist_cap = max_iter + 4;
        if (int rc = run_stencil<double>(SMode::MATVEC_RES, a, s_b * 3 + vcb)) return rc;
    }
    int cyc = 0;
    int t = 1;
    bool stopped = false;
    for (; t <= max_iter && !stopped; t++) {
        SArgs<double> a = base_sargs<double>(op);
        a.omega = omega;
        a.precond = precond;
        a.in0 = X + (size_t)slot(t) * ld;
        a.in1 = b;
        a.in2 = Y + (size_t)slot(t) * ld;
        a.bvec = b;
        a.out0 = X + (size_t)slot(t + 1) * ld;
        a.out1 = Y + (size_t)slot(t + 1) * ld;
        a.trace = 1;
        a.partials = part;
        a.counter = cnt + 1;
        a.ctrl = ctrl;
        a.fin.fin = FIN_AAR_RES;
        a.fin.it = t + 1;
        a.fin.check = 0;
        a.fin.hist = hist;
        a.fin.hist_cap = max_iter + 4;
        if (int rc = run_stencil<double>(SMode::PGR_RICH, a, s_b * 5 + 2 * vcb)) return rc;
        if (t % p == 0) {                                           // PGR.c:111
            // history by age relative to the newest iterate x_{t+1}: the
            // column of age k belongs to iteration t-k (PGR.c:105-109)
            GramArgs<double> g;
            PUpdArgs<double> u;
            const int nv = std::min(m, t);
            for (int k = 0; k <= m; k++) g.slot[k] = u.slot[k] = t + 1 - k >= 1 ? slot(t + 1 - k) : 0;
            g.n = u.n = n;
            g.m = u.m = m;
            g.nvalid = u.nvalid = nv;
            g.it = u.it = t;
            g.kind = l2 ? 2 : 1;
            g.F = X;
            g.Y = Y;
            g.b = b;
            g.ld = ld;
            g.fslot = slot(t + 1);
            g.nblocks = gram_blocks(n, m);
            g.partials = part;
            g.counter = cnt + 2;
            g.ctrl = ctrl;
            prof_begin("gram");
            HC(run_gram<double>(g));
            prof_end(s_b * ((l2 ? 1 : 2) * (nv + 1) + 1));
            prof_begin("ls");
            HC(launch_ls(ctrl, m, 0, C.stream));
            prof_end(0);
            u.X = X;
            u.Y = Y;
            u.b = b;
            u.ld = ld;
            u.nblocks = pgr_update_blocks(n, m, nb);
            if ((size_t)u.nblocks > need) return fail(AAR_HIP_ERR_ARG, "PGR: grid too large for the partials");
            u.partials = part;
            u.counter = cnt + 3;
            u.ctrl = ctrl;
            u.fin.fin = FIN_PGR_RES;
            u.fin.it = t + 1;
            u.fin.check = 1;
            u.fin.hist = nullptr;
            prof_begin("pgr_update");
            HC(run_pgr_update(u));
            prof_end(s_b * (2 * (nv + 1) + 1 + 2));
            C.bytes_acc += s_b * ((l2 ? 1 : 2) * (nv + 1) + 1) + s_b * (2 * (nv + 1) + 3);
            if (int rc = poll_record(ctrl, cyc)) return rc;
            if (poll_check(cyc - 2)) stopped = true;
            cyc++;
        }
    }
    HC(hipEventRecord(C.t1, C.stream));
    Ctrl &hc = C.ctrl_h[0];
    if (int rc = finish_ctrl(ctrl, &hc)) return rc;
    float ms = 0.f;
    HC(hipEventElapsedTime(&ms, C.t0, C.t1));
    prof_flush();
    int iterations, loop_iter;
    if (hc.done) {
        iterations = hc.final_iter;
        loop_iter = hc.final_iter + 1;
    } else {
        iterations = max_iter;
        loop_iter = max_iter + 1;
    }
    HC(hipMemcpyAsync(x, X + (size_t)slot(iterations + 1) * ld, sizeof(double) * n, hipMemcpyDeviceToDevice,
                      C.stream));
    HC(hipStreamSynchronize(C.stream));
    const int hl = std::min(iterations + 1, max_iter + 4);
    if (int rc = store_history(hist, hl)) return rc;
    fill_stats(st, hc, iterations, loop_iter, max_iter, ms, C.bytes_acc, hl);
    if (hc.status == 1) return fail(AAR_HIP_ERR_SVD, "no convergence in 250 svd iterations");
    if (hc.status == 2) return fail(AAR_HIP_ERR_NONFINITE, "residual is not finite");
    return 0;
}

// ---------------------------------------------------------------------------
// CG (CG.c:33-175), FP64
// ---------------------------------------------------------------------------
int cg_solve(const Op &op, double *x, const double *b, int max_iter, double tol, int precond,
             aar_hip_stats *st)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!x || !b) return fail(AAR_HIP_ERR_ARG, "null vector");
    if (max_iter < 0) return fail(AAR_HIP_ERR_ARG, "max_iter must be >= 0");
    const long n = op.n;
    // r, d[2], q (padded slots, as the rings)
    const long ld = ring_ld(n, sizeof(double));
    HC(ensure(C.ring64x, sizeof(double) * ld * 4));
    HC(ensure(C.hist0, sizeof(double) * (size_t)(max_iter + 4)));
    const int nb = nblocks_for(n);
    const size_t need = std::max<size_t>((size_t)stencil_partials_needed<double>(op), (size_t)nb * GRAM_NA_MAX);
    HC(ensure(C.part0, sizeof(double) * need));
    double *r = (double *)C.ring64x.p, *dd[2] = {r + ld, r + 2 * ld}, *q = r + 3 * ld;
    double *hist = (double *)C.hist0.p, *part = (double *)C.part0.p;
    Ctrl *ctrl = C.ctrl;
    unsigned *cnt = C.counters;
    const double s_b = 8.0 * n;
    const double vcb = op.c ? s_b : 0.0;    // the coefficient field's reads
    C.bytes_acc = 0.0;

    HC(launch_zero_ctrl(ctrl, C.stream));
    HC(hipEventRecord(C.t0, C.stream));
    {
        DotArgs<double> d;
        d.n = n;
        d.v = b;
        d.nblocks = nb;
        d.partials = part;
        d.counter = cnt + 0;
        d.ctrl = ctrl;
        d.fin.fin = FIN_NORM_B;
        d.fin.tol = tol;
        HC(run_dot<double>(d));
    }
    // r = b - A x; d = r / diag; delta = r.d; relres = ||r|| (CG.c:65-84)
    {
        SArgs<double> a = base_sargs<double>(op);
        a.precond = precond;
        a.in0 = x;
        a.bvec = b;
        a.out0 = r;
        a.out1 = dd[0];
        a.partials = part;
        a.counter = cnt + 1;
        a.ctrl = ctrl;
        a.fin.fin = FIN_CG_BETA;
        a.fin.it = 0;
        a.fin.hist = hist;
        a.fin.hist_cap = max_iter + 4;
        if (int rc = run_stencil<double>(SMode::CG_RESID, a, s_b * 4 + vcb)) return rc;
    }
    int k = 0;
    bool stopped = false;
    int cyc = 0;
    for (; k < max_iter && !stopped; k++) {
        const int cur = (k + 1) & 1, old = k & 1;
        {
            SArgs<double> a = base_sargs<double>(op);
            a.precond = precond;
            a.cg_first = (k == 0);
            a.cg_beta = &ctrl->beta;
            a.in0 = r;
            a.in1 = dd[old];
            a.out0 = dd[cur];
            a.out1 = q;
            a.partials = part;
            a.counter = cnt + 2;
            a.ctrl = ctrl;
            a.fin.fin = FIN_CG_ALPHA;
            if (int rc = run_stencil<double>(SMode::CG_Q, a, s_b * 4 + 2 * vcb)) return rc;
        }
        const bool restart = (k % 50) == 0;                          // CG.c:109
        {
            CgUpdArgs u;
            u.n = n;
            u.x = x;
            u.r = r;
            u.d = dd[cur];
            u.q = q;
            u.diag = op.diag;
            u.c = op.c;
            u.precond = precond;
            u.update_r = restart ? 0 : 1;
            u.nblocks = cg_update_blocks(n);
            if ((size_t)u.nblocks * 2 > need) return fail(AAR_HIP_ERR_ARG, "CG: grid too large for the partials");
            u.partials = part;
            u.counter = cnt + 3;
            u.ctrl = ctrl;
            u.fin.fin = FIN_CG_BETA;
            u.fin.it = k + 1;
            u.fin.hist = hist;
            u.fin.hist_cap = max_iter + 4;
            prof_begin("cg_update");
            HC(run_cg_update(u));
            prof_end(s_b * (restart ? 3 : 6) + (restart ? 0 : vcb));
            C.bytes_acc += s_b * (restart ? 3 : 6) + (restart ? 0 : vcb);
        }
        if (restart) {
            SArgs<double> a = base_sargs<double>(op);
            a.precond = precond;
            a.in0 = x;
            a.bvec = b;
            a.out0 = r;
            a.partials = part;
            a.counter = cnt + 1;
            a.ctrl = ctrl;
            a.fin.fin = FIN_CG_BETA;
            a.fin.it = k + 1;
            a.fin.hist = hist;
            a.fin.hist_cap = max_iter + 4;
            if (int rc = run_stencil<double>(SMode::CG_RESID, a, s_b * 3 + vcb)) return rc;
        }
        if ((k + 1) % 8 == 0) {
            if (int rc = poll_record(ctrl, cyc)) return rc;
            if (poll_check(cyc - 2)) stopped = true;
            cyc++;
        }
    }
    HC(hipEventRecord(C.t1, C.stream));
    Ctrl &hc = C.ctrl_h[0];
    if (int rc = finish_ctrl(ctrl, &hc)) return rc;
    float ms = 0.f;
    HC(hipEventElapsedTime(&ms, C.t0, C.t1));
    prof_flush();
    const int iterations = hc.done ? hc.final_iter : max_iter;
    const int hl = std::min(iterations + 1, max_iter + 4);
    if (int rc = store_history(hist, hl)) return rc;
    fill_stats(st, hc, iterations, iterations, max_iter, ms, C.bytes_acc, hl);
    if (hc.status == 2) return fail(AAR_HIP_ERR_NONFINITE, "residual is not finite");
    return 0;
}

// ---------------------------------------------------------------------------
// matvec
// ---------------------------------------------------------------------------
template <typename T>
int matvec(const Op &op, const T *x, T *y)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!x || !y) return fail(AAR_HIP_ERR_ARG, "null vector");
    if (op.c && sizeof(T) != 8) return fail(AAR_HIP_ERR_ARG, "the variable-coefficient matvec is FP64");
    SArgs<T> a = base_sargs<T>(op);
    a.in0 = x;
    a.out0 = y;
    const double by = (op.c ? 3.0 : 2.0) * sizeof(T) * op.n;
    if (int rc = run_stencil<T>(SMode::MATVEC, a, by)) return rc;
    HC(hipStreamSynchronize(C.stream));
    prof_flush();
    return 0;
}
template int matvec<double>(const Op &, const double *, double *);
template int matvec<float>(const Op &, const float *, float *);

// ---------------------------------------------------------------------------
// Complex Helmholtz AAR (extension; AAR.c:37-177 control flow with the
// Hermitian Gram of PETSC_COMPLEX/AAR_Complex.c:207-246) by the same residual
// recurrence as aar_rec: with the Jacobi diagonal D(x) = (a*3*c0 + c(x)) +
// i eta, a Richardson step changes g = D^{-1}(b - A u) by
// g' = g - omega D^{-1} A g, pointwise.
// ---------------------------------------------------------------------------
int make_hop(const aar_hip_helm *h, HOp &op)
{
    if (!h) return fail(AAR_HIP_ERR_ARG, "null operator");
    if (h->nx < 1 || h->ny < 1 || h->nz < 1) return fail(AAR_HIP_ERR_ARG, "grid sizes must be positive");
    if ((double)h->nx * h->ny * h->nz > 1073741823.0) return fail(AAR_HIP_ERR_ARG, "grid too large");
    double c[R + 1];
    if (int rc = stencil_weights(h->fdn, nullptr, c, &op.fdn)) return rc;
    op.nx = h->nx;
    op.ny = h->ny;
    op.nz = h->nz;
    op.n = (long)h->nx * h->ny * h->nz;
    for (int o = 0; o <= R; o++) op.w[o] = h->a * c[o];
    op.dlap = h->a * (3 * c[0]);
    op.c = h->c;
    op.c0 = h->c0;
    op.eta = h->eta;
    return 0;
}

static HArgs base_hargs(const HOp &op)
{
    HArgs a;
    a.nx = op.nx;
    a.ny = op.ny;
    a.nz = op.nz;
    for (int o = 0; o <= R; o++) a.w[o] = op.w[o];
    a.dlap = op.dlap;
    a.c = op.c;
    a.c0 = op.c0;
    a.eta = op.eta;
    return a;
}

static const char *hmode_name(HMode m)
{
    switch (m) {
    case HMode::MATVEC: return "helm_matvec";
    case HMode::TRUE_RES: return "helm_true";
    case HMode::REC: return "helm_rec";
    }
    return "helm";
}

static int run_hstencil(HMode mode, const HArgs &a, double bytes)
{
    Context &C = ctx();
    prof_begin(hmode_name(mode));
    hipError_t e = launch_hstencil(mode, a, C.stream);
    prof_end(bytes);
    if (e == hipSuccess && a.partials)
        e = launch_finish(a.partials, hstencil_items(a), 1, a.fin, a.ctrl, C.stream);
    if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("helmholtz stencil launch: ") + hipGetErrorString(e));
    C.bytes_acc += bytes;
    return 0;
}

// y = A u on the library stream, no wait; skipped once gate->done is set
int helm_matvec_async(const HOp &op, const double *u, double *y, Ctrl *gate)
{
    HArgs a = base_hargs(op);
    a.in0 = (const double2 *)u;
    a.out0 = (double2 *)y;
    a.ctrl = gate;
    return run_hstencil(HMode::MATVEC, a, 32.0 * op.n + (op.c ? 8.0 * op.n : 0.0));
}

int helm_matvec(const HOp &op, const double *u, double *y)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!u || !y) return fail(AAR_HIP_ERR_ARG, "null vector");
    HArgs a = base_hargs(op);
    a.in0 = (const double2 *)u;
    a.out0 = (double2 *)y;
    if (int rc = run_hstencil(HMode::MATVEC, a, 32.0 * op.n + (op.c ? 8.0 * op.n : 0.0))) return rc;
    HC(hipStreamSynchronize(C.stream));
    prof_flush();
    return 0;
}

int helm_solve(const HOp &op, double *u, const double *b, double omega, double beta, int m, int p, int max_iter,
               double tol, aar_hip_stats *st)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!u || !b) return fail(AAR_HIP_ERR_ARG, "null vector");
    if (m < 1 || m > MAXM_CPLX) return fail(AAR_HIP_ERR_ARG, "complex solver: history m must be in 1..16");
    if (p < 1) return fail(AAR_HIP_ERR_ARG, "p must be >= 1");
    if (max_iter < 0) return fail(AAR_HIP_ERR_ARG, "max_iter must be >= 0");
    const long n = op.n;
    const int S = m + 2;
    const int NC = (p <= m) ? m / p + 1 : 0;
    const long ld = ring_ld(n, 16);
    HC(ensure(C.ringc, 16 * (size_t)ld * S));
    if (NC > 0) HC(ensure(C.corrc, 16 * (size_t)ld * NC));
    HC(ensure(C.hist0, sizeof(double) * (size_t)(max_iter + 4)));
    const int nb = nblocks_for(n);
    HArgs probe = base_hargs(op);
    const size_t need = std::max<size_t>((size_t)hstencil_items(probe), (size_t)nb * HGRAM_NA_MAX);
    HC(ensure(C.part0, sizeof(double) * need));
    double2 *G = (double2 *)C.ringc.p;
    double2 *CR = NC > 0 ? (double2 *)C.corrc.p : nullptr;
    double2 *x = (double2 *)u;
    const double2 *bb = (const double2 *)b;
    double *hist = (double *)C.hist0.p, *part = (double *)C.part0.p;
    Ctrl *ctrl = C.ctrl;
    unsigned *cnt = C.counters;
    auto gs = [&](long k) { return (int)(k % S); };
    auto gptr = [&](long k) { return G + (size_t)gs(k) * ld; };
    auto is_and = [&](long k) { return k > 1 && k % p == 0; };
    auto next_and = [&](long k) { long q = (k / p + 1) * p; return q > 1 ? q : q + p; };
    const double sv = 16.0 * n, sc = op.c ? 8.0 * n : 0.0;
    C.bytes_acc = 0.0;

    auto true_res = [&](long k, int check) -> int {
        HArgs a = base_hargs(op);
        a.in0 = x;
        a.bvec = bb;
        a.out0 = gptr(k);
        a.partials = part;
        a.ctrl = ctrl;
        a.fin.fin = FIN_AAR_RES;
        a.fin.it = (int)k;
        a.fin.check = check;
        a.fin.hist = hist;
        a.fin.hist_cap = max_iter + 4;
        return run_hstencil(HMode::TRUE_RES, a, 3 * sv + sc);
    };
    auto anchor = [&](HAncArgs &hu, double bytes) -> int {
        prof_begin("helm_anchor_update");
        hipError_t e = launch_hanchor_update(hu, C.stream);
        prof_end(bytes);
        if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("helmholtz update: ") + hipGetErrorString(e));
        C.bytes_acc += bytes;
        return 0;
    };

    HC(launch_zero_ctrl(ctrl, C.stream));
    HC(hipEventRecord(C.t0, C.stream));
    {
        DotArgs<double> d;              // ||b||^2 over the 2n reals
        d.n = 2 * n;
        d.v = b;
        d.nblocks = nb;
        d.partials = part;
        d.counter = cnt + 0;
        d.ctrl = ctrl;
        d.fin.fin = FIN_NORM_B;
        d.fin.tol = tol;
        d.fin.check = 1;
        HC(run_dot<double>(d));
        C.bytes_acc += sv;
    }
    if (int rc = true_res(1, 0)) return rc;
    long first = 1;
    int cyc = 0;
    long it = 1;
    bool stopped = false;
    for (; it <= max_iter && !stopped; it++) {
        if (!is_and(it)) {
            HArgs a = base_hargs(op);
            a.omega = omega;
            a.in0 = gptr(it);
            a.out0 = gptr(it + 1);
            a.trace = 1;
            a.partials = part;
            a.ctrl = ctrl;
            a.fin.fin = FIN_AAR_RES;
            a.fin.it = (int)it + 1;
            a.fin.check = 0;
            a.fin.hist = hist;
            a.fin.hist_cap = max_iter + 4;
            if (int rc = run_hstencil(HMode::REC, a, 2 * sv + sc)) return rc;
            if (it < next_and(it) - m) {
                HAncArgs hu;
                hu.n = n;
                hu.x = x;
                hu.G = G;
                hu.ld = ld;
                hu.m = m;
                hu.ng = 1;
                hu.gslot[0] = gs(it);
                hu.xlo = 0;
                hu.xhi = 1;
                hu.omega = omega;
                hu.ctrl = ctrl;
                if (int rc = anchor(hu, 3 * sv)) return rc;
                first = it + 1;
            }
            continue;
        }
        const int nv = (int)std::min<long>(m, it - 1);
        {
            HGramArgs g;
            for (int k = 0; k <= m; k++) g.slot[k] = it - k >= 1 ? gs(it - k) : 0;
            g.n = n;
            g.m = m;
            g.nvalid = nv;
            g.it = (int)it;
            g.F = G;
            g.ld = ld;
            g.nblocks = nb;
            g.partials = part;
            g.ctrl = ctrl;
            prof_begin("helm_gram");
            HC(launch_hgram(g, C.stream));
            prof_end(sv * (nv + 1));
            C.bytes_acc += sv * (nv + 1);
        }
        prof_begin("ls");
        HC(launch_ls(ctrl, 2 * m, C.complex_ref_order ? 0 : LS_TEXTBOOK, C.stream));
        prof_end(0);
        {
            HAncArgs hu;
            const long jlo = std::min(first, it - nv);
            hu.n = n;
            hu.x = x;
            hu.G = G;
            hu.C = CR;
            hu.ld = ld;
            hu.m = m;
            hu.ng = (int)(it - jlo + 1);
            for (int i = 0; i < hu.ng; i++) hu.gslot[i] = gs(jlo + i);
            hu.xlo = (int)(first - jlo);
            hu.xhi = (int)(it - jlo);
            hu.with_beta = 1;
            hu.ncol = nv;
            hu.nc = 0;
            for (int t = 0; t < nv; t++) {
                const long j = it - t;
                hu.col_ref[t] = (int)((j - 2) % m);
                hu.col_j[t] = (int)(j - jlo);
                hu.col_c[t] = -1;
                if (is_and(j - 1)) {
                    hu.cslot[hu.nc] = (int)(((j - 1) / p) % NC);
                    hu.col_c[t] = hu.nc++;
                }
            }
            hu.cout = NC > 0 ? (int)((it / p) % NC) : -1;
            hu.omega = omega;
            hu.beta = beta;
            hu.ctrl = ctrl;
            if (int rc = anchor(hu, sv * (2 + hu.ng + hu.nc + (hu.cout >= 0 ? 1 : 0)))) return rc;
        }
        first = it + 1;
        if (int rc = true_res(it + 1, 1)) return rc;
        if (int rc = poll_record(ctrl, cyc)) return rc;
        if (poll_check(cyc - 2)) stopped = true;
        cyc++;
    }
    if (!stopped && first <= max_iter) {
        HAncArgs hu;
        hu.n = n;
        hu.x = x;
        hu.G = G;
        hu.ld = ld;
        hu.m = m;
        hu.ng = (int)(max_iter - first + 1);
        for (int i = 0; i < hu.ng; i++) hu.gslot[i] = gs(first + i);
        hu.xlo = 0;
        hu.xhi = hu.ng;
        hu.omega = omega;
        hu.ctrl = ctrl;
        if (int rc = anchor(hu, sv * (2 + hu.ng))) return rc;
    }
    HC(hipEventRecord(C.t1, C.stream));
    Ctrl &hc = C.ctrl_h[0];
    if (int rc = finish_ctrl(ctrl, &hc)) return rc;
    float ms = 0.f;
    HC(hipEventElapsedTime(&ms, C.t0, C.t1));
    prof_flush();
    int iterations, loop_iter;
    if (hc.done) {
        iterations = hc.final_iter;
        loop_iter = hc.final_iter + 1;
    } else {
        iterations = max_iter;
        loop_iter = max_iter + 1;
    }
    const int hl = std::min(iterations + 1, max_iter + 4);
    if (int rc = store_history(hist, hl)) return rc;
    fill_stats(st, hc, iterations, loop_iter, max_iter, ms, C.bytes_acc, hl);
    if (hc.status == 1) return fail(AAR_HIP_ERR_SVD, "no convergence in 250 svd iterations");
    if (hc.status == 2) return fail(AAR_HIP_ERR_NONFINITE, "residual is not finite");
    return 0;
}

// ---------------------------------------------------------------------------
// FP64 residual refinement around FP32 AAR (extension)
// ---------------------------------------------------------------------------
int aar_refined(const Op &op, double *x, const double *b, double omega, double beta, int m, int p,
                int max_outer, int max_inner, double tol, double inner_tol, int precond,
                aar_hip_stats *st)
{
    if (op.c)
        return fail(AAR_HIP_ERR_ARG, "the refined solver has no variable-coefficient form (use AAR mixed)");
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (!x || !b) return fail(AAR_HIP_ERR_ARG, "null vector");
    const long n = op.n;
    HC(ensure(C.vec32a, sizeof(float) * n));   // r (FP32 RHS of the correction)
    HC(ensure(C.vec32b, sizeof(float) * n));   // e (FP32 correction)
    HC(ensure(C.hist0, sizeof(double) * (size_t)(max_outer + 4)));
    const int nb = nblocks_for(n);
    const size_t need = std::max<size_t>((size_t)stencil_partials_needed<double>(op), (size_t)nb * GRAM_NA_MAX);
    HC(ensure(C.part0, sizeof(double) * need));
    float *r32 = (float *)C.vec32a.p, *e32 = (float *)C.vec32b.p;
    double *part = (double *)C.part0.p;
    Ctrl *ctrl = C.ctrl;
    unsigned *cnt = C.counters;
    const double s_b = 8.0 * n;
    Op op32 = op;
    std::vector<double> hist_out;
    double bytes = 0.0, ms_total = 0.0;
    int inner_total = 0, outer = 0;
    const bool prof_was = C.prof;

    hipEvent_t e0, e1;
    HC(hipEventCreate(&e0));
    HC(hipEventCreate(&e1));
    HC(launch_zero_ctrl(ctrl, C.stream));
    HC(hipEventRecord(e0, C.stream));
    {
        DotArgs<double> d;
        d.n = n;
        d.v = b;
        d.nblocks = nb;
        d.partials = part;
        d.counter = cnt + 0;
        d.ctrl = ctrl;
        d.fin.fin = FIN_NORM_B;
        d.fin.tol = tol;
        HC(run_dot<double>(d));
        bytes += s_b;
    }
    double rhs_norm = 0.0, rel = 0.0;
    int status = 0;
    for (outer = 0;; outer++) {
        SArgs<double> a = base_sargs<double>(op);
        a.in0 = x;
        a.bvec = b;
        a.out32 = r32;
        a.partials = part;
        a.counter = cnt + 1;
        a.ctrl = ctrl;
        a.fin.fin = FIN_SUMS;
        if (int rc = run_stencil<double>(SMode::RESID, a, s_b * 2 + 4.0 * n)) return rc;
        bytes += s_b * 2 + 4.0 * n;
        Ctrl &hc = C.ctrl_h[0];
        if (int rc = finish_ctrl(ctrl, &hc)) return rc;
        rhs_norm = hc.rhs_norm;
        const double rn = std::sqrt(hc.red[0]);
        rel = rhs_norm > 0 ? rn / rhs_norm : rn;
        hist_out.push_back(rn);
        if (!(rel == rel)) { status = AAR_HIP_ERR_NONFINITE; break; }
        if (rel <= tol || outer >= max_outer) break;
        const double tin = std::max(inner_tol, 0.5 * tol / rel);
        HC(launch_fill_f32(e32, 0.0f, n, C.stream));
        aar_hip_stats ist;
        int rc = aar_solve<float>(op32, e32, r32, omega, beta, m, p, max_inner, tin, precond, 1, &ist, 0, 0);
        if (rc) return rc;
        inner_total += ist.iterations;
        bytes += ist.bytes + 4.0 * n;
        HC(launch_axpy_f64_f32(x, e32, n, C.stream));
        bytes += s_b * 2 + 4.0 * n;
    }
    HC(hipEventRecord(e1, C.stream));
    HC(hipStreamSynchronize(C.stream));
    float ms = 0.f;
    HC(hipEventElapsedTime(&ms, e0, e1));
    ms_total = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)prof_was;
    C.hist_host = hist_out;
    aar_hip_stats s;
    std::memset(&s, 0, sizeof(s));
    s.iterations = inner_total;
    s.inner_iterations = inner_total;
    s.outer_steps = outer;
    s.relres = rel;
    s.rhs_norm = rhs_norm;
    s.converged = rel <= tol;
    s.printed_converged = s.converged;
    s.status = status;
    s.hist_len = (int)hist_out.size();
    s.solve_ms = ms_total;
    s.bytes = bytes;
    C.last = s;
    if (st) *st = s;
    if (status) return fail(status, "residual is not finite");
    return 0;
}

// ---------------------------------------------------------------------------
// pinv on device (host matrices)
// ---------------------------------------------------------------------------
int pinv_host(const double *A, const double *b, double *x, int m, double *w_out, double *v_out)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    if (m < 1 || m > MAXM_LS) return fail(AAR_HIP_ERR_ARG, "pinv: m must be in 1..32");
    Ctrl &hc = C.ctrl_h[1];
    std::memset(&hc, 0, sizeof(Ctrl));
    // allredvec layout: A[i][j] at G[j*m + i], b at G[m*m + j]
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) hc.G[j * m + i] = A[i * m + j];
    for (int j = 0; j < m; j++) hc.G[m * m + j] = b[j];
    Ctrl *dc = C.ctrl + 1;
    HC(hipMemcpyAsync(dc, &hc, sizeof(Ctrl), hipMemcpyHostToDevice, C.stream));
    HC(launch_ls(dc, m, 0, C.stream));
    HC(hipMemcpyAsync(&hc, dc, sizeof(Ctrl), hipMemcpyDeviceToHost, C.stream));
    HC(hipStreamSynchronize(C.stream));
    for (int j = 0; j < m; j++) x[j] = hc.s[j];
    if (w_out)
        for (int j = 0; j < m; j++) w_out[j] = hc.w[j];
    if (v_out)
        for (int j = 0; j < m * m; j++) v_out[j] = hc.V[j];
    if (hc.status == 1) return fail(AAR_HIP_ERR_SVD, "no convergence in 250 svd iterations");
    return 0;
}

void release_all()
{
    Context &C = ctx();
    DevBuf *bufs[] = {&C.ring64x, &C.ring64f, &C.ring32x, &C.ring32f, &C.hist0, &C.hist1,
                      &C.part0, &C.part1, &C.vec32a, &C.vec32b, &C.vec64a, &C.vec64b,
                      &C.ringc, &C.corrc, &C.vch, &C.kctl, &C.kvec, &C.kilu};
    for (DevBuf *b : bufs) {
        if (b->p) (void)hipFree(b->p);
        b->p = nullptr;
        b->bytes = 0;
    }
}

}  // namespace aar
