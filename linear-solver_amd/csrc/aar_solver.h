// aar_solver.h -- host-side context and solver entry points (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <string>
#include <vector>

#include "aar_common.h"
#include "aar_hip.h"
#include "aar_kernels.h"

namespace aar {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct ProfEntry {
    int count = 0;
    double ms = 0.0;
    double bytes = 0.0;
};

// launch shape of the stencil sweeps for one grid and element size
struct Shape {
    int lz = 0, wy = 8, py = 4;
    double us = 0.0;          // measured time of one recurrence sweep (tuned shapes)
};

struct PendingEv {
    std::string name;
    hipEvent_t a = nullptr, b = nullptr;
    double bytes = 0.0;
};

constexpr int NPOLL = 64;

// One per process: the library is single-GPU and, like the reference,
// not re-entrant across threads.
struct Context {
    bool init = false;
    hipStream_t own = nullptr, stream = nullptr;
    bool stream_set = false;     // aar_hip_set_stream() chose `stream` (may be the null stream)
    Ctrl *ctrl = nullptr;        // [2]: 0 outer / plain solves, 1 inner FP32 solves
    Ctrl *ctrl_h = nullptr;      // pinned host mirror [2]
    int *poll_h = nullptr;       // pinned polling slots
    hipEvent_t poll_ev[NPOLL] = {};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    unsigned *counters = nullptr;   // last-block counters (64)
    DevBuf ring64x, ring64f, ring32x, ring32f;
    DevBuf hist0, hist1, part0, part1;
    DevBuf vec32a, vec32b, vec64a, vec64b;
    DevBuf ringc, corrc;         // complex Helmholtz g ring and corrections
    DevBuf vch;                  // variable coefficient: the recurrence field h(x)
    DevBuf kctl, kvec, kilu;     // Krylov comparators: control block, basis / work vectors, ILU(0) factors
    std::string err;
    int lz = 0, wy = 8, py = 4;
    long ring_pad = 512;           // extra elements per ring slot (multiple of 32)
    int hist_blocks = 0;           // grid of the reducing history kernels (0: automatic)
    bool two_pass = true;          // reductions finish in a one-block second kernel
    bool prof = false;
    bool tuning_set = false;       // aar_hip_set_tuning() fixed the shape for every grid
    bool autotune = false;         // tune a grid's shape on its first solve
    bool complex_ref_order = false;  // complex solver's SVD in the tools.c order (default: textbook)
    int user_fdn = 0;              // registered stencil weights (aar_hip_set_stencil): order / 2, 0 = none
    double user_w[R + 1] = {};
    std::map<std::array<int, 4>, Shape> tuned;        // (nx, ny, nz, bytes) -> shape
    std::map<std::array<int, 4>, Shape> tuned_batch;  // lz for the batched 8x4 launch
    std::map<std::string, ProfEntry> profile;
    std::vector<PendingEv> pending;
    std::vector<hipEvent_t> evpool;
    aar_hip_stats last = {};
    std::vector<double> hist_host;
    double bytes_acc = 0.0;
};

struct Op {
    int nx = 0, ny = 0, nz = 0;
    long n = 0;
    int fdn = R;              // stencil order / 2: weights w[o] = 0 for o > fdn
    double w[R + 1] = {};
    double sigma = 0.0;
    double diag = 1.0;
    // variable coefficient (device, n doubles): A = a*Lap + c(x) with the
    // pointwise Jacobi diagonal diag + c(x); sigma is 0 then
    const double *c = nullptr;
};

Context &ctx();
int context_init();
hipError_t ensure(DevBuf &b, size_t bytes);
int auto_lz(const Op &op, int wy, int py, size_t es);
long ring_ld(long n, size_t es);
int nblocks_for(long n);
// grid of the register Gram (m <= 8) on n points
int gram_blocks(long n, int m);
template <typename T>
SArgs<T> base_sargs(const Op &op);
template <typename T>
hipError_t run_dot(DotArgs<T> d);
int poll_record(Ctrl *ctrl, int cyc);
Shape stencil_shape(int nx, int ny, int nz, size_t es);
int autotune_stencil(int nx, int ny, int nz, size_t es, bool batch_shape, Shape *out);
bool poll_check(int cyc);
void prof_begin(const char *name);
void prof_end(double bytes);
int fail(int code, const std::string &msg);
void lap_coefficient(double *c, int fdn);
int make_op(const aar_hip_grid *g, const double *coeff, Op &op);
int stencil_weights(int fdn_in, const double *coeff, double (&c)[R + 1], int *fdn_out);

template <typename T>
int aar_solve(const Op &op, T *x, const T *b, double omega, double beta, int m, int p, int max_iter,
              double tol, int precond, int ci, aar_hip_stats *st, int keep_history, int ref_start = 1);
int aar_mixed(const Op &op, double *x, const double *b, double omega, double beta, int m, int p,
              int max_iter, double tol, int precond, aar_hip_stats *st);
int pgr_solve(const Op &op, double *x, const double *b, double omega, int m, int p, int max_iter,
              double tol, int precond, int l2, aar_hip_stats *st);
int cg_solve(const Op &op, double *x, const double *b, int max_iter, double tol, int precond,
             aar_hip_stats *st);
int gmres_solve(const Op &op, double *x, const double *b, int restart, int max_it, double tol, int precond,
                aar_hip_stats *st);
int bicg_solve(const Op &op, double *x, const double *b, int max_it, double tol, int precond, aar_hip_stats *st);
int kspcg_solve(const Op &op, double *x, const double *b, int max_it, double tol, int precond, aar_hip_stats *st);
int ilu0_apply(const Op &op, const double *r, double *z, int trans);
int aar_petsc_solve(const Op &op, double *x, const double *b, double omega, double beta, int m, int p,
                    int max_iter, double tol, int precond, aar_hip_stats *st);
int pgr_petsc_solve(const Op &op, double *x, const double *b, double omega, int m, int p, int max_iter, double tol,
                    int precond, int l2, aar_hip_stats *st);
template <typename T>
int run_stencil(SMode mode, SArgs<T> a, double bytes);
template <typename T>
int matvec(const Op &op, const T *x, T *y);
int aar_refined(const Op &op, double *x, const double *b, double omega, double beta, int m, int p,
                int max_outer, int max_inner, double tol, double inner_tol, int precond,
                aar_hip_stats *st);
// complex Helmholtz operator: a*Lap + (c(x) + i eta), complex128 interleaved
struct HOp {
    int nx = 0, ny = 0, nz = 0;
    long n = 0;
    int fdn = R;
    double w[R + 1] = {};
    double dlap = 0.0;
    const double *c = nullptr;
    double c0 = 0.0, eta = 0.0;
};
int make_hop(const aar_hip_helm *h, HOp &op);
int helm_matvec(const HOp &op, const double *u, double *y);
int helm_matvec_async(const HOp &op, const double *u, double *y, Ctrl *gate);
int helm_gmres_solve(const HOp &op, double *u, const double *b, int restart, int max_it, double tol, int precond,
                     aar_hip_stats *st);
int helm_lgmres_solve(const HOp &op, double *u, const double *b, int restart, int aug, int max_it, double tol,
                      int precond, aar_hip_stats *st);
int helm_bicg_solve(const HOp &op, double *u, const double *b, int max_it, double tol, int precond,
                    aar_hip_stats *st);
int helm_ilu0_apply(const HOp &op, const double *r, double *z, int trans);
int helm_solve(const HOp &op, double *u, const double *b, double omega, double beta, int m, int p, int max_iter,
               double tol, aar_hip_stats *st);
template <typename Tx, typename Tg>
int batch_solve(const aar_hip_instance *ins, int count, double omega, double beta, int m, int p, int max_iter,
                double tol, int precond, aar_hip_stats *stats);
int pinv_host(const double *A, const double *b, double *x, int m, double *w_out, double *v_out);
void release_all();
void prof_flush();

}  // namespace aar
