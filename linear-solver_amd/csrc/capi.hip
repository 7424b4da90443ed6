// capi.hip -- the C ABI of libaar_hip.
//
// Part 1: the reference's own entry points (MPI_REAL/AAR.h, PGR.h, PL2R.h,
// CG.h, system.h, tools.h), so that the reference driver main.c relinks
// against libaar_hip unchanged.  Host arrays in and out; every numerical
// step runs on the GPU.  Single rank: the decomposition routines produce
// the one-rank layout and no MPI call is made.
// Part 2: the device-resident extension API (include/aar_hip.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "aar_common.h"
#include "aar_kernels.h"
#include "aar_solver.h"
#include "poisson_abi.h"

using namespace aar;

#define HCR(x, ret)                                                                            \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                \
            fail(AAR_HIP_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));            \
            return ret;                                                                        \
        }                                                                                      \
    } while (0)

extern "C" {
void Lap_Vec_mult(POISSON *system, double a, double *phi, double *Lap_phi, aar_mpi_comm comm_laplacian);
void Precondition(double diag, double *res, int Np);
void Lap_Vec_mult_vc(POISSON *system, double a, double *phi, double *Lap_phi, aar_mpi_comm comm_laplacian);
void Precondition_vc(double diag, double *res, int Np);
void Lap_coefficient(double *coeff_lap, int FDn);
}

namespace {

// device staging for the host-array entry points
struct Staging {
    double *x = nullptr, *b = nullptr;
    size_t cap = 0;
};
Staging g_stage;

int stage(long n)
{
    if (int rc = context_init()) return rc;
    const size_t bytes = sizeof(double) * (size_t)n;
    if (g_stage.cap >= bytes) return 0;
    if (g_stage.x) (void)hipFree(g_stage.x);
    if (g_stage.b) (void)hipFree(g_stage.b);
    g_stage.x = g_stage.b = nullptr;
    g_stage.cap = 0;
    HCR(hipMalloc(&g_stage.x, bytes), AAR_HIP_ERR_HIP);
    HCR(hipMalloc(&g_stage.b, bytes), AAR_HIP_ERR_HIP);
    g_stage.cap = bytes;
    return 0;
}

// the coefficient field registered for the Lap_Vec_mult_vc /
// Precondition_vc slots (aar_hip_set_coefficient)
DevBuf g_coef;
long g_coef_n = 0;

void stage_release()
{
    if (g_stage.x) (void)hipFree(g_stage.x);
    if (g_stage.b) (void)hipFree(g_stage.b);
    g_stage = Staging();
    if (g_coef.p) (void)hipFree(g_coef.p);
    g_coef = DevBuf();
    g_coef_n = 0;
}

// The reference's operator slot: only libaar_hip's own device operators
// run (on the GPU) -- the FD12 stencil (Lap_Vec_mult, Precondition) or the
// variable-coefficient operator with the registered c(x) (Lap_Vec_mult_vc,
// Precondition_vc); a foreign callback is not executed on the host.
int check_callbacks(const char *who, aar_op_cb op, aar_pc_cb pc, int *precond, int *vc)
{
    const bool opv = op == &Lap_Vec_mult_vc;
    if (op != &Lap_Vec_mult && !opv) {
        std::printf("%s: operator callback is not libaar_hip's Lap_Vec_mult or Lap_Vec_mult_vc; refusing to "
                    "run it (register a device operator instead)\n", who);
        fail(AAR_HIP_ERR_CALLBACK, "foreign operator callback");
        return AAR_HIP_ERR_CALLBACK;
    }
    const aar_pc_cb own = opv ? &Precondition_vc : &Precondition;
    if (pc != nullptr && pc != own) {
        std::printf("%s: preconditioner callback is not libaar_hip's %s; refusing to run it\n", who,
                    opv ? "Precondition_vc (the operator is Lap_Vec_mult_vc)" : "Precondition");
        fail(AAR_HIP_ERR_CALLBACK, "foreign preconditioner callback");
        return AAR_HIP_ERR_CALLBACK;
    }
    *precond = pc != nullptr;
    *vc = opv;
    return 0;
}

// variable-coefficient operator (extension): A = a*Lap + c(x), Jacobi
// diagonal a*(3*coeff[0]) + c(x); g->sigma must be 0 (a shift goes into c)
int make_op_vc(const aar_hip_grid *g, const double *c, Op &op)
{
    if (int rc = make_op(g, nullptr, op)) return rc;
    if (!c) return fail(AAR_HIP_ERR_ARG, "null coefficient field");
    if (g->sigma != 0.0) return fail(AAR_HIP_ERR_ARG, "sigma must be 0 with a coefficient field (add it to c)");
    op.c = c;
    return 0;
}

// the registered coefficient for an n-point grid
int registered_coefficient(long n, const double **c)
{
    if (g_coef_n != n || !g_coef.p)
        return fail(AAR_HIP_ERR_ARG, "no coefficient field of this grid's size is registered "
                                     "(aar_hip_set_coefficient)");
    *c = (const double *)g_coef.p;
    return 0;
}

int op_from_system(POISSON *sys, double a, int Np, Op &op)
{
    if (!sys) return fail(AAR_HIP_ERR_ARG, "null POISSON");
    aar_hip_grid g;
    g.nx = sys->psize[0];
    g.ny = sys->psize[1];
    g.nz = sys->psize[2];
    g.fdn = sys->FDn;
    g.a = a;
    g.sigma = 0.0;
    if ((long)g.nx * g.ny * g.nz != (long)Np) return fail(AAR_HIP_ERR_ARG, "Np != psize product");
    return make_op(&g, sys->coeff_lap, op);
}

void report(const char *name, const aar_hip_stats &s)
{
    // AAR.c:151-159 (PGR.c:145-153, PL2R.c:145-152, CG.c:157-165)
    if (s.printed_converged)
        std::printf("%s converged!:  Iterations = %d, Relative Residual = %g, Time = %.4f sec\n", name,
                    s.iterations, s.relres, s.solve_ms * 1e-3);
    if (s.printed_warning) {
        std::printf("WARNING: %s exceeded maximum iterations.\n", name);
        std::printf("%s:  Iterations = %d, Relative Residual = %g, Time = %.4f sec \n", name, s.iterations,
                    s.relres, s.solve_ms * 1e-3);
    }
    std::fflush(stdout);
}

// With AAR_HIP_DEBUG set (non-empty, not "0"), the trace the reference
// prints when built with -DDEBUG: the relative residual of the start and
// of every iteration (AAR.c:84-86/125-126/141-143, PGR.c:83-85/135-138,
// PL2R.c:83-85/134-137), or CG's absolute residuals (CG.c:141-142).
static bool debug_trace_on()
{
    const char *e = std::getenv("AAR_HIP_DEBUG");
    return e && *e && std::strcmp(e, "0") != 0;
}

static void debug_trace(const char *name, const aar_hip_stats &s)
{
    if (!debug_trace_on()) return;
    const std::vector<double> &h = ctx().hist_host;
    const bool cg = std::strcmp(name, "CG") == 0;
    const int n = std::min<int>((int)h.size(), s.hist_len);
    // CG's history starts with the initial residual, which CG.c does not print
    for (int k = cg ? 1 : 0; k < n; k++) {
        if (cg) std::printf("norm of res: %g\n", h[k]);
        else std::printf("Relative Residual: %g\n", s.rhs_norm > 0 ? h[k] / s.rhs_norm : h[k]);
    }
}

void report_failure(const char *name, int rc)
{
    std::printf("%s: libaar_hip error %d: %s\n", name, rc, ctx().err.c_str());
    std::fflush(stdout);
}

// One HIP call of a host-array entry point: a failure is recorded
// (aar_hip_last_error) and printed as the entry's failure line.
bool hip_ok(hipError_t e, const char *name, const char *what)
{
    if (e == hipSuccess) return true;
    fail(AAR_HIP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    report_failure(name, AAR_HIP_ERR_HIP);
    return false;
}

// run one of the four solvers on host arrays
template <class F>
void host_solve(const char *name, POISSON *sys, aar_op_cb opcb, aar_pc_cb pccb, double a, double *x,
                double *rhs, int Np, F &&run)
{
    int precond = 0, vc = 0;
    if (check_callbacks(name, opcb, pccb, &precond, &vc)) return;
    Op op;
    if (int rc = op_from_system(sys, a, Np, op)) { report_failure(name, rc); return; }
    if (vc)
        if (int rc = registered_coefficient(op.n, &op.c)) { report_failure(name, rc); return; }
    if (int rc = stage(op.n)) { report_failure(name, rc); return; }
    Context &C = ctx();
    const size_t bytes = sizeof(double) * (size_t)op.n;
    if (hipMemcpyAsync(g_stage.x, x, bytes, hipMemcpyHostToDevice, C.stream) != hipSuccess ||
        hipMemcpyAsync(g_stage.b, rhs, bytes, hipMemcpyHostToDevice, C.stream) != hipSuccess) {
        fail(AAR_HIP_ERR_HIP, "H2D copy failed");
        report_failure(name, AAR_HIP_ERR_HIP);
        return;
    }
    aar_hip_stats st;
    const int rc = run(op, g_stage.x, g_stage.b, precond, &st);
    if (rc && rc != AAR_HIP_ERR_SVD && rc != AAR_HIP_ERR_NONFINITE) { report_failure(name, rc); return; }
    if (hipMemcpyAsync(x, g_stage.x, bytes, hipMemcpyDeviceToHost, C.stream) != hipSuccess ||
        hipStreamSynchronize(C.stream) != hipSuccess) {
        fail(AAR_HIP_ERR_HIP, "D2H copy failed");
        report_failure(name, AAR_HIP_ERR_HIP);
        return;
    }
    debug_trace(name, st);
    if (rc == AAR_HIP_ERR_SVD) {
        // tools.c:268-271: the reference prints this and ends the process
        std::printf("no convergence in %d svd iterations \n", 250);
        std::fflush(stdout);
        std::exit(1);
    }
    report(name, st);
}

// host matrices -> device slots for the stand-alone Anderson / Galerkin /
// L2 steps (DIRECT mode of the history kernels): slot j+1 holds column j,
// slot m+1 of F holds f, slot 0 holds zeros (the update's base point and,
// for PGR/PL2R, the A x of f = b - A x with b = f).
struct StepBufs {
    double *X = nullptr, *F = nullptr;
};

int step_upload(double **DX, double **DF, const double *f, int m, int n, StepBufs &sb)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    const size_t bytes = sizeof(double) * (size_t)n;
    const size_t need = bytes * (size_t)(m + 2);
    static DevBuf bx, bf;
    auto ens = [&](DevBuf &b) -> int {
        if (b.bytes >= need) return 0;
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
        HCR(hipMalloc(&b.p, need), AAR_HIP_ERR_HIP);
        b.bytes = need;
        return 0;
    };
    if (int rc = ens(bx)) return rc;
    if (int rc = ens(bf)) return rc;
    sb.X = (double *)bx.p;
    sb.F = (double *)bf.p;
    HCR(launch_fill_f64(sb.X, 0.0, n, C.stream), AAR_HIP_ERR_HIP);
    HCR(launch_fill_f64(sb.F, 0.0, n, C.stream), AAR_HIP_ERR_HIP);
    for (int j = 0; j < m; j++) {
        if (DX) HCR(hipMemcpyAsync(sb.X + (size_t)(j + 1) * n, DX[j], bytes, hipMemcpyHostToDevice, C.stream), AAR_HIP_ERR_HIP);
        HCR(hipMemcpyAsync(sb.F + (size_t)(j + 1) * n, DF[j], bytes, hipMemcpyHostToDevice, C.stream), AAR_HIP_ERR_HIP);
    }
    HCR(hipMemcpyAsync(sb.F + (size_t)(m + 1) * n, f, bytes, hipMemcpyHostToDevice, C.stream), AAR_HIP_ERR_HIP);
    return 0;
}

// kind 0: AAR (f in F slot m+1); kind 1: PGR (X = DX, F = DF, f as b);
// kind 2: PL2R.  Leaves G/g/s in ctrl[1] (host mirror after sync).
int step_gram_ls(int kind, const StepBufs &sb, int m, int n)
{
    Context &C = ctx();
    Ctrl *dc = C.ctrl + 1;
    static DevBuf part;
    const int nb = (int)std::max(1L, std::min(2048L, ((long)n + 4095) / 4096));
    if (part.bytes < sizeof(double) * nb * GRAM_NA_MAX) {
        if (part.p) (void)hipFree(part.p);
        part.p = nullptr;
        part.bytes = 0;
        HCR(hipMalloc(&part.p, sizeof(double) * nb * GRAM_NA_MAX), AAR_HIP_ERR_HIP);
        part.bytes = sizeof(double) * nb * GRAM_NA_MAX;
    }
    HCR(launch_zero_ctrl(dc, C.stream), AAR_HIP_ERR_HIP);
    GramArgs<double> g;
    g.n = n;
    g.m = m;
    g.kind = kind;
    g.ld = n;
    g.direct = 1;
    g.nvalid = m;
    for (int j = 0; j < m; j++) g.slot[j] = j + 1;
    g.nblocks = nb;
    g.partials = (double *)part.p;
    g.counter = C.counters + 32;
    g.ctrl = dc;
    if (kind == 0) {
        g.F = sb.F;
        g.fslot = m + 1;
    } else {
        g.F = sb.X;
        g.Y = sb.F;
        g.b = sb.F + (size_t)(m + 1) * n;   // f = f - 0
        g.fslot = 0;
    }
    HCR(launch_gram<double>(g, C.stream), AAR_HIP_ERR_HIP);
    HCR(launch_ls(dc, m, 0, C.stream), AAR_HIP_ERR_HIP);
    return 0;
}

hipError_t step_outputs(int m, double **FtF, double *allredvec, double *Ftf, double *svec)
{
    Context &C = ctx();
    Ctrl &hc = C.ctrl_h[1];
    hipError_t e = hipMemcpyAsync(&hc, C.ctrl + 1, sizeof(Ctrl), hipMemcpyDeviceToHost, C.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(C.stream);
    if (e != hipSuccess) return e;
    // AAR.c:212-221: FtF[i][j] = allredvec[j*m + i]; Ftf[j] = allredvec[m*m + j]
    if (allredvec) std::memcpy(allredvec, hc.G, sizeof(double) * (m * m + m));
    if (FtF)
        for (int j = 0; j < m; j++)
            for (int i = 0; i < m; i++) FtF[i][j] = hc.G[j * m + i];
    if (Ftf)
        for (int j = 0; j < m; j++) Ftf[j] = hc.G[m * m + j];
    if (svec)
        for (int j = 0; j < m; j++) svec[j] = hc.s[j];
    if (hc.status == 1) {
        // tools.c:268-271 ends the process when the SVD does not converge
        std::printf("no convergence in %d svd iterations \n", 250);
        std::fflush(stdout);
        std::exit(1);
    }
    return hipSuccess;
}

}  // namespace

// ===========================================================================
// Part 1: reference entry points
// ===========================================================================
extern "C" {

// MPI_REAL/system.c:22-64
void CheckInputs(POISSON *system, int argc, char **argv)
{
    std::printf("*************************************************************************** \n");
    std::printf("                     AAR, PGR and PL2R linear solver\n");
    std::printf("*************************************************************************** \n");
    time_t now = time(nullptr);
    std::printf("Starting time: %s \n", ctime(&now));
    if (argc != 6) {
        std::printf("Wrong inputs");
        std::exit(-1);
    }
    system->solver_tol = atof(argv[1]);
    system->m = (int)atof(argv[2]);
    system->p = (int)atof(argv[3]);
    system->omega = atof(argv[4]);
    system->beta = atof(argv[5]);
    std::printf("***************************************************************************\n");
    std::printf("                           INPUT PARAMETERS                                \n");
    std::printf("***************************************************************************\n");
    std::printf("FD Order    : %d\n", system->FDn * 2);
    std::printf("Solver      : AAR\n");
    std::printf("Solver      : PGR\n");
    std::printf("Solver      : PL2R\n");
    std::printf("Solver      : CG\n");
    std::printf("solver_tol  : %e \n", system->solver_tol);
    std::printf("m           : %d\n", system->m);
    std::printf("p           : %d\n", system->p);
    std::printf("omega       : %lf\n", system->omega);
    std::printf("beta        : %lf\n", system->beta);
    std::printf("***************************************************************************\n");
}

// MPI_REAL/system.c:368-399, one rank
void Processor_Domain(int ssize[3], int psize[3], int np[3], int coords[3], int rem[3], aar_mpi_comm comm,
                      aar_mpi_comm *cart)
{
    for (int i = 0; i < 3; i++) {
        np[i] = 1;
        coords[i] = 0;
        rem[i] = 0;
        psize[i] = ssize[i];
    }
    *cart = comm;
}

// MPI_REAL/system.c:238-360, one rank: no neighbours and no ghost traffic
void Comm_topologies(POISSON *system, int FDn, int psize[3], int coords[3], int rem[3], int np[3],
                     aar_mpi_comm cart, aar_mpi_comm *comm_laplacian, int *send_neighs, int *rec_neighs,
                     int *send_counts, int *rec_counts, int send_layers[6], int rec_layers[6], int *n_in,
                     int *n_out)
{
    (void)FDn; (void)psize; (void)coords; (void)rem; (void)np;
    (void)send_neighs; (void)rec_neighs; (void)send_counts; (void)rec_counts;
    for (int i = 0; i < 6; i++) send_layers[i] = rec_layers[i] = 0;
    *n_in = *n_out = 0;
    *comm_laplacian = cart;
    system->scounts = (int *)calloc(1, sizeof(int));
    system->rcounts = (int *)calloc(1, sizeof(int));
    system->sdispls = (int *)calloc(1, sizeof(int));
    system->rdispls = (int *)calloc(1, sizeof(int));
}

// MPI_REAL/system.c:407-415
void Max_layer(int ssize[3], int np[3], int FDn, int *max_layer)
{
    for (int i = 0; i < 3; i++) *max_layer += 2 * (FDn / (ssize[i] / np[i]) + 1);
}

// MPI_REAL/system.c:578-587
void Find_size_dir(int rem, int coords, int psize, int *small, int *large)
{
    if (coords < rem) {
        *large = psize;
        *small = psize - 1;
    } else {
        *large = psize + 1;
        *small = psize;
    }
}

// MPI_REAL/system.c:552-569
void Get_block_origin_global_coords(int coords[3], int rem[3], int psize[3], int g_origin[3])
{
    for (int i = 0; i < 3; i++) {
        int small, large;
        Find_size_dir(rem[i], coords[i], psize[i], &small, &large);
        g_origin[i] = 0;
        for (int j = 0; j < coords[i]; j++) g_origin[i] += (j < rem[i]) ? large : small;
    }
}

// MPI_REAL/system.c:506-513
void Vec_copy(int *a, int *b, int n)
{
    for (int i = 0; i < n; i++) a[i] = b[i];
}

// MPI_REAL/system.c:480-498 (accumulates into coeff_lap[0] as the reference does)
void Lap_coefficient(double *coeff_lap, int FDn)
{
    for (int p = 1; p <= FDn; p++) coeff_lap[0] += -(2.0 / (p * p));
    double tmp[64];
    lap_coefficient(tmp, FDn < 63 ? FDn : 63);
    for (int p = 1; p <= FDn && p < 64; p++) coeff_lap[p] = tmp[p];
}

// MPI_REAL/system.c:425-471: allocations, FD weights and the seeded RHS
// (srand(global index + 1), 2*rand()/RAND_MAX, global mean removed), phi = 1.
void Initialize(POISSON *system, int max_layer)
{
    system->send_neighs = (int *)calloc(max_layer, sizeof(int));
    system->rec_neighs = (int *)calloc(max_layer, sizeof(int));
    system->send_counts = (int *)calloc(max_layer, sizeof(int));
    system->rec_counts = (int *)calloc(max_layer, sizeof(int));
    system->coeff_lap = (double *)calloc(system->FDn + 1, sizeof(double));
    Lap_coefficient(system->coeff_lap, system->FDn);
    const int *ps = system->psize, *ss = system->ssize;
    const long nn = (long)ps[0] * ps[1] * ps[2];
    system->phi = (double *)calloc(nn, sizeof(double));
    system->rhs = (double *)calloc(nn, sizeof(double));
    system->Lap_phi = (double *)calloc(nn, sizeof(double));
    int o[3];
    Get_block_origin_global_coords(system->coords, system->rem, system->psize, o);
    double sum = 0.0;
    for (int k = 0; k < ps[2]; k++)
        for (int j = 0; j < ps[1]; j++)
            for (int i = 0; i < ps[0]; i++) {
                const int g = (o[0] + i) + (o[1] + j) * ss[0] + (o[2] + k) * ss[0] * ss[1];
                srand(g + 1);
                const double v = 2 * (double)(rand()) / (double)(RAND_MAX);
                system->rhs[i + j * ps[0] + (long)k * ps[0] * ps[1]] = v;
                sum += v;
            }
    sum /= (double)((long)ss[0] * ss[1] * ss[2]);
    for (long i = 0; i < nn; i++) {
        system->rhs[i] -= sum;
        system->phi[i] = 1.0;
    }
}

// MPI_REAL/system.c:522-543
void Deallocate_memory(POISSON *system)
{
    free(system->send_neighs);
    free(system->rec_neighs);
    free(system->send_counts);
    free(system->rec_counts);
    free(system->scounts);
    free(system->sdispls);
    free(system->rcounts);
    free(system->rdispls);
    free(system->coeff_lap);
    free(system->phi);
    free(system->rhs);
    free(system->Lap_phi);
    free(system);
    release_all();
    stage_release();
}

// MPI_REAL/system.c:80-213: Lap_phi = a * Lap * phi, on the GPU
void Lap_Vec_mult(POISSON *system, double a, double *phi, double *Lap_phi, aar_mpi_comm)
{
    Op op;
    const int np = system->psize[0] * system->psize[1] * system->psize[2];
    if (int rc = op_from_system(system, a, np, op)) { report_failure("Lap_Vec_mult", rc); return; }
    if (int rc = stage(op.n)) { report_failure("Lap_Vec_mult", rc); return; }
    Context &C = ctx();
    const size_t bytes = sizeof(double) * (size_t)op.n;
    const char *nm = "Lap_Vec_mult";
    if (!hip_ok(hipMemcpyAsync(g_stage.x, phi, bytes, hipMemcpyHostToDevice, C.stream), nm, "H2D copy")) return;
    if (int rc = matvec<double>(op, g_stage.x, g_stage.b)) { report_failure(nm, rc); return; }
    if (!hip_ok(hipMemcpyAsync(Lap_phi, g_stage.b, bytes, hipMemcpyDeviceToHost, C.stream), nm, "D2H copy")) return;
    (void)hip_ok(hipStreamSynchronize(C.stream), nm, "synchronize");
}

// res /= d on the GPU, in place on host memory (Np doubles); c = nullptr:
// d = diag, else d = diag + c(x) pointwise (device field)
static void host_divide(const char *nm, double diag, const double *c, double *res, int Np)
{
    if (Np <= 0) return;
    if (int rc = stage(Np)) { report_failure(nm, rc); return; }
    Context &C = ctx();
    const size_t bytes = sizeof(double) * (size_t)Np;
    if (!hip_ok(hipMemcpyAsync(g_stage.x, res, bytes, hipMemcpyHostToDevice, C.stream), nm, "H2D copy")) return;
    const hipError_t e = c ? launch_div_vc(g_stage.x, diag, c, Np, C.stream) : launch_div_f64(g_stage.x, diag, Np, C.stream);
    if (!hip_ok(e, nm, "launch")) return;
    if (!hip_ok(hipMemcpyAsync(res, g_stage.x, bytes, hipMemcpyDeviceToHost, C.stream), nm, "D2H copy")) return;
    (void)hip_ok(hipStreamSynchronize(C.stream), nm, "synchronize");
}

// MPI_REAL/system.c:222-227: res /= diag, on the GPU
void Precondition(double diag, double *res, int Np)
{
    host_divide("Precondition", diag, nullptr, res, Np);
}

// Variable-coefficient operator slots (no MPI_REAL counterpart: the
// device operator behind the reference's callback slot, README:25):
// Lap_phi = a * Lap * phi + c(x) phi with the registered c(x), on the GPU
void Lap_Vec_mult_vc(POISSON *system, double a, double *phi, double *Lap_phi, aar_mpi_comm)
{
    Op op;
    const int np = system->psize[0] * system->psize[1] * system->psize[2];
    if (int rc = op_from_system(system, a, np, op)) { report_failure("Lap_Vec_mult_vc", rc); return; }
    if (int rc = registered_coefficient(op.n, &op.c)) { report_failure("Lap_Vec_mult_vc", rc); return; }
    if (int rc = stage(op.n)) { report_failure("Lap_Vec_mult_vc", rc); return; }
    Context &C = ctx();
    const size_t bytes = sizeof(double) * (size_t)op.n;
    const char *nm = "Lap_Vec_mult_vc";
    if (!hip_ok(hipMemcpyAsync(g_stage.x, phi, bytes, hipMemcpyHostToDevice, C.stream), nm, "H2D copy")) return;
    if (int rc = matvec<double>(op, g_stage.x, g_stage.b)) { report_failure(nm, rc); return; }
    if (!hip_ok(hipMemcpyAsync(Lap_phi, g_stage.b, bytes, hipMemcpyDeviceToHost, C.stream), nm, "D2H copy")) return;
    (void)hip_ok(hipStreamSynchronize(C.stream), nm, "synchronize");
}

// res /= (diag + c(x)) pointwise with the registered c(x), on the GPU
// (diag = a*(3*coeff[0]), as the solvers pass it, AAR.c:94)
void Precondition_vc(double diag, double *res, int Np)
{
    const double *c = nullptr;
    if (int rc = registered_coefficient(Np, &c)) { report_failure("Precondition_vc", rc); return; }
    host_divide("Precondition_vc", diag, c, res, Np);
}

// MPI_REAL/tools.c:345-352
double pythag(double a, double b)
{
    const double absa = fabs(a), absb = fabs(b);
    if (absa > absb) return absa * sqrt(1.0 + (double)(absb * absb / (absa * absa)));
    return absb == 0.0 ? 0.0 : absb * sqrt(1.0 + (double)(absa * absa / (absb * absb)));
}

// MPI_REAL/tools.c:113-337 for square matrices (m == n <= 16), on the GPU:
// a <- V (the reference overwrites U with V), w <- truncated singular values.
void SingularValueDecomp(double **a, int m, int n, double *w, double **v)
{
    if (m != n || n < 1 || n > MAXM_LS) {
        std::printf("SingularValueDecomp: libaar_hip supports square matrices up to 16 x 16\n");
        return;
    }
    std::vector<double> A((size_t)n * n), b(n, 0.0), x(n), wv(n), V((size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) A[(size_t)i * n + j] = a[i][j];
    if (int rc = pinv_host(A.data(), b.data(), x.data(), n, wv.data(), V.data())) {
        if (rc == AAR_HIP_ERR_SVD) {
            std::printf("no convergence in %d svd iterations \n", 250);
            std::exit(1);   // tools.c:270
        }
        report_failure("SingularValueDecomp", rc);
        return;
    }
    for (int i = 0; i < n; i++) {
        w[i] = wv[i];
        for (int j = 0; j < n; j++) {
            v[i][j] = V[(size_t)i * n + j];
            a[i][j] = V[(size_t)i * n + j];
        }
    }
}

// MPI_REAL/tools.c:52-105, on the GPU
void PseudoInverseTimesVec(double **Am, double *b, double *x, int m)
{
    if (m < 1 || m > MAXM_LS) {
        std::printf("PseudoInverseTimesVec: libaar_hip supports m <= 32\n");
        return;
    }
    std::vector<double> A((size_t)m * m);
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) A[(size_t)i * m + j] = Am[i][j];
    if (int rc = pinv_host(A.data(), b, x, m, nullptr, nullptr)) {
        if (rc == AAR_HIP_ERR_SVD) {
            std::printf("no convergence in %d svd iterations \n", 250);
            std::exit(1);
        }
        report_failure("PseudoInverseTimesVec", rc);
    }
}

// sum of u*v (v = nullptr: u*u) over host arrays, on the GPU with the
// library's deterministic two-pass reduction; 0 or an AAR_HIP_ERR_* code
static int host_dot(const char *nm, const double *u, const double *v, int len, double *out)
{
    *out = 0.0;
    if (len <= 0) return 0;
    if (int rc = stage(len)) { report_failure(nm, rc); return rc; }
    Context &C = ctx();
    const size_t bytes = sizeof(double) * (size_t)len;
    if (!hip_ok(hipMemcpyAsync(g_stage.x, u, bytes, hipMemcpyHostToDevice, C.stream), nm, "H2D copy"))
        return AAR_HIP_ERR_HIP;
    if (v && !hip_ok(hipMemcpyAsync(g_stage.b, v, bytes, hipMemcpyHostToDevice, C.stream), nm, "H2D copy"))
        return AAR_HIP_ERR_HIP;
    static DevBuf part;
    const int nb = (int)std::max(1L, std::min(2048L, ((long)len + 4095) / 4096));
    if (!hip_ok(ensure(part, sizeof(double) * nb), nm, "allocation")) return AAR_HIP_ERR_HIP;
    Ctrl *dc = C.ctrl + 1;
    if (!hip_ok(launch_zero_ctrl(dc, C.stream), nm, "launch")) return AAR_HIP_ERR_HIP;
    DotArgs<double> d;
    d.n = len;
    d.v = g_stage.x;
    d.w = v ? g_stage.b : nullptr;
    d.nblocks = nb;
    d.partials = (double *)part.p;
    d.counter = C.counters + 40;
    d.ctrl = dc;
    d.fin.fin = FIN_SUMS;
    if (!hip_ok(launch_dot<double>(d, C.stream), nm, "launch")) return AAR_HIP_ERR_HIP;
    Ctrl &hc = C.ctrl_h[1];
    if (!hip_ok(hipMemcpyAsync(&hc, dc, sizeof(Ctrl), hipMemcpyDeviceToHost, C.stream), nm, "D2H copy"))
        return AAR_HIP_ERR_HIP;
    if (!hip_ok(hipStreamSynchronize(C.stream), nm, "synchronize")) return AAR_HIP_ERR_HIP;
    *out = hc.red[0];
    return 0;
}

// MPI_REAL/tools.c:18-28 (the sum in a fixed order on the GPU, then sqrt);
// on a failure the line is printed and *ResVal is NaN
void Vector2Norm(double *Vec, int len, double *ResVal, aar_mpi_comm)
{
    double s;
    *ResVal = host_dot("Vector2Norm", Vec, nullptr, len, &s) ? NAN : sqrt(s);
}

// MPI_REAL/tools.c:35-43
void VectorDotProduct(double *Vec1, double *Vec2, int len, double *ResVal, aar_mpi_comm)
{
    double s;
    *ResVal = host_dot("VectorDotProduct", Vec1, Vec2, len, &s) ? NAN : s;
}

// MPI_REAL/AAR.c:37-177
void AAR(POISSON *system, aar_op_cb Lap, aar_pc_cb Pre, double a, double *x, double *rhs, double omega,
         double beta, int m, int p, int max_iter, double tol, int Np, aar_mpi_comm comm)
{
    (void)comm;
    host_solve("AAR", system, Lap, Pre, a, x, rhs, Np,
               [&](const Op &op, double *dx, double *db, int precond, aar_hip_stats *st) {
                   return aar_solve<double>(op, dx, db, omega, beta, m, p, max_iter, tol, precond, 0, st, 1);
               });
}

// MPI_REAL/PGR.c:35-173
void PGR(POISSON *system, aar_op_cb Lap, aar_pc_cb Pre, double a, double *x, double *rhs, double omega, int m,
         int p, int max_iter, double tol, int Np, aar_mpi_comm comm)
{
    (void)comm;
    host_solve("PGR", system, Lap, Pre, a, x, rhs, Np,
               [&](const Op &op, double *dx, double *db, int precond, aar_hip_stats *st) {
                   return pgr_solve(op, dx, db, omega, m, p, max_iter, tol, precond, 0, st);
               });
}

// MPI_REAL/PL2R.c:35-172
void PL2R(POISSON *system, aar_op_cb Lap, aar_pc_cb Pre, double a, double *x, double *rhs, double omega, int m,
          int p, int max_iter, double tol, int Np, aar_mpi_comm comm)
{
    (void)comm;
    host_solve("PL2R", system, Lap, Pre, a, x, rhs, Np,
               [&](const Op &op, double *dx, double *db, int precond, aar_hip_stats *st) {
                   return pgr_solve(op, dx, db, omega, m, p, max_iter, tol, precond, 1, st);
               });
}

// MPI_REAL/CG.c:33-175
void CG(POISSON *system, aar_op_cb Lap, aar_pc_cb Pre, double a, double *x, double *b, int max_iter, double tol,
        int Np, aar_mpi_comm comm)
{
    if (comm == AAR_MPI_COMM_NULL) return;   // CG.c:38
    host_solve("CG", system, Lap, Pre, a, x, b, Np,
               [&](const Op &op, double *dx, double *db, int precond, aar_hip_stats *st) {
                   return cg_solve(op, dx, db, max_iter, tol, precond, st);
               });
}

// MPI_REAL/AAR.c:185-232 on host arrays; the Gram, the solve and am_vec on the GPU
void AndersonExtrapolation(double **DX, double **DF, double *f, double beta, int m, int N, double *am_vec,
                           double **FtF, double *allredvec, double *Ftf, double *svec, aar_mpi_comm)
{
    if (m < 1 || m > MAXM) { std::printf("AndersonExtrapolation: libaar_hip supports m <= 32\n"); return; }
    StepBufs sb;
    if (int rc = step_upload(DX, DF, f, m, N, sb)) { report_failure("AndersonExtrapolation", rc); return; }
    if (int rc = step_gram_ls(0, sb, m, N)) { report_failure("AndersonExtrapolation", rc); return; }
    Context &C = ctx();
    // x_new = (0 + beta*0) - am  ->  am = -x_new (exact)
    UpdArgs<double> u;
    u.n = N;
    u.m = m;
    u.beta = beta;
    u.X = sb.X;
    u.F = sb.F;
    u.ld = N;
    u.direct = 1;
    u.nvalid = m;
    for (int j = 0; j < m; j++) u.slot[j] = j + 1;
    u.xslot = 0;
    u.fslot = 0;
    u.oslot = 0;   // written after all reads of slot 0 in the same pass (pointwise)
    u.ctrl = C.ctrl + 1;
    const char *nm = "AndersonExtrapolation";
    if (!hip_ok(launch_aar_update<double>(u, C.stream), nm, "launch")) return;
    std::vector<double> tmp(N);
    if (!hip_ok(hipMemcpyAsync(tmp.data(), sb.X, sizeof(double) * (size_t)N, hipMemcpyDeviceToHost, C.stream), nm,
                "D2H copy"))
        return;
    if (!hip_ok(step_outputs(m, FtF, allredvec, Ftf, svec), nm, "D2H copy")) return;
    for (int k = 0; k < N; k++) am_vec[k] = -tmp[k];
}

// MPI_REAL/PGR.c:181-218
void Galerkin_Richardson(double **DX, double **DF, double *f, int m, int Np, double **XtF, double *allredvec,
                         double *Xtf, double *svec, aar_mpi_comm)
{
    if (m < 1 || m > MAXM) { std::printf("Galerkin_Richardson: libaar_hip supports m <= 32\n"); return; }
    StepBufs sb;
    if (int rc = step_upload(DX, DF, f, m, Np, sb)) { report_failure("Galerkin_Richardson", rc); return; }
    if (int rc = step_gram_ls(1, sb, m, Np)) { report_failure("Galerkin_Richardson", rc); return; }
    (void)hip_ok(step_outputs(m, XtF, allredvec, Xtf, svec), "Galerkin_Richardson", "D2H copy");
}

// MPI_REAL/PL2R.c:180-218
void L2_Richardson(double **DF, double *f, int m, int Np, double **FtF, double *allredvec, double *Ftf,
                   double *svec, aar_mpi_comm)
{
    if (m < 1 || m > MAXM) { std::printf("L2_Richardson: libaar_hip supports m <= 32\n"); return; }
    StepBufs sb;
    if (int rc = step_upload(nullptr, DF, f, m, Np, sb)) { report_failure("L2_Richardson", rc); return; }
    if (int rc = step_gram_ls(2, sb, m, Np)) { report_failure("L2_Richardson", rc); return; }
    (void)hip_ok(step_outputs(m, FtF, allredvec, Ftf, svec), "L2_Richardson", "D2H copy");
}

// ===========================================================================
// Part 2: device-resident extension API (include/aar_hip.h)
// ===========================================================================

int aar_hip_matvec(const aar_hip_grid *g, const double *x, double *y)
{
    Op op;
    if (int rc = make_op(g, nullptr, op)) return rc;
    return matvec<double>(op, x, y);
}

int aar_hip_set_coefficient(const double *c, long n)
{
    if (int rc = context_init()) return rc;
    if (!c || n <= 0) {
        g_coef_n = 0;
        return 0;
    }
    Context &C = ctx();
    HCR(ensure(g_coef, sizeof(double) * (size_t)n), AAR_HIP_ERR_HIP);
    HCR(hipMemcpyAsync(g_coef.p, c, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, C.stream), AAR_HIP_ERR_HIP);
    HCR(hipStreamSynchronize(C.stream), AAR_HIP_ERR_HIP);
    g_coef_n = n;
    return 0;
}

int aar_hip_matvec_vc(const aar_hip_grid *g, const double *c, const double *x, double *y)
{
    Op op;
    if (int rc = make_op_vc(g, c, op)) return rc;
    return matvec<double>(op, x, y);
}

int aar_hip_aar_vc(const aar_hip_grid *g, const double *c, double *x, const double *b, double omega, double beta,
                   int m, int p, int max_iter, double tol, int precond, int mixed, aar_hip_stats *st)
{
    Op op;
    if (int rc = make_op_vc(g, c, op)) return rc;
    if (mixed) return aar_mixed(op, x, b, omega, beta, m, p, max_iter, tol, precond, st);
    return aar_solve<double>(op, x, b, omega, beta, m, p, max_iter, tol, precond, 0, st, 1);
}

int aar_hip_pgr_vc(const aar_hip_grid *g, const double *c, double *x, const double *b, double omega, int m, int p,
                   int max_iter, double tol, int precond, int l2, aar_hip_stats *st)
{
    Op op;
    if (int rc = make_op_vc(g, c, op)) return rc;
    return pgr_solve(op, x, b, omega, m, p, max_iter, tol, precond, l2, st);
}

int aar_hip_cg_vc(const aar_hip_grid *g, const double *c, double *x, const double *b, int max_iter, double tol,
                  int precond, aar_hip_stats *st)
{
    Op op;
    if (int rc = make_op_vc(g, c, op)) return rc;
    return cg_solve(op, x, b, max_iter, tol, precond, st);
}

int aar_hip_gmres(const aar_hip_grid *g, const double *c, double *x, const double *b, int restart, int max_iter,
                  double tol, int precond, aar_hip_stats *st)
{
    Op op;
    if (int rc = c ? make_op_vc(g, c, op) : make_op(g, nullptr, op)) return rc;
    return gmres_solve(op, x, b, restart, max_iter, tol, precond, st);
}

int aar_hip_bicg(const aar_hip_grid *g, const double *c, double *x, const double *b, int max_iter, double tol,
                 int precond, aar_hip_stats *st)
{
    Op op;
    if (int rc = c ? make_op_vc(g, c, op) : make_op(g, nullptr, op)) return rc;
    return bicg_solve(op, x, b, max_iter, tol, precond, st);
}

int aar_hip_kspcg(const aar_hip_grid *g, const double *c, double *x, const double *b, int max_iter, double tol,
                  int precond, aar_hip_stats *st)
{
    Op op;
    if (int rc = c ? make_op_vc(g, c, op) : make_op(g, nullptr, op)) return rc;
    return kspcg_solve(op, x, b, max_iter, tol, precond, st);
}

int aar_hip_aar_petsc(const aar_hip_grid *g, const double *c, double *x, const double *b, double omega, double beta,
                      int m, int p, int max_iter, double tol, int precond, aar_hip_stats *st)
{
    Op op;
    if (int rc = c ? make_op_vc(g, c, op) : make_op(g, nullptr, op)) return rc;
    return aar_petsc_solve(op, x, b, omega, beta, m, p, max_iter, tol, precond, st);
}

int aar_hip_pgr_petsc(const aar_hip_grid *g, const double *c, double *x, const double *b, double omega, int m, int p,
                      int max_iter, double tol, int precond, int l2, aar_hip_stats *st)
{
    Op op;
    if (int rc = c ? make_op_vc(g, c, op) : make_op(g, nullptr, op)) return rc;
    return pgr_petsc_solve(op, x, b, omega, m, p, max_iter, tol, precond, l2, st);
}

int aar_hip_ilu0_apply(const aar_hip_grid *g, const double *c, const double *r, double *z, int trans)
{
    Op op;
    if (int rc = c ? make_op_vc(g, c, op) : make_op(g, nullptr, op)) return rc;
    return ilu0_apply(op, r, z, trans);
}

int aar_hip_matvec_f32(const aar_hip_grid *g, const float *x, float *y)
{
    Op op;
    if (int rc = make_op(g, nullptr, op)) return rc;
    return matvec<float>(op, x, y);
}

int aar_hip_aar(const aar_hip_grid *g, double *x, const double *b, double omega, double beta, int m, int p,
                int max_iter, double tol, int precond, aar_hip_stats *st)
{
    Op op;
    if (int rc = make_op(g, nullptr, op)) return rc;
    return aar_solve<double>(op, x, b, omega, beta, m, p, max_iter, tol, precond, 0, st, 1);
}

int aar_hip_aar_f32(const aar_hip_grid *g, float *x, const float *b, double omega, double beta, int m, int p,
                    int max_iter, double tol, int precond, aar_hip_stats *st)
{
    Op op;
    if (int rc = make_op(g, nullptr, op)) return rc;
    return aar_solve<float>(op, x, b, omega, beta, m, p, max_iter, tol, precond, 0, st, 1);
}

int aar_hip_aar_mixed(const aar_hip_grid *g, double *x, const double *b, double omega, double beta, int m,
                      int p, int max_iter, double tol, int precond, aar_hip_stats *st)
{
    Op op;
    if (int rc = make_op(g, nullptr, op)) return rc;
    return aar_mixed(op, x, b, omega, beta, m, p, max_iter, tol, precond, st);
}

int aar_hip_pgr(const aar_hip_grid *g, double *x, const double *b, double omega, int m, int p, int max_iter,
                double tol, int precond, int l2, aar_hip_stats *st)
{
    Op op;
    if (int rc = make_op(g, nullptr, op)) return rc;
    return pgr_solve(op, x, b, omega, m, p, max_iter, tol, precond, l2, st);
}

int aar_hip_cg(const aar_hip_grid *g, double *x, const double *b, int max_iter, double tol, int precond,
               aar_hip_stats *st)
{
    Op op;
    if (int rc = make_op(g, nullptr, op)) return rc;
    return cg_solve(op, x, b, max_iter, tol, precond, st);
}

int aar_hip_aar_refined(const aar_hip_grid *g, double *x, const double *b, double omega, double beta, int m,
                        int p, int max_outer, int max_inner, double tol, double inner_tol, int precond,
                        aar_hip_stats *st)
{
    Op op;
    if (int rc = make_op(g, nullptr, op)) return rc;
    if (m < 1 || m > MAXM) return fail(AAR_HIP_ERR_ARG, "history m must be in 1..32");
    return aar_refined(op, x, b, omega, beta, m, p, max_outer, max_inner, tol, inner_tol, precond, st);
}

int aar_hip_aar_batch(const aar_hip_instance *inst, int count, double omega, double beta, int m, int p,
                      int max_iter, double tol, int precond, int mixed, aar_hip_stats *stats)
{
    if (mixed) return batch_solve<double, float>(inst, count, omega, beta, m, p, max_iter, tol, precond, stats);
    return batch_solve<double, double>(inst, count, omega, beta, m, p, max_iter, tol, precond, stats);
}

int aar_hip_helm_matvec(const aar_hip_helm *h, const double *u, double *y)
{
    HOp op;
    if (int rc = make_hop(h, op)) return rc;
    return helm_matvec(op, u, y);
}

int aar_hip_helm_aar(const aar_hip_helm *h, double *u, const double *b, double omega, double beta, int m, int p,
                     int max_iter, double tol, int adjoint, aar_hip_stats *st)
{
    HOp op;
    if (int rc = make_hop(h, op)) return rc;
    if (adjoint) op.eta = -op.eta;        // A^H = a Lap + (c - i eta)
    return helm_solve(op, u, b, omega, beta, m, p, max_iter, tol, st);
}

int aar_hip_helm_gmres(const aar_hip_helm *h, double *u, const double *b, int restart, int max_iter, double tol,
                       int precond, aar_hip_stats *st)
{
    HOp op;
    if (int rc = make_hop(h, op)) return rc;
    return helm_gmres_solve(op, u, b, restart, max_iter, tol, precond, st);
}

int aar_hip_helm_lgmres(const aar_hip_helm *h, double *u, const double *b, int restart, int aug, int max_iter,
                        double tol, int precond, aar_hip_stats *st)
{
    HOp op;
    if (int rc = make_hop(h, op)) return rc;
    return helm_lgmres_solve(op, u, b, restart, aug, max_iter, tol, precond, st);
}

int aar_hip_helm_bicg(const aar_hip_helm *h, double *u, const double *b, int max_iter, double tol, int precond,
                      aar_hip_stats *st)
{
    HOp op;
    if (int rc = make_hop(h, op)) return rc;
    return helm_bicg_solve(op, u, b, max_iter, tol, precond, st);
}

int aar_hip_helm_ilu0_apply(const aar_hip_helm *h, const double *r, double *z, int trans)
{
    HOp op;
    if (int rc = make_hop(h, op)) return rc;
    return helm_ilu0_apply(op, r, z, trans);
}

int aar_hip_helm_grad_c(long n, const double *lambda, const double *u, double *dc)
{
    if (int rc = context_init()) return rc;
    hipError_t e = launch_helm_grad_c(n, (const double2 *)lambda, (const double2 *)u, dc, ctx().stream);
    if (e != hipSuccess) return fail(AAR_HIP_ERR_ARG, std::string("gradient: ") + hipGetErrorString(e));
    e = hipStreamSynchronize(ctx().stream);
    if (e != hipSuccess) return fail(AAR_HIP_ERR_HIP, std::string("gradient: ") + hipGetErrorString(e));
    return 0;
}

int aar_hip_pinv(const double *A, const double *b, double *x, int m, double *w_out, double *v_out)
{
    if (!A || !b || !x) return fail(AAR_HIP_ERR_ARG, "null pointer");
    return pinv_host(A, b, x, m, w_out, v_out);
}

int aar_hip_history(double *out, int cap)
{
    const auto &h = ctx().hist_host;
    const int n = (int)std::min<size_t>(h.size(), (size_t)std::max(cap, 0));
    if (out && n > 0) std::memcpy(out, h.data(), sizeof(double) * n);
    return n;
}

int aar_hip_last_stats(aar_hip_stats *st)
{
    if (!st) return fail(AAR_HIP_ERR_ARG, "null stats");
    *st = ctx().last;
    return 0;
}

int aar_hip_set_stencil(const double *coeff, int fdn)
{
    Context &C = ctx();
    if (!coeff) {
        C.user_fdn = 0;
        return 0;
    }
    if (fdn < 1 || fdn > R) return fail(AAR_HIP_ERR_ARG, "set_stencil: fdn must be 1..6");
    for (int o = 0; o <= R; o++) C.user_w[o] = o <= fdn ? coeff[o] : 0.0;
    C.user_fdn = fdn;
    return 0;
}

int aar_hip_set_stream(void *stream)
{
    Context &C = ctx();
    if (int rc = context_init()) return rc;
    C.stream = (hipStream_t)stream;
    C.stream_set = true;
    return 0;
}

int aar_hip_set_tuning(int lz, int wy, int py)
{
    if (py != 2 && py != 4 && py != 8) return fail(AAR_HIP_ERR_ARG, "py must be 2, 4 or 8");
    if (wy != 4 && wy != 8) return fail(AAR_HIP_ERR_ARG, "wy must be 4 or 8");
    if (py == 8 && wy != 4) return fail(AAR_HIP_ERR_ARG, "py = 8 needs wy = 4");
    if (lz < 0) return fail(AAR_HIP_ERR_ARG, "lz must be >= 0");
    ctx().lz = lz;
    ctx().wy = wy;
    ctx().py = py;
    // the defaults (0, 8, 4) hand the choice back to the automatic/tuned shapes
    ctx().tuning_set = (lz != 0 || wy != 8 || py != 4);
    return 0;
}

int aar_hip_last_system(int n, double *G, double *w)
{
    if (n < 1 || n > MAXM_LS) return fail(AAR_HIP_ERR_ARG, "system size must be in 1..32");
    const Ctrl &c = ctx().ctrl_h[0];
    if (G)
        for (int k = 0; k < n * n + n; k++) G[k] = c.G[k];
    if (w)
        for (int k = 0; k < n; k++) w[k] = c.w[k];
    return 0;
}

int aar_hip_autotune(int nx, int ny, int nz, int fp32, int batch_shape, int *lz, int *wy, int *py, double *us)
{
    if (nx < 1 || ny < 1 || nz < 1) return fail(AAR_HIP_ERR_ARG, "grid sizes must be positive");
    Shape sh;
    if (int rc = autotune_stencil(nx, ny, nz, fp32 ? 4 : 8, batch_shape != 0, &sh)) return rc;
    if (lz) *lz = sh.lz;
    if (wy) *wy = sh.wy;
    if (py) *py = sh.py;
    if (us) *us = sh.us;
    return 0;
}

int aar_hip_set_option(const char *name, long value)
{
    Context &C = ctx();
    const std::string k = name ? name : "";
    if (k == "ring_pad") {
        if (value < 0 || value % 32) return fail(AAR_HIP_ERR_ARG, "ring_pad must be a non-negative multiple of 32");
        C.ring_pad = value;
    } else if (k == "hist_blocks") {
        if (value < 0 || value > 16384) return fail(AAR_HIP_ERR_ARG, "hist_blocks must be in 0..16384");
        C.hist_blocks = (int)value;
    } else if (k == "two_pass") {
        C.two_pass = value != 0;
    } else if (k == "autotune") {
        C.autotune = value != 0;
    } else if (k == "complex_pinv_reference_order") {
        C.complex_ref_order = value != 0;
    } else {
        return fail(AAR_HIP_ERR_ARG, "unknown option " + k);
    }
    return 0;
}

int aar_hip_profile(int enable)
{
    ctx().prof = enable != 0;
    return 0;
}

int aar_hip_profile_reset(void)
{
    ctx().profile.clear();
    return 0;
}

int aar_hip_profile_query(const char *name, int *count, double *total_ms, double *bytes)
{
    auto it = ctx().profile.find(name ? name : "");
    if (it == ctx().profile.end()) {
        if (count) *count = 0;
        if (total_ms) *total_ms = 0.0;
        if (bytes) *bytes = 0.0;
        return 0;
    }
    if (count) *count = it->second.count;
    if (total_ms) *total_ms = it->second.ms;
    if (bytes) *bytes = it->second.bytes;
    return 0;
}

int aar_hip_release(void)
{
    release_all();
    stage_release();
    return 0;
}

const char *aar_hip_last_error(void) { return ctx().err.c_str(); }

}  // extern "C"
