// batch.hip -- many independent AAR solves (uneven grids, screened Poisson
// operators, several right-hand sides) advanced together (no MPI_REAL
// counterpart; each solve follows MPI_REAL/AAR.c:37-177 exactly as
// aar_rec in solver.hip does).
//
// Every instance shares (m, p, omega, beta, tol, max_iter), so all of them
// run the same iteration schedule.  One Anderson cycle (p iterations) is a
// fixed sequence of batched kernels: (p-1) recurrence sweeps, the Gram and
// its finisher, the m x m solves, the anchor update, the true residual and
// its finisher.  Each batched kernel reads a work table that maps its
// blocks to (instance, tile or chunk) and a per-instance descriptor; ring
// slots, history columns and coefficient tables are derived on the device
// from an iteration clock in device memory, so the cycle is captured once
// into a hipGraph and replayed until every instance has converged (or
// max_iter).  Converged instances' blocks return at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "aar_solver.h"
#include "stencil_body.h"

namespace aar {

namespace {

constexpr int BWY = 8, BPY = 4;        // launch shape of the batched stencils
constexpr int CHUNK = 4096;            // elements per block of the streaming kernels

template <typename Tx, typename Tg>
struct BInst {
    SArgs<Tx> tr;              // true-residual stencil: geometry, weights, x, b, partials, ctrl
    SArgs<Tg> rc;              // recurrence stencil: geometry, folded weights
    long n = 0, ld = 0;
    int S = 0, NC = 0;
    Tx *x = nullptr;
    Tg *G = nullptr;
    Tg *C = nullptr;
    Ctrl *ctrl = nullptr;
    double *hist = nullptr;
    int hist_cap = 0;
    double *spart = nullptr;   // stencil partials
    double *gpart = nullptr;   // Gram partials [nchunks][NA]
    int nchunks = 0;
    int tr_items = 0;
};

struct WorkItem {
    int inst;
    int idx;
};

struct BShared {               // batch-wide values
    int m, p, nvalid_cap;
    double omega, beta;
};

// -------------------------------------------------------------------------
// stencils: REC (g_{it+1} from g_it) and TRUE (g_{it+1} from x, after the
// Anderson step at it), it = *clock + jrel
// -------------------------------------------------------------------------
template <typename T, typename Tx, typename Tg, SMode MODE>
__global__ __launch_bounds__(64 * BWY) void bstencil_kernel(const BInst<Tx, Tg> *inst, const WorkItem *work,
                                                            int nwork, const int *clock, int jrel)
{
    // the instance's arguments, staged in LDS (raw storage: __shared__
    // objects cannot have initialisers)
    __shared__ __align__(16) unsigned char sa_raw[sizeof(SArgs<T>)];
    SArgs<T> &sa = *reinterpret_cast<SArgs<T> *>(sa_raw);
    const WorkItem wi = work[xcd_item(blockIdx.x, nwork)];
    const BInst<Tx, Tg> &I = inst[wi.inst];
    if (I.ctrl->done) return;
    if (threadIdx.x == 0) {
        const long it = (long)*clock + jrel;
        if constexpr (MODE == SMode::AAR_REC) {
            sa = I.rc;
            sa.in0 = (const T *)(I.G + (size_t)(it % I.S) * I.ld);
            sa.out0 = (T *)(I.G + (size_t)((it + 1) % I.S) * I.ld);
            sa.partials = nullptr;
        } else {
            sa = I.tr;
            Tg *o = I.G + (size_t)((it + 1) % I.S) * I.ld;
            if constexpr (sizeof(Tx) == sizeof(Tg)) sa.out0 = (T *)o;
            else sa.out32 = (float *)o;
        }
    }
    __syncthreads();
    // FP32 with the plain index keeps the sweep within 128 VGPRs (two
    // blocks per CU with the batch's per-block argument staging)
    stencil_body<T, BWY, BPY, MODE, sizeof(T) == 4 ? 0 : -1>(sa, wi.idx);
}

// per-instance finish of the true residual: relres check at the Anderson
// iteration it = *clock + jrel (FIN_AAR_RES of iterate it+1)
template <typename Tx, typename Tg>
__global__ __launch_bounds__(BLOCK) void bfinish_kernel(const BInst<Tx, Tg> *inst, const int *clock, int jrel,
                                                        int max_iter, int *ndone)
{
    __shared__ double scratch[BLOCK / 64];
    __shared__ double out[1];
    const BInst<Tx, Tg> &I = inst[blockIdx.x];
    if (I.ctrl->done) return;
    final_sum(I.spart, I.tr_items, 1, out, scratch);
    if (threadIdx.x == 0) {
        FinArgs f = {};
        f.fin = FIN_AAR_RES;
        f.it = *clock + jrel + 1;
        f.check = 1;
        f.max_iter = max_iter;
        f.hist = I.hist;
        f.hist_cap = I.hist_cap;
        finalize(f, I.ctrl, out, 1);
        if (I.ctrl->done) atomicAdd(ndone, 1);
    }
}

// -------------------------------------------------------------------------
// Gram of the g ring of each instance at the Anderson iteration k
// -------------------------------------------------------------------------
template <typename Tg, int MM>
__global__ __launch_bounds__(BLOCK) void bgram_kernel(const BInst<double, Tg> *instd, const BInst<float, Tg> *instf,
                                                      const WorkItem *work, const int *clock, int jrel,
                                                      BShared sh)
{
    constexpr int NG = MM * (MM + 1) / 2;
    constexpr int NA = NG + MM;
    constexpr int NS = MM + 1;
    __shared__ double scratch[NA * (BLOCK / 64)];
    const WorkItem wi = work[blockIdx.x];
    Ctrl *ctrl;
    const Tg *G;
    long n, ld;
    int S;
    double *gpart;
    if (instd) {
        const auto &I = instd[wi.inst];
        ctrl = I.ctrl; G = I.G; n = I.n; ld = I.ld; S = I.S; gpart = I.gpart;
    } else {
        const auto &I = instf[wi.inst];
        ctrl = I.ctrl; G = I.G; n = I.n; ld = I.ld; S = I.S; gpart = I.gpart;
    }
    if (ctrl->done) return;
    const long k = (long)*clock + jrel;
    const int nv = (int)min((long)sh.m, k - 1);
    const Tg *pd[NS];
    bool use[NS];
#pragma unroll
    for (int q = 0; q < NS; q++) {
        use[q] = q <= nv;
        pd[q] = G + (size_t)(use[q] ? ((k - q) % S) : 0) * ld;
    }
    double acc[NA];
#pragma unroll
    for (int q = 0; q < NA; q++) acc[q] = 0.0;
    const long e0 = (long)wi.idx * CHUNK, e1 = min(n, e0 + CHUNK);
    for (long e = e0 + threadIdx.x; e < e1; e += BLOCK) {
        double v[NS];
#pragma unroll
        for (int q = 0; q < NS; q++) v[q] = use[q] ? (double)pd[q][e] : 0.0;
        double d[MM];
#pragma unroll
        for (int q = 0; q < MM; q++) d[q] = q < nv ? v[q] - v[q + 1] : 0.0;
        int t = 0;
#pragma unroll
        for (int j = 0; j < MM; j++) {
#pragma unroll
            for (int q = 0; q <= j; q++) acc[t++] += d[q] * d[j];
            acc[NG + j] += d[j] * v[0];
        }
    }
    block_reduce_to<NA>(acc, gpart + (size_t)wi.idx * NA, scratch);
}

// per instance: fixed-order sum of the chunk partials, laid out as the
// reference's allredvec (AAR.c:192-212) for the m x m solve
template <typename Tg, int MM>
__global__ __launch_bounds__(BLOCK) void bgram_finish_kernel(const BInst<double, Tg> *instd,
                                                             const BInst<float, Tg> *instf, const int *clock,
                                                             int jrel, BShared sh)
{
    constexpr int NG = MM * (MM + 1) / 2;
    constexpr int NA = NG + MM;
    __shared__ double scratch[BLOCK / 64];
    __shared__ double sums[NA];
    Ctrl *ctrl;
    const double *gpart;
    int nchunks;
    if (instd) {
        const auto &I = instd[blockIdx.x];
        ctrl = I.ctrl; gpart = I.gpart; nchunks = I.nchunks;
    } else {
        const auto &I = instf[blockIdx.x];
        ctrl = I.ctrl; gpart = I.gpart; nchunks = I.nchunks;
    }
    if (ctrl->done) return;
    final_sum(gpart, nchunks, NA, sums, scratch);
    if (threadIdx.x != 0) return;
    const int k = *clock + jrel, m = sh.m;
    const int nv = min(m, k - 1);
    double *Gm = ctrl->G;
    for (int q = 0; q < m * m + m; q++) Gm[q] = 0.0;
    auto col = [&](int age) { return ((k - 2 - age) % m + m) % m; };     // AAR.c:97
    for (int j = 0; j < nv; j++)
        for (int q = 0; q <= j; q++) {
            const double v = sums[j * (j + 1) / 2 + q];
            Gm[col(j) * m + col(q)] = v;
            Gm[col(q) * m + col(j)] = v;
        }
    for (int j = 0; j < nv; j++) Gm[m * m + col(j)] = sums[NG + j];
}

// -------------------------------------------------------------------------
// anchor update at the Anderson iteration k = *clock + jrel (mode 0), the
// closing x += omega sum_{first <= j <= k} g_j of a solve that ends on
// Richardson steps (mode 1), or (mode 2, p > m+1 only) x += omega g_j of a
// Richardson residual j = *clock + jrel that no Anderson window reads: it
// is folded into the anchor before its ring slot is reused, so the
// Richardson terms of the next update start after the folded ones.
// -------------------------------------------------------------------------
template <typename Tx, typename Tg>
__global__ __launch_bounds__(BLOCK) void banchor_kernel(const BInst<Tx, Tg> *inst, const WorkItem *work,
                                                        const int *clock, int jrel, int mode, BShared sh)
{
    __shared__ double s_cg[NGMAX], s_cc[NCMAX];
    __shared__ int s_gs[NGMAX], s_cs[NCMAX], s_ng, s_nc, s_xlo, s_xhi, s_cout;
    const WorkItem wi = work[blockIdx.x];
    const BInst<Tx, Tg> &I = inst[wi.inst];
    if (I.ctrl->done) return;
    if (threadIdx.x == 0) {
        const int m = sh.m, p = sh.p, S = I.S, NC = I.NC;
        const long base = *clock;
        const long k = base + jrel;
        for (int i = 0; i < NGMAX; i++) s_cg[i] = 0.0;
        for (int c = 0; c < NCMAX; c++) s_cc[c] = 0.0;
        // first Richardson increment not yet in the anchor: the iterate after
        // the last Anderson step; iteration 1 is a Richardson step even for
        // p = 1 (AAR.c:109), whose cycles then start at base = 1.  With
        // p > m+1 the cycle's residuals base+1 .. base+p-m-1 are folded in
        // (mode 2) right after the sweep that follows them.
        const int nfold_cycle = p > m + 1 ? p - m - 1 : 0;
        const int nfold = mode == 0 ? nfold_cycle : (mode == 1 ? min(jrel, nfold_cycle) : 0);
        const long first = ((p == 1 && base == 1) ? 1 : base + 1) + nfold;
        if (mode == 2) {
            s_gs[0] = (int)(k % S);
            s_ng = 1;
            s_xlo = 0;
            s_xhi = 1;
            s_nc = 0;
            s_cout = -1;
        } else if (mode == 0) {
            const int nv = (int)min((long)m, k - 1);
            const long jlo = min(first, k - nv);
            const int ng = (int)(k - jlo + 1);
            for (int i = 0; i < ng; i++) s_gs[i] = (int)((jlo + i) % S);
            s_ng = ng;
            s_xlo = (int)(first - jlo);
            s_xhi = (int)(k - jlo);
            s_cg[ng - 1] = sh.beta;
            int nc = 0;
            for (int t = 0; t < nv; t++) {
                const long j = k - t;
                const double gam = I.ctrl->s[(int)((j - 2) % m)];
                const int ij = (int)(j - jlo);
                s_cg[ij] -= sh.beta * gam;
                const bool corr = (j - 1) > 1 && (j - 1) % p == 0;
                if (corr) {
                    s_cs[nc] = (int)(((j - 1) / p) % NC);
                    s_cc[nc] -= gam;
                    nc++;
                    s_cg[ij - 1] += sh.beta * gam;
                } else {
                    s_cg[ij - 1] += (sh.beta - sh.omega) * gam;
                }
            }
            s_nc = nc;
            s_cout = NC > 0 ? (int)((k / p) % NC) : -1;
        } else {
            const long k0 = first;
            const int ng = (int)(k - k0 + 1);
            for (int i = 0; i < ng; i++) s_gs[i] = (int)((k0 + i) % S);
            s_ng = ng;
            s_xlo = 0;
            s_xhi = ng;
            s_nc = 0;
            s_cout = -1;
        }
    }
    __syncthreads();
    const int ng = s_ng, nc = s_nc, xlo = s_xlo, xhi = s_xhi;
    Tg *co = s_cout >= 0 ? I.C + (size_t)s_cout * I.ld : nullptr;
    const double om = sh.omega;
    const long e0 = (long)wi.idx * CHUNK, e1 = min(I.n, e0 + CHUNK);
    for (long e = e0 + threadIdx.x; e < e1; e += BLOCK) {
        double sx = 0.0, sc = 0.0;
        for (int i = 0; i < ng; i++) {
            const double g = (double)I.G[(size_t)s_gs[i] * I.ld + e];
            if (i >= xlo && i < xhi) sx += g;
            sc += s_cg[i] * g;
        }
        for (int c = 0; c < nc; c++) sc += s_cc[c] * (double)I.C[(size_t)s_cs[c] * I.ld + e];
        I.x[e] = (Tx)((double)I.x[e] + (om * sx + sc));
        if (co) co[e] = (Tg)sc;
    }
}

__global__ void bclock_kernel(int *clock, int p) { if (threadIdx.x == 0 && blockIdx.x == 0) *clock += p; }
__global__ void bset_kernel(int *v, int value) { if (threadIdx.x == 0 && blockIdx.x == 0) *v = value; }

template <typename T>
hipError_t bsp_alloc(T **p, size_t count, std::vector<void *> &owned)
{
    hipError_t e = hipMalloc((void **)p, sizeof(T) * std::max<size_t>(count, 1));
    if (e == hipSuccess) owned.push_back(*p);
    return e;
}

}  // namespace

#define BHC(x)                                                                                    \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            rc = fail(AAR_HIP_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));           \
            goto done;                                                                            \
        }                                                                                         \
    } while (0)

template <typename Tx, typename Tg>
int batch_solve(const aar_hip_instance *ins, int count, double omega, double beta, int m, int p, int max_iter,
                double tol, int precond, aar_hip_stats *stats)
{
    Context &C = ctx();
    if (int rc0 = context_init()) return rc0;
    if (!ins || count < 1) return fail(AAR_HIP_ERR_ARG, "empty batch");
    if (m < 1 || m > MAXM_BATCH) return fail(AAR_HIP_ERR_ARG, "batched solves: history m must be in 1..16");
    if (p < 1) return fail(AAR_HIP_ERR_ARG, "p must be >= 1");
    if (max_iter < 0) return fail(AAR_HIP_ERR_ARG, "max_iter must be >= 0");
    int rc = 0;
    std::vector<void *> owned;
    std::vector<BInst<Tx, Tg>> hI(count);
    std::vector<Op> ops(count);
    std::vector<WorkItem> wtr, wrc, wch;
    std::vector<Ctrl *> hctrl(count);
    std::vector<Ctrl> hc(count);
    BInst<Tx, Tg> *dI = nullptr;
    WorkItem *dtr = nullptr, *drc = nullptr, *dch = nullptr;
    Ctrl *ctrls = nullptr;
    Ctrl **dctrl = nullptr;
    int *clock = nullptr, *ndone = nullptr, *ndone_h = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const int S = m + 2;
    const int NC = (p <= m) ? m / p + 1 : 0;
    // Gram partials per chunk: the accumulators of the instantiation run
    const int MMI = m == 7 ? 7 : (m <= MAXM_FAST ? MAXM_FAST : MAXM_BATCH);
    const int NA = MMI * (MMI + 1) / 2 + MMI;
    BShared sh = {m, p, m, omega, beta};
    // graphs cannot be captured on the legacy default stream that framework
    // callers bind: the batch runs on the library's own stream, ordered
    // with the caller's stream by events on both sides
    hipStream_t st = C.own;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    float ms = 0.f;
    int cyc = 0;
    long base = (p == 1) ? 1 : 0;
    bool stopped = false;

    BHC(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
    BHC(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
    BHC(hipEventRecord(ev_in, C.stream));
    BHC(hipStreamWaitEvent(st, ev_in, 0));
    BHC(bsp_alloc(&ctrls, count, owned));
    BHC(bsp_alloc(&dctrl, count, owned));
    BHC(bsp_alloc(&clock, 1, owned));
    BHC(hipHostMalloc((void **)&ndone_h, sizeof(int) * 64, hipHostMallocDefault));
    BHC(bsp_alloc(&ndone, 1, owned));
    for (int i = 0; i < count; i++) {
        aar_hip_grid g = {ins[i].nx, ins[i].ny, ins[i].nz, 6, ins[i].a, ins[i].sigma};
        if ((rc = make_op(&g, nullptr, ops[i]))) goto done;
        if (!ins[i].x || !ins[i].b) {
            rc = fail(AAR_HIP_ERR_ARG, "null vector in batch");
            goto done;
        }
        const Op &op = ops[i];
        BInst<Tx, Tg> &I = hI[i];
        I.n = op.n;
        I.ld = ring_ld(op.n, sizeof(Tg));
        I.S = S;
        I.NC = NC;
        I.x = (Tx *)ins[i].x;
        I.ctrl = ctrls + i;
        hctrl[i] = ctrls + i;
        I.hist_cap = max_iter + 4;
        BHC(bsp_alloc(&I.G, (size_t)I.ld * S, owned));
        if (NC > 0) BHC(bsp_alloc(&I.C, (size_t)I.ld * NC, owned));
        BHC(bsp_alloc(&I.hist, (size_t)I.hist_cap, owned));
        // stencil args (auto z-chunk for this grid and the batch launch shape)
        // slab depth: the one fixed by aar_hip_set_tuning (its wy, py do not
        // apply: the batch launch shape is 8 x 4), else tuned for the batched
        // launch where available, else automatic; stencil_geometry clamps
        // it to the instance's nz
        auto batch_lz = [&](size_t es) {
            if (C.tuning_set && C.lz > 0) return C.lz;
            auto t = C.tuned_batch.find({op.nx, op.ny, op.nz, (int)es});
            return t != C.tuned_batch.end() ? t->second.lz : auto_lz(op, BWY, BPY, es);
        };
        SArgs<Tx> tr = base_sargs<Tx>(op);
        tr.lz = batch_lz(sizeof(Tx));
        SArgs<Tg> rcg = base_sargs<Tg>(op);
        rcg.lz = batch_lz(sizeof(Tg));
        const int ntr = stencil_items<Tx>(tr, BWY, BPY), nrc = stencil_items<Tg>(rcg, BWY, BPY);
        stencil_geometry<Tx>(tr, BWY, BPY);
        stencil_geometry<Tg>(rcg, BWY, BPY);
        tr.omega = rcg.omega = omega;
        tr.precond = rcg.precond = precond;
        {
            const double c = precond ? omega / op.diag : omega;
            rcg.cf[0] = 1.0 - c * (3.0 * op.w[0] + op.sigma);
            for (int o = 1; o <= R; o++) rcg.cf[o] = -c * op.w[o];
        }
        tr.in0 = I.x;
        tr.bvec = (const Tx *)ins[i].b;
        tr.ctrl = I.ctrl;
        rcg.ctrl = I.ctrl;
        I.nchunks = (int)((op.n + CHUNK - 1) / CHUNK);
        const int nbd = nblocks_for(op.n);
        BHC(bsp_alloc(&I.spart, (size_t)std::max({ntr, nrc, nbd}) * 2, owned));
        BHC(bsp_alloc(&I.gpart, (size_t)I.nchunks * NA, owned));
        tr.partials = I.spart;
        I.tr = tr;
        I.rc = rcg;
        I.tr_items = ntr;
        for (int t = 0; t < ntr; t++) wtr.push_back({i, t});
        for (int t = 0; t < nrc; t++) wrc.push_back({i, t});
        for (int t = 0; t < I.nchunks; t++) wch.push_back({i, t});
    }
    BHC(bsp_alloc(&dI, count, owned));
    BHC(bsp_alloc(&dtr, wtr.size(), owned));
    BHC(bsp_alloc(&drc, wrc.size(), owned));
    BHC(bsp_alloc(&dch, wch.size(), owned));
    BHC(hipMemcpy(dI, hI.data(), sizeof(BInst<Tx, Tg>) * count, hipMemcpyHostToDevice));
    BHC(hipMemcpy(dtr, wtr.data(), sizeof(WorkItem) * wtr.size(), hipMemcpyHostToDevice));
    BHC(hipMemcpy(drc, wrc.data(), sizeof(WorkItem) * wrc.size(), hipMemcpyHostToDevice));
    BHC(hipMemcpy(dch, wch.data(), sizeof(WorkItem) * wch.size(), hipMemcpyHostToDevice));
    BHC(hipMemcpy(dctrl, hctrl.data(), sizeof(Ctrl *) * count, hipMemcpyHostToDevice));

    {
        // kernels of the batch, launched on `st` (eagerly or under capture)
        const BInst<double, Tg> *ID = nullptr;
        const BInst<float, Tg> *IF = nullptr;
        if constexpr (sizeof(Tx) == 8) ID = (const BInst<double, Tg> *)dI;
        else IF = (const BInst<float, Tg> *)dI;
        auto rec = [&](int jrel) {
            hipLaunchKernelGGL((bstencil_kernel<Tg, Tx, Tg, SMode::AAR_REC>), dim3((unsigned)wrc.size()),
                               dim3(64 * BWY), 0, st, dI, drc, (int)wrc.size(), clock, jrel);
            // p > m+1: the residual of iteration base + jrel is read by no
            // Anderson window (AAR.c:97-102 keeps the last m columns)
            if (p > m + 1 && jrel < p - m)
                hipLaunchKernelGGL((banchor_kernel<Tx, Tg>), dim3((unsigned)wch.size()), dim3(BLOCK), 0, st, dI, dch,
                                   clock, jrel, 2, sh);
        };
        auto anderson = [&](int jrel) {
            auto gram = [&](auto mm) {
                constexpr int M = decltype(mm)::value;
                hipLaunchKernelGGL((bgram_kernel<Tg, M>), dim3((unsigned)wch.size()), dim3(BLOCK), 0, st, ID, IF, dch,
                                   clock, jrel, sh);
                hipLaunchKernelGGL((bgram_finish_kernel<Tg, M>), dim3(count), dim3(BLOCK), 0, st, ID, IF, clock, jrel,
                                   sh);
            };
            if (m == 7) gram(std::integral_constant<int, 7>());
            else if (m <= MAXM_FAST) gram(std::integral_constant<int, MAXM_FAST>());
            else gram(std::integral_constant<int, MAXM_BATCH>());
            (void)launch_ls_batch(dctrl, count, m, st);
            hipLaunchKernelGGL((banchor_kernel<Tx, Tg>), dim3((unsigned)wch.size()), dim3(BLOCK), 0, st, dI, dch, clock,
                               jrel, 0, sh);
            hipLaunchKernelGGL((bstencil_kernel<Tx, Tx, Tg, SMode::AAR_TRUE>), dim3((unsigned)wtr.size()),
                               dim3(64 * BWY), 0, st, dI, dtr, (int)wtr.size(), clock, jrel);
            hipLaunchKernelGGL((bfinish_kernel<Tx, Tg>), dim3(count), dim3(BLOCK), 0, st, dI, clock, jrel, max_iter,
                               ndone);
        };

        // prologue, per instance: ||b||, the tolerance rule of AAR.c:76-79
        // and the true residual of x0 (single-solve kernels)
        for (int i = 0; i < count; i++) BHC(launch_zero_ctrl(ctrls + i, st));
        hipLaunchKernelGGL(bset_kernel, dim3(1), dim3(64), 0, st, ndone, 0);
        hipLaunchKernelGGL(bset_kernel, dim3(1), dim3(64), 0, st, clock, 0);
        BHC(hipEventRecord(C.t0, st));
        for (int i = 0; i < count; i++) {
            const BInst<Tx, Tg> &I = hI[i];
            DotArgs<Tx> d;
            d.n = I.n;
            d.v = (const Tx *)ins[i].b;
            d.nblocks = nblocks_for(I.n);
            d.partials = I.spart;
            d.counter = nullptr;
            d.ctrl = I.ctrl;
            d.fin.fin = FIN_NORM_B;
            d.fin.tol = tol;
            d.fin.check = 1;
            BHC(run_dot<Tx>(d));
            SArgs<Tx> a = I.tr;
            Tg *o = I.G + (size_t)(1 % S) * I.ld;
            if constexpr (sizeof(Tx) == sizeof(Tg)) a.out0 = (Tx *)o;
            else a.out32 = (float *)o;
            a.counter = nullptr;
            a.fin.fin = FIN_AAR_RES;
            a.fin.it = 1;
            a.fin.check = 0;
            a.fin.hist = I.hist;
            a.fin.hist_cap = I.hist_cap;
            BHC(launch_stencil<Tx>(SMode::AAR_TRUE, a, BWY, BPY, st));
            BHC(launch_finish(a.partials, I.tr_items, 1, a.fin, a.ctrl, st));
        }
        if (p == 1 && max_iter >= 1) {       // iteration 1 is never an Anderson step (AAR.c:109)
            rec(1);
            hipLaunchKernelGGL(bset_kernel, dim3(1), dim3(64), 0, st, clock, 1);
        }
        BHC(hipGetLastError());

        // one Anderson cycle: iterations base+1 .. base+p, captured once
        BHC(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        for (int j = 1; j < p; j++) rec(j);
        anderson(p);
        hipLaunchKernelGGL(bclock_kernel, dim3(1), dim3(64), 0, st, clock, p);
        BHC(hipStreamEndCapture(st, &graph));
        BHC(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));

        for (; base + p <= max_iter && !stopped; base += p) {
            BHC(hipGraphLaunch(exec, st));
            const int s = cyc % 64;
            BHC(hipMemcpyAsync(&ndone_h[s], ndone, sizeof(int), hipMemcpyDeviceToHost, st));
            BHC(hipEventRecord(C.poll_ev[s], st));
            if (cyc >= 2) {
                const int s2 = (cyc - 2) % 64;
                if (hipEventSynchronize(C.poll_ev[s2]) != hipSuccess || ndone_h[s2] >= count) stopped = true;
            }
            cyc++;
        }
        const long first_end = (p == 1 && base == 1) ? 1 : base + 1;
        if (!stopped && first_end <= max_iter) {
            // the last iterations are Richardson steps; then x = x_a + omega sum g
            const int rem = (int)(max_iter - base);
            for (int j = 1; j <= rem; j++) rec(j);
            hipLaunchKernelGGL((banchor_kernel<Tx, Tg>), dim3((unsigned)wch.size()), dim3(BLOCK), 0, st, dI, dch,
                               clock, rem, 1, sh);
            BHC(hipGetLastError());
        }
        BHC(hipEventRecord(C.t1, st));
        BHC(hipMemcpyAsync(hc.data(), ctrls, sizeof(Ctrl) * count, hipMemcpyDeviceToHost, st));
        BHC(hipStreamSynchronize(st));
        BHC(hipEventElapsedTime(&ms, C.t0, C.t1));
    }
    for (int i = 0; i < count; i++) {
        const Ctrl &c = hc[i];
        aar_hip_stats s;
        std::memset(&s, 0, sizeof(s));
        const int iterations = c.done ? c.final_iter : max_iter;
        const int loop_iter = c.done ? c.final_iter + 1 : max_iter + 1;
        s.iterations = iterations;
        s.rhs_norm = c.rhs_norm;
        s.relres = c.rhs_norm > 0 ? c.relres / c.rhs_norm : c.relres;
        s.converged = c.relres <= c.tol_abs;
        s.printed_converged = (loop_iter < max_iter && c.relres <= c.tol_abs);
        s.printed_warning = (loop_iter >= max_iter);
        s.status = c.status == 1 ? AAR_HIP_ERR_SVD : (c.status == 2 ? AAR_HIP_ERR_NONFINITE : 0);
        s.solve_ms = ms;
        if (stats) stats[i] = s;
        if (i == 0) C.last = s;
    }
done:
    if (ev_out) {
        (void)hipEventRecord(ev_out, st);
        (void)hipStreamWaitEvent(C.stream, ev_out, 0);
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipStreamSynchronize(st);
    for (void *q : owned) (void)hipFree(q);
    if (ndone_h) (void)hipHostFree(ndone_h);
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    return rc;
}

template int batch_solve<double, double>(const aar_hip_instance *, int, double, double, int, int, int, double, int,
                                         aar_hip_stats *);
template int batch_solve<double, float>(const aar_hip_instance *, int, double, double, int, int, int, double, int,
                                        aar_hip_stats *);

}  // namespace aar
