// MLMG solve orchestration (MLMGT::solve / oneIter / mgVcycle /
// actualBottomSolve restated for one GPU).  Differences from the reference's
// GPU design: no host synchronisation inside a V-cycle (the bottom BiCGStab
// iterates on the device), reductions land in device memory, and one outer
// iteration is captured into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

#include "amrex/AMReX_MLMG_Core.H"
#include "amrex/AMReX_MultiFabUtil.H"
#include "amrex/AMReX_PlotFileUtil.H"
#include "mgx_amr.H"
#include "mgx_impl.H"
#include "mgx_fused.H"
#include "mgx_kernels.H"
#include "mgx_smoother.H"

namespace mgx {

using amrex::DeviceBuffer;

template <class T>
static T* P(const std::unique_ptr<DeviceBuffer>& b) {
    return b ? static_cast<T*>(b->ptr) : nullptr;
}

SolverImpl::~SolverImpl() {
    release_graph();
    if (hpin) (void)hipHostFree(hpin);
    if (hpin_status) (void)hipHostFree(hpin_status);
}

void SolverImpl::release_graph() {
    if (amr_exec) (void)hipGraphExecDestroy(amr_exec);
    if (amr_graph) (void)hipGraphDestroy(amr_graph);
    amr_exec = nullptr;
    amr_graph = nullptr;
    for (int q = 0; q < kGraphs; ++q) {
        if (exec[q]) (void)hipGraphExecDestroy(exec[q]);
        if (graph[q]) (void)hipGraphDestroy(graph[q]);
        exec[q] = nullptr;
        graph[q] = nullptr;
        graph_ok[q] = false;
    }
}

void SolverImpl::sync_linop() {
    if (!lp.prepared) lp.prepare();
    if (graph_epoch != lp.epoch) {
        release_graph();
        graph_epoch = lp.epoch;
    }
}

double* SolverImpl::sol_buf(int parity) {
    return static_cast<double*>((parity == 0 || !lp.dense) ? sol->ptr : sol2->ptr);
}

// MLMGT::solve (AMReX_MLMG.H:367-376): Default becomes the operator's default;
// hypre and PETSc bottom solvers are not part of this library.
void SolverImpl::check_bottom() {
    using amrex::BottomSolver;
    if (mg.bottom_solver == BottomSolver::Default) mg.bottom_solver = mg.linop.getDefaultBottomSolver();
    if (mg.bottom_solver == BottomSolver::hypre || mg.bottom_solver == BottomSolver::petsc)
        throw amrex::MLMG::error("MLMG: the hypre and PETSc bottom solvers are not available in this library");
}

void SolverImpl::start_bottom(hipStream_t s) {
    check_bottom();
    if (!allocated) return;
    set_val<int>(P<int>(status) + 2, 1, static_cast<int>(mg.bottom_solver), s);
    hpin_status[2] = static_cast<int>(mg.bottom_solver);  // no copy into it is in flight between solves
}

void SolverImpl::end_bottom() {
    using amrex::BottomSolver;
    const BottomSolver now = static_cast<BottomSolver>(hpin_status[2]);
    if ((mg.bottom_solver == BottomSolver::bicgcg || mg.bottom_solver == BottomSolver::cgbicg) &&
        (now == BottomSolver::bicgstab || now == BottomSolver::cg))
        mg.bottom_solver = now;
}

void SolverImpl::prepare_cycles() {
    check_bottom();
    if (lp.kind == 0 && (std::isnan(lp.a_scalar) || std::isnan(lp.b_scalar)))
        throw amrex::MLMG::error("MLABecLaplacian: setScalars must be called before solve");
    sync_linop();
    if (!allocated || alloc_prec != mg.precision) allocate();
    refresh_ops();
    start_bottom(amrex::Gpu::gpuStream());
}

double* SolverImpl::spare_sol() { return static_cast<double*>(sol_at == 0 ? sol2->ptr : sol->ptr); }

template <class V>
static void alloc_cycle(Cycle<V>& C, const LinOpImpl& lp, V* res0) {
    const size_t nl = lp.lev.size();
    C.own.clear();
    C.cor.assign(nl, nullptr);
    C.res.assign(nl, nullptr);
    C.rescor.assign(nl, nullptr);
    C.tmp.assign(nl, nullptr);
    C.cur.assign(nl, nullptr);
    C.hold.clear();
    C.op.resize(nl);
    auto mk = [&](int64_t n) {
        C.own.emplace_back(new DeviceBuffer(sizeof(V) * (size_t)n));
        return static_cast<V*>(C.own.back()->ptr);
    };
    for (size_t l = 0; l < nl; ++l) {
        const int64_t n = lp.lev[l].plan.nalloc;
        C.cor[l] = mk(n);
        C.res[l] = (l == 0 && res0) ? res0 : mk(n);
        C.rescor[l] = mk(n);
        C.tmp[l] = mk(n);
        C.cur[l] = C.cor[l];
        hipStream_t s = amrex::Gpu::gpuStream();
        set_val<V>(C.tmp[l], n, V(0), s);
        set_val<V>(C.cor[l], n, V(0), s);
        set_val<V>(C.res[l], n, V(0), s);
        set_val<V>(C.rescor[l], n, V(0), s);
    }
}

void SolverImpl::allocate() {
    release_graph();
    const auto& P0 = lp.lev[0].plan;
    hipStream_t s = amrex::Gpu::gpuStream();
    sol_at = -1;
    sol.reset(new DeviceBuffer(sizeof(double) * (size_t)P0.nalloc));
    sol2.reset(new DeviceBuffer(sizeof(double) * (size_t)P0.nalloc));
    set_val<double>(P<double>(sol2), P0.nalloc, 0.0, s);
    rhs.reset(new DeviceBuffer(sizeof(double) * (size_t)P0.nalloc));
    res.reset(new DeviceBuffer(sizeof(double) * (size_t)P0.nalloc));
    set_val<double>(P<double>(sol), P0.nalloc, 0.0, s);
    set_val<double>(P<double>(rhs), P0.nalloc, 0.0, s);
    set_val<double>(P<double>(res), P0.nalloc, 0.0, s);
    alloc_prec = mg.precision;
    if (mg.precision == amrex::MLMG::Precision::Double) {
        alloc_cycle<double>(c64, lp, P<double>(res));
        c32 = Cycle<float>{};
    } else {
        alloc_cycle<float>(c32, lp, nullptr);
        c64 = Cycle<double>{};
    }
    const auto& PB = lp.lev.back().plan;
    bot_work.reset(new DeviceBuffer(sizeof(double) * 6 * (size_t)PB.nalloc));
    bot_x.reset(new DeviceBuffer(sizeof(double) * (size_t)PB.nalloc));
    bot_b.reset(new DeviceBuffer(sizeof(double) * (size_t)PB.nalloc));
    status.reset(new DeviceBuffer(sizeof(int) * 4));
    int64_t np = 0;
    for (const auto& L : lp.lev) np = std::max(np, reduce_partials(L.plan.dev, (int)L.plan.mbc));
    if (lp.dense) np = std::max(np, update_residual_partials(P0.dev));
    scr.ensure(np);
    if (!hpin) MGX_HIP_CHECK(hipHostMalloc((void**)&hpin, sizeof(double) * 8, hipHostMallocDefault));
    if (!hpin_status) MGX_HIP_CHECK(hipHostMalloc((void**)&hpin_status, sizeof(int) * 4, hipHostMallocDefault));
    allocated = true;
}

// Operator scalars and coefficient pointers of every level in the cycle precision.
void SolverImpl::refresh_ops() {
    if (alloc_prec == amrex::MLMG::Precision::Double) {
        for (size_t l = 0; l < lp.lev.size(); ++l) c64.op[l] = lp.coef<double>((int)l);
    } else {
        for (size_t l = 0; l < lp.lev.size(); ++l) c32.op[l] = lp.coef<float>((int)l);
    }
}

template <class V>
void SolverImpl::smooth(Cycle<V>& C, int l, bool skip_first_fill, hipStream_t s) {
    // linop.smooth (AMReX_MLCellLinOp.H:1205-1217): ghosts refreshed before each colour
    const auto& L = lp.lev[(size_t)l].plan;
    if (tcycle()) {
        // the adjoint cycle's transposed sweep reads no ghosts
        for (int rb = 0; rb < 2; ++rb) gsrb_T<V>(L.dev, C.op[(size_t)l], C.cor[l], C.res[l], rb, s);
        return;
    }
    for (int rb = 0; rb < 2; ++rb) {
        if (!(skip_first_fill && rb == 0)) fill_ghosts<V>(L.dev, C.cor[l], nullptr, 0, s);
        gsrb<V>(L.dev, (int)L.mbc, C.op[(size_t)l], C.cor[l], C.res[l], rb, s);
    }
}

template <class V>
void SolverImpl::bottom(Cycle<V>& C, hipStream_t s) {
    const int nb = (int)lp.lev.size() - 1;
    const auto& L = lp.lev[(size_t)nb].plan;
    const bool sing = lp.singular && lp.enforce_solvable;
    if (mg.bottom_solver == amrex::BottomSolver::smoother) {
        set_val<V>(C.cor[nb], L.nalloc, V(0), s);
        for (int i = 0; i < mg.nuf; ++i) smooth(C, nb, i == 0, s);
        return;
    }
    const OpCoef<double> op = lp.coef<double>(nb);
    double* x;
    const double* b;
    if constexpr (std::is_same<V, double>::value) {
        x = C.cor[nb];
        b = C.res[nb];
    } else {
        // the bottom solve of the mixed form runs in FP64
        convert<double, float>(P<double>(bot_b), C.res[nb], L.nalloc, s);
        x = P<double>(bot_x);
        b = P<double>(bot_b);
    }
    bottom_solve<double>(L.dev, op, x, b, P<double>(bot_work), mg.bottom_reltol, mg.bottom_abstol, mg.bottom_maxiter,
                            sing ? 1 : 0, lp.volinv_bot, mg.nub, mg.nuf, P<int>(status), tcycle() ? 1 : 0, s);
    if constexpr (!std::is_same<V, double>::value) convert<float, double>(C.cor[nb], x, L.nalloc, s);
}

// Two consecutive sweeps of level l run as one launch (gsrb_sweep2): FP32
// sweeps and FP64 Poisson sweeps on levels the 64 x 16 tiles cover (FP64
// ABecLaplacian needs more registers than the march keeps).
static bool g_pair_sweeps = true;
void set_pair_sweeps(bool on) { g_pair_sweeps = on; }
bool pair_sweeps_enabled() { return g_pair_sweeps; }
template <class V>
bool SolverImpl::pair_sweeps(int l) const {
    if (!g_pair_sweeps || !fused(l)) return false;
    if (sizeof(V) == 8 && lp.kind == 0) return false;
    return gsrb_sweep2_pays(lp.lev[(size_t)l].plan.dev);
}
template bool SolverImpl::pair_sweeps<double>(int) const;
template bool SolverImpl::pair_sweeps<float>(int) const;

// The adjoint's cycles work on L^T (sweeps, residuals of corrections, the
// bottom solve) where its red-black sweep is defined: dense levels and
// maxorder <= 3 (a row of L^T then reads the other colour only).  Otherwise
// they are the forward cycles and only the outer residual is b - L^T x.
bool SolverImpl::tcycle() const { return transpose && lp.dense && lp.maxorder <= 3; }

bool SolverImpl::fused(int l) const {
    const auto& L = lp.lev[(size_t)l].plan;
    if (tcycle()) return false;
    if (!lp.dense || L.dev.nbox != 1 || lp.maxorder > 3) return false;
    for (int d = 0; d < 3; ++d)
        if (L.dev.periodic[d] && (L.dev.n[d] % 2) != 0) return false;  // colours must match across the wrap
    return true;
}

// mgVcycle (AMReX_MLMG.H:1308-1415) for amrlev 0: input res[0], output the
// correction of level 0 in C.cur[0].  Dense levels are smoothed by the fused
// out-of-place sweep: the zero initial correction is folded into the first
// pre-smoothing sweep and the prolongation into the first post-smoothing sweep.
// warm: the top level starts from the correction in C.cur[top] instead of
// zero (the second V-cycle of a mixed-precision refinement step).
template <class V>
void SolverImpl::vcycle(Cycle<V>& C, hipStream_t s, int top, bool plain_top, bool warm) {
    const int nb = (int)lp.lev.size() - 1;
    // plain_top: the top level's correction must end with the ghosts the
    // reference's last boundary fill leaves there (read by the F-cycle's
    // interpolation to the next finer level): its last post-smoothing sweep
    // runs in place with a fill before each colour, every other sweep fused
    auto fast = [&](int l) { return fused(l); };
    for (int l = top; l < nb; ++l) {
        const auto& L = lp.lev[(size_t)l].plan;
        const bool from_cur = warm && l == top;
        V* x = C.cor[l];
        if (fast(l) && mg.nu1 > 0) {
            V* a = from_cur ? C.cur[l] : C.cor[l];
            V* b = (a == C.cor[l]) ? C.tmp[l] : C.cor[l];
            for (int i = 0; i < mg.nu1;) {
                const int zero = (i == 0 && !from_cur) ? 1 : 0;
                if (i + 1 < mg.nu1 && pair_sweeps<V>(l)) {
                    gsrb_sweep2<V>(L.dev, L.hbox[0], C.op[(size_t)l], a, b, C.res[l], zero, nullptr, nullptr, s);
                    i += 2;
                } else {
                    gsrb_sweep<V>(L.dev, C.op[(size_t)l], a, b, C.res[l], zero, nullptr, nullptr, s);
                    ++i;
                }
                std::swap(a, b);
            }
            x = a;
        } else {
            if (!from_cur) set_val<V>(x, L.nalloc, V(0), s);
            else if (C.cur[l] != x) convert<V, V>(x, C.cur[l], L.nalloc, s);
            for (int i = 0; i < mg.nu1; ++i) smooth(C, l, i == 0 && !from_cur, s);
        }
        C.cur[l] = x;
        const auto& Lc = lp.lev[(size_t)l + 1].plan;
        if (tcycle()) {
            // rescor = res - L^T(cor), then restriction
            resid_T<V>(L.dev, C.op[(size_t)l], C.rescor[l], x, C.res[l], s);
            restrict_avg<V>(Lc.dev, (int)Lc.mbc, L.dev, C.res[l + 1], C.rescor[l], lp.same_layout[(size_t)l + 1], s);
        } else if (lp.dense && Lc.dev.nbox == 1 && L.dev.nbox == 1) {
            // res[l+1] = average_down(res - L(cor)), the fine residual never stored
            resid_restrict<V>(L.dev, C.op[(size_t)l], x, C.res[l], Lc.dev, C.res[l + 1], s);
        } else {
            // rescor = res - L(cor), homogeneous BC (computeResOfCorrection)
            fill_ghosts<V>(L.dev, x, nullptr, 0, s);
            apply<V>(L.dev, (int)L.mbc, C.op[(size_t)l], C.rescor[l], x, C.res[l], s);
            restrict_avg<V>(Lc.dev, (int)Lc.mbc, L.dev, C.res[l + 1], C.rescor[l], lp.same_layout[(size_t)l + 1], s);
        }
    }
    bottom(C, s);
    C.cur[nb] = C.cor[nb];
    for (int l = nb - 1; l >= top; --l) {
        const auto& L = lp.lev[(size_t)l].plan;
        const auto& Lc = lp.lev[(size_t)l + 1].plan;
        V* x = C.cur[l];
        const bool last_plain = plain_top && l == top;
        const int nfused = last_plain ? mg.nu2 - 1 : mg.nu2;
        if (fast(l) && nfused > 0 && fused(l + 1) && lp.same_layout[(size_t)l + 1]) {
            V* a = x;
            V* b = (x == C.cor[l]) ? C.tmp[l] : C.cor[l];
            for (int i = 0; i < nfused;) {
                const V* crse = i == 0 ? C.cur[l + 1] : nullptr;
                if (i + 1 < nfused && pair_sweeps<V>(l)) {
                    gsrb_sweep2<V>(L.dev, L.hbox[0], C.op[(size_t)l], a, b, C.res[l], 0, crse, &Lc.hbox[0], s);
                    i += 2;
                } else {
                    gsrb_sweep<V>(L.dev, C.op[(size_t)l], a, b, C.res[l], 0, crse, &Lc.dev, s);
                    ++i;
                }
                std::swap(a, b);
            }
            if (last_plain) {
                // the fused buffers write valid cells only: their edge and
                // corner ghosts are 0, as the in-place correction's are
                if (a != C.cor[l]) convert<V, V>(C.cor[l], a, L.nalloc, s);
                smooth(C, l, false, s);
                a = C.cor[l];
            }
            C.cur[l] = a;
        } else {
            if (x != C.cor[l]) {
                // bring the correction back to the in-place buffer
                convert<V, V>(C.cor[l], x, L.nalloc, s);
                x = C.cor[l];
            }
            prolong_add<V>(L.dev, (int)L.mbc, Lc.dev, x, C.cur[l + 1], lp.same_layout[(size_t)l + 1], s);
            for (int i = 0; i < mg.nu2; ++i) smooth(C, l, false, s);
            C.cur[l] = x;
        }
    }
}

// mgFcycle (AMReX_MLMG.H:1422-1457) for amrlev 0: input res[0], output the
// correction of level 0 in C.cur[0].  The residual is averaged down to every
// level (avgDownResMG) and solved at the bottom; on the way up each level's
// correction starts from the trilinear interpolation of the coarser one
// (interpAssign), a V-cycle from that level reduces the residual of that
// start, and the two are added on the valid cells.  The interpolation reads
// coarse ghosts outside physical faces as the reference does: what the coarse
// level's last boundary fill left there when the two levels share their
// BoxArray lineage, else 0.  To leave the same values there, the top level of
// each of these V-cycles runs the reference's in-place two-colour sweeps (a
// boundary fill before each colour) instead of the fused out-of-place sweep.
template <class V>
void SolverImpl::fcycle(Cycle<V>& C, hipStream_t s) {
    const int nb = (int)lp.lev.size() - 1;
    if (C.hold.size() != lp.lev.size()) {
        C.hold.assign(lp.lev.size(), nullptr);
        for (int l = 0; l < nb; ++l) {
            C.own.emplace_back(new DeviceBuffer(sizeof(V) * (size_t)lp.lev[(size_t)l].plan.nalloc));
            C.hold[(size_t)l] = static_cast<V*>(C.own.back()->ptr);
            set_val<V>(C.hold[(size_t)l], lp.lev[(size_t)l].plan.nalloc, V(0), s);
        }
    }
    for (int l = 1; l <= nb; ++l) {
        const auto& Lc = lp.lev[(size_t)l].plan;
        restrict_avg<V>(Lc.dev, (int)Lc.mbc, lp.lev[(size_t)l - 1].plan.dev, C.res[l], C.res[l - 1],
                        lp.same_layout[(size_t)l], s);
    }
    bottom(C, s);
    C.cur[nb] = C.cor[nb];
    for (int l = nb - 1; l >= 0; --l) {
        const auto& L = lp.lev[(size_t)l].plan;
        const auto& Lc = lp.lev[(size_t)l + 1].plan;
        V* hold = C.hold[(size_t)l];
        fmg_interp<V>(L.dev, Lc.dev, hold, C.cur[l + 1], lp.agglomerated[(size_t)l + 1] ? 0 : 1, s);
        // res = res - L(hold), homogeneous BC (computeResOfCorrection, then LocalCopy into res)
        if (tcycle()) {
            resid_T<V>(L.dev, C.op[(size_t)l], C.res[l], hold, C.res[l], s);
        } else {
            fill_ghosts<V>(L.dev, hold, nullptr, 0, s);
            apply<V>(L.dev, (int)L.mbc, C.op[(size_t)l], C.res[l], hold, C.res[l], s);
        }
        vcycle(C, s, l, true);
        add_valid<V>(L.dev, C.cur[l], hold, s);
    }
}

// r = b - L(x) with the inhomogeneous level BC (solutionResidual, MLCellLinOp.H:1220-1235).
void SolverImpl::residual(double* r, double* x, const double* b, hipStream_t s) {
    const auto& L = lp.lev[0].plan;
    fill_ghosts<double>(L.dev, x, P<double>(lp.bcv), lp.bcv ? 1 : 0, s);
    apply<double>(L.dev, (int)L.mbc, lp.coef<double>(0), r, x, b, s);
}

double SolverImpl::norminf(const double* x, int l, hipStream_t s) {
    const auto& L = lp.lev[(size_t)l].plan;
    reduce<double>(L.dev, (int)L.mbc, 0, x, nullptr, scr.partials(), scr.scalar(4), s);
    MGX_HIP_CHECK(hipMemcpyAsync(hpin + 4, scr.scalar(4), sizeof(double), hipMemcpyDeviceToHost, s));
    MGX_HIP_CHECK(hipStreamSynchronize(s));
    return hpin[4];
}

// Residual of the initial guess into the V-cycle's level-0 right-hand side;
// scalar(0) = max|r|, scalar(1) = sum r (dense levels; else scalar(0) only).
template <class V>
void SolverImpl::initial_residual(Cycle<V>& C, hipStream_t s) {
    const auto& L0 = lp.lev[0].plan;
    if (lp.dense) {
        update_residual<V, V>(L0.dev, lp.coef<double>(0), P<double>(sol), static_cast<const V*>(nullptr), nullptr,
                              P<double>(rhs), P<double>(lp.bcv), lp.bcv ? 1 : 0, C.res[0], scr.partials(),
                              scr.scalar(0), transpose ? 1 : 0, s);
    } else {
        residual(P<double>(res), P<double>(sol), P<double>(rhs), s);
        reduce<double>(L0.dev, (int)L0.mbc, 0, P<double>(res), nullptr, scr.partials(), scr.scalar(0), s);
        reduce<double>(L0.dev, (int)L0.mbc, 1, P<double>(res), nullptr, scr.partials(), scr.scalar(1), s);
        if constexpr (!std::is_same<V, double>::value) convert<V, double>(C.res[0], P<double>(res), L0.nalloc, s);
    }
}

// One outer iteration (oneIter + computeResidual + ResNormInf), enqueued
// without host synchronisation; the residual norm and the bottom status end in
// pinned host memory.  On dense levels the update, the residual, its norm and
// the downcast to the cycle precision are one pass that reads sol_buf(parity)
// and writes sol_buf(1 - parity).
void SolverImpl::iteration(hipStream_t s, int parity, bool fmg, int ncycles) {
    const auto& L0 = lp.lev[0].plan;
    const bool dbl = alloc_prec == amrex::MLMG::Precision::Double;
    if (lp.singular && lp.enforce_solvable) {
        // makeSolvable(res) with the FP64 sum of the last residual (scalar 1)
        if (dbl) add_scaled_dev_scalar<double>(L0.dev, (int)L0.mbc, c64.res[0], scr.scalar(1), lp.volinv0, s);
        else add_scaled_dev_scalar<float>(L0.dev, (int)L0.mbc, c32.res[0], scr.scalar(1), lp.volinv0, s);
    }
    if (lp.dense) {
        double* sin = sol_buf(parity);
        double* sout = sol_buf(1 - parity);
        const double* bcv = P<double>(lp.bcv);
        if (dbl) {
            if (fmg) fcycle(c64, s);
            else vcycle(c64, s);
            update_residual<double, double>(L0.dev, lp.coef<double>(0), sin, c64.cur[0], sout, P<double>(rhs), bcv,
                                            bcv ? 1 : 0, c64.res[0], scr.partials(), scr.scalar(0), transpose ? 1 : 0, s);
        } else {
            if (fmg) fcycle(c32, s);
            else vcycle(c32, s);
            if (ncycles == 2) vcycle(c32, s, 0, false, true);
            update_residual<float, float>(L0.dev, lp.coef<double>(0), sin, c32.cur[0], sout, P<double>(rhs), bcv,
                                          bcv ? 1 : 0, c32.res[0], scr.partials(), scr.scalar(0), transpose ? 1 : 0, s);
        }
    } else {
        if (dbl) {
            if (fmg) fcycle(c64, s);
            else vcycle(c64, s);
            axpy<double>(P<double>(sol), c64.cur[0], 1.0, L0.nalloc, s);
        } else {
            if (fmg) fcycle(c32, s);
            else vcycle(c32, s);
            if (ncycles == 2) vcycle(c32, s, 0, false, true);
            add_convert<double, float>(P<double>(sol), c32.cur[0], L0.nalloc, s);
        }
        residual(P<double>(res), P<double>(sol), P<double>(rhs), s);
        reduce<double>(L0.dev, (int)L0.mbc, 0, P<double>(res), nullptr, scr.partials(), scr.scalar(0), s);
        reduce<double>(L0.dev, (int)L0.mbc, 1, P<double>(res), nullptr, scr.partials(), scr.scalar(1), s);
        if (!dbl) convert<float, double>(c32.res[0], P<double>(res), L0.nalloc, s);
    }
    MGX_HIP_CHECK(hipMemcpyAsync(hpin, scr.scalar(0), sizeof(double), hipMemcpyDeviceToHost, s));
    MGX_HIP_CHECK(hipMemcpyAsync(hpin_status, P<int>(status), 3 * sizeof(int), hipMemcpyDeviceToHost, s));
}

// ---------------------------------------------------------------------------
// Two-level AMR composite solve (MLPoisson, MLABecLaplacian): MLMGT::oneIter for AMR levels
// 0 and 1 (AMReX_MLMG.H:1228-1293), with computeResWithCrseSolFineCor,
// computeResWithCrseCorFineCor and interpCorrection (AMReX_MLMG.H:1652-1747).
// The coarse V-cycle and the fine level's smoothing sweeps run in the cycle
// precision; the fine solution, residuals, reflux and averaging stay FP64.
// ---------------------------------------------------------------------------
void SolverImpl::allocate_amr() {
    const auto& PF = lp.fine->plan;
    const auto& P0 = lp.lev[0].plan;
    hipStream_t s = amrex::Gpu::gpuStream();
    for (auto* b : {&f_sol, &f_rhs, &f_res, &f_cor, &f_bsol, &f_bcor}) {
        b->reset(new DeviceBuffer(sizeof(double) * (size_t)PF.nalloc));
        set_val<double>(static_cast<double*>((*b)->ptr), PF.nalloc, 0.0, s);
    }
    c_cor64.reset(new DeviceBuffer(sizeof(double) * (size_t)P0.nalloc));
    set_val<double>(P<double>(c_cor64), P0.nalloc, 0.0, s);
    // FP32 fine smoothing of the mixed form
    f_cor32.reset(new DeviceBuffer(sizeof(float) * (size_t)PF.nalloc));
    f_res32.reset(new DeviceBuffer(sizeof(float) * (size_t)PF.nalloc));
    set_val<float>(P<float>(f_cor32), PF.nalloc, 0.0f, s);
    set_val<float>(P<float>(f_res32), PF.nalloc, 0.0f, s);
    int64_t np = std::max<int64_t>(masked_max_partials(), reduce_partials(PF.dev, (int)PF.mbc));
    np = std::max(np, scr.cap);
    scr.ensure(np);
}

// f_res = rhs1 - L1 sol1 with the coarse/fine boundary interpolated from sol0
// (updateSolBC + solutionResidual); leaves f_bsol and the ghosts of sol1 current.
void SolverImpl::amr_fine_residual(hipStream_t s) {
    const auto& F = *lp.fine;
    const auto& P0 = lp.lev[0].plan;
    fine_bndry(F.plan.dev, P0.dev, P<double>(sol), F.d_pbc ? P<double>(F.d_pbc) : nullptr, P<double>(f_bsol), s);
    fine_fill<double>(F.plan.dev, P<double>(f_sol), P<double>(f_bsol), F.cc, lp.maxorder, s);
    apply<double>(F.plan.dev, (int)F.plan.mbc, lp.fine_coef<double>(), P<double>(f_res), P<double>(f_sol),
                  P<double>(f_rhs), s);
}

// res = rhs0 - L0 sol0 with the coarse level's boundary data (inhomogeneous, as
// computeResWithCrseSolFineCor's solutionResidual, AMReX_MLMG.H:1662-1690),
// refluxed against the fine solution (ghosts from f_bsol).
void SolverImpl::amr_coarse_residual(hipStream_t s) {
    const auto& F = *lp.fine;
    const auto& P0 = lp.lev[0].plan;
    update_residual<double, double>(P0.dev, lp.coef<double>(0), P<double>(sol), static_cast<const double*>(nullptr),
                                    nullptr, P<double>(rhs), P<double>(lp.bcv), lp.bcv ? 1 : 0, P<double>(res),
                                    scr.partials(), scr.scalar(6), 0, s);
    fine_fill<double>(F.plan.dev, P<double>(f_sol), P<double>(f_bsol), F.cc, lp.maxorder, s);
    reflux(P0.dev, F.plan.dev, static_cast<const RefluxCell*>(F.d_reflux->ptr), (int)F.hreflux.size(), P<double>(res),
           P<double>(sol), P<double>(f_sol), lp.reflux_coef(), s);
}

void SolverImpl::amr_iteration(hipStream_t s, bool fmg) {
    const auto& F = *lp.fine;
    const auto& PF = F.plan;
    const auto& P0 = lp.lev[0].plan;
    const bool dbl = alloc_prec == amrex::MLMG::Precision::Double;
    const OpCoef<double> fop = lp.fine_coef<double>();
    double* sol1 = P<double>(f_sol);
    double* res1 = P<double>(f_res);
    double* cor1 = P<double>(f_cor);
    // cor1 = 0, nu1 sweeps against res1, homogeneous coarse/fine boundary; the
    // mixed form sweeps in FP32 on an FP32 copy of res1 and widens the result
    auto mini_cycle = [&]() {
        if (dbl) {
            set_val<double>(cor1, PF.nalloc, 0.0, s);
            for (int i = 0; i < mg.nu1; ++i)
                for (int rb = 0; rb < 2; ++rb) fine_smooth_colour<double>(PF.dev, fop, cor1, res1, rb, F.cc, lp.maxorder, s);
            return;
        }
        float* c32 = P<float>(f_cor32);
        float* r32 = P<float>(f_res32);
        const OpCoef<float> fop32 = lp.fine_coef<float>();
        convert<float, double>(r32, res1, PF.nalloc, s);
        set_val<float>(c32, PF.nalloc, 0.0f, s);
        for (int i = 0; i < mg.nu1; ++i)
            for (int rb = 0; rb < 2; ++rb) fine_smooth_colour<float>(PF.dev, fop32, c32, r32, rb, F.cc, lp.maxorder, s);
        convert<double, float>(cor1, c32, PF.nalloc, s);
    };
    // ---- AMR level 1, down
    mini_cycle();
    axpy<double>(sol1, cor1, 1.0, PF.nalloc, s);
    // computeResWithCrseSolFineCor: coarse residual, fine residual of the correction, reflux, average down
    update_residual<double, double>(P0.dev, lp.coef<double>(0), P<double>(sol), static_cast<const double*>(nullptr),
                                    nullptr, P<double>(rhs), P<double>(lp.bcv), lp.bcv ? 1 : 0, P<double>(res),
                                    scr.partials(), scr.scalar(6), 0, s);
    fine_fill<double>(PF.dev, cor1, nullptr, F.cc, lp.maxorder, s);
    apply<double>(PF.dev, (int)PF.mbc, fop, res1, cor1, res1, s);
    fine_fill<double>(PF.dev, sol1, P<double>(f_bsol), F.cc, lp.maxorder, s);
    reflux(P0.dev, PF.dev, static_cast<const RefluxCell*>(F.d_reflux->ptr), (int)F.hreflux.size(), P<double>(res),
           P<double>(sol), sol1, lp.reflux_coef(), s);
    fine_avgdown<double>(PF.dev, P0.dev, P<double>(res), res1, s);
    // ---- coarsest AMR level: makeSolvable, V-cycle, sol0 += cor0
    if (lp.singular && lp.enforce_solvable) {
        reduce<double>(P0.dev, (int)P0.mbc, 1, P<double>(res), nullptr, scr.partials(), scr.scalar(7), s);
        add_scaled_dev_scalar<double>(P0.dev, (int)P0.mbc, P<double>(res), scr.scalar(7), lp.volinv0, s);
    }
    const double* cor0;
    if (dbl) {
        if (fmg) fcycle(c64, s);  // c64.res[0] is res
        else vcycle(c64, s);
        cor0 = c64.cur[0];
        axpy<double>(P<double>(sol), c64.cur[0], 1.0, P0.nalloc, s);
    } else {
        convert<float, double>(c32.res[0], P<double>(res), P0.nalloc, s);
        if (fmg) fcycle(c32, s);
        else vcycle(c32, s);
        convert<double, float>(P<double>(c_cor64), c32.cur[0], P0.nalloc, s);
        cor0 = P<double>(c_cor64);
        axpy<double>(P<double>(sol), cor0, 1.0, P0.nalloc, s);
    }
    // ---- AMR level 1, up: interpolated correction, its residual with the
    // coarse correction as boundary data, miniCycle
    fine_interp(PF.dev, P0.dev, cor1, cor0, s);
    axpy<double>(sol1, cor1, 1.0, PF.nalloc, s);
    fine_bndry(PF.dev, P0.dev, cor0, nullptr, P<double>(f_bcor), s);
    fine_fill<double>(PF.dev, cor1, P<double>(f_bcor), F.cc, lp.maxorder, s);
    apply<double>(PF.dev, (int)PF.mbc, fop, res1, cor1, res1, s);
    mini_cycle();
    axpy<double>(sol1, cor1, 1.0, PF.nalloc, s);
    // averageDownAndSync, then the fine residual and its norm
    fine_avgdown<double>(PF.dev, P0.dev, P<double>(sol), sol1, s);
    amr_fine_residual(s);
    reduce<double>(PF.dev, (int)PF.mbc, 0, res1, nullptr, scr.partials(), scr.scalar(0), s);
    MGX_HIP_CHECK(hipMemcpyAsync(hpin, scr.scalar(0), sizeof(double), hipMemcpyDeviceToHost, s));
    MGX_HIP_CHECK(hipMemcpyAsync(hpin_status, P<int>(status), 3 * sizeof(int), hipMemcpyDeviceToHost, s));
}

// MLMGT::solve for two AMR levels (AMReX_MLMG.H:356-541): prepareForSolve
// (average down solution and rhs, makeSolvable over both levels),
// computeMLResidual and the composite norms (coarse cells under the fine level
// masked), then oneIter until the fine residual and, once that passes, the
// refluxed coarse residual meet the target.
amrex::Real SolverImpl::solve_amr(const amrex::Vector<amrex::MultiFab*>& a_sol,
                                  const amrex::Vector<amrex::MultiFab const*>& a_rhs, amrex::Real a_tol_rel,
                                  amrex::Real a_tol_abs) {
    using clock = std::chrono::steady_clock;
    using amrex::MLMG;
    const auto t0 = clock::now();
    hipStream_t s = amrex::Gpu::gpuStream();
    if (a_sol.size() != 2 || a_rhs.size() != 2) throw MLMG::error("MLMG::solve: two AMR levels expected");
    if (mg.final_fill_bc) throw MLMG::error("MLMG::setFinalFillBC: not implemented for two AMR levels");
    check_bottom();
    sync_linop();
    const bool fresh = !allocated || alloc_prec != mg.precision;
    if (fresh) allocate();
    if (fresh || !f_sol) allocate_amr();
    sol_at = -1;
    const std::array<double, 8> key{{(double)mg.nu1, (double)mg.nu2, (double)mg.nuf, (double)mg.nub,
                                     (double)mg.bottom_solver, mg.bottom_reltol, mg.bottom_abstol,
                                     (double)mg.bottom_maxiter}};
    if (amr_exec && key != graph_key) release_graph();
    graph_key = key;
    refresh_ops();
    start_bottom(amrex::Gpu::gpuStream());
    const auto& P0 = lp.lev[0].plan;
    const auto& F = *lp.fine;
    const auto& PF = F.plan;
    for (int l = 0; l < 2; ++l) {
        const auto& ref = l == 0 ? P0.ref_ba : PF.ref_ba;
        if (!one_comp_on(*a_sol[(size_t)l], ref) || !one_comp_on(*a_rhs[(size_t)l], ref))
            throw MLMG::error("MLMG::solve: BoxArray mismatch (one component on the operator's BoxArray expected)");
    }
    auto in = [&](const amrex::MultiFab& m, const LevelPlan& Pl, double* dst) {
        set_val<double>(dst, Pl.nalloc, 0.0, s);
        copy_user_to_level<double, double>(Pl.dev, dst, m.dataPtr(), static_cast<const UserBox*>(m.devBoxes()),
                                           (int)m.size(), m.maxBoxCells(), m.nGrow(), s);
    };
    in(*a_sol[0], P0, P<double>(sol));
    in(*a_rhs[0], P0, P<double>(rhs));
    in(*a_sol[1], PF, P<double>(f_sol));
    in(*a_rhs[1], PF, P<double>(f_rhs));
    // prepareForSolve: averageDownSolutionRHS, makeSolvable
    fine_avgdown<double>(PF.dev, P0.dev, P<double>(sol), P<double>(f_sol), s);
    fine_avgdown<double>(PF.dev, P0.dev, P<double>(rhs), P<double>(f_rhs), s);
    if (lp.singular && lp.enforce_solvable) {
        reduce<double>(P0.dev, (int)P0.mbc, 1, P<double>(rhs), nullptr, scr.partials(), scr.scalar(3), s);
        add_scaled_dev_scalar<double>(P0.dev, (int)P0.mbc, P<double>(rhs), scr.scalar(3), lp.volinv0, s);
        add_scaled_dev_scalar<double>(PF.dev, (int)PF.mbc, P<double>(f_rhs), scr.scalar(3), lp.volinv0, s);
    }
    const unsigned char* cov = static_cast<const unsigned char*>(F.d_covered->ptr);
    auto host = [&](double* dev) {
        MGX_HIP_CHECK(hipMemcpyAsync(hpin + 5, dev, sizeof(double), hipMemcpyDeviceToHost, s));
        MGX_HIP_CHECK(hipStreamSynchronize(s));
        return hpin[5];
    };
    // computeMLResidual(1) and the composite norms
    amr_fine_residual(s);
    amr_coarse_residual(s);
    masked_max(P0.dev, cov, P<double>(res), scr.partials(), scr.scalar(8), s);
    const double r0c = host(scr.scalar(8));
    reduce<double>(PF.dev, (int)PF.mbc, 0, P<double>(f_res), nullptr, scr.partials(), scr.scalar(8), s);
    const double r0f = host(scr.scalar(8));
    masked_max(P0.dev, cov, P<double>(rhs), scr.partials(), scr.scalar(8), s);
    const double b0c = host(scr.scalar(8));
    reduce<double>(PF.dev, (int)PF.mbc, 0, P<double>(f_rhs), nullptr, scr.partials(), scr.scalar(8), s);
    const double b0f = host(scr.scalar(8));
    const double resnorm0 = std::max(r0c, r0f), rhsnorm0 = std::max(b0c, b0f);
    mg.m_init_resnorm0 = resnorm0;
    mg.m_rhsnorm0 = rhsnorm0;
    const bool use_bnorm = mg.always_use_bnorm || rhsnorm0 >= resnorm0;
    const double max_norm = use_bnorm ? rhsnorm0 : resnorm0;
    const double res_target = std::max(a_tol_abs, std::max(a_tol_rel, 1.e-16) * max_norm);
    mg.m_iter_fine_resnorm0.clear();
    mg.m_niters_cg.clear();
    mg.m_iters = 0;
    double composite = resnorm0;
    const auto t_iter = clock::now();
    if (resnorm0 > res_target) {
        bool converged = false;
        const int niters = mg.do_fixed_number_of_iters ? mg.do_fixed_number_of_iters : mg.max_iters;
        for (int iter = 0; iter < niters; ++iter) {
            if (iter < mg.max_fmg_iters) {
                amr_iteration(s, true);  // the F-cycle iterations run uncaptured
            } else if (mg.use_graph) {
                if (!amr_exec) {
                    MGX_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    amr_iteration(s);
                    MGX_HIP_CHECK(hipStreamEndCapture(s, &amr_graph));
                    MGX_HIP_CHECK(hipGraphInstantiate(&amr_exec, amr_graph, nullptr, nullptr, 0));
                }
                MGX_HIP_CHECK(hipGraphLaunch(amr_exec, s));
            } else {
                amr_iteration(s);
            }
            MGX_HIP_CHECK(hipStreamSynchronize(s));
            const double fine = hpin[0];
            mg.m_niters_cg.push_back(hpin_status[1]);
            last_bottom_status = hpin_status[0];
            mg.m_iter_fine_resnorm0.push_back(fine);
            composite = fine;
            mg.m_iters = iter + 1;
            mg.m_nvcycles = iter + 1;
            if (fine <= res_target) {
                amr_coarse_residual(s);
                masked_max(P0.dev, cov, P<double>(res), scr.partials(), scr.scalar(8), s);
                const double crse = host(scr.scalar(8));
                composite = std::max(fine, crse);
                if (mg.verbose >= 2)
                    std::printf("MLMG: Iteration %3d Fine resid/max_norm = %.6e Crse resid/max_norm = %.6e\n", iter + 1,
                                fine / max_norm, crse / max_norm);
                if (crse <= res_target) { converged = true; break; }
            } else if (mg.verbose >= 2) {
                std::printf("MLMG: Iteration %3d Fine resid/max_norm = %.6e\n", iter + 1, fine / max_norm);
            }
            if (composite > 1.e20 * max_norm || std::isnan(composite)) {
                mg.m_final_resnorm0 = composite;
                throw MLMG::error("MLMG blew up.");
            }
        }
        if (!converged && mg.do_fixed_number_of_iters == 0) {
            mg.m_final_resnorm0 = composite;
            throw MLMG::error("MLMG failed to converge.");
        }
    }
    end_bottom();
    mg.m_final_resnorm0 = composite;
    mg.timer[1] = std::chrono::duration<double>(clock::now() - t_iter).count();
    auto out = [&](amrex::MultiFab& m, const LevelPlan& Pl, const double* src) {
        copy_level_to_user<double, double>(Pl.dev, src, m.dataPtr(), static_cast<const UserBox*>(m.devBoxes()),
                                           (int)m.size(), m.maxBoxCells(), m.nGrow(), s);
    };
    out(*a_sol[0], P0, P<double>(sol));
    out(*a_sol[1], PF, P<double>(f_sol));
    sol_at = 0;
    ++mg.m_solve_called;
    MGX_HIP_CHECK(hipStreamSynchronize(s));
    mg.timer[0] = std::chrono::duration<double>(clock::now() - t0).count();
    return composite;
}

// MLCellLinOp::compGrad (AMReX_MLCellLinOp.H:1392-1437) and
// MLCellABecLap::getFluxes (AMReX_MLCellABecLap.H:272-289) on one AMR level:
// applyBC(Inhomogeneous, Solution) on the solution, then per face
//   gradient  dxinv * (x(i) - x(i-1))
//   flux      ((-b_scalar*dxinv) * b) * (x(i) - x(i-1))  [MLABecLaplacian]
//             dxinv * (x(i) - x(i-1))                    [MLPoisson, b_scalar -1]
// times 1/b_scalar when that is not 1.  The fine level of a two-level
// hierarchy takes its coarse/fine ghosts from the boundary values of the last
// solve's coarse solution, as the reference's m_bndry_sol does (also when a
// solution is passed in).
void SolverImpl::face_values(int amrlev, const std::array<amrex::MultiFab*, 3>& out, const amrex::MultiFab* usol,
                             bool flux, hipStream_t s) {
    using amrex::MLMG;
    const char* who = flux ? "MLMG::getFluxes" : "MLMG::getGradSolution";
    const bool fine = amrlev == 1;
    const LevelPlan& Pl = fine ? lp.fine->plan : lp.lev[0].plan;
    for (int d = 0; d < 3; ++d) {
        if (!out[(size_t)d]) throw MLMG::error(std::string(who) + ": null face MultiFab");
        if (!one_comp_on(*out[(size_t)d], amrex::convert(Pl.ref_ba, amrex::IntVect::TheDimensionVector(d))))
            throw MLMG::error(std::string(who) + ": face BoxArray is not the level's BoxArray converted to faces");
    }
    if (sol_at < 0 && (!usol || fine))
        throw MLMG::error(std::string(who) + ": no solution of a forward solve to take it from");
    double* x;
    if (usol) {
        if (!one_comp_on(*usol, Pl.ref_ba)) throw MLMG::error(std::string(who) + ": solution BoxArray mismatch (one component on the operator's BoxArray expected)");
        x = fine ? P<double>(f_cor) : spare_sol();
        set_val<double>(x, Pl.nalloc, 0.0, s);
        copy_user_to_level<double, double>(Pl.dev, x, usol->dataPtr(), static_cast<const UserBox*>(usol->devBoxes()),
                                           (int)usol->size(), usol->maxBoxCells(), usol->nGrow(), s);
    } else {
        x = fine ? P<double>(f_sol) : P<double>(sol_at == 1 ? sol2 : sol);
    }
    if (fine) fine_fill<double>(Pl.dev, x, P<double>(f_bsol), lp.fine->cc, lp.maxorder, s);
    else fill_ghosts<double>(Pl.dev, x, P<double>(lp.bcv), lp.bcv ? 1 : 0, s);
    const double bscalar = lp.kind == 0 ? lp.b_scalar : -1.0;
    const double betainv = 1.0 / bscalar;
    for (int d = 0; d < 3; ++d) {
        amrex::MultiFab& f = *out[(size_t)d];
        const double dxi = Pl.geom.InvCellSize(d);
        const bool abec = flux && lp.kind == 0;
        const double* b = abec ? P<double>(fine ? lp.fine->b64[(size_t)d] : lp.lev[0].b64[(size_t)d]) : nullptr;
        mgx::face_values<double>(Pl.dev, x, b, static_cast<const UserBox*>(f.devBoxes()), (int)f.size(), f.maxBoxCells(),
                                 f.nGrow(), d, abec ? bscalar * dxi : dxi, (flux && betainv != 1.0) ? 1 : 0, betainv,
                                 f.dataPtr(), s);
    }
}

template void SolverImpl::vcycle<double>(Cycle<double>&, hipStream_t, int, bool, bool);
template void SolverImpl::vcycle<float>(Cycle<float>&, hipStream_t, int, bool, bool);

}  // namespace mgx

namespace amrex {

MLMG::MLMG(MLLinOp& a_lp) : linop(a_lp), m_impl(new mgx::SolverImpl(*this, a_lp.impl())) {}
MLMG::~MLMG() = default;

Real MLMG::solve(std::initializer_list<MultiFab*> a_sol, std::initializer_list<MultiFab const*> a_rhs, Real tol_rel,
                 Real tol_abs, const char* ck) {
    return solve(Vector<MultiFab*>(a_sol), Vector<MultiFab const*>(a_rhs), tol_rel, tol_abs, ck);
}

Real MLMG::solve(const Vector<MultiFab*>& a_sol, const Vector<MultiFab const*>& a_rhs, Real a_tol_rel, Real a_tol_abs,
                 const char* checkpoint_file) {
    if (checkpoint_file != nullptr) checkPoint(a_sol, a_rhs, a_tol_rel, a_tol_abs, checkpoint_file);  // MLMG.H:361-365
    return solve_impl(a_sol, a_rhs, a_tol_rel, a_tol_abs, false);
}

void MLMG::checkPoint(const Vector<MultiFab*>& a_sol, const Vector<MultiFab const*>& a_rhs, Real a_tol_rel,
                      Real a_tol_abs, const char* a_file_name) const {
    const std::string file_name(a_file_name);
    const int namr = linop.NAMRLevels();
    if ((int)a_sol.size() < namr || (int)a_rhs.size() < namr)
        throw error("MLMG::checkPoint: one solution and one rhs per AMR level expected");
    UtilCreateCleanDirectory(file_name, false);
    {
        std::ofstream hs(file_name + "/Header", std::ios::binary | std::ios::trunc);
        if (!hs) throw error("MLMG::checkPoint: cannot open " + file_name + "/Header");
        hs.precision(17);
        hs << linop.name() << "\n"
           << "a_tol_rel = " << a_tol_rel << "\n"
           << "a_tol_abs = " << a_tol_abs << "\n"
           << "verbose = " << verbose << "\n"
           << "max_iters = " << max_iters << "\n"
           << "nu1 = " << nu1 << "\n"
           << "nu2 = " << nu2 << "\n"
           << "nuf = " << nuf << "\n"
           << "nub = " << nub << "\n"
           << "max_fmg_iters = " << max_fmg_iters << "\n"
           << "bottom_solver = " << static_cast<int>(bottom_solver) << "\n"
           << "bottom_verbose = " << bottom_verbose << "\n"
           << "bottom_maxiter = " << bottom_maxiter << "\n"
           << "bottom_reltol = " << bottom_reltol << "\n"
           << "always_use_bnorm = " << always_use_bnorm << "\n"
           << "namrlevs = " << namr << "\n"
           << "finest_amr_lev = " << namr - 1 << "\n"
           << "linop_prepared = " << linop.impl().prepared << "\n"
           << "solve_called = " << m_solve_called << "\n";
        if (!hs) throw error("MLMG::checkPoint: write of " + file_name + "/Header failed");
        for (int l = 0; l < namr; ++l) UtilCreateCleanDirectory(file_name + "/Level_" + std::to_string(l), false);
    }
    for (int l = 0; l < namr; ++l) {
        VisMF::Write(*a_sol[(size_t)l], file_name + "/Level_" + std::to_string(l) + "/sol");
        VisMF::Write(*a_rhs[(size_t)l], file_name + "/Level_" + std::to_string(l) + "/rhs");
    }
    // linop.checkPoint(file_name + "/linop"): not implemented for these operators
    throw error("MLLinOp:checkPoint: not implemented");
}

Real MLMG::solveAdjoint(const Vector<MultiFab*>& a_sol, const Vector<MultiFab const*>& a_rhs, Real a_tol_rel,
                        Real a_tol_abs) {
    return solve_impl(a_sol, a_rhs, a_tol_rel, a_tol_abs, true);
}

// solve (transpose false) and solveAdjoint (transpose true) share the loop of
// MLMGT::solve; the adjoint differs only in its fine residual b - L^T x, with
// homogeneous boundary data, and keeps its own captured iterations.
Real MLMG::solve_impl(const Vector<MultiFab*>& a_sol, const Vector<MultiFab const*>& a_rhs, Real a_tol_rel,
                      Real a_tol_abs, bool transpose) {
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    auto& S = *m_impl;
    auto& lp = linop.impl();
    hipStream_t s = Gpu::gpuStream();
    if (lp.namr == 2) {
        if (transpose) throw error("MLMG::solveAdjoint: one AMR level expected");
        return S.solve_amr(a_sol, a_rhs, a_tol_rel, a_tol_abs);
    }
    if (a_sol.size() != 1 || a_rhs.size() != 1) throw error("MLMG::solve: one AMR level expected");
    S.check_bottom();
    if (lp.kind == 0 && (std::isnan(lp.a_scalar) || std::isnan(lp.b_scalar)))
        throw error("MLABecLaplacian: setScalars must be called before solve");
    S.sync_linop();
    if (!S.allocated || S.alloc_prec != precision) S.allocate();
    S.sol_at = -1;
    if (transpose && !lp.dense)
        throw error("MLMG::solveAdjoint: the adjoint solve needs the dense storage of the finest level");
    S.transpose = transpose;
    // the captured iteration bakes in these settings; recapture when they change
    const std::array<double, 8> key{{(double)nu1, (double)nu2, (double)nuf, (double)nub, (double)bottom_solver,
                                     bottom_reltol, bottom_abstol, (double)bottom_maxiter}};
    if (std::any_of(S.graph_ok, S.graph_ok + mgx::SolverImpl::kGraphs, [](bool b) { return b; }) && key != S.graph_key)
        S.release_graph();
    S.graph_key = key;
    S.refresh_ops();
    S.start_bottom(s);
    const auto& L0 = lp.lev[0].plan;
    const MultiFab& urhs = *a_rhs[0];
    MultiFab& usol = *a_sol[0];
    if (!one_comp_on(urhs, L0.ref_ba) || !one_comp_on(usol, L0.ref_ba)) throw error("MLMG::solve: BoxArray mismatch (one component on the operator's BoxArray expected)");
    double* sol = static_cast<double*>(S.sol->ptr);
    double* rhs = static_cast<double*>(S.rhs->ptr);
    double* res = static_cast<double*>(S.res->ptr);
    // prepareForSolve (MLMG.H:937-1116): copies, solvability fix
    mgx::set_val<double>(sol, L0.nalloc, 0.0, s);
    mgx::copy_user_to_level<double, double>(L0.dev, sol, usol.dataPtr(), static_cast<const mgx::UserBox*>(usol.devBoxes()), (int)usol.size(), usol.maxBoxCells(), usol.nGrow(), s);
    mgx::copy_user_to_level<double, double>(L0.dev, rhs, urhs.dataPtr(), static_cast<const mgx::UserBox*>(urhs.devBoxes()), (int)urhs.size(), urhs.maxBoxCells(), urhs.nGrow(), s);
    (void)res;
    if (lp.singular && lp.enforce_solvable) {
        mgx::reduce<double>(L0.dev, (int)L0.mbc, 1, rhs, nullptr, S.scr.partials(), S.scr.scalar(3), s);
        mgx::add_scaled_dev_scalar<double>(L0.dev, (int)L0.mbc, rhs, S.scr.scalar(3), lp.volinv0, s);
    }
    if (precision == Precision::Double) S.initial_residual(S.c64, s);
    else S.initial_residual(S.c32, s);
    MGX_HIP_CHECK(hipMemcpyAsync(S.hpin, S.scr.scalar(0), sizeof(double), hipMemcpyDeviceToHost, s));
    MGX_HIP_CHECK(hipStreamSynchronize(s));
    const double resnorm0 = S.hpin[0];
    const double rhsnorm0 = S.norminf(rhs, 0, s);
    m_init_resnorm0 = resnorm0;
    m_rhsnorm0 = rhsnorm0;
    if (verbose >= 1) std::printf("MLMG: Initial rhs               = %.17g\nMLMG: Initial residual (resid0) = %.17g\n", rhsnorm0, resnorm0);
    const bool use_bnorm = always_use_bnorm || rhsnorm0 >= resnorm0;
    const double max_norm = use_bnorm ? rhsnorm0 : resnorm0;
    const char* norm_name = use_bnorm ? "bnorm" : "resid0";
    const double res_target = std::max(a_tol_abs, std::max(a_tol_rel, 1.e-16) * max_norm);
    m_iter_fine_resnorm0.clear();
    m_niters_cg.clear();
    m_iters = 0;
    m_nvcycles = 0;
    double composite = resnorm0;
    double rho = 0.1;  // contraction per V-cycle assumed before the first measured step
    const auto t_iter = clock::now();
    if (resnorm0 <= res_target) {
        if (verbose >= 1) std::printf("MLMG: No iterations needed\n");
    } else {
        bool converged = false;
        const int niters = do_fixed_number_of_iters ? do_fixed_number_of_iters : max_iters;
        for (int iter = 0; iter < niters; ++iter) {
            const int par = lp.dense ? (iter & 1) : 0;
            const bool fmg = iter < max_fmg_iters;
            const int ncyc = fmg ? 1 : refinement_cycles(composite, rho, res_target);
            const int gi = 4 * (ncyc - 1) + (transpose ? 2 : 0) + par;
            if (fmg) {
                S.iteration(s, par, true);  // the F-cycle iterations run uncaptured
            } else if (use_graph) {
                if (!S.graph_ok[gi]) {
                    MGX_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    S.iteration(s, par, false, ncyc);
                    MGX_HIP_CHECK(hipStreamEndCapture(s, &S.graph[gi]));
                    MGX_HIP_CHECK(hipGraphInstantiate(&S.exec[gi], S.graph[gi], nullptr, nullptr, 0));
                    S.graph_ok[gi] = true;
                }
                MGX_HIP_CHECK(hipGraphLaunch(S.exec[gi], s));
            } else {
                S.iteration(s, par, false, ncyc);
            }
            MGX_HIP_CHECK(hipStreamSynchronize(s));
            const double fine = S.hpin[0];
            if (composite > 0 && fine > 0) rho = std::pow(fine / composite, 1.0 / ncyc);
            m_nvcycles += ncyc;
            m_niters_cg.push_back(S.hpin_status[1]);
            S.last_bottom_status = S.hpin_status[0];
            m_iter_fine_resnorm0.push_back(fine);
            composite = fine;
            m_iters = iter + 1;
            if (verbose >= 2)
                std::printf("MLMG: Iteration %3d Fine resid/%s = %.6e\n", iter + 1, norm_name, fine / max_norm);
            if (fine <= res_target) {
                converged = true;
                if (verbose >= 1)
                    std::printf("MLMG: Final Iter. %d resid, resid/%s = %.6e, %.6e\n", iter + 1, norm_name, fine, fine / max_norm);
                break;
            }
            // NaN never compares greater, so test it explicitly (the reference would iterate to max_iters)
            if (fine > 1.e20 * max_norm || std::isnan(fine)) {
                m_final_resnorm0 = composite;
                throw error("MLMG blew up.");
            }
        }
        if (!converged && do_fixed_number_of_iters == 0) {
            m_final_resnorm0 = composite;
            if (verbose > 0)
                std::printf("MLMG: Failed to converge after %d iterations. resid, resid/%s = %.6e, %.6e\n", max_iters,
                            norm_name, composite, composite / max_norm);
            throw error("MLMG failed to converge.");
        }
    }
    S.end_bottom();
    m_final_resnorm0 = composite;
    timer[1] = std::chrono::duration<double>(clock::now() - t_iter).count();
    sol = S.sol_buf(lp.dense ? (m_iters & 1) : 0);
    mgx::copy_level_to_user<double, double>(L0.dev, sol, usol.dataPtr(), static_cast<const mgx::UserBox*>(usol.devBoxes()), (int)usol.size(), usol.maxBoxCells(), usol.nGrow(), s);
    // Ghost cells of the returned solution (MLMG.H:970-995, :515-523).  With
    // one ghost layer the reference solves in the caller's MultiFab itself, so
    // its face ghosts come back as the last residual's boundary fill left them
    // (closures with the level's boundary data, periodic images, neighbours)
    // and its edges and corners as the caller left them.  With more layers it
    // solves in a copy whose ghosts start at zero and, under setFinalFillBC,
    // copies back the first layer: faces filled, edges and corners zero.
    if (!transpose && usol.nGrow() >= 1 && (final_fill_bc || usol.nGrow() == 1)) {
        mgx::fill_ghosts<double>(L0.dev, sol, lp.bcv ? static_cast<double*>(lp.bcv->ptr) : nullptr, lp.bcv ? 1 : 0, s);
        int64_t max_face = 0, max_edge = 0;
        for (const Box& b : usol.boxArray().boxes()) {
            for (int d = 0; d < 3; ++d)
                max_face = std::max<int64_t>(max_face, b.numPts() / b.length(d));
            max_edge = std::max<int64_t>(max_edge, 4 * ((int64_t)b.length(0) + b.length(1) + b.length(2)) + 8);
        }
        const auto* ub = static_cast<const mgx::UserBox*>(usol.devBoxes());
        mgx::face_ghosts_to_user(L0.dev, sol, usol.dataPtr(), ub, (int)usol.size(), max_face, usol.nGrow(), s);
        if (usol.nGrow() >= 2) mgx::zero_user_edge_ghosts(usol.dataPtr(), ub, (int)usol.size(), max_edge, usol.nGrow(), s);
    }
    if (!transpose) S.sol_at = (lp.dense && (m_iters & 1)) ? 1 : 0;
    ++m_solve_called;
    MGX_HIP_CHECK(hipStreamSynchronize(s));
    timer[0] = std::chrono::duration<double>(clock::now() - t0).count();
    if (verbose >= 1) std::printf("MLMG: Timers: Solve = %.6f Iter = %.6f Bottom = %.6f\n", timer[0], timer[1], timer[2]);
    return composite;
}

void MLMG::apply(const Vector<MultiFab*>& out, const Vector<MultiFab*>& in) {
    auto& S = *m_impl;
    auto& lp = linop.impl();
    hipStream_t s = Gpu::gpuStream();
    S.sync_linop();
    if (!S.allocated) S.allocate();
    const auto& L0 = lp.lev[0].plan;
    if (out.size() != 1 || in.size() != 1 || !one_comp_on(*out[0], L0.ref_ba) || !one_comp_on(*in[0], L0.ref_ba))
        throw error("MLMG::apply: BoxArray mismatch (one component on the operator's BoxArray expected)");
    double* x = S.spare_sol();
    double* y = static_cast<double*>(S.res->ptr);
    mgx::set_val<double>(x, L0.nalloc, 0.0, s);
    mgx::copy_user_to_level<double, double>(L0.dev, x, in[0]->dataPtr(), static_cast<const mgx::UserBox*>(in[0]->devBoxes()), (int)in[0]->size(), in[0]->maxBoxCells(), in[0]->nGrow(), s);
    mgx::fill_ghosts<double>(L0.dev, x, static_cast<double*>(lp.bcv ? lp.bcv->ptr : nullptr), lp.bcv ? 1 : 0, s);
    mgx::apply<double>(L0.dev, (int)L0.mbc, lp.coef<double>(0), y, x, nullptr, s);
    mgx::copy_level_to_user<double, double>(L0.dev, y, out[0]->dataPtr(), static_cast<const mgx::UserBox*>(out[0]->devBoxes()), (int)out[0]->size(), out[0]->maxBoxCells(), out[0]->nGrow(), s);
    MGX_HIP_CHECK(hipStreamSynchronize(s));
}

// MLMG::getGradSolution (AMReX_MLMG.H:556-575): per AMR level, the face
// gradients of the last solution.
void MLMG::getGradSolution(const Vector<std::array<MultiFab*, 3>>& a_grad_sol, Location) {
    face_values_impl(a_grad_sol, nullptr, false);
}

// MLMG::getFluxes (AMReX_MLMG.H:585-660): per AMR level, the face fluxes of
// the last solution or of a_sol.
void MLMG::getFluxes(const Vector<std::array<MultiFab*, 3>>& a_flux, Location) {
    face_values_impl(a_flux, nullptr, true);
}

void MLMG::getFluxes(const Vector<std::array<MultiFab*, 3>>& a_flux, const Vector<MultiFab*>& a_sol, Location) {
    face_values_impl(a_flux, &a_sol, true);
}

void MLMG::getFluxes(const Vector<MultiFab*>& a_flux, Location) { cell_fluxes_impl(a_flux, nullptr); }

void MLMG::getFluxes(const Vector<MultiFab*>& a_flux, const Vector<MultiFab*>& a_sol, Location) {
    cell_fluxes_impl(a_flux, &a_sol);
}

// The face fluxes into face MultiFabs on each level's BoxArray, then
// average_face_to_cellcenter into components 0-2 (MLMG.H:708-733).
void MLMG::cell_fluxes_impl(const Vector<MultiFab*>& a_flux, const Vector<MultiFab*>* a_sol) {
    auto& lp = linop.impl();
    const int n = linop.NAMRLevels();
    if ((int)a_flux.size() != n) throw error("MLMG::getFluxes: one entry per AMR level expected");
    if (!lp.prepared) lp.prepare();
    std::vector<std::array<std::unique_ptr<MultiFab>, 3>> faces((size_t)n);
    Vector<std::array<MultiFab*, 3>> fp((size_t)n);
    for (int l = 0; l < n; ++l) {
        const BoxArray& ba = (l == 1 ? lp.fine->plan : lp.lev[0].plan).ref_ba;
        if (!a_flux[(size_t)l] || a_flux[(size_t)l]->boxArray() != ba || a_flux[(size_t)l]->nComp() < 3)
            throw error("MLMG::getFluxes: three components on the level's BoxArray expected");
        for (int d = 0; d < 3; ++d) {
            const BoxArray fba = amrex::convert(ba, IntVect::TheDimensionVector(d));
            faces[(size_t)l][(size_t)d].reset(new MultiFab(fba, DistributionMapping(fba), 1, 0));
            fp[(size_t)l][(size_t)d] = faces[(size_t)l][(size_t)d].get();
        }
    }
    face_values_impl(fp, a_sol, true);
    for (int l = 0; l < n; ++l)
        average_face_to_cellcenter(*a_flux[(size_t)l], 0,
                                   {fp[(size_t)l][0], fp[(size_t)l][1], fp[(size_t)l][2]});
}

void MLMG::face_values_impl(const Vector<std::array<MultiFab*, 3>>& out, const Vector<MultiFab*>* a_sol, bool flux) {
    auto& S = *m_impl;
    hipStream_t s = Gpu::gpuStream();
    const int n = linop.NAMRLevels();
    if ((int)out.size() != n || (a_sol && (int)a_sol->size() != n))
        throw error(flux ? "MLMG::getFluxes: one entry per AMR level expected"
                         : "MLMG::getGradSolution: one entry per AMR level expected");
    S.sync_linop();
    if (!S.allocated) S.allocate();
    for (int l = 0; l < n; ++l) S.face_values(l, out[(size_t)l], a_sol ? (*a_sol)[(size_t)l] : nullptr, flux, s);
    MGX_HIP_CHECK(hipStreamSynchronize(s));
}

void MLMG::compResidual(const Vector<MultiFab*>& a_res, const Vector<MultiFab*>& a_sol, const Vector<MultiFab const*>& a_rhs) {
    auto& S = *m_impl;
    auto& lp = linop.impl();
    hipStream_t s = Gpu::gpuStream();
    S.sync_linop();
    if (!S.allocated) S.allocate();
    const auto& L0 = lp.lev[0].plan;
    if (a_res.size() != 1 || a_sol.size() != 1 || a_rhs.size() != 1 || !one_comp_on(*a_res[0], L0.ref_ba) ||
        !one_comp_on(*a_sol[0], L0.ref_ba) || !one_comp_on(*a_rhs[0], L0.ref_ba))
        throw error("MLMG::compResidual: BoxArray mismatch (one component on the operator's BoxArray expected)");
    double* x = S.spare_sol();
    double* b = static_cast<double*>(S.rhs->ptr);
    double* r = static_cast<double*>(S.res->ptr);
    mgx::set_val<double>(x, L0.nalloc, 0.0, s);
    mgx::copy_user_to_level<double, double>(L0.dev, x, a_sol[0]->dataPtr(), static_cast<const mgx::UserBox*>(a_sol[0]->devBoxes()), (int)a_sol[0]->size(), a_sol[0]->maxBoxCells(), a_sol[0]->nGrow(), s);
    mgx::copy_user_to_level<double, double>(L0.dev, b, a_rhs[0]->dataPtr(), static_cast<const mgx::UserBox*>(a_rhs[0]->devBoxes()), (int)a_rhs[0]->size(), a_rhs[0]->maxBoxCells(), a_rhs[0]->nGrow(), s);
    S.residual(r, x, b, s);
    mgx::copy_level_to_user<double, double>(L0.dev, r, a_res[0]->dataPtr(), static_cast<const mgx::UserBox*>(a_res[0]->devBoxes()), (int)a_res[0]->size(), a_res[0]->maxBoxCells(), a_res[0]->nGrow(), s);
    MGX_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace amrex
