// MLMG solve orchestration (MLMGT::solve / oneIter / mgVcycle /
// actualBottomSolve restated for one GPU).  Differences from the reference's
// GPU design: no host synchronisation inside a V-cycle (the bottom BiCGStab
// iterates on the device), reductions land in device memory, and one outer
// iteration is captured into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>

#include "amrex/AMReX_MLMG_Core.H"
#include "mgx_impl.H"
#include "mgx_kernels.H"

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
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    graph_ok = false;
}

template <class V>
static void alloc_cycle(Cycle<V>& C, const LinOpImpl& lp, V* res0) {
    const size_t nl = lp.lev.size();
    C.own.clear();
    C.cor.assign(nl, nullptr);
    C.res.assign(nl, nullptr);
    C.rescor.assign(nl, nullptr);
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
        hipStream_t s = amrex::Gpu::gpuStream();
        set_val<V>(C.cor[l], n, V(0), s);
        set_val<V>(C.res[l], n, V(0), s);
        set_val<V>(C.rescor[l], n, V(0), s);
    }
}

void SolverImpl::allocate() {
    release_graph();
    const auto& P0 = lp.lev[0].plan;
    hipStream_t s = amrex::Gpu::gpuStream();
    sol.reset(new DeviceBuffer(sizeof(double) * (size_t)P0.nalloc));
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
    bottom_bicgstab<double>(L.dev, op, x, b, P<double>(bot_work), mg.bottom_reltol, mg.bottom_abstol, mg.bottom_maxiter,
                            sing ? 1 : 0, lp.volinv_bot, mg.nub, mg.nuf, P<int>(status), s);
    if constexpr (!std::is_same<V, double>::value) convert<float, double>(C.cor[nb], x, L.nalloc, s);
}

// mgVcycle (AMReX_MLMG.H:1308-1415) for amrlev 0: input res[0], output cor[0].
template <class V>
void SolverImpl::vcycle(Cycle<V>& C, hipStream_t s) {
    const int nb = (int)lp.lev.size() - 1;
    for (int l = 0; l < nb; ++l) {
        const auto& L = lp.lev[(size_t)l].plan;
        set_val<V>(C.cor[l], L.nalloc, V(0), s);
        for (int i = 0; i < mg.nu1; ++i) smooth(C, l, i == 0, s);
        // rescor = res - L(cor), homogeneous BC (computeResOfCorrection)
        fill_ghosts<V>(L.dev, C.cor[l], nullptr, 0, s);
        apply<V>(L.dev, (int)L.mbc, C.op[(size_t)l], C.rescor[l], C.cor[l], C.res[l], s);
        const auto& Lc = lp.lev[(size_t)l + 1].plan;
        restrict_avg<V>(Lc.dev, (int)Lc.mbc, L.dev, C.res[l + 1], C.rescor[l], lp.same_layout[(size_t)l + 1], s);
    }
    bottom(C, s);
    for (int l = nb - 1; l >= 0; --l) {
        const auto& L = lp.lev[(size_t)l].plan;
        const auto& Lc = lp.lev[(size_t)l + 1].plan;
        prolong_add<V>(L.dev, (int)L.mbc, Lc.dev, C.cor[l], C.cor[l + 1], lp.same_layout[(size_t)l + 1], s);
        for (int i = 0; i < mg.nu2; ++i) smooth(C, l, false, s);
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

// One outer iteration (oneIter + computeResidual + ResNormInf), enqueued
// without host synchronisation; the residual norm and the bottom status end in
// pinned host memory.
void SolverImpl::iteration(hipStream_t s) {
    const auto& L0 = lp.lev[0].plan;
    if (lp.singular && lp.enforce_solvable) {
        reduce<double>(L0.dev, (int)L0.mbc, 1, P<double>(res), nullptr, scr.partials(), scr.scalar(2), s);
        add_scaled_dev_scalar<double>(L0.dev, (int)L0.mbc, P<double>(res), scr.scalar(2), lp.volinv0, s);
    }
    if (alloc_prec == amrex::MLMG::Precision::Double) {
        vcycle(c64, s);
        axpy<double>(P<double>(sol), c64.cor[0], 1.0, L0.nalloc, s);
    } else {
        convert<float, double>(c32.res[0], P<double>(res), L0.nalloc, s);
        vcycle(c32, s);
        add_convert<double, float>(P<double>(sol), c32.cor[0], L0.nalloc, s);
    }
    residual(P<double>(res), P<double>(sol), P<double>(rhs), s);
    reduce<double>(L0.dev, (int)L0.mbc, 0, P<double>(res), nullptr, scr.partials(), scr.scalar(0), s);
    MGX_HIP_CHECK(hipMemcpyAsync(hpin, scr.scalar(0), sizeof(double), hipMemcpyDeviceToHost, s));
    MGX_HIP_CHECK(hipMemcpyAsync(hpin_status, P<int>(status), 2 * sizeof(int), hipMemcpyDeviceToHost, s));
}

}  // namespace mgx

namespace amrex {

MLMG::MLMG(MLLinOp& a_lp) : linop(a_lp), m_impl(new mgx::SolverImpl(*this, a_lp.impl())) {}
MLMG::~MLMG() = default;

Real MLMG::solve(std::initializer_list<MultiFab*> a_sol, std::initializer_list<MultiFab const*> a_rhs, Real tol_rel,
                 Real tol_abs, const char* ck) {
    return solve(Vector<MultiFab*>(a_sol), Vector<MultiFab const*>(a_rhs), tol_rel, tol_abs, ck);
}

Real MLMG::solve(const Vector<MultiFab*>& a_sol, const Vector<MultiFab const*>& a_rhs, Real a_tol_rel, Real a_tol_abs,
                 const char*) {
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    auto& S = *m_impl;
    auto& lp = linop.impl();
    hipStream_t s = Gpu::gpuStream();
    if (a_sol.size() != 1 || a_rhs.size() != 1) throw error("MLMG::solve: one AMR level expected");
    if (bottom_solver == BottomSolver::Default) bottom_solver = linop.getDefaultBottomSolver();
    if (bottom_solver != BottomSolver::bicgstab && bottom_solver != BottomSolver::smoother)
        throw error("MLMG::solve: only the bicgstab and smoother bottom solvers are implemented");
    if (lp.kind == 0 && (std::isnan(lp.a_scalar) || std::isnan(lp.b_scalar)))
        throw error("MLABecLaplacian: setScalars must be called before solve");
    if (!lp.prepared) {
        lp.prepare();
        S.release_graph();
    }
    if (!S.allocated || S.alloc_prec != precision) S.allocate();
    // the captured iteration bakes in these settings; recapture when they change
    const std::array<double, 8> key{{(double)nu1, (double)nu2, (double)nuf, (double)nub, (double)bottom_solver,
                                     bottom_reltol, bottom_abstol, (double)bottom_maxiter}};
    if (S.graph_ok && key != S.graph_key) S.release_graph();
    S.graph_key = key;
    S.refresh_ops();
    const auto& L0 = lp.lev[0].plan;
    const MultiFab& urhs = *a_rhs[0];
    MultiFab& usol = *a_sol[0];
    if (urhs.boxArray() != L0.ba || usol.boxArray() != L0.ba) throw error("MLMG::solve: BoxArray mismatch");
    double* sol = static_cast<double*>(S.sol->ptr);
    double* rhs = static_cast<double*>(S.rhs->ptr);
    double* res = static_cast<double*>(S.res->ptr);
    // prepareForSolve (MLMG.H:937-1116): copies, solvability fix
    mgx::set_val<double>(sol, L0.nalloc, 0.0, s);
    mgx::copy_user_to_level<double, double>(L0.dev, (int)L0.mbc, sol, usol.dataPtr(), usol.devOffsets(), usol.nGrow(), s);
    mgx::copy_user_to_level<double, double>(L0.dev, (int)L0.mbc, rhs, urhs.dataPtr(), urhs.devOffsets(), urhs.nGrow(), s);
    if (lp.singular && lp.enforce_solvable) {
        mgx::reduce<double>(L0.dev, (int)L0.mbc, 1, rhs, nullptr, S.scr.partials(), S.scr.scalar(3), s);
        mgx::add_scaled_dev_scalar<double>(L0.dev, (int)L0.mbc, rhs, S.scr.scalar(3), lp.volinv0, s);
    }
    S.residual(res, sol, rhs, s);
    const double resnorm0 = S.norminf(res, 0, s);
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
    double composite = resnorm0;
    const auto t_iter = clock::now();
    if (resnorm0 <= res_target) {
        if (verbose >= 1) std::printf("MLMG: No iterations needed\n");
    } else {
        bool converged = false;
        const int niters = do_fixed_number_of_iters ? do_fixed_number_of_iters : max_iters;
        for (int iter = 0; iter < niters; ++iter) {
            if (use_graph) {
                if (!S.graph_ok) {
                    MGX_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    S.iteration(s);
                    MGX_HIP_CHECK(hipStreamEndCapture(s, &S.graph));
                    MGX_HIP_CHECK(hipGraphInstantiate(&S.exec, S.graph, nullptr, nullptr, 0));
                    S.graph_ok = true;
                }
                MGX_HIP_CHECK(hipGraphLaunch(S.exec, s));
            } else {
                S.iteration(s);
            }
            MGX_HIP_CHECK(hipStreamSynchronize(s));
            const double fine = S.hpin[0];
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
    m_final_resnorm0 = composite;
    timer[1] = std::chrono::duration<double>(clock::now() - t_iter).count();
    mgx::copy_level_to_user<double, double>(L0.dev, (int)L0.mbc, sol, usol.dataPtr(), usol.devOffsets(), usol.nGrow(), s);
    MGX_HIP_CHECK(hipStreamSynchronize(s));
    timer[0] = std::chrono::duration<double>(clock::now() - t0).count();
    if (verbose >= 1) std::printf("MLMG: Timers: Solve = %.6f Iter = %.6f Bottom = %.6f\n", timer[0], timer[1], timer[2]);
    return composite;
}

void MLMG::apply(const Vector<MultiFab*>& out, const Vector<MultiFab*>& in) {
    auto& S = *m_impl;
    auto& lp = linop.impl();
    hipStream_t s = Gpu::gpuStream();
    if (!lp.prepared) { lp.prepare(); S.release_graph(); }
    if (!S.allocated) S.allocate();
    const auto& L0 = lp.lev[0].plan;
    if (out.size() != 1 || in.size() != 1 || out[0]->boxArray() != L0.ba || in[0]->boxArray() != L0.ba)
        throw error("MLMG::apply: BoxArray mismatch");
    double* x = static_cast<double*>(S.sol->ptr);
    double* y = static_cast<double*>(S.res->ptr);
    mgx::set_val<double>(x, L0.nalloc, 0.0, s);
    mgx::copy_user_to_level<double, double>(L0.dev, (int)L0.mbc, x, in[0]->dataPtr(), in[0]->devOffsets(), in[0]->nGrow(), s);
    mgx::fill_ghosts<double>(L0.dev, x, static_cast<double*>(lp.bcv ? lp.bcv->ptr : nullptr), lp.bcv ? 1 : 0, s);
    mgx::apply<double>(L0.dev, (int)L0.mbc, lp.coef<double>(0), y, x, nullptr, s);
    mgx::copy_level_to_user<double, double>(L0.dev, (int)L0.mbc, y, out[0]->dataPtr(), out[0]->devOffsets(), out[0]->nGrow(), s);
    MGX_HIP_CHECK(hipStreamSynchronize(s));
}

void MLMG::compResidual(const Vector<MultiFab*>& a_res, const Vector<MultiFab*>& a_sol, const Vector<MultiFab const*>& a_rhs) {
    auto& S = *m_impl;
    auto& lp = linop.impl();
    hipStream_t s = Gpu::gpuStream();
    if (!lp.prepared) { lp.prepare(); S.release_graph(); }
    if (!S.allocated) S.allocate();
    const auto& L0 = lp.lev[0].plan;
    if (a_res.size() != 1 || a_sol.size() != 1 || a_rhs.size() != 1 || a_res[0]->boxArray() != L0.ba ||
        a_sol[0]->boxArray() != L0.ba || a_rhs[0]->boxArray() != L0.ba)
        throw error("MLMG::compResidual: BoxArray mismatch");
    double* x = static_cast<double*>(S.sol->ptr);
    double* b = static_cast<double*>(S.rhs->ptr);
    double* r = static_cast<double*>(S.res->ptr);
    mgx::set_val<double>(x, L0.nalloc, 0.0, s);
    mgx::copy_user_to_level<double, double>(L0.dev, (int)L0.mbc, x, a_sol[0]->dataPtr(), a_sol[0]->devOffsets(), a_sol[0]->nGrow(), s);
    mgx::copy_user_to_level<double, double>(L0.dev, (int)L0.mbc, b, a_rhs[0]->dataPtr(), a_rhs[0]->devOffsets(), a_rhs[0]->nGrow(), s);
    S.residual(r, x, b, s);
    mgx::copy_level_to_user<double, double>(L0.dev, (int)L0.mbc, r, a_res[0]->dataPtr(), a_res[0]->devOffsets(), a_res[0]->nGrow(), s);
    MGX_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace amrex
