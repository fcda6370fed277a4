// GMRES with the MLMG V-cycle as right preconditioner, on the GPU.
//
// The Krylov loop follows GMRES<V,M> (AMReX_GMRES.H:199-413): restarted
// GMRES(m) with classical Gram-Schmidt applied twice, Givens rotations on the
// host and the solution built from the back-substituted coefficients through
// one more preconditioner application; GMRESMLMGT (AMReX_GMRES_MLMG.H:111-231)
// supplies the vector space (level-0 cell arrays, dot products over the valid
// cells), the operator (homogeneous BC) and the preconditioner (V-cycles of
// the MLMG object).  Vectors stay in HBM; only the Hessenberg column crosses
// to the host, one transfer per Gram-Schmidt pass.
//
// With MLMG::Precision::Mixed the preconditioner is FP32 V-cycles, which differ
// from one application to the next at the FP32 rounding level; the reference's
// solution update x += M(V y) assumes one fixed M, so the true residual would
// stall near 1e-5 of the initial one.  In that precision the loop keeps the
// preconditioned vectors z_j = M(v_j) and updates x += Z y (flexible GMRES,
// Saad 1993), which meets the FP64 tolerance; in FP64 it is the reference's
// algorithm.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

#include "amrex/AMReX_GMRES_MLMG.H"
#include "mgx_impl.H"
#include "mgx_kernels.H"
#include "mgx_krylov.H"

namespace mgx {

using amrex::DeviceBuffer;

struct GmresImpl {
    amrex::MLMG& mg;
    int verbose = 0, maxiter = 2000, restart = 30, precond_niters = 1;
    bool use_precond = true, prop_zero = false;
    int its = 0, status = -1;
    double res = std::numeric_limits<double>::max();
    double rtol = 0.0, atol = 0.0;
    amrex::Vector<double> hist;
    // level-0 FP64 vectors, allocated on first use and kept with the object
    std::vector<std::unique_ptr<DeviceBuffer>> vv, zz;  // Krylov vectors; preconditioned ones (flexible)
    bool flexible = false;
    std::unique_ptr<DeviceBuffer> tmp_rhs, tmp_lhs, xsol, brhs, bres, xcor, partials, ddots;
    double* hdots = nullptr;  // pinned
    int64_t nalloc = 0;
    // Hessenberg matrices (0:rs+1, 0:rs), Givens rotations and the rotated rhs
    std::vector<double> hh, hes, grs, cc, ss;

    explicit GmresImpl(amrex::MLMG& m) : mg(m) {}
    ~GmresImpl() {
        if (hdots) (void)hipHostFree(hdots);
    }

    SolverImpl& S() { return mg.impl(); }
    LinOpImpl& lp() { return S().lp; }
    const LevelPlan& L0() { return lp().lev[0].plan; }
    double& H(std::vector<double>& t, int i, int j) { return t[(size_t)i * (size_t)(restart + 1) + (size_t)j]; }

    double* vec(std::unique_ptr<DeviceBuffer>& b, hipStream_t s) {
        if (!b) {
            b.reset(new DeviceBuffer(sizeof(double) * (size_t)nalloc));
            set_val<double>(static_cast<double*>(b->ptr), nalloc, 0.0, s);
        }
        return static_cast<double*>(b->ptr);
    }
    double* V(int j, hipStream_t s) {
        while ((int)vv.size() <= j) vv.emplace_back();
        return vec(vv[(size_t)j], s);
    }
    double* Z(int j, hipStream_t s) {
        while ((int)zz.size() <= j) zz.emplace_back();
        return vec(zz[(size_t)j], s);
    }

    // ---- vector space of GMRESMLMGT (AMReX_GMRES_MLMG.H:120-185)
    void copy(double* dst, const double* src, hipStream_t s) { convert<double, double>(dst, src, nalloc, s); }
    // out[j] = <y, xs[j]> over the valid cells, j < k
    void dots(const double* y, const double* const* xs, int k, double* out, hipStream_t s) {
        double* dd = static_cast<double*>(ddots->ptr);
        for (int c0 = 0; c0 < k; c0 += kMaxKrylov) {
            KrylovSet set{};
            const int kc = std::min(kMaxKrylov, k - c0);
            for (int j = 0; j < kc; ++j) set.x[j] = xs[c0 + j];
            multi_dot(L0().dev, y, set, kc, static_cast<double*>(partials->ptr), dd + c0, s);
        }
        MGX_HIP_CHECK(hipMemcpyAsync(hdots, dd, sizeof(double) * (size_t)k, hipMemcpyDeviceToHost, s));
        MGX_HIP_CHECK(hipStreamSynchronize(s));
        for (int j = 0; j < k; ++j) out[j] = hdots[j];
    }
    double norm2(const double* x, hipStream_t s) {
        double d = 0.0;
        dots(x, &x, 1, &d, s);
        return std::sqrt(d);
    }
    // y = a*y + sum_j c[j] xs[j]
    void lincomb(double* y, double a, const double* const* xs, const double* c, int k, hipStream_t s) {
        if (k == 0) {
            if (a != 1.0) mgx::scale(y, a, nalloc, s);
            return;
        }
        for (int c0 = 0; c0 < k; c0 += kMaxKrylov) {
            KrylovSet set{};
            KrylovCoef cs{};
            const int kc = std::min(kMaxKrylov, k - c0);
            for (int j = 0; j < kc; ++j) {
                set.x[j] = xs[c0 + j];
                cs.c[j] = c[c0 + j];
            }
            multi_axpy(y, c0 == 0 ? a : 1.0, set, cs, kc, nalloc, s);
        }
    }
    // lhs = L(rhs), homogeneous BC (MLLinOp::apply, correction state); fills rhs's ghosts
    void apply(double* lhs, const double* rhs, hipStream_t s) {
        double* x = const_cast<double*>(rhs);
        fill_ghosts<double>(L0().dev, x, nullptr, 0, s);
        mgx::apply<double>(L0().dev, (int)L0().mbc, lp().coef<double>(0), lhs, x, nullptr, s);
    }
    // GMRESMLMGT::precond (AMReX_GMRES_MLMG.H:187-211).  Returns where the
    // result is: lhs, or (lhs == nullptr, FP64) the cycle's own correction buffer,
    // valid until the next V-cycle.  One FP64 V-cycle reads its right-hand side
    // in place.
    const double* precond(double* lhs, const double* rhs, hipStream_t s) {
        if (!use_precond) {
            if (!lhs) return rhs;
            copy(lhs, rhs, s);
            return lhs;
        }
        SolverImpl& So = S();
        const auto& L = L0();
        const bool dbl = So.alloc_prec == amrex::MLMG::Precision::Double;
        if (!dbl && !lhs) lhs = vec(tmp_lhs, s);
        for (int ic = 0; ic < precond_niters; ++ic) {
            if (dbl) {
                auto& C = So.c64;
                double* res0 = C.res[0];
                if (ic == 0 && precond_niters == 1) {
                    C.res[0] = const_cast<double*>(rhs);  // read only by the V-cycle
                } else if (ic == 0) {
                    convert<double, double>(C.res[0], rhs, nalloc, s);
                } else {  // computeResOfCorrection(0, 0), then res = rescor
                    fill_ghosts<double>(L.dev, C.cur[0], nullptr, 0, s);
                    mgx::apply<double>(L.dev, (int)L.mbc, C.op[0], C.rescor[0], C.cur[0], C.res[0], s);
                    convert<double, double>(C.res[0], C.rescor[0], nalloc, s);
                }
                So.vcycle(C, s);
                C.res[0] = res0;
                if (precond_niters == 1 && !lhs) return C.cur[0];
                if (!lhs) lhs = vec(tmp_lhs, s);
                if (ic == 0) convert<double, double>(lhs, C.cur[0], nalloc, s);
                else axpy<double>(lhs, C.cur[0], 1.0, nalloc, s);
            } else {
                auto& C = So.c32;
                if (ic == 0) {
                    convert<float, double>(C.res[0], rhs, nalloc, s);
                } else {
                    fill_ghosts<float>(L.dev, C.cur[0], nullptr, 0, s);
                    mgx::apply<float>(L.dev, (int)L.mbc, C.op[0], C.rescor[0], C.cur[0], C.res[0], s);
                    convert<float, float>(C.res[0], C.rescor[0], nalloc, s);
                }
                So.vcycle(C, s);
                if (ic == 0) convert<double, float>(lhs, C.cur[0], nalloc, s);
                else add_convert<double, float>(lhs, C.cur[0], nalloc, s);
            }
        }
        return lhs;
    }

    // ---- GMRES<V,M> (AMReX_GMRES.H:199-413)
    bool converged(double r0, double r) const { return (r < r0 * rtol) || (r < atol); }

    void gram_schmidt(int it, hipStream_t s) {
        double* v1 = V(it + 1, s);
        std::vector<const double*> xs((size_t)it + 1);
        for (int j = 0; j <= it; ++j) xs[(size_t)j] = V(j, s);
        for (int j = 0; j <= it; ++j) H(hh, j, it) = H(hes, j, it) = 0.0;
        std::vector<double> lhh((size_t)it + 1), neg((size_t)it + 1);
        for (int cnt = 0; cnt < 2; ++cnt) {
            dots(v1, xs.data(), it + 1, lhh.data(), s);
            for (int j = 0; j <= it; ++j) neg[(size_t)j] = -lhh[(size_t)j];
            lincomb(v1, 1.0, xs.data(), neg.data(), it + 1, s);
            for (int j = 0; j <= it; ++j) {
                H(hh, j, it) += lhh[(size_t)j];
                H(hes, j, it) -= lhh[(size_t)j];
            }
        }
    }

    void update_hessenberg(int it, bool happyend) {
        for (int j = 1; j <= it; ++j) {
            const double tt = H(hh, j - 1, it);
            H(hh, j - 1, it) = cc[(size_t)j - 1] * tt + ss[(size_t)j - 1] * H(hh, j, it);
            H(hh, j, it) = cc[(size_t)j - 1] * H(hh, j, it) - ss[(size_t)j - 1] * tt;
        }
        if (!happyend) {
            const double tt = std::sqrt(H(hh, it, it) * H(hh, it, it) + H(hh, it + 1, it) * H(hh, it + 1, it));
            cc[(size_t)it] = H(hh, it, it) / tt;
            ss[(size_t)it] = H(hh, it + 1, it) / tt;
            grs[(size_t)it + 1] = -(ss[(size_t)it] * grs[(size_t)it]);
            grs[(size_t)it] = cc[(size_t)it] * grs[(size_t)it];
            H(hh, it, it) = cc[(size_t)it] * H(hh, it, it) + ss[(size_t)it] * H(hh, it + 1, it);
            res = std::abs(grs[(size_t)it + 1]);
        } else {
            res = 0.0;
        }
    }

    void build_solution(double* xx, int it, hipStream_t s) {
        if (it < 0) return;
        if (H(hh, it, it) != 0.0) grs[(size_t)it] /= H(hh, it, it);
        else grs[(size_t)it] = 0.0;
        for (int ii = 1; ii <= it; ++ii) {
            const int k = it - ii;
            double tt = grs[(size_t)k];
            for (int j = k + 1; j <= it; ++j) tt -= H(hh, k, j) * grs[(size_t)j];
            grs[(size_t)k] = tt / H(hh, k, k);
        }
        std::vector<const double*> xs((size_t)it + 1);
        if (flexible) {  // x += Z y
            for (int j = 0; j <= it; ++j) xs[(size_t)j] = Z(j, s);
            lincomb(xx, 1.0, xs.data(), grs.data(), it + 1, s);
            return;
        }
        for (int j = 0; j <= it; ++j) xs[(size_t)j] = V(j, s);
        double* tr = vec(tmp_rhs, s);
        lincomb(tr, 0.0, xs.data(), grs.data(), it + 1, s);
        axpy<double>(xx, precond(nullptr, tr, s), 1.0, nalloc, s);
    }

    void compute_residual(double* rr, const double* xx, const double* bb, hipStream_t s) {
        double* tl = vec(tmp_lhs, s);
        double* tr = vec(tmp_rhs, s);
        copy(tl, xx, s);
        apply(tr, tl, s);
        const double* xs[2] = {bb, tr};
        const double c[2] = {1.0, -1.0};
        lincomb(rr, 0.0, xs, c, 2, s);
    }

    void cycle(double* xx, double& rnorm0, hipStream_t s) {
        double* v0 = V(0, s);
        res = norm2(v0, s);
        grs[0] = res;
        if (res == 0.0) {
            status = 0;
            return;
        }
        mgx::scale(v0, 1.0 / res, nalloc, s);
        if (its == 0) rnorm0 = res;
        status = converged(rnorm0, res) ? 0 : -1;
        int it = 0;
        while (it < restart && its < maxiter) {
            if (verbose > 1) std::printf("GMRES: iter = %d, residual = %.6g, %.6g (rel.)\n", its, res, res / rnorm0);
            if (status == 0) break;
            double* vn = V(it + 1, s);
            apply(vn, precond(flexible ? Z(it, s) : nullptr, V(it, s), s), s);
            gram_schmidt(it, s);
            const double tt = norm2(vn, s);
            const bool happyend = tt < 1.e-99;
            if (!happyend) mgx::scale(vn, 1.0 / tt, nalloc, s);
            H(hh, it + 1, it) = tt;
            H(hes, it + 1, it) = tt;
            update_hessenberg(it, happyend);
            ++it;
            ++its;
            hist.push_back(res);
            status = converged(rnorm0, res) ? 0 : -1;
            if (happyend) break;
        }
        if (verbose > 1 && (status != 0 || its >= maxiter))
            std::printf("GMRES: iter = %d, residual = %.6g, %.6g (rel.)\n", its, res, res / rnorm0);
        build_solution(xx, it - 1, s);
    }

    // GMRES<V,M>::solve: x = 0, then cycles until converged or maxiter
    void gmres(double* x, const double* b, hipStream_t s) {
        const size_t rs = (size_t)restart;
        hh.assign((rs + 2) * (rs + 1), 0.0);
        hes.assign(hh.size(), 0.0);
        grs.assign(rs + 2, 0.0);
        cc.assign(rs + 1, 0.0);
        ss.assign(rs + 1, 0.0);
        copy(V(0, s), b, s);
        set_val<double>(x, nalloc, 0.0, s);
        its = 0;
        status = -1;
        hist.clear();
        double rnorm0 = 0.0;
        cycle(x, rnorm0, s);
        while (status == -1 && its < maxiter) {
            compute_residual(V(0, s), x, b, s);
            cycle(x, rnorm0, s);
        }
        if (status == -1 && its >= maxiter) status = 1;
    }

    void solve(amrex::MultiFab& usol, const amrex::MultiFab& urhs, double a_rtol, double a_atol) {
        using clock = std::chrono::steady_clock;
        const auto t0 = clock::now();
        if (lp().namr != 1) throw amrex::MLMG::error("GMRESMLMG: only single-level solves are supported");
        if (restart < 1) throw amrex::MLMG::error("GMRESMLMG: the restart length must be positive");
        S().prepare_cycles();
        const auto& L = L0();
        if (!one_comp_on(usol, L.ref_ba) || !one_comp_on(urhs, L.ref_ba))
            throw amrex::MLMG::error("GMRESMLMG::solve: BoxArray mismatch (one component on the operator's BoxArray expected)");
        hipStream_t s = amrex::Gpu::gpuStream();
        flexible = use_precond && S().alloc_prec == amrex::MLMG::Precision::Mixed;
        if (nalloc != L.nalloc) {
            vv.clear();
            zz.clear();
            for (auto* b : {&tmp_rhs, &tmp_lhs, &xsol, &brhs, &bres, &xcor}) b->reset();
            nalloc = L.nalloc;
        }
        if (!partials) partials.reset(new DeviceBuffer(sizeof(double) * (size_t)multi_dot_partials(L.dev)));
        const size_t nd = (size_t)std::max(restart + 2, kMaxKrylov);
        if (!ddots || ddots->bytes < sizeof(double) * nd) {
            ddots.reset(new DeviceBuffer(sizeof(double) * nd));
            if (hdots) MGX_HIP_CHECK(hipHostFree(hdots));
            MGX_HIP_CHECK(hipHostMalloc((void**)&hdots, sizeof(double) * nd, hipHostMallocDefault));
        }
        rtol = a_rtol;
        atol = a_atol;
        double* x = vec(xsol, s);
        double* b = vec(brhs, s);
        set_val<double>(x, nalloc, 0.0, s);
        set_val<double>(b, nalloc, 0.0, s);
        copy_user_to_level<double, double>(L.dev, x, usol.dataPtr(), static_cast<const UserBox*>(usol.devBoxes()),
                                           (int)usol.size(), usol.maxBoxCells(), usol.nGrow(), s);
        copy_user_to_level<double, double>(L.dev, b, urhs.dataPtr(), static_cast<const UserBox*>(urhs.devBoxes()),
                                           (int)urhs.size(), urhs.maxBoxCells(), urhs.nGrow(), s);
        if (prop_zero) {
            gmres(x, b, s);  // L(x) = rhs from x = 0
        } else {
            // res = L(sol) - rhs with the level's boundary data (MLMG::apply), L(cor) = res, sol -= cor
            double* r = vec(bres, s);
            fill_ghosts<double>(L.dev, x, P(lp().bcv), lp().bcv ? 1 : 0, s);
            mgx::apply<double>(L.dev, (int)L.mbc, lp().coef<double>(0), r, x, nullptr, s);
            axpy<double>(r, b, -1.0, nalloc, s);
            double* c = vec(xcor, s);
            gmres(c, r, s);
            axpy<double>(x, c, -1.0, nalloc, s);
        }
        copy_level_to_user<double, double>(L.dev, x, usol.dataPtr(), static_cast<const UserBox*>(usol.devBoxes()),
                                           (int)usol.size(), usol.maxBoxCells(), usol.nGrow(), s);
        MGX_HIP_CHECK(hipStreamSynchronize(s));
        if (verbose > 0)
            std::printf("GMRES: Solve Time = %.6f\n", std::chrono::duration<double>(clock::now() - t0).count());
    }

    static const double* P(const std::unique_ptr<DeviceBuffer>& b) {
        return b ? static_cast<const double*>(b->ptr) : nullptr;
    }
};

}  // namespace mgx

namespace amrex {

// GMRESMLMGT::GMRESMLMGT (AMReX_GMRES_MLMG.H:111-119): quiet MLMG, single level.
GMRESMLMG::GMRESMLMG(MLMG& mlmg) : m_impl(new mgx::GmresImpl(mlmg)) {
    if (mlmg.numAMRLevels() != 1) throw MLMG::error("GMRESMLMG: only single-level solves are supported");
    mlmg.setVerbose(0);
    mlmg.setBottomVerbose(0);
}
GMRESMLMG::~GMRESMLMG() = default;

void GMRESMLMG::solve(MultiFab& a_sol, const MultiFab& a_rhs, Real a_tol_rel, Real a_tol_abs) {
    m_impl->solve(a_sol, a_rhs, a_tol_rel, a_tol_abs);
}
void GMRESMLMG::setVerbose(int v) { m_impl->verbose = v; }
void GMRESMLMG::setMaxIters(int niters) { m_impl->maxiter = niters; }
void GMRESMLMG::setRestartLength(int rl) { m_impl->restart = rl; }
int GMRESMLMG::getNumIters() const { return m_impl->its; }
int GMRESMLMG::getStatus() const { return m_impl->status; }
Real GMRESMLMG::getResidualNorm() const { return m_impl->res; }
void GMRESMLMG::setPropertyOfZero(bool b) { m_impl->prop_zero = b; }
bool GMRESMLMG::usePrecond(bool new_flag) {
    const bool old = m_impl->use_precond;
    m_impl->use_precond = new_flag;
    return old;
}
void GMRESMLMG::setPrecondNumIters(int precond_niters) { m_impl->precond_niters = precond_niters; }
const Vector<Real>& GMRESMLMG::getResidualHistory() const { return m_impl->hist; }

}  // namespace amrex
