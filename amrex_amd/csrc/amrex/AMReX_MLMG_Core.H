// MLMG / MLLinOp / MLABecLaplacian / MLPoisson call surface on MI355X.
//
// Public names and argument meaning mirror the reference:
//   MLMGT::solve, setters, getters          AMReX_MLMG.H:35-228
//   LPInfo, BottomSolver                    AMReX_MLLinOp.H:30-81
//   setDomainBC / setLevelBC / setMaxOrder  AMReX_MLLinOp.H:143-156, 232, 247
//   MLABecLaplacian setScalars / setACoeffs / setBCoeffs  AMReX_MLABecLaplacian.H:25-145
//   MLPoisson                               AMReX_MLPoisson.H:26-30
//   LinOpBCType                             AMReX_LO_BCTYPES.H:21-35
// The work runs in HIP kernels of this library (mgx_kernels.hip) on the
// stream returned by amrex::Gpu::gpuStream().  Single AMR level.
#pragma once
#include <array>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../mgx_device.H"
#include "AMReX_Base.H"

namespace mgx {
struct LinOpImpl;
struct SolverImpl;
}  // namespace mgx

namespace amrex {

enum struct LinOpBCType : int {
    interior = 0,
    Dirichlet = 101,
    Neumann = 102,
    reflect_odd = 103,
    Marshak = 104,
    SanchezPomraning = 105,
    inflow = 106,
    inhomogNeumann = 107,
    Robin = 108,
    symmetry = 109,
    Periodic = 200,
    bogus = 1729
};

enum class BottomSolver : int { Default, smoother, bicgstab, cg, bicgcg, cgbicg, hypre, petsc };

struct LPInfo {
    bool do_agglomeration = true;
    bool do_consolidation = true;
    bool do_semicoarsening = false;
    int agg_grid_size = -1;
    int con_grid_size = -1;
    int con_ratio = 2;
    int con_strategy = 3;
    bool has_metric_term = true;
    int max_coarsening_level = 30;
    int max_semicoarsening_level = 0;
    int semicoarsening_direction = -1;
    int hidden_direction = -1;

    LPInfo& setAgglomeration(bool x) noexcept { do_agglomeration = x; return *this; }
    LPInfo& setConsolidation(bool x) noexcept { do_consolidation = x; return *this; }
    LPInfo& setSemicoarsening(bool x) noexcept { do_semicoarsening = x; return *this; }
    LPInfo& setAgglomerationGridSize(int x) noexcept { agg_grid_size = x; return *this; }
    LPInfo& setConsolidationGridSize(int x) noexcept { con_grid_size = x; return *this; }
    LPInfo& setConsolidationRatio(int x) noexcept { con_ratio = x; return *this; }
    LPInfo& setConsolidationStrategy(int x) noexcept { con_strategy = x; return *this; }
    LPInfo& setMetricTerm(bool x) noexcept { has_metric_term = x; return *this; }
    LPInfo& setMaxCoarseningLevel(int n) noexcept { max_coarsening_level = n; return *this; }
    LPInfo& setMaxSemicoarseningLevel(int n) noexcept { max_semicoarsening_level = n; return *this; }
    LPInfo& setSemicoarseningDirection(int n) noexcept { semicoarsening_direction = n; return *this; }
    LPInfo& setHiddenDirection(int n) noexcept { hidden_direction = n; return *this; }
    // GPU defaults of the reference (AMReX_MLLinOp.H:66-80)
    static constexpr int getDefaultAgglomerationGridSize() { return 32; }
    static constexpr int getDefaultConsolidationGridSize() { return 32; }
};

class MLLinOp {
  public:
    enum struct BCMode { Homogeneous, Inhomogeneous };
    // Where face quantities live (AMReX_MLLinOp.H:87); without embedded
    // boundaries the centroid of a face is its centre.
    enum struct Location { FaceCenter, FaceCentroid, CellCenter, CellCentroid };
    virtual ~MLLinOp();
    MLLinOp(const MLLinOp&) = delete;
    MLLinOp& operator=(const MLLinOp&) = delete;

    void setDomainBC(const std::array<LinOpBCType, 3>& lobc, const std::array<LinOpBCType, 3>& hibc);
    // Ghost cells of levelbcdata at physical faces hold inhomogeneous Dirichlet
    // values; nullptr means homogeneous.
    void setLevelBC(int amrlev, const MultiFab* levelbcdata);
    // Level-0 boxes that leave part of the domain uncovered take Dirichlet (or
    // homogeneous Neumann) values on their coarse/fine faces, interpolated
    // from crse (cell data on the domain coarsened by crse_ratio; nullptr:
    // zero) when setLevelBC is called afterwards (AMReX_MLLinOp.H:200-212).
    void setCoarseFineBC(const MultiFab* crse, int crse_ratio, LinOpBCType bc_type = LinOpBCType::Dirichlet);
    void setMaxOrder(int maxorder);
    int getMaxOrder() const;
    void setEnforceSingularSolvable(bool o);
    bool getEnforceSingularSolvable() const;
    void setGaussSeidel(bool flag);
    int NAMRLevels() const;
    int NMGLevels(int amrlev) const;
    const Geometry& Geom(int amrlev, int mglev = 0) const;
    const BoxArray& boxArray(int amrlev, int mglev = 0) const;
    bool isSingular(int amrlev) const;
    virtual std::string name() const = 0;
    virtual BottomSolver getDefaultBottomSolver() const { return BottomSolver::bicgstab; }

    mgx::LinOpImpl& impl() { return *m_impl; }
    const mgx::LinOpImpl& impl() const { return *m_impl; }

  protected:
    MLLinOp(const Vector<Geometry>& a_geom, const Vector<BoxArray>& a_grids,
            const Vector<DistributionMapping>& a_dmap, const LPInfo& a_info, int kind);
    std::unique_ptr<mgx::LinOpImpl> m_impl;
};

class MLABecLaplacian : public MLLinOp {
  public:
    MLABecLaplacian(const Vector<Geometry>& a_geom, const Vector<BoxArray>& a_grids,
                    const Vector<DistributionMapping>& a_dmap, const LPInfo& a_info = LPInfo(), int ncomp = 1);
    void setScalars(Real a, Real b);
    void setACoeffs(int amrlev, const MultiFab& alpha);
    void setACoeffs(int amrlev, Real alpha);
    void setBCoeffs(int amrlev, const std::array<MultiFab const*, 3>& beta);
    void setBCoeffs(int amrlev, Real beta);
    // Extension (no reference counterpart): gradient of a functional J through
    // the solve L(a, b) phi = rhs, given the adjoint lam = L^{-T} dJ/dphi
    // (MLMG::solveAdjoint).  grad_a (cell MultiFab, may be null) receives
    // dJ/da = -alpha*lam*phi; grad_b[d] (face MultiFabs on the same BoxArray,
    // entries may be null) receive dJ/db_d = -lam . d(L phi)/db_d, which
    // includes the boundary closure and the inhomogeneous level BC of phi.
    // Values overwrite the outputs, or are added to them when accumulate.
    void compCoeffGradient(int amrlev, const MultiFab& lam, const MultiFab& phi, MultiFab* grad_a,
                           const std::array<MultiFab*, 3>& grad_b, bool accumulate);
    std::string name() const override { return "MLABecLaplacian"; }
};

class MLPoisson : public MLLinOp {
  public:
    MLPoisson(const Vector<Geometry>& a_geom, const Vector<BoxArray>& a_grids,
              const Vector<DistributionMapping>& a_dmap, const LPInfo& a_info = LPInfo());
    std::string name() const override { return "MLPoisson"; }
};

class MLMG {
  public:
    using BottomSolver = amrex::BottomSolver;
    using Location = MLLinOp::Location;
    // Arithmetic of the solve.  Double: the reference algorithm in FP64.
    // Mixed: an FP32 V-cycle (smoothing, residual of the correction,
    // restriction, prolongation) inside the FP64 outer iteration, whose
    // residual, correction update and bottom solve stay FP64.
    enum class Precision { Double, Mixed };

    class error : public std::runtime_error {
      public:
        using std::runtime_error::runtime_error;
    };

    explicit MLMG(MLLinOp& a_lp);
    ~MLMG();
    MLMG(const MLMG&) = delete;
    MLMG& operator=(const MLMG&) = delete;

    Real solve(const Vector<MultiFab*>& a_sol, const Vector<MultiFab const*>& a_rhs, Real a_tol_rel,
               Real a_tol_abs, const char* checkpoint_file = nullptr);
    Real solve(std::initializer_list<MultiFab*> a_sol, std::initializer_list<MultiFab const*> a_rhs,
               Real a_tol_rel, Real a_tol_abs, const char* checkpoint_file = nullptr);
    // Extension (no reference counterpart): solve L^T sol = rhs with
    // homogeneous boundary data, the adjoint of solve, by the same V-cycles
    // with the fine residual rhs - L^T sol.  Same convergence test and
    // precision policy as solve.  Needs the dense storage of the finest level.
    Real solveAdjoint(const Vector<MultiFab*>& a_sol, const Vector<MultiFab const*>& a_rhs, Real a_tol_rel,
                      Real a_tol_abs);
    // out = L(in) with the inhomogeneous level BC (MLMG.H:114, 860-932)
    void apply(const Vector<MultiFab*>& out, const Vector<MultiFab*>& in);
    // Face gradients dxinv*(sol(i) - sol(i-1)) of the last solution on each
    // AMR level (MLMG.H:54-58); a_grad_sol[lev][d] is defined on the level's
    // BoxArray converted to faces of direction d.  Ghost faces are left as
    // they are.
    void getGradSolution(const Vector<std::array<MultiFab*, 3>>& a_grad_sol, Location a_loc = Location::FaceCenter);
    // Face fluxes of the last solution (MLMG.H:65-68) or of a_sol (MLMG.H:73-76):
    // -b*grad(sol) for MLABecLaplacian (the b_scalar factor divided out),
    // -grad(sol) for MLPoisson.  The solution's ghost cells are filled
    // internally; a_sol is not written.
    void getFluxes(const Vector<std::array<MultiFab*, 3>>& a_flux, Location a_loc = Location::FaceCenter);
    void getFluxes(const Vector<std::array<MultiFab*, 3>>& a_flux, const Vector<MultiFab*>& a_sol,
                   Location a_loc = Location::FaceCenter);
    // Cell-centred fluxes (MLMG.H:667-745): the face fluxes averaged to the
    // cell centres into components 0, 1, 2 of a_flux[amrlev], which lives on
    // the level's BoxArray; its other components and ghost cells are left.
    void getFluxes(const Vector<MultiFab*>& a_flux, Location a_loc = Location::CellCenter);
    void getFluxes(const Vector<MultiFab*>& a_flux, const Vector<MultiFab*>& a_sol,
                   Location a_loc = Location::CellCenter);
    // MLMG::checkPoint (MLMG.H:1979-2032): a Header with the solver settings and
    // Level_<l>/sol, Level_<l>/rhs in VisMF format under file_name; then the
    // operator's part, which MLABecLaplacian and MLPoisson do not implement
    // (MLLinOp::checkPoint aborts, AMReX_MLLinOp.H:745-747): MLMG::error.
    void checkPoint(const Vector<MultiFab*>& a_sol, const Vector<MultiFab const*>& a_rhs, Real a_tol_rel,
                    Real a_tol_abs, const char* file_name) const;
    // res = rhs - L(sol) (MLMG.H:100)
    void compResidual(const Vector<MultiFab*>& a_res, const Vector<MultiFab*>& a_sol,
                      const Vector<MultiFab const*>& a_rhs);

    void setVerbose(int v) noexcept { verbose = v; }
    void setMaxIter(int n) noexcept { max_iters = n; }
    void setMaxFmgIter(int n) noexcept { max_fmg_iters = n; }
    void setFixedIter(int nit) noexcept { do_fixed_number_of_iters = nit; }
    void setPreSmooth(int n) noexcept { nu1 = n; }
    void setPostSmooth(int n) noexcept { nu2 = n; }
    void setFinalSmooth(int n) noexcept { nuf = n; }
    void setBottomSmooth(int n) noexcept { nub = n; }
    void setBottomSolver(BottomSolver s) noexcept { bottom_solver = s; }
    // no counterpart in the reference: the bottom solver in force (after a
    // solve, Default resolved and a bicgcg/cgbicg switch taken over)
    BottomSolver getBottomSolver() const noexcept { return bottom_solver; }
    void setBottomVerbose(int v) noexcept { bottom_verbose = v; }
    void setBottomMaxIter(int n) noexcept { bottom_maxiter = n; }
    void setBottomTolerance(Real t) noexcept { bottom_reltol = t; }
    void setBottomToleranceAbs(Real t) noexcept { bottom_abstol = t; }
    void setAlwaysUseBNorm(int flag) noexcept { always_use_bnorm = flag; }
    void setFinalFillBC(int flag) noexcept { final_fill_bc = flag; }
    void setThrowException(bool t) noexcept { throw_exception = t; }
    // Extensions of this library.
    void setPrecision(Precision p) noexcept { precision = p; }
    Precision getPrecision() const noexcept { return precision; }
    void setUseGraph(bool g) noexcept { use_graph = g; }
    // Mixed precision: FP32 V-cycles per FP64 refinement step (sol += cor,
    // r = rhs - L sol).  1 (default): one, as the FP64 form iterates.  2: two,
    // the second continuing from the first's correction.  0: two, or one when
    // one V-cycle at the last measured contraction per V-cycle is predicted to
    // reach the target.  The F-cycle iterations and the FP64 form run one.
    void setRefinementVcycles(int n) {
        if (n < 0 || n > 2) throw error("MLMG::setRefinementVcycles: 0, 1 or 2 expected");
        refine_vcycles = n;
    }
    int getNumVcycles() const noexcept { return m_nvcycles; }

    int getNumIters() const noexcept { return m_iters; }
    Vector<Real> const& getResidualHistory() const noexcept { return m_iter_fine_resnorm0; }
    Real getInitRHS() const noexcept { return m_rhsnorm0; }
    Real getInitResidual() const noexcept { return m_init_resnorm0; }
    Real getFinalResidual() const noexcept { return m_final_resnorm0; }
    Vector<int> const& getNumCGIters() const noexcept { return m_niters_cg; }
    int numAMRLevels() const noexcept { return linop.NAMRLevels(); }
    // Timers of the last solve: solve, iterations, bottom (seconds)
    std::array<double, 3> getTimers() const noexcept { return timer; }

    mgx::SolverImpl& impl() { return *m_impl; }

  private:
    friend struct mgx::SolverImpl;
    Real solve_impl(const Vector<MultiFab*>& a_sol, const Vector<MultiFab const*>& a_rhs, Real a_tol_rel,
                    Real a_tol_abs, bool transpose);
    void face_values_impl(const Vector<std::array<MultiFab*, 3>>& out, const Vector<MultiFab*>* a_sol, bool flux);
    void cell_fluxes_impl(const Vector<MultiFab*>& a_flux, const Vector<MultiFab*>* a_sol);
    // V-cycles of the next refinement step (fine: the current residual norm,
    // rho: the last measured contraction per V-cycle)
    int refinement_cycles(double fine, double rho, double target) const noexcept {
        if (precision != Precision::Mixed) return 1;
        if (refine_vcycles == 1 || refine_vcycles == 2) return refine_vcycles;
        return fine * rho <= target ? 1 : 2;
    }
    MLLinOp& linop;
    std::unique_ptr<mgx::SolverImpl> m_impl;

    bool throw_exception = false;
    int verbose = 1;
    int max_iters = 200;
    int do_fixed_number_of_iters = 0;
    int nu1 = 2, nu2 = 2, nuf = 8, nub = 0;
    int max_fmg_iters = 0;
    BottomSolver bottom_solver = BottomSolver::Default;
    int bottom_verbose = 0;
    int bottom_maxiter = 200;
    Real bottom_reltol = 1.e-4;
    Real bottom_abstol = -1.0;
    int always_use_bnorm = 0;
    int final_fill_bc = 0;
    Precision precision = Precision::Double;
    bool use_graph = true;
    int refine_vcycles = 1;

    int m_iters = 0;
    int m_nvcycles = 0;
    Long m_solve_called = 0;
    Real m_rhsnorm0 = 0, m_init_resnorm0 = 0, m_final_resnorm0 = 0;
    Vector<Real> m_iter_fine_resnorm0;
    Vector<int> m_niters_cg;
    std::array<double, 3> timer{{0, 0, 0}};
};

}  // namespace amrex
