// GMRES with MLMG as preconditioner: the call surface of amrex::GMRESMLMG
// (AMReX_GMRES_MLMG.H:18-109) and of the GMRES settings it forwards to
// (AMReX_GMRES.H:44-108), for one AMR level.
//
// Right-preconditioned restarted GMRES on L(cor) = L(sol) - rhs with the
// homogeneous operator, then sol -= cor (AMReX_GMRES_MLMG.H:211-231).  The
// Krylov vectors are FP64 arrays in the level-0 layout in HBM; the
// preconditioner is setPrecondNumIters V-cycles of the MLMG object in its own
// precision.  With MLMG::Precision::Mixed (FP32 V-cycles, which vary at the
// FP32 rounding level) the solution is built from the stored preconditioned
// vectors (flexible GMRES) so that it meets the FP64 tolerance.
#pragma once

#include <memory>

#include "AMReX_MLMG_Core.H"

namespace mgx {
struct GmresImpl;
}

namespace amrex {

class GMRESMLMG {
  public:
    explicit GMRESMLMG(MLMG& mlmg);
    ~GMRESMLMG();
    GMRESMLMG(const GMRESMLMG&) = delete;
    GMRESMLMG& operator=(const GMRESMLMG&) = delete;

    // a_sol: initial guess in, solution out (valid cells); a_rhs: right-hand side.
    void solve(MultiFab& a_sol, const MultiFab& a_rhs, Real a_tol_rel, Real a_tol_abs);

    void setVerbose(int v);
    void setMaxIters(int niters);
    // GMRES::setRestartLength (the reference reaches it through getGMRES()).
    void setRestartLength(int rl);
    int getNumIters() const;
    // GMRES::getStatus: 0 converged, 1 iteration limit reached, -1 not run.
    int getStatus() const;
    // 2-norm of the last GMRES residual estimate.
    Real getResidualNorm() const;
    // Only for operators with L(0) = 0 (no inhomogeneous boundary data): GMRES
    // then solves L(sol) = rhs directly from a zero initial guess.
    void setPropertyOfZero(bool b);
    // Returns the previous setting.
    bool usePrecond(bool new_flag);
    void setPrecondNumIters(int precond_niters);
    // Extension: the residual estimate after each iteration of the last solve.
    const Vector<Real>& getResidualHistory() const;

  private:
    std::unique_ptr<mgx::GmresImpl> m_impl;
};

}  // namespace amrex
