// Fused residual kernels of dense MG levels (mgx_fused.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mgx_device.H"

namespace mgx {

// Doubles of the partials buffer update_residual needs for level L.
int64_t update_residual_partials(const LevelDev& L);

// sol_out = sol + cor (cor may be null: no update, sol_out may be null), then
// res = rhs - L(sol_out) with the level BC (inhomogeneous Dirichlet values in
// bcv when inhomog), stored as TR; out[0] = max|res|, out[1] = sum res (FP64).
// With transpose, res = rhs - L^T(sol_out) with homogeneous BC (bcv unused).
// sol and sol_out must be different arrays.
template <class TC, class TR>
void update_residual(const LevelDev& L, const OpCoef<double>& op, const double* sol, const TC* cor, double* sol_out,
                     const double* rhs, const double* bcv, int inhomog, TR* res, double* partials, double* out,
                     int transpose, hipStream_t s);

// Coefficient gradients of a functional through the solve L phi = rhs, given
// the adjoint lam = L^{-T} dJ/dphi; lam and phi in the dense level-0 layout.
// grad_b: faces of direction d of a user face FAB set (nodal boxes, ghost ung,
// max_faces = largest box); grad_a: cells of a user cell FAB set.  The values
// overwrite out, or are added to it when accumulate.
void grad_b(const LevelDev& L, const OpCoef<double>& op, const double* lam, const double* phi, const double* bcv,
            int inhomog, const UserBox* ub, int nbox, int64_t max_faces, int ung, int d, double* out, int accumulate,
            hipStream_t s);
void grad_a(const LevelDev& L, double alpha, const double* lam, const double* phi, const UserBox* ub, int nbox,
            int64_t max_cells, int ung, double* out, int accumulate, hipStream_t s);

// crse = average_down(res - L(cor)) with homogeneous BC (Lc: the next coarser
// dense level).
template <class T>
void resid_restrict(const LevelDev& Lf, const OpCoef<T>& op, const T* cor, const T* res, const LevelDev& Lc, T* crse,
                    hipStream_t s);

}  // namespace mgx
