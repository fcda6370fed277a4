// Pipelined z-march residual passes of dense MG levels (mgx_march.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mgx_device.H"

namespace mgx {

// The level can use the march kernels: one FAB covering the domain, closures
// reaching at most two interior cells, at least 32 x 8 cells per plane.
bool march_ok(const LevelDev& L);
// Length of the (max, sum) partials array the march update writes.
int64_t march_partials(const LevelDev& L, int kind);

// sol_out = sol + cor (cor may be null); res = rhs - L(sol_out) with the level
// BC; per-block (max|res|, sum res) into partials.
template <class TC, class TR>
void march_update_residual(const LevelDev& L, const OpCoef<double>& op, const double* sol, const TC* cor,
                           double* sol_out, const double* rhs, const double* bcv, int inhomog, TR* res,
                           double* partials, hipStream_t s);
// crse = average of (res - L(cor)) over the 8 children, homogeneous BC (even extents).
template <class T>
void march_resid_restrict(const LevelDev& Lf, const OpCoef<T>& op, const T* cor, const T* res, const LevelDev& Lc,
                          T* crse, hipStream_t s);

}  // namespace mgx
