// Vector operations of GMRES over one MG level's cell arrays (mgx_krylov.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mgx_device.H"

namespace mgx {

constexpr int kMaxKrylov = 32;  // vectors one fused pass can take (restart length + 2)

struct KrylovSet {
    const double* x[kMaxKrylov];
};
struct KrylovCoef {
    double c[kMaxKrylov];
};

// Partials buffer (doubles) multi_dot needs on level L.
int64_t multi_dot_partials(const LevelDev& L);
// out_dev[j] = sum over the valid cells of y * xs.x[j], j < k <= kMaxKrylov.
void multi_dot(const LevelDev& L, const double* y, const KrylovSet& xs, int k, double* partials, double* out_dev,
               hipStream_t s);
// y = a*y + sum_j cs.c[j] * xs.x[j] over n elements (a == 0: y is not read).
void multi_axpy(double* y, double a, const KrylovSet& xs, const KrylovCoef& cs, int k, int64_t n, hipStream_t s);
// x *= f over n elements.
void scale(double* x, double f, int64_t n, hipStream_t s);

}  // namespace mgx
