// Fused red+black GSRB sweep of a dense MG level (mgx_smoother.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mgx_device.H"

namespace mgx {

// pout = one red+black sweep of pin against rhs (homogeneous BC), out of place.
// zero_in: pin is taken as zero (first sweep after the correction is cleared).
// crse (with its level Lc): the coarse correction is added to pin on the fly
// (piecewise-constant prolongation fused into the sweep).  L must be dense
// (one box covering the domain) with even extents in periodic directions.
template <class T>
void gsrb_sweep(const LevelDev& L, const OpCoef<T>& op, const T* pin, T* pout, const T* rhs, int zero_in,
                const T* crse, const LevelDev* Lc, hipStream_t s);

}  // namespace mgx
