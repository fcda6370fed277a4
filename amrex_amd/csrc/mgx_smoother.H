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

// Two consecutive sweeps (the first one as gsrb_sweep with zero_in / crse,
// the second plain) in one launch, pin -> pout; B (and Bc) the host copies
// of the level's (and the coarse level's) box.  For dense levels that
// gsrb_sweep2_fits: extents multiples of the 64 x 16 tile in x and y, at
// least 8 planes, even periodic extents.  gsrb_sweep2_pays: the level also
// has work for at least 256 work-groups of chunks of 16 planes or more.
bool gsrb_sweep2_fits(const LevelDev& L);
bool gsrb_sweep2_pays(const LevelDev& L);
template <class T>
void gsrb_sweep2(const LevelDev& L, const BoxDev& B, const OpCoef<T>& op, const T* pin, T* pout, const T* rhs,
                 int zero_in, const T* crse, const BoxDev* Bc, hipStream_t s);

}  // namespace mgx
