// Reductions over the valid cells of a user MultiFab (one FAB per box, ghost-
// grown, x fastest, components one after another in each FAB): per component,
// the sum and the largest magnitude.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mgx_device.H"

namespace mgx {

// Upper bounds of one call: work-groups of the first pass and components.
constexpr int kMfReduceBlocks = 1024;
constexpr int kMfReduceMaxComp = 16;
// Doubles of partials one call needs at most (2 per component per work-group).
constexpr int64_t kMfReducePartials = (int64_t)2 * kMfReduceMaxComp * kMfReduceBlocks;

// out_dev[2c] = sum, out_dev[2c+1] = max |x| over the valid cells of component
// comp0 + c, c < ncomp; partials_dev holds kMfReducePartials doubles.  The
// grid depends only on nvalid (the MultiFab's valid cell count), so the same
// MultiFab reduces in the same order on every call.
template <class T>
void mf_reduce(const T* data, const UserBox* boxes_dev, int nbox, int ngrow, int64_t nvalid, int comp0, int ncomp,
               double* partials_dev, double* out_dev, hipStream_t s);

// Device address of the current GPU's file-scope partials array
// (kMfReducePartials doubles), looked up on the first call on each GPU.  It
// serves the reductions that allocate nothing; every such call on one GPU
// shares it, so calls that overlap on one GPU race.
double* mf_reduce_scratch();

}  // namespace mgx
