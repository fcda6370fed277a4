// HBM streaming probe: the practical bandwidth ceiling the stencil kernels are
// measured against.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mgx {

// mode 0: dst = src; 1: read src only; 2: write dst only.  nbytes is a
// multiple of 16 and both pointers are 16-byte aligned.  Asynchronous.
void hbm_probe(int mode, const void* src, void* dst, int64_t nbytes, double* sink, hipStream_t s);

}  // namespace mgx
