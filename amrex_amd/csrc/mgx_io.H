// Device helpers of the on-disk interchange (mgx_io.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mgx {

// One component of one FAB: element (i, j, k) at off + (i-lo0) + nx*((j-lo1) + ny*(k-lo2)).
struct IoFab {
    int lo[3];
    int64_t nx, ny;
    int64_t off;
};
// A box of cells (inclusive bounds).
struct IoRegion {
    int lo[3], hi[3];
};

constexpr int kIoBlocksPerBox = 512;

// Per-box min and max over each region: partials[(b * nblk + blk) * 2 + {0, 1}],
// nblk = io_minmax_blocks(max_region).
void box_minmax(const double* data, const IoFab* fabs_dev, const IoRegion* reg_dev, int nbox, int64_t max_region,
                double* partials_dev, hipStream_t s);
int io_minmax_blocks(int64_t max_region);
// dst = src over each box's region.
void copy_region(double* dst, const IoFab* dfab_dev, const double* src, const IoFab* sfab_dev, const IoRegion* reg_dev,
                 int nbox, int64_t max_region, hipStream_t s);

// Component d of a cell = the mean of the face values of direction d on its
// two sides (amrex_avg_fc_to_cc, AMReX_MultiFabUtil_3D_C.H:39-49), over each
// box's region: dst layout dfab[3*b + d], faces f.f[d] in layout sfab[3*b + d].
struct FaceSrc {
    const double* f[3];
};
void face_to_cell(double* dst, const IoFab* dfab_dev, FaceSrc f, const IoFab* sfab_dev, const IoRegion* reg_dev,
                  int nbox, int64_t max_region, hipStream_t s);

// Cells of each box's region `whole` outside its region `valid` = vals[box]
// (FArrayBox::setComplement of one component).
void set_complement(double* data, const IoFab* fabs_dev, const IoRegion* valid_dev, const IoRegion* whole_dev,
                    const double* vals_dev, int nbox, int64_t max_region, hipStream_t s);

}  // namespace mgx
