// Device helpers of the on-disk interchange (VisMF, plotfiles): per-box
// minima and maxima of one component over the valid cells (the VisMF header
// records them, AMReX_VisMF.cpp:699-735) and box-by-box copies of a region
// between two FAB layouts (MultiFab::Copy; ghost-free copies for plotfiles),
// and the ghost fill of VisMF::Write with set_ghost.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "mgx_io.H"

namespace mgx {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int64_t fab_index(const IoFab& F, int i, int j, int k) {
    return F.off + (int64_t)(i - F.lo[0]) + F.nx * ((int64_t)(j - F.lo[1]) + F.ny * (int64_t)(k - F.lo[2]));
}

// partials[(box * gridDim.x + blk) * 2 + {0, 1}] = min, max over this block's share of the box
__global__ __launch_bounds__(kBlock) void k_box_minmax(const double* data, const IoFab* fabs, const IoRegion* reg,
                                                       double* partials) {
    __shared__ double smin[kBlock / 64], smax[kBlock / 64];
    const IoFab& F = fabs[blockIdx.y];
    const IoRegion& R = reg[blockIdx.y];
    const int nx = R.hi[0] - R.lo[0] + 1, ny = R.hi[1] - R.lo[1] + 1, nz = R.hi[2] - R.lo[2] + 1;
    const int64_t n = (int64_t)nx * ny * nz;
    double mn = DBL_MAX, mx = -DBL_MAX;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlock) {
        const int i = R.lo[0] + (int)(q % nx), j = R.lo[1] + (int)((q / nx) % ny), k = R.lo[2] + (int)(q / ((int64_t)nx * ny));
        const double v = data[fab_index(F, i, j, k)];
        mn = fmin(mn, v);
        mx = fmax(mx, v);
    }
    for (int o = 32; o > 0; o >>= 1) {
        mn = fmin(mn, __shfl_down(mn, o, 64));
        mx = fmax(mx, __shfl_down(mx, o, 64));
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        smin[wid] = mn;
        smax[wid] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) {
            mn = fmin(mn, smin[w]);
            mx = fmax(mx, smax[w]);
        }
        const int64_t p = ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 2;
        partials[p] = mn;
        partials[p + 1] = mx;
    }
}

// dst(region of box b) = src(region of box b), one component each
__global__ __launch_bounds__(kBlock) void k_copy_region(double* dst, const IoFab* dfab, const double* src,
                                                        const IoFab* sfab, const IoRegion* reg) {
    const IoFab& D = dfab[blockIdx.y];
    const IoFab& S = sfab[blockIdx.y];
    const IoRegion& R = reg[blockIdx.y];
    const int nx = R.hi[0] - R.lo[0] + 1, ny = R.hi[1] - R.lo[1] + 1, nz = R.hi[2] - R.lo[2] + 1;
    const int64_t n = (int64_t)nx * ny * nz;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlock) {
        const int i = R.lo[0] + (int)(q % nx), j = R.lo[1] + (int)((q / nx) % ny), k = R.lo[2] + (int)(q / ((int64_t)nx * ny));
        dst[fab_index(D, i, j, k)] = src[fab_index(S, i, j, k)];
    }
}

// every cell of box b's region `whole` outside its region `valid` = vals[b]
__global__ __launch_bounds__(kBlock) void k_set_complement(double* data, const IoFab* fabs, const IoRegion* valid,
                                                           const IoRegion* whole, const double* vals) {
    const IoFab& F = fabs[blockIdx.y];
    const IoRegion& V = valid[blockIdx.y];
    const IoRegion& R = whole[blockIdx.y];
    const double val = vals[blockIdx.y];
    const int nx = R.hi[0] - R.lo[0] + 1, ny = R.hi[1] - R.lo[1] + 1, nz = R.hi[2] - R.lo[2] + 1;
    const int64_t n = (int64_t)nx * ny * nz;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlock) {
        const int i = R.lo[0] + (int)(q % nx), j = R.lo[1] + (int)((q / nx) % ny), k = R.lo[2] + (int)(q / ((int64_t)nx * ny));
        const bool inside = i >= V.lo[0] && i <= V.hi[0] && j >= V.lo[1] && j <= V.hi[1] && k >= V.lo[2] && k <= V.hi[2];
        if (!inside) data[fab_index(F, i, j, k)] = val;
    }
}

// dst(comp of direction d) = (f_d(cell) + f_d(cell + e_d)) / 2 over the region of box blockIdx.y, d = blockIdx.z
__global__ __launch_bounds__(kBlock) void k_face_to_cell(double* dst, const IoFab* dfab, FaceSrc f, const IoFab* sfab,
                                                         const IoRegion* reg) {
    const int d = (int)blockIdx.z;
    const IoFab& D = dfab[3 * blockIdx.y + d];
    const IoFab& S = sfab[3 * blockIdx.y + d];
    const IoRegion& R = reg[blockIdx.y];
    const double* src = f.f[d];
    const int nx = R.hi[0] - R.lo[0] + 1, ny = R.hi[1] - R.lo[1] + 1, nz = R.hi[2] - R.lo[2] + 1;
    const int64_t n = (int64_t)nx * ny * nz;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlock) {
        const int i = R.lo[0] + (int)(q % nx), j = R.lo[1] + (int)((q / nx) % ny), k = R.lo[2] + (int)(q / ((int64_t)nx * ny));
        const double lo = src[fab_index(S, i, j, k)];
        const double hi = src[fab_index(S, i + (d == 0), j + (d == 1), k + (d == 2))];
        dst[fab_index(D, i, j, k)] = 0.5 * (lo + hi);
    }
}

inline unsigned blocks_for(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, kIoBlocksPerBox));
}

}  // namespace

void box_minmax(const double* data, const IoFab* fabs_dev, const IoRegion* reg_dev, int nbox, int64_t max_region,
                double* partials_dev, hipStream_t s) {
    hipLaunchKernelGGL(k_box_minmax, dim3(blocks_for(max_region), (unsigned)nbox), dim3(kBlock), 0, s, data, fabs_dev,
                       reg_dev, partials_dev);
}

int io_minmax_blocks(int64_t max_region) { return (int)blocks_for(max_region); }

void copy_region(double* dst, const IoFab* dfab_dev, const double* src, const IoFab* sfab_dev, const IoRegion* reg_dev,
                 int nbox, int64_t max_region, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_region, dim3(blocks_for(max_region), (unsigned)nbox), dim3(kBlock), 0, s, dst, dfab_dev,
                       src, sfab_dev, reg_dev);
}

void face_to_cell(double* dst, const IoFab* dfab_dev, FaceSrc f, const IoFab* sfab_dev, const IoRegion* reg_dev,
                  int nbox, int64_t max_region, hipStream_t s) {
    hipLaunchKernelGGL(k_face_to_cell, dim3(blocks_for(max_region), (unsigned)nbox, 3), dim3(kBlock), 0, s, dst,
                       dfab_dev, f, sfab_dev, reg_dev);
}

void set_complement(double* data, const IoFab* fabs_dev, const IoRegion* valid_dev, const IoRegion* whole_dev,
                    const double* vals_dev, int nbox, int64_t max_region, hipStream_t s) {
    hipLaunchKernelGGL(k_set_complement, dim3(blocks_for(max_region), (unsigned)nbox), dim3(kBlock), 0, s, data,
                       fabs_dev, valid_dev, whole_dev, vals_dev);
}

}  // namespace mgx
