// MultiFab reductions (FabArray::sum / FabArray::norminf over valid cells,
// AMReX_FabArray.H:3539-3627): a first pass of a fixed number of work-groups
// striding over every box, each writing its partial sum and maximum, and a
// one-work-group second pass over the partials.
#include "mgx_mfreduce.H"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <string>

// Partials of the reductions that allocate nothing: one array per GPU, held
// until the process ends.  Host code never passes this name to a kernel or a
// copy (on the host it is only a stand-in); it uses the device address
// mgx::mf_reduce_scratch() looks up.
static __device__ __align__(256) double scratch[mgx::kMfReducePartials];

namespace mgx {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ void block_sum_max(double& s, double& m, double* lds) {
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        m = fmax(m, __shfl_xor(m, o, 64));
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) {
        lds[w] = s;
        lds[8 + w] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
            s += lds[q];
            m = fmax(m, lds[8 + q]);
        }
    }
}

// partials layout: [c][0] sums of the G work-groups, then [c][1] their maxima
template <class T>
__global__ __launch_bounds__(kThreads) void k_mf_partials(const T* __restrict__ data, const UserBox* __restrict__ boxes,
                                                          int nbox, int ng, int comp0, int ncomp,
                                                          double* __restrict__ partials) {
    __shared__ double lds[16];
    const int G = gridDim.x;
    const int64_t stride = (int64_t)G * kThreads;
    const int64_t t0 = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    for (int c = 0; c < ncomp; ++c) {
        double s = 0.0, m = 0.0;
        for (int b = 0; b < nbox; ++b) {
            const UserBox B = boxes[b];
            const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
            const int64_t gx = nx + 2 * ng, gy = ny + 2 * ng, gz = nz + 2 * ng;
            const int64_t ncell = (int64_t)nx * ny * nz;
            const T* f = data + B.off + (int64_t)(comp0 + c) * gx * gy * gz;
            for (int64_t q = t0; q < ncell; q += stride) {
                const int64_t row = q / nx;
                const int i = (int)(q - row * nx), j = (int)(row % ny), k = (int)(row / ny);
                const double v = (double)f[(i + ng) + gx * ((j + ng) + gy * (int64_t)(k + ng))];
                s += v;
                m = fmax(m, fabs(v));
            }
        }
        block_sum_max(s, m, lds);
        if (threadIdx.x == 0) {
            partials[(2 * c) * G + blockIdx.x] = s;
            partials[(2 * c + 1) * G + blockIdx.x] = m;
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_mf_final(const double* __restrict__ partials, int G, int ncomp,
                                                       double* __restrict__ out) {
    __shared__ double lds[16];
    for (int c = 0; c < ncomp; ++c) {
        double s = 0.0, m = 0.0;
        for (int q = threadIdx.x; q < G; q += kThreads) {
            s += partials[(2 * c) * G + q];
            m = fmax(m, partials[(2 * c + 1) * G + q]);
        }
        block_sum_max(s, m, lds);
        if (threadIdx.x == 0) {
            out[2 * c] = s;
            out[2 * c + 1] = m;
        }
    }
}

}  // namespace

template <class T>
void mf_reduce(const T* data, const UserBox* boxes_dev, int nbox, int ngrow, int64_t nvalid, int comp0, int ncomp,
               double* partials_dev, double* out_dev, hipStream_t s) {
    if (ncomp < 1 || ncomp > kMfReduceMaxComp)
        throw std::runtime_error("MultiFab reduction: 1 to " + std::to_string(kMfReduceMaxComp) +
                                 " components per call, got " + std::to_string(ncomp));
    if (nbox < 1) throw std::runtime_error("MultiFab reduction: no boxes");
    const int G = (int)std::min<int64_t>(kMfReduceBlocks, std::max<int64_t>(1, (nvalid + 8 * kThreads - 1) / (8 * kThreads)));
    hipLaunchKernelGGL(k_mf_partials<T>, dim3(G), dim3(kThreads), 0, s, data, boxes_dev, nbox, ngrow, comp0, ncomp,
                       partials_dev);
    hipLaunchKernelGGL(k_mf_final, dim3(1), dim3(kThreads), 0, s, partials_dev, G, ncomp, out_dev);
}

double* mf_reduce_scratch() {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static double* addr[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) throw std::runtime_error(std::string("reduction scratch: hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("reduction scratch: device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("reduction scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
        addr[dev] = static_cast<double*>(p);
    }
    return addr[dev];
}

template void mf_reduce<double>(const double*, const UserBox*, int, int, int64_t, int, int, double*, double*,
                                hipStream_t);
template void mf_reduce<float>(const float*, const UserBox*, int, int, int64_t, int, int, double*, double*,
                               hipStream_t);

}  // namespace mgx
