// Vector kernels of GMRES on one MG level (gfx950): batched dot products over
// the valid cells and fused linear combinations over whole allocations.
//
// The reference's GMRES computes the Gram-Schmidt projections of a new Krylov
// vector as independent dot products and then applies them one increment at
// a time (AMReX_GMRES.H:354-375).  Here the k projections are one pass that
// reads the new vector once (k_multi_dot), and the k increments one pass that
// writes it once (k_multi_axpy), each element updated in the reference's order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mgx_krylov.H"

namespace mgx {

namespace {

constexpr int kBlock = 256;
constexpr int kTX = kTileX, kTY = kTileY, kTZ = kTileZRed;

__device__ __forceinline__ int64_t cidx(const BoxDev& B, int i, int j, int k) {
    return B.off + (i - B.lo[0] + 1) + B.sy * (j - B.lo[1] + 1) + B.sz * (k - B.lo[2] + 1);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// partials[j * nblk + blk] = sum over this block's cells of y * x_j
template <int K>
__global__ __launch_bounds__(kBlock) void k_multi_dot(LevelDev L, const double* y, KrylovSet xs, int k,
                                                      double* partials, int nblk) {
    __shared__ double lds[K][kBlock / 64];
    const BoxDev& B = L.box[blockIdx.y];
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
    const int ntx = (nx + kTX - 1) / kTX, nty = (ny + kTY - 1) / kTY, ntz = (nz + kTZ - 1) / kTZ;
    double acc[K];
#pragma unroll
    for (int j = 0; j < K; ++j) acc[j] = 0.0;
    int b = blockIdx.x;
    if (b < ntx * nty * ntz) {
        const int bx = b % ntx;
        b /= ntx;
        const int by = b % nty, bz = b / nty;
        const int i = B.lo[0] + bx * kTX + (int)(threadIdx.x % kTX);
        const int jj = B.lo[1] + by * kTY + (int)(threadIdx.x / kTX);
        const int k0 = B.lo[2] + bz * kTZ, k1 = min(k0 + kTZ, B.hi[2] + 1);
        if (i <= B.hi[0] && jj <= B.hi[1]) {
            for (int kk = k0; kk < k1; ++kk) {
                const int64_t c = cidx(B, i, jj, kk);
                const double yc = y[c];
#pragma unroll
                for (int j = 0; j < K; ++j)
                    if (j < k) acc[j] += yc * xs.x[j][c];
            }
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (j >= k) break;
        const double v = wave_sum(acc[j]);
        if (lane == 0) lds[j][wid] = v;
    }
    __syncthreads();
    const int blk = (int)(blockIdx.y * gridDim.x + blockIdx.x);
    if ((int)threadIdx.x < k) {
        double s = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) s += lds[threadIdx.x][w];
        partials[(int64_t)threadIdx.x * nblk + blk] = s;
    }
}

// out[j] = sum of partials[j * nblk .. j * nblk + nblk - 1], one block per j
__global__ __launch_bounds__(1024) void k_multi_dot_final(const double* partials, int nblk, double* out) {
    __shared__ double lds[1024 / 64];
    const double* p = partials + (int64_t)blockIdx.x * nblk;
    double v = 0.0;
    for (int q = threadIdx.x; q < nblk; q += 1024) v += p[q];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < 1024 / 64; ++w) s += lds[w];
        out[blockIdx.x] = s;
    }
}

// y = a * y + sum_j c_j x_j, element by element in the order j = 0, 1, ...
__global__ __launch_bounds__(kBlock) void k_multi_axpy(double* y, double a, KrylovSet xs, KrylovCoef cs, int k,
                                                       int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlock) {
        double v = (a == 0.0) ? 0.0 : a * y[q];
        for (int j = 0; j < k; ++j) v += cs.c[j] * xs.x[j][q];
        y[q] = v;
    }
}

__global__ __launch_bounds__(kBlock) void k_scale(double* x, double f, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlock) x[q] *= f;
}

inline dim3 flat(int64_t n) {
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 1 << 16)));
}

// tiles of 64 x 4 cells by kTileZRed planes per box, as the level reductions use
int dot_blocks(const LevelDev& L) { return std::max(L.mtiles[1], 1); }

}  // namespace

int64_t multi_dot_partials(const LevelDev& L) { return (int64_t)dot_blocks(L) * L.nbox * kMaxKrylov; }

void multi_dot(const LevelDev& L, const double* y, const KrylovSet& xs, int k, double* partials, double* out_dev,
               hipStream_t s) {
    if (k < 1 || k > kMaxKrylov) return;
    const int nbx = dot_blocks(L);
    const int nblk = nbx * L.nbox;
    const dim3 grid((unsigned)nbx, (unsigned)L.nbox);
    if (k <= 4) hipLaunchKernelGGL(k_multi_dot<4>, grid, dim3(kBlock), 0, s, L, y, xs, k, partials, nblk);
    else if (k <= 8) hipLaunchKernelGGL(k_multi_dot<8>, grid, dim3(kBlock), 0, s, L, y, xs, k, partials, nblk);
    else if (k <= 16) hipLaunchKernelGGL(k_multi_dot<16>, grid, dim3(kBlock), 0, s, L, y, xs, k, partials, nblk);
    else hipLaunchKernelGGL(k_multi_dot<kMaxKrylov>, grid, dim3(kBlock), 0, s, L, y, xs, k, partials, nblk);
    hipLaunchKernelGGL(k_multi_dot_final, dim3((unsigned)k), dim3(1024), 0, s, partials, nblk, out_dev);
}

void multi_axpy(double* y, double a, const KrylovSet& xs, const KrylovCoef& cs, int k, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_multi_axpy, flat(n), dim3(kBlock), 0, s, y, a, xs, cs, k, n);
}

void scale(double* x, double f, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_scale, flat(n), dim3(kBlock), 0, s, x, f, n);
}

}  // namespace mgx
