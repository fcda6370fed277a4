// Fused residual kernels of dense MG levels (one FAB covering the domain).
//
// Boundary ghosts are evaluated inline from the interior (the physical
// closure of mllinop_apply_bc_*, AMReX_MLLinOp_K.H:14-115, or the periodic
// image) instead of being written by a ghost-fill pass, so each kernel is a
// single streaming pass:
//   update_residual: sol_out = sol + cor; r = rhs - L(sol_out) with the
//       inhomogeneous level BC; per-block max|r| (and sum r); r stored in FP64
//       or directly as the FP32 right-hand side of the next V-cycle.  Covers
//       MLMGT::oneIter's LocalAdd (AMReX_MLMG.H:1264), computeResidual
//       (AMReX_MLMG.H:1652) and ResNormInf (AMReX_MLMG.H:1808).
//   resid_restrict: crse = average_down(res - L(cor)) with homogeneous BC:
//       computeResOfCorrection + restriction of mgVcycle (AMReX_MLMG.H:1330-1339)
//       without storing the fine residual of the correction.
//   update_residual with transpose: the same pass with r = rhs - L^T(sol_out),
//       homogeneous BC (the fine residual of the adjoint solve; no reference
//       counterpart), and grad_coeffs: dJ/da, dJ/db of a functional through the
//       solve, given the adjoint.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mgx_device.H"
#include "mgx_fused.H"
#include "mgx_march.H"
#include "mgx_transpose.H"

namespace mgx {
namespace {

constexpr int kFB = 256;

__device__ __forceinline__ int64_t didx(const BoxDev& B, int i, int j, int k) {
    return B.off + (i + 1) + B.sy * (j + 1) + B.sz * (k + 1);
}
// cell (i, j, k) of box B's ghost-grown FAB (didx when B starts at 0)
__device__ __forceinline__ int64_t bidx(const BoxDev& B, int i, int j, int k) {
    return B.off + (i - B.lo[0] + 1) + B.sy * (j - B.lo[1] + 1) + B.sz * (k - B.lo[2] + 1);
}
// box of the level holding the real cell (i, j, k)
__device__ __forceinline__ const BoxDev& box_at(const LevelDev& L, int i, int j, int k) {
    if (L.nbox == 1) return L.box[0];
    return L.box[L.box_of[L.slab_of[0][i] + L.nslab[0] * (L.slab_of[1][j] + L.nslab[1] * L.slab_of[2][k])]];
}

__device__ __forceinline__ double nanmax_d(double a, double b) { return (b > a || b != b) ? b : a; }

// Value of a field at (i,j,k) where at most one coordinate may lie one cell
// outside the domain; f(i,j,k) returns the value of a real cell.
template <class T, class F>
__device__ __forceinline__ T value_bc(const LevelDev& L, const BoxDev& B, F f, int i, int j, int k,
                                      const double* bcv, int inhomog) {
    int p[3] = {i, j, k};
    int d = -1, side = 0;
    for (int e = 0; e < 3; ++e) {
        if (p[e] < 0) { d = e; side = 0; }
        else if (p[e] >= L.n[e]) { d = e; side = 1; }
    }
    if (d < 0) return f(i, j, k);
    if (L.periodic[d]) {
        p[d] = side == 0 ? p[d] + L.n[d] : p[d] - L.n[d];
        return f(p[0], p[1], p[2]);
    }
    const int fc = d + 3 * side;
    const int s = side == 0 ? 1 : -1;
    auto at = [&](int m) {
        int q[3] = {p[0], p[1], p[2]};
        q[d] += m * s;
        return f(q[0], q[1], q[2]);
    };
    const int bct = B.bct[fc];
    if (bct == BC_NEUMANN) return at(1);
    if (bct == BC_REFLECT_ODD) return -at(1);
    const int NX = B.nxo[fc];
    T tmp = T(0.0);
    for (int m = 1; m < NX; ++m) tmp += at(m) * T(B.bcc[fc][m]);
    if (inhomog) tmp += static_cast<T>(bcv[bidx(B, i, j, k)]) * T(B.bcc[fc][0]);
    return tmp;
}

// L(x) at a real cell from its six neighbours (mlabeclap_adotx,
// AMReX_MLABecLap_3D_K.H:9-28; mlpoisson_adotx, AMReX_MLPoisson_3D_K.H:9-16).
template <class T>
__device__ __forceinline__ T apply_dense(const BoxDev& B, const OpCoef<T>& op, int n0, int n1, int i, int j, int k,
                                         T xc, T xm, T xp, T ym, T yp, T zm, T zp) {
    if (op.kind == 0) {
        const T dhx = op.dha[0], dhy = op.dha[1], dhz = op.dha[2];
        const int64_t fxi = B.foff[0] + i + (int64_t)(n0 + 1) * (j + (int64_t)n1 * k);
        const int64_t fyi = B.foff[1] + i + (int64_t)n0 * (j + (int64_t)(n1 + 1) * k);
        const int64_t fzi = B.foff[2] + i + (int64_t)n0 * (j + (int64_t)n1 * k);
        return op.alpha * op.a[didx(B, i, j, k)] * xc
            - dhx * (op.b[0][fxi + 1] * (xp - xc) - op.b[0][fxi] * (xc - xm))
            - dhy * (op.b[1][fyi + n0] * (yp - xc) - op.b[1][fyi] * (xc - ym))
            - dhz * (op.b[2][fzi + (int64_t)n0 * n1] * (zp - xc) - op.b[2][fzi] * (xc - zm));
    }
    const T dhx = op.dpo[0], dhy = op.dpo[1], dhz = op.dpo[2];
    return dhx * (xm - T(2.0) * xc + xp) + dhy * (ym - T(2.0) * xc + yp) + dhz * (zm - T(2.0) * xc + zp);
}

template <int NT>
__device__ __forceinline__ void block_max_sum(double& vmax, double& vsum, double* lds) {
    for (int o = 32; o > 0; o >>= 1) {
        vmax = nanmax_d(vmax, __shfl_xor(vmax, o, 64));
        vsum += __shfl_xor(vsum, o, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        lds[2 * wid] = vmax;
        lds[2 * wid + 1] = vsum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NT / 64; ++w) {
            vmax = nanmax_d(vmax, lds[2 * w]);
            vsum += lds[2 * w + 1];
        }
    }
}

template <class TC, class TR, bool TRANS>
__global__ __launch_bounds__(kFB) void k_update_residual(LevelDev L, OpCoef<double> op, const double* __restrict__ sol,
                                                         const TC* __restrict__ cor, double* __restrict__ sol_out,
                                                         const double* __restrict__ rhs, const double* __restrict__ bcv,
                                                         int inhomog, TR* __restrict__ res, double* __restrict__ partials) {
    __shared__ double lds[2 * kFB / 64];
    const BoxDev& B = L.box[0];
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    const int ntx = (n0 + kTileX - 1) / kTileX, nty = (n1 + kTileY - 1) / kTileY;
    const int b = blockIdx.x;
    const int i = (b % ntx) * kTileX + (int)(threadIdx.x % kTileX);
    const int j = ((b / ntx) % nty) * kTileY + (int)(threadIdx.x / kTileX);
    const int k0 = (b / (ntx * nty)) * kTileZ, k1 = min(k0 + kTileZ, n2);
    auto x = [&](int ii, int jj, int kk) -> double {
        const int64_t c = didx(B, ii, jj, kk);
        return cor ? sol[c] + static_cast<double>(cor[c]) : sol[c];
    };
    double vmax = 0.0, vsum = 0.0;
    if (i < n0 && j < n1) {
        for (int k = k0; k < k1; ++k) {
            const int64_t c = didx(B, i, j, k);
            const double xc = x(i, j, k);
            double y;
            if constexpr (TRANS) {
                y = lt::apply_dense_T<double>(L, B, op, x, i, j, k);
            } else {
                const double xm = value_bc<double>(L, B, x, i - 1, j, k, bcv, inhomog);
                const double xp = value_bc<double>(L, B, x, i + 1, j, k, bcv, inhomog);
                const double ym = value_bc<double>(L, B, x, i, j - 1, k, bcv, inhomog);
                const double yp = value_bc<double>(L, B, x, i, j + 1, k, bcv, inhomog);
                const double zm = value_bc<double>(L, B, x, i, j, k - 1, bcv, inhomog);
                const double zp = value_bc<double>(L, B, x, i, j, k + 1, bcv, inhomog);
                y = apply_dense<double>(B, op, n0, n1, i, j, k, xc, xm, xp, ym, yp, zm, zp);
            }
            const double r = rhs[c] + (-1.0) * y;
            if (sol_out) sol_out[c] = xc;
            res[c] = static_cast<TR>(r);
            vmax = nanmax_d(vmax, fabs(r));
            vsum += r;
        }
    }
    block_max_sum<kFB>(vmax, vsum, lds);
    if (threadIdx.x == 0) {
        partials[2 * (int64_t)blockIdx.x] = vmax;
        partials[2 * (int64_t)blockIdx.x + 1] = vsum;
    }
}

// Reduce the (max, sum) pairs of the blocks in fixed order: out[0] = max, out[1] = sum.
__global__ __launch_bounds__(1024) void k_reduce_pairs(const double* __restrict__ partials, int64_t n,
                                                       double* __restrict__ out) {
    __shared__ double lds[2 * 1024 / 64];
    double vmax = 0.0, vsum = 0.0;
    for (int64_t q = threadIdx.x; q < n; q += 1024) {
        vmax = nanmax_d(vmax, partials[2 * q]);
        vsum += partials[2 * q + 1];
    }
    block_max_sum<1024>(vmax, vsum, lds);
    if (threadIdx.x == 0) {
        out[0] = vmax;
        out[1] = vsum;
    }
}

template <class T>
__global__ __launch_bounds__(kFB) void k_resid_restrict(LevelDev Lf, OpCoef<T> op, const T* __restrict__ cor,
                                                        const T* __restrict__ res, LevelDev Lc, T* __restrict__ crse,
                                                        int tz) {
    const BoxDev& F = Lf.box[0];
    const BoxDev& C = Lc.box[0];
    const int n0 = Lf.n[0], n1 = Lf.n[1];
    const int m0 = Lc.n[0], m1 = Lc.n[1], m2 = Lc.n[2];
    const int ntx = (m0 + kTileX - 1) / kTileX, nty = (m1 + kTileY - 1) / kTileY;
    const int b = blockIdx.x;
    const int I = (b % ntx) * kTileX + (int)(threadIdx.x % kTileX);
    const int J = ((b / ntx) % nty) * kTileY + (int)(threadIdx.x / kTileX);
    const int K0 = (b / (ntx * nty)) * tz, K1 = min(K0 + tz, m2);
    if (I >= m0 || J >= m1) return;
    auto x = [&](int ii, int jj, int kk) -> T { return cor[didx(F, ii, jj, kk)]; };
    for (int K = K0; K < K1; ++K) {
        T c = 0;
        for (int kr = 0; kr < 2; ++kr)
            for (int jr = 0; jr < 2; ++jr)
                for (int ir = 0; ir < 2; ++ir) {
                    const int i = 2 * I + ir, j = 2 * J + jr, k = 2 * K + kr;
                    const T xc = x(i, j, k);
                    const T xm = value_bc<T>(Lf, F, x, i - 1, j, k, nullptr, 0);
                    const T xp = value_bc<T>(Lf, F, x, i + 1, j, k, nullptr, 0);
                    const T ym = value_bc<T>(Lf, F, x, i, j - 1, k, nullptr, 0);
                    const T yp = value_bc<T>(Lf, F, x, i, j + 1, k, nullptr, 0);
                    const T zm = value_bc<T>(Lf, F, x, i, j, k - 1, nullptr, 0);
                    const T zp = value_bc<T>(Lf, F, x, i, j, k + 1, nullptr, 0);
                    const T y = apply_dense<T>(F, op, n0, n1, i, j, k, xc, xm, xp, ym, yp, zm, zp);
                    c += res[didx(F, i, j, k)] + T(-1.0) * y;
                }
        crse[didx(C, I, J, K)] = (T(1.0) / T(8)) * c;
    }
}

// dJ/db_d on the faces of a user face FAB set (nodal boxes, ghost ung), from
// the adjoint lam and the solution phi in the level layout (dense or boxes; a
// boundary ghost closes with the box holding the cell beside it).  Row p of
// L(phi) holds w*(phi_p - phi_side) for each side, w = dh*b_face, so each face
// takes -lam_p*dh*(phi_p - phi_side) from each real cell p beside it; phi_side
// is the neighbour, the periodic image or the boundary ghost (value_bc, with
// the inhomogeneous Dirichlet data when given).
__global__ __launch_bounds__(kFB) void k_grad_b(LevelDev L, OpCoef<double> op, const double* __restrict__ lam,
                                                const double* __restrict__ phi, const double* __restrict__ bcv,
                                                int inhomog, const UserBox* __restrict__ ub, int ung, int d,
                                                double* __restrict__ out, int accumulate) {
    const UserBox& U = ub[blockIdx.y];
    const int nx = U.hi[0] - U.lo[0] + 1, ny = U.hi[1] - U.lo[1] + 1, nz = U.hi[2] - U.lo[2] + 1;
    const int64_t q = (int64_t)blockIdx.x * kFB + threadIdx.x;
    if (q >= (int64_t)nx * ny * nz) return;
    const int f[3] = {U.lo[0] + (int)(q % nx), U.lo[1] + (int)((q / nx) % ny), U.lo[2] + (int)(q / ((int64_t)nx * ny))};
    const int64_t ux = nx + 2 * ung, uy = ny + 2 * ung;
    const int64_t u = U.off + (f[0] - U.lo[0] + ung) + ux * ((f[1] - U.lo[1] + ung) + uy * (int64_t)(f[2] - U.lo[2] + ung));
    auto ph = [&](int i, int j, int k) -> double { return phi[bidx(box_at(L, i, j, k), i, j, k)]; };
    const double dh = op.dha[d];
    int c[3] = {f[0], f[1], f[2]};
    c[d] -= 1;  // cell below the face
    double g = 0.0;
    if (c[d] >= 0) {
        const BoxDev& B = box_at(L, c[0], c[1], c[2]);
        const double side = value_bc<double>(L, B, ph, f[0], f[1], f[2], bcv, inhomog);
        g += -lam[bidx(B, c[0], c[1], c[2])] * dh * (ph(c[0], c[1], c[2]) - side);
    }
    if (f[d] < L.n[d]) {
        const BoxDev& B = box_at(L, f[0], f[1], f[2]);
        const double side = value_bc<double>(L, B, ph, c[0], c[1], c[2], bcv, inhomog);
        g += -lam[bidx(B, f[0], f[1], f[2])] * dh * (ph(f[0], f[1], f[2]) - side);
    }
    out[u] = accumulate ? out[u] + g : g;
}

// dJ/da = -alpha * lam * phi on the cells of a user cell FAB set.
__global__ __launch_bounds__(kFB) void k_grad_a(LevelDev L, double alpha, const double* __restrict__ lam,
                                                const double* __restrict__ phi, const UserBox* __restrict__ ub, int ung,
                                                double* __restrict__ out, int accumulate) {
    const UserBox& U = ub[blockIdx.y];
    const int nx = U.hi[0] - U.lo[0] + 1, ny = U.hi[1] - U.lo[1] + 1, nz = U.hi[2] - U.lo[2] + 1;
    const int64_t q = (int64_t)blockIdx.x * kFB + threadIdx.x;
    if (q >= (int64_t)nx * ny * nz) return;
    const int i = U.lo[0] + (int)(q % nx), j = U.lo[1] + (int)((q / nx) % ny), k = U.lo[2] + (int)(q / ((int64_t)nx * ny));
    const int64_t ux = nx + 2 * ung, uy = ny + 2 * ung;
    const int64_t u = U.off + (i - U.lo[0] + ung) + ux * ((j - U.lo[1] + ung) + uy * (int64_t)(k - U.lo[2] + ung));
    const int64_t c = bidx(box_at(L, i, j, k), i, j, k);
    const double g = -alpha * lam[c] * phi[c];
    out[u] = accumulate ? out[u] + g : g;
}

inline unsigned ntiles(const LevelDev& L) {
    return (unsigned)(((L.n[0] + kTileX - 1) / kTileX) * ((L.n[1] + kTileY - 1) / kTileY) *
                      ((L.n[2] + kTileZ - 1) / kTileZ));
}

}  // namespace

int64_t update_residual_partials(const LevelDev& L) {
    // the Poisson update runs more work-groups than the ABecLaplacian one
    return std::max<int64_t>(2 * (int64_t)ntiles(L), march_ok(L) ? std::max(march_partials(L, 0), march_partials(L, 1)) : 0);
}

template <class TC, class TR>
void update_residual(const LevelDev& L, const OpCoef<double>& op, const double* sol, const TC* cor, double* sol_out,
                     const double* rhs, const double* bcv, int inhomog, TR* res, double* partials, double* out,
                     int transpose, hipStream_t s) {
    if (!transpose && march_ok(L)) {
        march_update_residual<TC, TR>(L, op, sol, cor, sol_out, rhs, bcv, inhomog, res, partials, s);
        hipLaunchKernelGGL(k_reduce_pairs, dim3(1), dim3(1024), 0, s, partials, march_partials(L, op.kind) / 2, out);
        return;
    }
    const unsigned nb = ntiles(L);
    if (transpose)
        hipLaunchKernelGGL((k_update_residual<TC, TR, true>), dim3(nb), dim3(kFB), 0, s, L, op, sol, cor, sol_out, rhs,
                           nullptr, 0, res, partials);
    else
        hipLaunchKernelGGL((k_update_residual<TC, TR, false>), dim3(nb), dim3(kFB), 0, s, L, op, sol, cor, sol_out, rhs,
                           bcv, inhomog, res, partials);
    hipLaunchKernelGGL(k_reduce_pairs, dim3(1), dim3(1024), 0, s, partials, (int64_t)nb, out);
}

void grad_b(const LevelDev& L, const OpCoef<double>& op, const double* lam, const double* phi, const double* bcv,
            int inhomog, const UserBox* ub, int nbox, int64_t max_faces, int ung, int d, double* out, int accumulate,
            hipStream_t s) {
    if (nbox <= 0 || max_faces <= 0) return;
    hipLaunchKernelGGL(k_grad_b, dim3((unsigned)((max_faces + kFB - 1) / kFB), (unsigned)nbox), dim3(kFB), 0, s, L, op,
                       lam, phi, bcv, inhomog, ub, ung, d, out, accumulate);
}

void grad_a(const LevelDev& L, double alpha, const double* lam, const double* phi, const UserBox* ub, int nbox,
            int64_t max_cells, int ung, double* out, int accumulate, hipStream_t s) {
    if (nbox <= 0 || max_cells <= 0) return;
    hipLaunchKernelGGL(k_grad_a, dim3((unsigned)((max_cells + kFB - 1) / kFB), (unsigned)nbox), dim3(kFB), 0, s, L,
                       alpha, lam, phi, ub, ung, out, accumulate);
}

template <class T>
void resid_restrict(const LevelDev& Lf, const OpCoef<T>& op, const T* cor, const T* res, const LevelDev& Lc, T* crse,
                    hipStream_t s) {
    if (march_ok(Lf) && ((Lf.n[0] | Lf.n[1] | Lf.n[2]) & 1) == 0) {
        march_resid_restrict<T>(Lf, op, cor, res, Lc, crse, s);
        return;
    }
    // one coarse plane per thread: small levels are latency-bound, not bandwidth-bound
    const int tz = 1;
    const unsigned nb = (unsigned)(((Lc.n[0] + kTileX - 1) / kTileX) * ((Lc.n[1] + kTileY - 1) / kTileY) *
                                   ((Lc.n[2] + tz - 1) / tz));
    hipLaunchKernelGGL(k_resid_restrict<T>, dim3(nb), dim3(kFB), 0, s, Lf, op, cor, res, Lc, crse, tz);
}

template void update_residual<double, double>(const LevelDev&, const OpCoef<double>&, const double*, const double*,
                                              double*, const double*, const double*, int, double*, double*, double*,
                                              int, hipStream_t);
template void update_residual<float, float>(const LevelDev&, const OpCoef<double>&, const double*, const float*,
                                            double*, const double*, const double*, int, float*, double*, double*,
                                            int, hipStream_t);
template void update_residual<float, double>(const LevelDev&, const OpCoef<double>&, const double*, const float*,
                                             double*, const double*, const double*, int, double*, double*, double*,
                                             int, hipStream_t);
template void resid_restrict<double>(const LevelDev&, const OpCoef<double>&, const double*, const double*,
                                     const LevelDev&, double*, hipStream_t);
template void resid_restrict<float>(const LevelDev&, const OpCoef<float>&, const float*, const float*,
                                    const LevelDev&, float*, hipStream_t);

}  // namespace mgx
