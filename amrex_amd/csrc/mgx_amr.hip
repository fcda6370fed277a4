// Coarse/fine kernels of the two-level AMR composite solve (MLPoisson or
// MLABecLaplacian, ratio 2):
// a dense coarse level (one FAB covering the periodic domain) and a fine level
// of boxes anywhere in that periodic domain, stored box by box (box-batched
// launches: blockIdx.y selects the box).  A fine position beyond a domain face
// stands for its periodic image.
//
//   fine_fill      face ghosts of the fine level: FillBoundary from the fine
//                  box that covers the ghost, the physical face's closure
//                  beyond a domain face, else the Dirichlet closure at a
//                  coarse/fine face (mllinop_apply_bc_*, AMReX_MLLinOp_K.H:
//                  14-115, bcl = 0.5*ratio*dx_f, MLMGBndry.H:145) with the
//                  boundary value interpolated from the coarse level
//   fine_gsrb      mlpoisson_gsrb (AMReX_MLPoisson_3D_K.H:155-196) or abec_gsrb
//                  (AMReX_MLABecLap_3D_K.H:210-264) with the undrrelxr weight
//                  where the face ghost is not covered
//   fine_bndry     interpbndrydata_{x,y,z}_o3 (AMReX_InterpBndryData_3D_K.H:
//                  23-118) with masks from the fine box layout
//   fine_avgdown   average_down ratio 2 (AMReX_MultiFabUtil_3D_C.H:377-396)
//   fine_avgdown_faces  average_down_faces of the fine b coefficients
//                  (averageDownCoeffsToCoarseAmrLevel, AMReX_MLABecLaplacian.H:694-712)
//   fine_interp    mlmg_lin_cc_interp_r2 (AMReX_MLMG_3D_K.H:9-38)
//   reflux         YAFluxRegister CrseAdd/FineAdd/Reflux for the operator's flux
//                  (AMReX_YAFluxRegister_3D_K.H; MLCellLinOp.H:1273-1344)
//   masked_max     max |x| over coarse cells not covered by the fine level
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mgx_amr.H"

namespace mgx {
namespace {

constexpr int kAB = 256;

__device__ __forceinline__ int64_t cix(const BoxDev& B, int i, int j, int k) {
    return B.off + (i - B.lo[0] + 1) + B.sy * (j - B.lo[1] + 1) + B.sz * (k - B.lo[2] + 1);
}
// face (i, j, k) along d of box B: the box's face FAB, x fastest, no ghosts
__device__ __forceinline__ int64_t fix(const BoxDev& B, int d, int i, int j, int k) {
    const int64_t fx = B.hi[0] - B.lo[0] + 1 + (d == 0);
    const int64_t fy = B.hi[1] - B.lo[1] + 1 + (d == 1);
    return B.foff[d] + (i - B.lo[0]) + fx * ((j - B.lo[1]) + fy * (k - B.lo[2]));
}
__device__ __forceinline__ int wrapi(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}
// fine box covering the periodic image of a fine position, -1 when none does,
// -2 beyond a non-periodic domain face
__device__ __forceinline__ int fine_owner(const LevelDev& L, int i, int j, int k) {
    if ((!L.periodic[0] && (i < 0 || i >= L.n[0])) || (!L.periodic[1] && (j < 0 || j >= L.n[1])) ||
        (!L.periodic[2] && (k < 0 || k >= L.n[2])))
        return -2;
    i = wrapi(i, L.n[0]);
    j = wrapi(j, L.n[1]);
    k = wrapi(k, L.n[2]);
    return L.box_of[L.slab_of[0][i] + L.nslab[0] * (L.slab_of[1][j] + L.nslab[1] * L.slab_of[2][k])];
}
// coarse value with periodic wrap (dense coarse level: box 0 covers the
// domain); 0 beyond a non-periodic face, as interpCorrection's zeroed copy of
// valid cells and periodic images reads there (AMReX_MLMG.H:1719-1750)
template <class T>
__device__ __forceinline__ T cval(const LevelDev& Lc, const T* c, int i, int j, int k) {
    if ((!Lc.periodic[0] && (i < 0 || i >= Lc.n[0])) || (!Lc.periodic[1] && (j < 0 || j >= Lc.n[1])) ||
        (!Lc.periodic[2] && (k < 0 || k >= Lc.n[2])))
        return T(0);
    const BoxDev& C = Lc.box[0];
    return c[cix(C, wrapi(i, Lc.n[0]), wrapi(j, Lc.n[1]), wrapi(k, Lc.n[2]))];
}
__device__ __forceinline__ int coarsen2(int i) { return i >= 0 ? i / 2 : -((-i + 1) / 2); }

// face ghost q of face f of box B -> (i, j, k)
__device__ __forceinline__ void ghost_of(const BoxDev& B, int f, int q, int* p) {
    const int d = f % 3, side = f / 3;
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    const int n1 = B.hi[t1] - B.lo[t1] + 1;
    p[d] = side == 0 ? B.lo[d] - 1 : B.hi[d] + 1;
    p[t1] = B.lo[t1] + q % n1;
    p[t2] = B.lo[t2] + q / n1;
}
__device__ __forceinline__ int face_cells(const BoxDev& B, int f) {
    const int d = f % 3;
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    return (B.hi[t1] - B.lo[t1] + 1) * (B.hi[t2] - B.lo[t2] + 1);
}
__device__ __forceinline__ int nx_of(const BoxDev& B, int f, int maxorder) {
    const int d = f % 3;
    return min(B.hi[d] - B.lo[d] + 2, maxorder);  // min(box length + 1, maxorder)
}
// undrrelxr of a box face cell whose ghost lies beyond a physical face
// (AMReX_MLLinOp_K.H:329-406)
__device__ __forceinline__ double phys_cf(const CfCoef& cc, int f, int NX) {
    switch (cc.bct[f]) {
        case BC_DIRICHLET: return cc.p[NX][1];
        case BC_NEUMANN: return 1.0;
        case BC_REFLECT_ODD: return 1.0;
        default: return 0.0;
    }
}

template <class T>
__global__ __launch_bounds__(kAB) void k_fine_fill(LevelDev L, T* phi, const T* bcv, CfCoef cc, int maxorder) {
    const BoxDev& B = L.box[blockIdx.y];
    const int f = blockIdx.z;
    const int nq = face_cells(B, f);
    const int d = f % 3;
    const int s = f < 3 ? 1 : -1;
    const int NX = nx_of(B, f, maxorder);
    for (int q = blockIdx.x * kAB + threadIdx.x; q < nq; q += gridDim.x * kAB) {
        int p[3];
        ghost_of(B, f, q, p);
        const int64_t g = cix(B, p[0], p[1], p[2]);
        const int o = fine_owner(L, p[0], p[1], p[2]);
        if (o >= 0) {
            phi[g] = phi[cix(L.box[o], wrapi(p[0], L.n[0]), wrapi(p[1], L.n[1]), wrapi(p[2], L.n[2]))];
            continue;
        }
        const int64_t st = (d == 0) ? 1 : (d == 1 ? B.sy : B.sz);
        // beyond a physical face: that face's closure (mllinop_apply_bc_*)
        if (o == -2 && cc.bct[f] == BC_NEUMANN) { phi[g] = phi[g + s * st]; continue; }
        if (o == -2 && cc.bct[f] == BC_REFLECT_ODD) { phi[g] = -phi[g + s * st]; continue; }
        const double* w = o == -2 ? cc.p[NX] : cc.c[NX];
        T tmp = T(0.0);
        for (int m = 1; m < NX; ++m) tmp += phi[g + s * m * st] * T(w[m]);
        if (bcv) tmp += bcv[g] * T(w[0]);
        phi[g] = tmp;
    }
}

template <class T>
__global__ __launch_bounds__(kAB) void k_fine_gsrb(LevelDev L, OpCoef<T> op, T* phi, const T* rhs, int rb, CfCoef cc,
                                                   int maxorder) {
    const BoxDev& B = L.box[blockIdx.y];
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
    constexpr T omega = T(1.15);
    for (int64_t q = (int64_t)blockIdx.x * kAB + threadIdx.x; q < (int64_t)nx * ny * nz; q += (int64_t)gridDim.x * kAB) {
        const int i = B.lo[0] + (int)(q % nx), j = B.lo[1] + (int)((q / nx) % ny), k = B.lo[2] + (int)(q / ((int64_t)nx * ny));
        if (((i + j + k + rb) & 1) != 0) continue;
        const int p[3] = {i, j, k};
        T cf[6];
        for (int f = 0; f < 6; ++f) {
            const int dd = f % 3;
            const bool at = f < 3 ? p[dd] == B.lo[dd] : p[dd] == B.hi[dd];
            int g[3] = {i, j, k};
            g[dd] += f < 3 ? -1 : 1;
            const int o = at ? fine_owner(L, g[0], g[1], g[2]) : 0;
            cf[f] = o == -1 ? T(cc.c[nx_of(B, f, maxorder)][1]) : (o == -2 ? T(phys_cf(cc, f, nx_of(B, f, maxorder))) : T(0.0));
        }
        const int64_t c = cix(B, i, j, k);
        const int64_t sy = B.sy, sz = B.sz;
        if (op.kind == 0) {
            // abec_gsrb (AMReX_MLABecLap_3D_K.H:210-264) with the box's coefficients
            const T dhx = op.dhs[0], dhy = op.dhs[1], dhz = op.dhs[2];
            const T bxm = op.b[0][fix(B, 0, i, j, k)], bxp = op.b[0][fix(B, 0, i + 1, j, k)];
            const T bym = op.b[1][fix(B, 1, i, j, k)], byp = op.b[1][fix(B, 1, i, j + 1, k)];
            const T bzm = op.b[2][fix(B, 2, i, j, k)], bzp = op.b[2][fix(B, 2, i, j, k + 1)];
            const T gamma = op.alpha * op.a[c] + dhx * (bxm + bxp) + dhy * (bym + byp) + dhz * (bzm + bzp);
            const T g_m_d = gamma - (dhx * (bxm * cf[0] + bxp * cf[3]) + dhy * (bym * cf[1] + byp * cf[4]) +
                                     dhz * (bzm * cf[2] + bzp * cf[5]));
            const T rho = dhx * (bxm * phi[c - 1] + bxp * phi[c + 1]) + dhy * (bym * phi[c - sy] + byp * phi[c + sy]) +
                          dhz * (bzm * phi[c - sz] + bzp * phi[c + sz]);
            const T res = rhs[c] - (gamma * phi[c] - rho);
            phi[c] = phi[c] + omega / g_m_d * res;
            continue;
        }
        const T dhx = op.dpo[0], dhy = op.dpo[1], dhz = op.dpo[2];
        const T gamma = T(-2.) * (dhx + dhy + dhz);
        const T g_m_d = gamma + dhx * (cf[0] + cf[3]) + dhy * (cf[1] + cf[4]) + dhz * (cf[2] + cf[5]);
        const T res = rhs[c] - gamma * phi[c] - dhx * (phi[c - 1] + phi[c + 1]) - dhy * (phi[c - sy] + phi[c + sy])
                    - dhz * (phi[c - sz] + phi[c + sz]);
        phi[c] = phi[c] + omega / g_m_d * res;
    }
}

// One GSRB colour with the face ghosts formed as it goes: what fine_fill with
// homogeneous boundary data followed by k_fine_gsrb computes, in one pass over
// the colour's cells only.  With closure order at most 3 a ghost reads the
// cell itself (its value before the update) and the next cell inward, and a
// covered ghost reads another box's cell; both are of the other colour (the
// fine domain has even extents), so nothing this colour writes is read.
template <class T>
__global__ __launch_bounds__(kAB) void k_fine_gsrb_h(LevelDev L, OpCoef<T> op, T* phi, const T* rhs, int rb, CfCoef cc,
                                                     int maxorder) {
    const BoxDev& B = L.box[blockIdx.y];
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
    const int npx = (nx + 1) >> 1;  // colour cells per x row, at most
    const uint32_t tot = (uint32_t)npx * (uint32_t)ny * (uint32_t)nz;
    constexpr T omega = T(1.15);
    const int64_t st[3] = {1, B.sy, B.sz};
    for (uint32_t q = blockIdx.x * kAB + threadIdx.x; q < tot; q += gridDim.x * kAB) {
        const uint32_t r = q / (uint32_t)npx;
        const int m = (int)(q - r * (uint32_t)npx);
        const int j = B.lo[1] + (int)(r % (uint32_t)ny), k = B.lo[2] + (int)(r / (uint32_t)ny);
        const int i = B.lo[0] + 2 * m + ((B.lo[0] + j + k + rb) & 1);
        if (i > B.hi[0]) continue;
        const int p[3] = {i, j, k};
        const int64_t c = cix(B, i, j, k);
        const T pc = phi[c];
        T nb[6], cf[6];
        for (int f = 0; f < 6; ++f) {
            const int d = f % 3;
            const int out = f < 3 ? -1 : 1;  // step from the cell to its neighbour across face f
            const bool at = f < 3 ? p[d] == B.lo[d] : p[d] == B.hi[d];
            cf[f] = T(0.0);
            if (!at) {
                nb[f] = phi[c + out * st[d]];
                continue;
            }
            int g[3] = {i, j, k};
            g[d] += out;
            const int o = fine_owner(L, g[0], g[1], g[2]);
            if (o >= 0) {
                nb[f] = phi[cix(L.box[o], wrapi(g[0], L.n[0]), wrapi(g[1], L.n[1]), wrapi(g[2], L.n[2]))];
                continue;
            }
            const int NX = nx_of(B, f, maxorder);
            if (o == -2 && cc.bct[f] == BC_NEUMANN) {
                nb[f] = pc;
            } else if (o == -2 && cc.bct[f] == BC_REFLECT_ODD) {
                nb[f] = -pc;
            } else {
                // sum over m >= 1 of w_m times the m-th cell inward from the ghost
                const double* w = o == -2 ? cc.p[NX] : cc.c[NX];
                T tmp = T(0.0);
                for (int mm = 1; mm < NX; ++mm) tmp += (mm == 1 ? pc : phi[c - (mm - 1) * out * st[d]]) * T(w[mm]);
                nb[f] = tmp;
            }
            cf[f] = o == -1 ? T(cc.c[NX][1]) : T(phys_cf(cc, f, NX));
        }
        if (op.kind == 0) {
            const T dhx = op.dhs[0], dhy = op.dhs[1], dhz = op.dhs[2];
            const T bxm = op.b[0][fix(B, 0, i, j, k)], bxp = op.b[0][fix(B, 0, i + 1, j, k)];
            const T bym = op.b[1][fix(B, 1, i, j, k)], byp = op.b[1][fix(B, 1, i, j + 1, k)];
            const T bzm = op.b[2][fix(B, 2, i, j, k)], bzp = op.b[2][fix(B, 2, i, j, k + 1)];
            const T gamma = op.alpha * op.a[c] + dhx * (bxm + bxp) + dhy * (bym + byp) + dhz * (bzm + bzp);
            const T g_m_d = gamma - (dhx * (bxm * cf[0] + bxp * cf[3]) + dhy * (bym * cf[1] + byp * cf[4]) +
                                     dhz * (bzm * cf[2] + bzp * cf[5]));
            const T rho = dhx * (bxm * nb[0] + bxp * nb[3]) + dhy * (bym * nb[1] + byp * nb[4]) +
                          dhz * (bzm * nb[2] + bzp * nb[5]);
            const T res = rhs[c] - (gamma * pc - rho);
            phi[c] = pc + omega / g_m_d * res;
            continue;
        }
        const T dhx = op.dpo[0], dhy = op.dpo[1], dhz = op.dpo[2];
        const T gamma = T(-2.) * (dhx + dhy + dhz);
        const T g_m_d = gamma + dhx * (cf[0] + cf[3]) + dhy * (cf[1] + cf[4]) + dhz * (cf[2] + cf[5]);
        const T res = rhs[c] - gamma * pc - dhx * (nb[0] + nb[3]) - dhy * (nb[1] + nb[4]) - dhz * (nb[2] + nb[5]);
        phi[c] = pc + omega / g_m_d * res;
    }
}

__global__ __launch_bounds__(kAB) void k_fine_bndry(LevelDev L, LevelDev Lc, const double* crse, const double* phys,
                                                    double* bcv) {
    const BoxDev& B = L.box[blockIdx.y];
    const int f = blockIdx.z;
    const int nq = face_cells(B, f);
    const int d = f % 3;
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    for (int q = blockIdx.x * kAB + threadIdx.x; q < nq; q += gridDim.x * kAB) {
        int p[3];
        ghost_of(B, f, q, p);
        const int64_t g = cix(B, p[0], p[1], p[2]);
        const int o = fine_owner(L, p[0], p[1], p[2]);
        if (o >= 0) { bcv[g] = 0.0; continue; }
        if (o == -2) { bcv[g] = phys ? phys[g] : 0.0; continue; }
        const int cc[3] = {coarsen2(p[0]), coarsen2(p[1]), coarsen2(p[2])};
        auto C = [&](int a1, int a2) {
            int c[3] = {cc[0], cc[1], cc[2]};
            c[t1] += a1;
            c[t2] += a2;
            return cval(Lc, crse, c[0], c[1], c[2]);
        };
        auto M = [&](int o1, int o2) {  // 1: inside the domain and not covered by a fine box
            int c[3] = {p[0], p[1], p[2]};
            c[t1] += o1;
            c[t2] += o2;
            return fine_owner(L, c[0], c[1], c[2]) == -1 ? 1 : 0;
        };
        const double c0 = C(0, 0);
        int lo = M(-2, 0) ? -1 : 0, hi = M(2, 0) ? 1 : 0;
        double fac = (hi == lo + 1) ? 1.0 : 0.5;
        const double d1 = fac * (C(hi, 0) - C(lo, 0));
        const double d11 = (hi == lo + 2) ? 0.5 * (C(1, 0) - 2. * c0 + C(-1, 0)) : 0.;
        lo = M(0, -2) ? -1 : 0;
        hi = M(0, 2) ? 1 : 0;
        fac = (hi == lo + 1) ? 1.0 : 0.5;
        const double d2 = fac * (C(0, hi) - C(0, lo));
        const double d22 = (hi == lo + 2) ? 0.5 * (C(0, 1) - 2. * c0 + C(0, -1)) : 0.;
        const double d12 = (M(-2, -2) && M(2, -2) && M(-2, 2) && M(2, 2))
                               ? 0.25 * (C(1, 1) - C(-1, 1) + C(-1, -1) - C(1, -1))
                               : 0.0;
        const double y1 = -0.5 + (p[t1] - cc[t1] * 2 + 0.5) / 2;
        const double y2 = -0.5 + (p[t2] - cc[t2] * 2 + 0.5) / 2;
        bcv[g] = c0 + y1 * d1 + (y1 * y1) * d11 + y2 * d2 + (y2 * y2) * d22 + y1 * y2 * d12;
    }
}

template <class T>
__global__ __launch_bounds__(kAB) void k_fine_avgdown(LevelDev L, LevelDev Lc, T* crse, const T* fine) {
    const BoxDev& B = L.box[blockIdx.y];
    const BoxDev& C = Lc.box[0];
    const int mx = (B.hi[0] - B.lo[0] + 1) / 2, my = (B.hi[1] - B.lo[1] + 1) / 2, mz = (B.hi[2] - B.lo[2] + 1) / 2;
    for (int64_t q = (int64_t)blockIdx.x * kAB + threadIdx.x; q < (int64_t)mx * my * mz; q += (int64_t)gridDim.x * kAB) {
        const int I = B.lo[0] / 2 + (int)(q % mx), J = B.lo[1] / 2 + (int)((q / mx) % my),
                  K = B.lo[2] / 2 + (int)(q / ((int64_t)mx * my));
        const int64_t f0 = cix(B, 2 * I, 2 * J, 2 * K);
        T c = 0;
        for (int kr = 0; kr < 2; ++kr)
            for (int jr = 0; jr < 2; ++jr)
                for (int ir = 0; ir < 2; ++ir) c += fine[f0 + ir + jr * B.sy + kr * B.sz];
        crse[cix(C, I, J, K)] = (T(1.0) / T(8)) * c;
    }
}

// amrex_avgdown_faces (AMReX_MultiFabUtil_3D_C.H:171-220) of fine box q along
// d onto the dense coarse face array: every coarse face of the coarsened box,
// its boundary faces included, is the mean of its 4 fine faces; on a periodic
// direction a face on a domain face is written to its image as well.  When the
// box spans the periodic domain along d, face n[d] (the later one) sets both.
__global__ __launch_bounds__(kAB) void k_fine_avgdown_faces(LevelDev L, LevelDev Lc, int q, int d, double* bc,
                                                            const double* bf) {
    const BoxDev& B = L.box[q];
    const BoxDev& C = Lc.box[0];
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    int lo[3], m[3];
    for (int e = 0; e < 3; ++e) {
        lo[e] = B.lo[e] / 2;
        m[e] = (B.hi[e] - B.lo[e] + 1) / 2 + (e == d);
    }
    const int n = Lc.n[d];
    const bool per = Lc.periodic[d] != 0, span = lo[d] == 0 && lo[d] + m[d] - 1 == n;
    for (int64_t e = (int64_t)blockIdx.x * kAB + threadIdx.x; e < (int64_t)m[0] * m[1] * m[2];
         e += (int64_t)gridDim.x * kAB) {
        int F[3];
        F[0] = lo[0] + (int)(e % m[0]);
        F[1] = lo[1] + (int)((e / m[0]) % m[1]);
        F[2] = lo[2] + (int)(e / ((int64_t)m[0] * m[1]));
        if (per && span && F[d] == 0) continue;
        double v = 0.0;
        for (int a2 = 0; a2 < 2; ++a2)
            for (int a1 = 0; a1 < 2; ++a1) {
                int f[3];
                f[d] = 2 * F[d];
                f[t1] = 2 * F[t1] + a1;
                f[t2] = 2 * F[t2] + a2;
                v += bf[fix(B, d, f[0], f[1], f[2])];
            }
        v /= 4.0;
        bc[fix(C, d, F[0], F[1], F[2])] = v;
        if (per && (F[d] == 0 || F[d] == n)) {
            int g[3] = {F[0], F[1], F[2]};
            g[d] = F[d] == 0 ? n : 0;
            bc[fix(C, d, g[0], g[1], g[2])] = v;
        }
    }
}

__global__ __launch_bounds__(kAB) void k_fine_interp(LevelDev L, LevelDev Lc, double* fine, const double* crse) {
    const BoxDev& B = L.box[blockIdx.y];
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
    for (int64_t q = (int64_t)blockIdx.x * kAB + threadIdx.x; q < (int64_t)nx * ny * nz; q += (int64_t)gridDim.x * kAB) {
        const int i = B.lo[0] + (int)(q % nx), j = B.lo[1] + (int)((q / nx) % ny), k = B.lo[2] + (int)(q / ((int64_t)nx * ny));
        const int ic = i / 2, jc = j / 2, kc = k / 2;
        const int io = 2 * (i - ic * 2) - 1, jo = 2 * (j - jc * 2) - 1, ko = 2 * (k - kc * 2) - 1;
        auto c = [&](int a, int b, int e) { return cval(Lc, crse, a, b, e); };
        fine[cix(B, i, j, k)] = 0.421875 * c(ic, jc, kc) + 0.140625 * c(ic + io, jc, kc) + 0.140625 * c(ic, jc + jo, kc)
                              + 0.140625 * c(ic, jc, kc + ko) + 0.046875 * c(ic, jc + jo, kc + ko)
                              + 0.046875 * c(ic + io, jc, kc + ko) + 0.046875 * c(ic + io, jc + jo, kc)
                              + 0.015625 * c(ic + io, jc + jo, kc + ko);
    }
}

// One coarse cell next to the fine region: res += sum over its coarse/fine
// faces of +-(F_crse - <F_fine>)/dx_c (side of the fine region above: +).
// The fluxes are FFlux's: dxinv*diff for MLPoisson, -b_scalar*dxinv*b*diff for
// an ABecLaplacian (mlabeclap_flux_*, AMReX_MLABecLap_3D_K.H:79-96) with the
// coarse level's face coefficients and the fine box's.
__global__ __launch_bounds__(kAB) void k_reflux(LevelDev Lc, LevelDev Lf, const RefluxCell* cells, int ncell,
                                                double* res, const double* sol0, const double* sol1, RefluxCoef rc) {
    const int e = blockIdx.x * kAB + threadIdx.x;
    if (e >= ncell) return;
    const RefluxCell& R = cells[e];
    const BoxDev& C = Lc.box[0];
    const int c[3] = {R.i, R.j, R.k};
    const bool abec = rc.bc[0] != nullptr;
    double d = 0.0;
    for (int f = 0; f < 6; ++f) {
        const int q = R.fbox[f];
        if (q < 0) continue;
        const int dir = f % 3, side = f / 3;  // side 1: fine region on the high side
        const int t1 = (dir == 0) ? 1 : 0, t2 = (dir == 2) ? 1 : 2;
        int hi[3] = {c[0], c[1], c[2]};
        if (side) hi[dir] += 1;
        int lo[3] = {hi[0], hi[1], hi[2]};
        lo[dir] -= 1;
        const double dxc = (double)Lc.n[dir], dxf = (double)Lf.n[dir];  // unit-cube dxinv
        const double bc = abec ? rc.bc[dir][fix(C, dir, hi[0], hi[1], hi[2])] : 1.0;
        const double fc = rc.scale * dxc * bc * (cval(Lc, sol0, hi[0], hi[1], hi[2]) - cval(Lc, sol0, lo[0], lo[1], lo[2]));
        const BoxDev& F = Lf.box[q];
        // the coarse/fine face in the frame of box q: its low face when the fine
        // region is above, else its high face (2*hi[dir] unless the face is the
        // periodic image across a domain face)
        const int fi = side ? F.lo[dir] : F.hi[dir] + 1;
        double fsum = 0.0;
        for (int a2 = 0; a2 < 2; ++a2)
            for (int a1 = 0; a1 < 2; ++a1) {
                int h2[3], l2[3];
                h2[dir] = fi;
                l2[dir] = fi - 1;
                h2[t1] = l2[t1] = 2 * c[t1] + a1;
                h2[t2] = l2[t2] = 2 * c[t2] + a2;
                const double bf = abec ? rc.bf[dir][fix(F, dir, h2[0], h2[1], h2[2])] : 1.0;
                fsum += rc.scale * dxf * bf * (sol1[cix(F, h2[0], h2[1], h2[2])] - sol1[cix(F, l2[0], l2[1], l2[2])]);
            }
        const double dtdx_f = dxf / 8.0;
        if (side) d += dxc * fc - dtdx_f * fsum;
        else d += -dxc * fc + dtdx_f * fsum;
    }
    res[cix(C, c[0], c[1], c[2])] += d;
}

__device__ __forceinline__ double nanmax2(double a, double b) { return (b > a || b != b) ? b : a; }

__global__ __launch_bounds__(kAB) void k_masked_max(LevelDev Lc, const unsigned char* covered, const double* x,
                                                    double* partials) {
    __shared__ double lds[kAB / 64];
    const BoxDev& C = Lc.box[0];
    const int n0 = Lc.n[0], n1 = Lc.n[1], n2 = Lc.n[2];
    const int64_t nc = (int64_t)n0 * n1 * n2;
    double m = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kAB + threadIdx.x; q < nc; q += (int64_t)gridDim.x * kAB) {
        if (covered[q]) continue;
        const int i = (int)(q % n0), j = (int)((q / n0) % n1), k = (int)(q / ((int64_t)n0 * n1));
        m = nanmax2(m, fabs(x[cix(C, i, j, k)]));
    }
    for (int o = 32; o > 0; o >>= 1) m = nanmax2(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kAB / 64; ++w) m = nanmax2(m, lds[w]);
        partials[blockIdx.x] = m;
    }
}
__global__ __launch_bounds__(256) void k_max_final(const double* partials, int n, double* out) {
    __shared__ double lds[4];
    double m = 0.0;
    for (int q = threadIdx.x; q < n; q += 256) m = nanmax2(m, partials[q]);
    for (int o = 32; o > 0; o >>= 1) m = nanmax2(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) m = nanmax2(m, lds[w]);
        out[0] = m;
    }
}

struct I3 {
    int v[3];
};

__global__ __launch_bounds__(kAB) void k_cf_window(const double* __restrict__ user, const UserBox* __restrict__ ub,
                                                   int nbox, int ung, I3 wlo, I3 wn, I3 nc, I3 per,
                                                   double* __restrict__ win) {
    const int64_t q = (int64_t)blockIdx.x * kAB + threadIdx.x;
    if (q >= (int64_t)wn.v[0] * wn.v[1] * wn.v[2]) return;
    int c[3] = {wlo.v[0] + (int)(q % wn.v[0]), wlo.v[1] + (int)((q / wn.v[0]) % wn.v[1]),
                wlo.v[2] + (int)(q / ((int64_t)wn.v[0] * wn.v[1]))};
    for (int d = 0; d < 3; ++d) {
        if (c[d] >= 0 && c[d] < nc.v[d]) continue;
        if (!per.v[d]) {
            win[q] = 0.0;
            return;
        }
        c[d] = wrapi(c[d], nc.v[d]);
    }
    for (int b = 0; b < nbox; ++b) {
        const UserBox& U = ub[b];
        if (c[0] < U.lo[0] || c[0] > U.hi[0] || c[1] < U.lo[1] || c[1] > U.hi[1] || c[2] < U.lo[2] || c[2] > U.hi[2])
            continue;
        const int64_t ux = U.hi[0] - U.lo[0] + 1 + 2 * ung, uy = U.hi[1] - U.lo[1] + 1 + 2 * ung;
        win[q] = user[U.off + (c[0] - U.lo[0] + ung) + ux * ((c[1] - U.lo[1] + ung) + uy * (int64_t)(c[2] - U.lo[2] + ung))];
        return;
    }
    win[q] = 0.0;
}

__device__ __forceinline__ int coarsen_r(int i, int r) { return i >= 0 ? i / r : -((-i + r - 1) / r); }

__global__ __launch_bounds__(kAB) void k_cf_bndry(LevelDev L, const double* __restrict__ win, I3 wlo, I3 wn, int r,
                                                  I3 rlo, I3 rhi, double* __restrict__ bcv) {
    const BoxDev& B = L.box[blockIdx.y];
    const int f = blockIdx.z;
    if (!((B.cfmask >> f) & 1)) return;
    const int nq = face_cells(B, f);
    const int d = f % 3;
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    for (int q = blockIdx.x * kAB + threadIdx.x; q < nq; q += gridDim.x * kAB) {
        int p[3];
        ghost_of(B, f, q, p);
        const int cc[3] = {coarsen_r(p[0], r), coarsen_r(p[1], r), coarsen_r(p[2], r)};
        // coarse value at tangential offsets (a1, a2) from the ghost's coarse cell
        auto C = [&](int a1, int a2) {
            int c[3] = {cc[0], cc[1], cc[2]};
            c[t1] += a1;
            c[t2] += a2;
            return win[(c[0] - wlo.v[0]) + (int64_t)wn.v[0] * ((c[1] - wlo.v[1]) + (int64_t)wn.v[1] * (c[2] - wlo.v[2]))];
        };
        // BndryData mask of the ghost position offset by (o1, o2) fine cells
        // tangentially: 1 not covered (inside the domain or a periodic image,
        // outside the region), else 0 (covered, or beyond a physical face)
        auto M = [&](int o1, int o2) {
            int c[3] = {p[0], p[1], p[2]};
            c[t1] += o1;
            c[t2] += o2;
            for (int e = 0; e < 3; ++e) {
                if (c[e] >= 0 && c[e] < L.n[e]) continue;
                if (!L.periodic[e]) return 0;
                c[e] = wrapi(c[e], L.n[e]);
            }
            const bool in = c[0] >= rlo.v[0] && c[0] <= rhi.v[0] && c[1] >= rlo.v[1] && c[1] <= rhi.v[1] &&
                            c[2] >= rlo.v[2] && c[2] <= rhi.v[2];
            return in ? 0 : 1;
        };
        const double c0 = C(0, 0);
        int lo = M(-r, 0) ? -1 : 0, hi = M(r, 0) ? 1 : 0;
        double fac = (hi == lo + 1) ? 1.0 : 0.5;
        const double d1 = fac * (C(hi, 0) - C(lo, 0));
        const double d11 = (hi == lo + 2) ? 0.5 * (C(1, 0) - 2. * c0 + C(-1, 0)) : 0.;
        lo = M(0, -r) ? -1 : 0;
        hi = M(0, r) ? 1 : 0;
        fac = (hi == lo + 1) ? 1.0 : 0.5;
        const double d2 = fac * (C(0, hi) - C(0, lo));
        const double d22 = (hi == lo + 2) ? 0.5 * (C(0, 1) - 2. * c0 + C(0, -1)) : 0.;
        const double d12 = (M(-r, -r) && M(r, -r) && M(-r, r) && M(r, r))
                               ? 0.25 * (C(1, 1) - C(-1, 1) + C(-1, -1) - C(1, -1))
                               : 0.0;
        const double y1 = -0.5 + (p[t1] - cc[t1] * r + 0.5) / r;
        const double y2 = -0.5 + (p[t2] - cc[t2] * r + 0.5) / r;
        bcv[cix(B, p[0], p[1], p[2])] = c0 + y1 * d1 + (y1 * y1) * d11 + y2 * d2 + (y2 * y2) * d22 + y1 * y2 * d12;
    }
}

// blocks per box face: one thread per ghost cell of the largest face
inline unsigned face_blocks(const LevelDev& L) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((L.max_face + kAB - 1) / kAB, 64));
}
inline dim3 fine_grid(const LevelDev& L, int64_t per_box) {
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((per_box + kAB - 1) / kAB, 2048)), (unsigned)L.nbox, 1);
}

}  // namespace

constexpr int kMaskedBlocks = 1024;
int64_t masked_max_partials() { return kMaskedBlocks; }

template <class T>
void fine_fill(const LevelDev& Lf, T* phi, const T* bcv, const CfCoef& cc, int maxorder, hipStream_t s) {
    hipLaunchKernelGGL(k_fine_fill<T>, dim3(face_blocks(Lf), Lf.nbox, 6), dim3(kAB), 0, s, Lf, phi, bcv, cc, maxorder);
}
template <class T>
void fine_gsrb(const LevelDev& Lf, const OpCoef<T>& op, T* phi, const T* rhs, int rb, const CfCoef& cc, int maxorder,
               hipStream_t s) {
    hipLaunchKernelGGL(k_fine_gsrb<T>, fine_grid(Lf, Lf.mbc), dim3(kAB), 0, s, Lf, op, phi, rhs, rb, cc, maxorder);
}
template <class T>
void fine_smooth_colour(const LevelDev& Lf, const OpCoef<T>& op, T* phi, const T* rhs, int rb, const CfCoef& cc,
                        int maxorder, hipStream_t s) {
    if (maxorder > 3) {
        fine_fill<T>(Lf, phi, nullptr, cc, maxorder, s);
        fine_gsrb<T>(Lf, op, phi, rhs, rb, cc, maxorder, s);
        return;
    }
    hipLaunchKernelGGL(k_fine_gsrb_h<T>, fine_grid(Lf, (Lf.mbc + 1) / 2), dim3(kAB), 0, s, Lf, op, phi, rhs, rb, cc,
                       maxorder);
}
void fine_bndry(const LevelDev& Lf, const LevelDev& Lc, const double* crse, const double* phys, double* bcv,
                hipStream_t s) {
    hipLaunchKernelGGL(k_fine_bndry, dim3(face_blocks(Lf), Lf.nbox, 6), dim3(kAB), 0, s, Lf, Lc, crse, phys, bcv);
}
template <class T>
void fine_avgdown(const LevelDev& Lf, const LevelDev& Lc, T* crse, const T* fine, hipStream_t s) {
    hipLaunchKernelGGL(k_fine_avgdown<T>, fine_grid(Lf, Lf.mbc / 8), dim3(kAB), 0, s, Lf, Lc, crse, fine);
}
void fine_interp(const LevelDev& Lf, const LevelDev& Lc, double* fine, const double* crse, hipStream_t s) {
    hipLaunchKernelGGL(k_fine_interp, fine_grid(Lf, Lf.mbc), dim3(kAB), 0, s, Lf, Lc, fine, crse);
}
void fine_avgdown_faces(const LevelDev& Lf, const LevelDev& Lc, int nbox, int d, double* bc, const double* bf,
                        hipStream_t s) {
    // one launch per box, in box order: a coarse face shared by two coarsened
    // boxes keeps the later box's mean
    for (int q = 0; q < nbox; ++q)
        hipLaunchKernelGGL(k_fine_avgdown_faces, fine_grid(Lf, Lf.mbf / 4 + 1).x, dim3(kAB), 0, s, Lf, Lc, q, d, bc, bf);
}
void reflux(const LevelDev& Lc, const LevelDev& Lf, const RefluxCell* cells, int ncell, double* res,
            const double* sol0, const double* sol1, const RefluxCoef& rc, hipStream_t s) {
    if (ncell <= 0) return;
    hipLaunchKernelGGL(k_reflux, dim3((unsigned)((ncell + kAB - 1) / kAB)), dim3(kAB), 0, s, Lc, Lf, cells, ncell, res,
                       sol0, sol1, rc);
}
void masked_max(const LevelDev& Lc, const unsigned char* covered, const double* x, double* partials, double* out,
                hipStream_t s) {
    hipLaunchKernelGGL(k_masked_max, dim3(kMaskedBlocks), dim3(kAB), 0, s, Lc, covered, x, partials);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, s, partials, kMaskedBlocks, out);
}

void cf_window(const double* user, const UserBox* ub, int nbox, int ung, const int* wlo, const int* wn, const int* nc,
               const int* periodic, double* win, hipStream_t s) {
    const int64_t n = (int64_t)wn[0] * wn[1] * wn[2];
    if (n <= 0) return;
    const I3 a{{wlo[0], wlo[1], wlo[2]}}, b{{wn[0], wn[1], wn[2]}}, c{{nc[0], nc[1], nc[2]}},
        p{{periodic[0], periodic[1], periodic[2]}};
    hipLaunchKernelGGL(k_cf_window, dim3((unsigned)((n + kAB - 1) / kAB)), dim3(kAB), 0, s, user, ub, nbox, ung, a, b, c,
                       p, win);
}
void cf_bndry(const LevelDev& L, const double* win, const int* wlo, const int* wn, int r, const int* rlo,
              const int* rhi, int64_t max_face, double* bcv, hipStream_t s) {
    const I3 a{{wlo[0], wlo[1], wlo[2]}}, b{{wn[0], wn[1], wn[2]}}, lo{{rlo[0], rlo[1], rlo[2]}},
        hi{{rhi[0], rhi[1], rhi[2]}};
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((max_face + kAB - 1) / kAB, 1024));
    hipLaunchKernelGGL(k_cf_bndry, dim3(gx, (unsigned)L.nbox, 6), dim3(kAB), 0, s, L, win, a, b, r, lo, hi, bcv);
}

template void fine_fill<double>(const LevelDev&, double*, const double*, const CfCoef&, int, hipStream_t);
template void fine_fill<float>(const LevelDev&, float*, const float*, const CfCoef&, int, hipStream_t);
template void fine_gsrb<float>(const LevelDev&, const OpCoef<float>&, float*, const float*, int, const CfCoef&, int,
                               hipStream_t);
template void fine_gsrb<double>(const LevelDev&, const OpCoef<double>&, double*, const double*, int, const CfCoef&, int,
                                hipStream_t);
template void fine_avgdown<double>(const LevelDev&, const LevelDev&, double*, const double*, hipStream_t);
template void fine_smooth_colour<float>(const LevelDev&, const OpCoef<float>&, float*, const float*, int, const CfCoef&,
                                        int, hipStream_t);
template void fine_smooth_colour<double>(const LevelDev&, const OpCoef<double>&, double*, const double*, int,
                                         const CfCoef&, int, hipStream_t);

}  // namespace mgx
