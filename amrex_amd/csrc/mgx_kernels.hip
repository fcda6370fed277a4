// HIP kernels of the MLMG hot path for CDNA4 (gfx950).
//
// Box-batched launches: blockIdx.y selects the box of the level, blockIdx.x
// and threadIdx.x walk that box's cells with x fastest, so one launch covers
// every box of a MultiFab (the reference launches one ParallelFor over all
// boxes the same way, AMReX_MFParallelForG.H:73-139).  Arithmetic of each
// per-cell update follows the reference kernels cited at each function.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <type_traits>

#include "mgx_kernels.H"
#include "mgx_transpose.H"


namespace mgx {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int64_t cidx(const BoxDev& B, int i, int j, int k) {
    return B.off + (i - B.lo[0] + 1) + B.sy * (j - B.lo[1] + 1) + B.sz * (k - B.lo[2] + 1);
}

__device__ __forceinline__ int64_t fidx(const BoxDev& B, int d, int i, int j, int k) {
    const int64_t fx = B.hi[0] - B.lo[0] + 1 + (d == 0);
    const int64_t fy = B.hi[1] - B.lo[1] + 1 + (d == 1);
    return B.foff[d] + (i - B.lo[0]) + fx * ((j - B.lo[1]) + fy * (k - B.lo[2]));
}

__device__ __forceinline__ int owner(const LevelDev& L, int i, int j, int k) {
    return L.box_of[L.slab_of[0][i] + L.nslab[0] * (L.slab_of[1][j] + L.nslab[1] * L.slab_of[2][k])];
}

// (i, j, k) offsets of element q of an nx x ny x nz box, x fastest; 32-bit
// divisions when the box has fewer than 2^31 cells (the 64-bit ones are
// emulated and cost more than the memory access of a copy).
__device__ __forceinline__ void split_index(int64_t q, int64_t n, int nx, int ny, int& i, int& j, int& k) {
    if (n <= 0x7fffffff) {
        const uint32_t qq = (uint32_t)q, r = qq / (uint32_t)nx;
        i = (int)(qq - r * (uint32_t)nx);
        j = (int)(r % (uint32_t)ny);
        k = (int)(r / (uint32_t)ny);
    } else {
        const int64_t r = q / nx;
        i = (int)(q - r * nx);
        j = (int)(r % ny);
        k = (int)(r / ny);
    }
}

// Valid cell q of box B (x fastest).  Returns false past the end of the box.
__device__ __forceinline__ bool cell_of(const BoxDev& B, int64_t q, int& i, int& j, int& k) {
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
    const int64_t n = (int64_t)nx * ny * nz;
    if (q >= n) return false;
    split_index(q, n, nx, ny, i, j, k);
    i += B.lo[0];
    j += B.lo[1];
    k += B.lo[2];
    return true;
}

inline dim3 box_grid(const LevelDev& L, int max_box_cells) {
    return dim3((unsigned)((max_box_cells + kBlock - 1) / kBlock), (unsigned)L.nbox, 1);
}

constexpr int kTX = kTileX, kTY = kTileY, kTZ = kTileZ;

inline dim3 tile_grid(const LevelDev& L, int which = 0) {
    return dim3((unsigned)std::max(L.mtiles[which], 1), (unsigned)L.nbox, 1);
}

// Cell (i, j) of this thread and the z range of tile blockIdx.x of box B.
struct TileCells {
    int i, j, k0, k1;
    bool ok;
};
__device__ __forceinline__ TileCells tile_cells(const BoxDev& B, int tz) {
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1, nz = B.hi[2] - B.lo[2] + 1;
    const int ntx = (nx + kTX - 1) / kTX, nty = (ny + kTY - 1) / kTY, ntz = (nz + tz - 1) / tz;
    TileCells t;
    int b = blockIdx.x;
    t.ok = b < ntx * nty * ntz;
    const int bx = b % ntx;
    b /= ntx;
    const int by = b % nty, bz = b / nty;
    t.i = B.lo[0] + bx * kTX + (int)(threadIdx.x % kTX);
    t.j = B.lo[1] + by * kTY + (int)(threadIdx.x / kTX);
    t.k0 = B.lo[2] + bz * tz;
    t.k1 = min(t.k0 + tz, B.hi[2] + 1);
    t.ok = t.ok && t.i <= B.hi[0] && t.j <= B.hi[1];
    return t;
}

inline dim3 flat_grid(int64_t n) {
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 1 << 20)), 1, 1);
}

// ---------------------------------------------------------------------------
// Per-cell device functions (shared by the level kernels and the bottom solve)
// ---------------------------------------------------------------------------

// Ghost cell q of face f of box b: the physical closure of mllinop_apply_bc_*
// (AMReX_MLLinOp_K.H:14-115) where the face is a domain boundary, otherwise a
// copy from the box (or periodic image) that owns the cell (FillBoundary with
// the cross stencil, AMReX_FabArray.H:3059-3130).
template <class T>
__device__ __forceinline__ void ghost_cell(const LevelDev& L, const BoxDev& B, int f, int q, T* phi,
                                           const T* bcv, int inhomog) {
    const int d = f % 3, side = f / 3;
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    const int n1 = B.hi[t1] - B.lo[t1] + 1;
    int idx[3];
    idx[d] = side == 0 ? B.lo[d] - 1 : B.hi[d] + 1;
    idx[t1] = B.lo[t1] + q % n1;
    idx[t2] = B.lo[t2] + q / n1;
    const int64_t cg = cidx(B, idx[0], idx[1], idx[2]);
    const int bct = B.bct[f];
    if (bct != BC_INTERIOR) {
        const int64_t step = (d == 0) ? 1 : (d == 1 ? B.sy : B.sz);
        const int64_t s = side == 0 ? step : -step;
        if (bct == BC_NEUMANN) {
            phi[cg] = phi[cg + s];
        } else if (bct == BC_REFLECT_ODD) {
            phi[cg] = -phi[cg + s];
        } else if (bct == BC_DIRICHLET) {
            const int NX = B.nxo[f];
            T tmp = T(0.0);
            for (int m = 1; m < NX; ++m) tmp += phi[cg + m * s] * T(B.bcc[f][m]);
            if (inhomog) tmp += bcv[cg] * T(B.bcc[f][0]);
            phi[cg] = tmp;
        }
    } else {
        int src[3] = {idx[0], idx[1], idx[2]};
        if (src[d] < 0) src[d] += L.n[d];
        if (src[d] >= L.n[d]) src[d] -= L.n[d];
        const BoxDev& S = L.box[owner(L, src[0], src[1], src[2])];
        phi[cg] = phi[cidx(S, src[0], src[1], src[2])];
    }
}

// One GSRB-SOR update (abec_gsrb, AMReX_MLABecLap_3D_K.H:210-264;
// mlpoisson_gsrb, AMReX_MLPoisson_3D_K.H:155-196).
template <class T>
__device__ __forceinline__ void gsrb_cell(const BoxDev& B, const OpCoef<T>& op, T* phi, const T* rhs,
                                          int i, int j, int k) {
    constexpr T omega = T(1.15);
    const int64_t c = cidx(B, i, j, k);
    const int64_t sy = B.sy, sz = B.sz;
    const T cf0 = (i == B.lo[0] && B.bct[0] != BC_INTERIOR) ? T(B.cf[0]) : T(0.0);
    const T cf1 = (j == B.lo[1] && B.bct[1] != BC_INTERIOR) ? T(B.cf[1]) : T(0.0);
    const T cf2 = (k == B.lo[2] && B.bct[2] != BC_INTERIOR) ? T(B.cf[2]) : T(0.0);
    const T cf3 = (i == B.hi[0] && B.bct[3] != BC_INTERIOR) ? T(B.cf[3]) : T(0.0);
    const T cf4 = (j == B.hi[1] && B.bct[4] != BC_INTERIOR) ? T(B.cf[4]) : T(0.0);
    const T cf5 = (k == B.hi[2] && B.bct[5] != BC_INTERIOR) ? T(B.cf[5]) : T(0.0);
    const T p = phi[c];
    if (op.kind == 0) {
        const T dhx = op.dhs[0], dhy = op.dhs[1], dhz = op.dhs[2];
        const T bxm = op.b[0][fidx(B, 0, i, j, k)], bxp = op.b[0][fidx(B, 0, i + 1, j, k)];
        const T bym = op.b[1][fidx(B, 1, i, j, k)], byp = op.b[1][fidx(B, 1, i, j + 1, k)];
        const T bzm = op.b[2][fidx(B, 2, i, j, k)], bzp = op.b[2][fidx(B, 2, i, j, k + 1)];
        const T gamma = op.alpha * op.a[c] + dhx * (bxm + bxp) + dhy * (bym + byp) + dhz * (bzm + bzp);
        const T g_m_d = gamma - (dhx * (bxm * cf0 + bxp * cf3) + dhy * (bym * cf1 + byp * cf4)
                                 + dhz * (bzm * cf2 + bzp * cf5));
        const T rho = dhx * (bxm * phi[c - 1] + bxp * phi[c + 1])
                    + dhy * (bym * phi[c - sy] + byp * phi[c + sy])
                    + dhz * (bzm * phi[c - sz] + bzp * phi[c + sz]);
        const T res = rhs[c] - (gamma * p - rho);
        phi[c] = p + omega / g_m_d * res;
    } else {
        const T dhx = op.dpo[0], dhy = op.dpo[1], dhz = op.dpo[2];
        const T gamma = T(-2.) * (dhx + dhy + dhz);
        const T g_m_d = gamma + dhx * (cf0 + cf3) + dhy * (cf1 + cf4) + dhz * (cf2 + cf5);
        const T res = rhs[c] - gamma * p - dhx * (phi[c - 1] + phi[c + 1])
                    - dhy * (phi[c - sy] + phi[c + sy]) - dhz * (phi[c - sz] + phi[c + sz]);
        phi[c] = p + omega / g_m_d * res;
    }
}

// L(x) at one cell (mlabeclap_adotx, AMReX_MLABecLap_3D_K.H:9-28;
// mlpoisson_adotx, AMReX_MLPoisson_3D_K.H:9-16).
template <class T>
__device__ __forceinline__ T apply_cell(const BoxDev& B, const OpCoef<T>& op, const T* x, int i, int j, int k,
                                        int64_t c) {
    const int64_t sy = B.sy, sz = B.sz;
    const T xc = x[c];
    if (op.kind == 0) {
        const T dhx = op.dha[0], dhy = op.dha[1], dhz = op.dha[2];
        return op.alpha * op.a[c] * xc
            - dhx * (op.b[0][fidx(B, 0, i + 1, j, k)] * (x[c + 1] - xc) - op.b[0][fidx(B, 0, i, j, k)] * (xc - x[c - 1]))
            - dhy * (op.b[1][fidx(B, 1, i, j + 1, k)] * (x[c + sy] - xc) - op.b[1][fidx(B, 1, i, j, k)] * (xc - x[c - sy]))
            - dhz * (op.b[2][fidx(B, 2, i, j, k + 1)] * (x[c + sz] - xc) - op.b[2][fidx(B, 2, i, j, k)] * (xc - x[c - sz]));
    } else {
        const T dhx = op.dpo[0], dhy = op.dpo[1], dhz = op.dpo[2];
        return dhx * (x[c - 1] - T(2.0) * xc + x[c + 1])
             + dhy * (x[c - sy] - T(2.0) * xc + x[c + sy])
             + dhz * (x[c - sz] - T(2.0) * xc + x[c + sz]);
    }
}

// Jacobi scaling of mlabeclap_normalize (AMReX_MLABecLap_3D_K.H:60-75).
template <class T>
__device__ __forceinline__ T diag_cell(const BoxDev& B, const OpCoef<T>& op, int i, int j, int k, int64_t c) {
    const T dhx = op.dha[0], dhy = op.dha[1], dhz = op.dha[2];
    return op.alpha * op.a[c]
        + dhx * (op.b[0][fidx(B, 0, i, j, k)] + op.b[0][fidx(B, 0, i + 1, j, k)])
        + dhy * (op.b[1][fidx(B, 1, i, j, k)] + op.b[1][fidx(B, 1, i, j + 1, k)])
        + dhz * (op.b[2][fidx(B, 2, i, j, k)] + op.b[2][fidx(B, 2, i, j, k + 1)]);
}

// ---------------------------------------------------------------------------
// Level kernels
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kBlock) void k_fill_ghosts(LevelDev L, T* phi, const T* bcv, int inhomog) {
    const int b = blockIdx.y;
    const int f = blockIdx.z;
    const BoxDev& B = L.box[b];
    const int d = f % 3;
    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
    const int nq = (B.hi[t1] - B.lo[t1] + 1) * (B.hi[t2] - B.lo[t2] + 1);
    for (int q = blockIdx.x * kBlock + threadIdx.x; q < nq; q += gridDim.x * kBlock)
        ghost_cell(L, B, f, q, phi, bcv, inhomog);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_gsrb(LevelDev L, OpCoef<T> op, T* phi, const T* rhs, int rb) {
    const BoxDev& B = L.box[blockIdx.y];
    const TileCells t = tile_cells(B, kTZ);
    if (!t.ok) return;
    for (int k = t.k0; k < t.k1; ++k)
        if (((t.i + t.j + k + rb) & 1) == 0) gsrb_cell(B, op, phi, rhs, t.i, t.j, k);
}

// The adjoint cycle's kernels on dense levels: one colour of the transposed
// sweep, and out = b - L^T x (homogeneous boundary data; ghosts unused).
// One thread per x pair of cells: in every plane it updates the pair's cell of
// the colour, so no lane idles on the other colour.
constexpr int kTPX = 2 * kTX;  // cells per tile row of the transposed sweep
template <class T>
__global__ __launch_bounds__(kBlock) void k_gsrb_T(LevelDev L, OpCoef<T> op, T* phi, const T* rhs, int rb) {
    const BoxDev& B = L.box[0];
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    const int ntx = (n0 + kTPX - 1) / kTPX, nty = (n1 + kTY - 1) / kTY;
    int b = blockIdx.x;
    const int bx = b % ntx;
    b /= ntx;
    const int by = b % nty, bz = b / nty;
    const int j = by * kTY + (int)(threadIdx.x / kTX);
    const int ip = bx * kTPX + 2 * (int)(threadIdx.x % kTX);  // even
    const int k0 = bz * kTZ, k1 = min(k0 + kTZ, n2);
    if (j >= n1) return;
    for (int k = k0; k < k1; ++k) {
        const int i = ip + ((j + k + rb) & 1);  // i + j + k + rb even
        if (i < n0) lt::gsrb_T_cell(L, B, op, phi, rhs, i, j, k);
    }
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_resid_T(LevelDev L, OpCoef<T> op, T* out, const T* x, const T* b) {
    const BoxDev& B = L.box[0];
    const TileCells t = tile_cells(B, kTZ);
    if (!t.ok) return;
    auto xv = [&](int ii, int jj, int kk) -> T { return x[lt::cell(B, ii, jj, kk)]; };
    for (int k = t.k0; k < t.k1; ++k) {
        const int64_t c = lt::cell(B, t.i, t.j, k);
        out[c] = b[c] + T(-1.0) * lt::apply_dense_T<T>(L, B, op, xv, t.i, t.j, k);
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_apply(LevelDev L, OpCoef<T> op, T* out, const T* x, const T* b) {
    const BoxDev& B = L.box[blockIdx.y];
    const TileCells t = tile_cells(B, kTZ);
    if (!t.ok) return;
    for (int k = t.k0; k < t.k1; ++k) {
        const int64_t c = cidx(B, t.i, t.j, k);
        const T y = apply_cell(B, op, x, t.i, t.j, k, c);
        out[c] = b ? b[c] + T(-1.0) * y : y;
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_normalize(LevelDev L, OpCoef<T> op, T* x) {
    const BoxDev& B = L.box[blockIdx.y];
    const TileCells t = tile_cells(B, kTZ);
    if (!t.ok) return;
    for (int k = t.k0; k < t.k1; ++k) {
        const int64_t c = cidx(B, t.i, t.j, k);
        x[c] /= diag_cell(B, op, t.i, t.j, k, c);
    }
}

// amrex_avgdown (AMReX_MultiFabUtil_3D_C.H:377-396), ratio 2.
template <class T>
__global__ __launch_bounds__(kBlock) void k_restrict(LevelDev Lc, LevelDev Lf, T* crse, const T* fine, int same) {
    const BoxDev& C = Lc.box[blockIdx.y];
    const TileCells t = tile_cells(C, kTZ);
    if (!t.ok) return;
    const int I = t.i, J = t.j;
    for (int K = t.k0; K < t.k1; ++K) {
        const BoxDev& F = Lf.box[same ? (int)blockIdx.y : owner(Lf, 2 * I, 2 * J, 2 * K)];
        const int64_t f0 = cidx(F, 2 * I, 2 * J, 2 * K);
        T c = 0;
        for (int kr = 0; kr < 2; ++kr)
            for (int jr = 0; jr < 2; ++jr)
                for (int ir = 0; ir < 2; ++ir) c += fine[f0 + ir + jr * F.sy + kr * F.sz];
        crse[cidx(C, I, J, K)] = (T(1.0) / T(8)) * c;
    }
}

// Piecewise-constant prolongation added to the fine correction
// (MLCellLinOpT::interpolation, AMReX_MLCellLinOp.H:955-999).
template <class T>
__global__ __launch_bounds__(kBlock) void k_prolong_add(LevelDev Lf, LevelDev Lc, T* fine, const T* crse, int same) {
    const BoxDev& F = Lf.box[blockIdx.y];
    const TileCells t = tile_cells(F, kTZ);
    if (!t.ok) return;
    for (int k = t.k0; k < t.k1; ++k) {
        const int ic = t.i >> 1, jc = t.j >> 1, kc = k >> 1;
        const BoxDev& C = Lc.box[same ? (int)blockIdx.y : owner(Lc, ic, jc, kc)];
        fine[cidx(F, t.i, t.j, k)] += crse[cidx(C, ic, jc, kc)];
    }
}

// Coarse value read by the F-cycle interpolation at (i, j, k), at most one cell
// outside the domain in each direction (MLCellLinOp::interpAssign,
// AMReX_MLCellLinOp.H:1002-1026).  Ghosts reached through periodic directions
// only take the periodic image.  A ghost outside exactly one physical face holds,
// when phys != 0, what the coarse array's ghost layer holds (the reference reads
// the coarse correction after FillBoundary, which leaves such ghosts as the last
// boundary fill wrote them) and 0 otherwise (the reference copies an
// agglomerated level into a zeroed array).  Ghosts outside a physical face and
// another face are 0: corrections are cleared with their ghosts at the start of
// each cycle and the face-only boundary fills never write them.
template <class T>
__device__ __forceinline__ T fmg_crse(const LevelDev& C, const T* c, int i, int j, int k, int phys) {
    int p[3] = {i, j, k}, q[3] = {i, j, k};
    int nout = 0, nphys = 0;
    for (int d = 0; d < 3; ++d) {
        if (p[d] >= 0 && p[d] < C.n[d]) continue;
        ++nout;
        if (C.periodic[d]) {
            p[d] += p[d] < 0 ? C.n[d] : -C.n[d];
            q[d] = p[d];
        } else {
            ++nphys;
            q[d] = p[d] < 0 ? 0 : C.n[d] - 1;
        }
    }
    if (nphys == 0) return c[cidx(C.box[owner(C, p[0], p[1], p[2])], p[0], p[1], p[2])];
    if (nout == 1 && phys) return c[cidx(C.box[owner(C, q[0], q[1], q[2])], i, j, k)];
    return T(0);
}

// fine = mlmg_lin_cc_interp_r2(crse) on the valid fine cells (AMReX_MLMG_3D_K.H:9-38):
// weights 27/64, 9/64, 3/64, 1/64 in the reference's order.
template <class T>
__global__ __launch_bounds__(kBlock) void k_fmg_interp(LevelDev Lf, LevelDev Lc, T* fine, const T* crse, int phys) {
    const BoxDev& F = Lf.box[blockIdx.y];
    const TileCells t = tile_cells(F, kTZ);
    if (!t.ok) return;
    const int ic = t.i >> 1, jc = t.j >> 1;
    const int io = 2 * (t.i - 2 * ic) - 1, jo = 2 * (t.j - 2 * jc) - 1;
    for (int k = t.k0; k < t.k1; ++k) {
        const int kc = k >> 1, ko = 2 * (k - 2 * kc) - 1;
        T v = T(0.421875) * fmg_crse(Lc, crse, ic, jc, kc, phys);
        v += T(0.140625) * fmg_crse(Lc, crse, ic + io, jc, kc, phys);
        v += T(0.140625) * fmg_crse(Lc, crse, ic, jc + jo, kc, phys);
        v += T(0.140625) * fmg_crse(Lc, crse, ic, jc, kc + ko, phys);
        v += T(0.046875) * fmg_crse(Lc, crse, ic, jc + jo, kc + ko, phys);
        v += T(0.046875) * fmg_crse(Lc, crse, ic + io, jc, kc + ko, phys);
        v += T(0.046875) * fmg_crse(Lc, crse, ic + io, jc + jo, kc, phys);
        v += T(0.015625) * fmg_crse(Lc, crse, ic + io, jc + jo, kc + ko, phys);
        fine[cidx(F, t.i, t.j, k)] = v;
    }
}

// y += x on the valid cells (LocalAdd with no ghost cells).
template <class T>
__global__ __launch_bounds__(kBlock) void k_add_valid(LevelDev L, T* y, const T* x) {
    const BoxDev& B = L.box[blockIdx.y];
    const TileCells t = tile_cells(B, kTZ);
    if (!t.ok) return;
    for (int k = t.k0; k < t.k1; ++k) {
        const int64_t c = cidx(B, t.i, t.j, k);
        y[c] += x[c];
    }
}

// amrex_avgdown_faces (AMReX_MultiFabUtil_3D_C.H:173-215), ratio 2.
template <class T>
__global__ __launch_bounds__(kBlock) void k_avgdown_faces(LevelDev Lc, LevelDev Lf, T* bc, const T* bf, int d,
                                                        int same) {
    const BoxDev& C = Lc.box[blockIdx.y];
    const int nx = C.hi[0] - C.lo[0] + 1 + (d == 0), ny = C.hi[1] - C.lo[1] + 1 + (d == 1),
              nz = C.hi[2] - C.lo[2] + 1 + (d == 2);
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= (int64_t)nx * ny * nz) return;
    const int I = C.lo[0] + (int)(q % nx), J = C.lo[1] + (int)((q / nx) % ny), K = C.lo[2] + (int)(q / ((int64_t)nx * ny));
    // owner of the fine cell just inside the face (any fine box holding these faces stores equal values)
    int fi = 2 * I, fj = 2 * J, fk = 2 * K;
    if (d == 0 && I == C.hi[0] + 1) fi -= 1;
    if (d == 1 && J == C.hi[1] + 1) fj -= 1;
    if (d == 2 && K == C.hi[2] + 1) fk -= 1;
    const BoxDev& F = Lf.box[same ? (int)blockIdx.y : owner(Lf, fi, fj, fk)];
    T c = T(0.);
    if (d == 0) {
        for (int kr = 0; kr < 2; ++kr) for (int jr = 0; jr < 2; ++jr) c += bf[fidx(F, 0, 2 * I, 2 * J + jr, 2 * K + kr)];
    } else if (d == 1) {
        for (int kr = 0; kr < 2; ++kr) for (int ir = 0; ir < 2; ++ir) c += bf[fidx(F, 1, 2 * I + ir, 2 * J, 2 * K + kr)];
    } else {
        for (int jr = 0; jr < 2; ++jr) for (int ir = 0; ir < 2; ++ir) c += bf[fidx(F, 2, 2 * I + ir, 2 * J + jr, 2 * K)];
    }
    bc[fidx(C, d, I, J, K)] = c * (T(1.0) / T(4));
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_set_val(T* x, int64_t n, T v) {
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlock) x[q] = v;
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_axpy(T* y, const T* x, T a, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlock) y[q] += a * x[q];
}
template <class TD, class TS>
__global__ __launch_bounds__(kBlock) void k_convert(TD* y, const TS* x, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlock) y[q] = static_cast<TD>(x[q]);
}
template <class TD, class TS>
__global__ __launch_bounds__(kBlock) void k_add_convert(TD* y, const TS* x, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlock) y[q] += static_cast<TD>(x[q]);
}
template <class T>
__global__ __launch_bounds__(kBlock) void k_add_scaled_scalar(LevelDev L, T* x, const double* s, double factor) {
    const BoxDev& B = L.box[blockIdx.y];
    const TileCells t = tile_cells(B, kTZ);
    if (!t.ok) return;
    // offset = sum * volinv computed in T, then rhs += -offset (fixSolvabilityByOffset)
    const T offset = static_cast<T>(s[0]) * static_cast<T>(factor);
    for (int k = t.k0; k < t.k1; ++k) x[cidx(B, t.i, t.j, k)] += -offset;
}

// Max that propagates NaN (fmax would drop it and let a NaN field look converged).
__device__ __forceinline__ double nanmax(double a, double b) { return (b > a || b != b) ? b : a; }

// Block reduction of a double (64-wide wavefronts, then LDS across waves).
__device__ __forceinline__ double wave_reduce(double v, int op) {
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = (op == 0) ? nanmax(v, w) : v + w;
    }
    return v;
}
template <int NT>
__device__ __forceinline__ double block_reduce(double v, int op, double* lds) {
    v = wave_reduce(v, op);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double r = (op == 0) ? 0.0 : 0.0;
    if (threadIdx.x == 0) {
        r = lds[0];
        for (int w = 1; w < NT / 64; ++w) r = (op == 0) ? nanmax(r, lds[w]) : r + lds[w];
        lds[NT / 64] = r;
    }
    __syncthreads();
    return lds[NT / 64];
}

constexpr int kTZred = kTileZRed;

template <class T>
__global__ __launch_bounds__(kBlock) void k_reduce_partial(LevelDev L, int op, const T* x, const T* y, double* partials) {
    __shared__ double lds[kBlock / 64 + 1];
    const BoxDev& B = L.box[blockIdx.y];
    const TileCells t = tile_cells(B, kTZred);
    double v = 0.0;
    if (t.ok) {
        for (int k = t.k0; k < t.k1; ++k) {
            const int64_t c = cidx(B, t.i, t.j, k);
            if (op == 0) v = nanmax(v, fabs((double)x[c]));
            else if (op == 1) v += (double)x[c];
            else v += (double)x[c] * (double)y[c];
        }
    }
    v = block_reduce<kBlock>(v, op == 0 ? 0 : 1, lds);
    if (threadIdx.x == 0) partials[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = v;
}

__global__ __launch_bounds__(1024) void k_reduce_final(const double* partials, int64_t n, int op, double* out) {
    __shared__ double lds[1024 / 64 + 1];
    double v = 0.0;
    for (int64_t q = threadIdx.x; q < n; q += 1024) v = (op == 0) ? nanmax(v, partials[q]) : v + partials[q];
    v = block_reduce<1024>(v, op == 0 ? 0 : 1, lds);
    if (threadIdx.x == 0) out[0] = v;
}

template <class TL, class TG>
__global__ __launch_bounds__(kBlock) void k_scatter_cells(LevelDev L, TL* lvl, const TG* g, int grown, int to_level) {
    const BoxDev& B = L.box[blockIdx.y];
    // grown: iterate over the ghost-grown box, global array carries one ghost layer
    const int gx = grown ? 1 : 0;
    const int nx = B.hi[0] - B.lo[0] + 1 + 2 * gx, ny = B.hi[1] - B.lo[1] + 1 + 2 * gx,
              nz = B.hi[2] - B.lo[2] + 1 + 2 * gx;
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= (int64_t)nx * ny * nz) return;
    const int i = B.lo[0] - gx + (int)(q % nx), j = B.lo[1] - gx + (int)((q / nx) % ny),
              k = B.lo[2] - gx + (int)(q / ((int64_t)nx * ny));
    const int64_t G0 = L.n[0] + 2 * gx, G1 = L.n[1] + 2 * gx;
    const int64_t gi = (i + gx) + G0 * ((j + gx) + G1 * (int64_t)(k + gx));
    if (to_level) lvl[cidx(B, i, j, k)] = static_cast<TL>(g[gi]);
    else {
        // gather writes only valid cells (ghost cells of the global array belong to no box)
        if (i < B.lo[0] || i > B.hi[0] || j < B.lo[1] || j > B.hi[1] || k < B.lo[2] || k > B.hi[2]) return;
        const_cast<TG*>(g)[gi] = static_cast<TG>(lvl[cidx(B, i, j, k)]);
    }
}

template <class TL, class TG>
__global__ __launch_bounds__(kBlock) void k_scatter_faces(LevelDev L, TL* lvl, TG* g, int d, int to_level) {
    const BoxDev& B = L.box[blockIdx.y];
    const int nx = B.hi[0] - B.lo[0] + 1 + (d == 0), ny = B.hi[1] - B.lo[1] + 1 + (d == 1),
              nz = B.hi[2] - B.lo[2] + 1 + (d == 2);
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= (int64_t)nx * ny * nz) return;
    const int i = B.lo[0] + (int)(q % nx), j = B.lo[1] + (int)((q / nx) % ny), k = B.lo[2] + (int)(q / ((int64_t)nx * ny));
    const int64_t G0 = L.n[0] + (d == 0), G1 = L.n[1] + (d == 1);
    const int64_t gi = i + G0 * (j + G1 * (int64_t)k);
    if (to_level) lvl[fidx(B, d, i, j, k)] = static_cast<TL>(g[gi]);
    else g[gi] = static_cast<TG>(lvl[fidx(B, d, i, j, k)]);
}

// User FAB set -> level storage (or back), valid cells of each user box.  The
// level storage may use another box layout; the owning level box is looked up.
template <class TL, class TU>
__global__ __launch_bounds__(kBlock) void k_copy_user(LevelDev L, TL* lvl, TU* user, const UserBox* ub, int ung,
                                                    int to_level) {
    const UserBox& U = ub[blockIdx.y];
    const int nx = U.hi[0] - U.lo[0] + 1, ny = U.hi[1] - U.lo[1] + 1, nz = U.hi[2] - U.lo[2] + 1;
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= n) return;
    int di, dj, dk;
    split_index(q, n, nx, ny, di, dj, dk);
    const int i = U.lo[0] + di, j = U.lo[1] + dj, k = U.lo[2] + dk;
    const int64_t ux = nx + 2 * ung, uy = ny + 2 * ung;
    const int64_t u = U.off + (di + ung) + ux * ((dj + ung) + uy * (int64_t)(dk + ung));
    // a level stored as one FAB owns every cell
    const int64_t c = cidx(L.box[L.nbox == 1 ? 0 : owner(L, i, j, k)], i, j, k);
    if (to_level) lvl[c] = static_cast<TL>(user[u]);
    else user[u] = static_cast<TU>(lvl[c]);
}

// Face ghosts (one layer, cross stencil) of each user box from a level whose
// ghosts are filled: a position inside the domain, or the periodic image of
// one, takes the valid cell of the level box that owns it; a position beyond a
// physical face takes the ghost of the level box owning the cell inside.
// Edge and corner ghosts of the user box are left.  blockIdx.z = face.
__global__ __launch_bounds__(kBlock) void k_face_ghosts_to_user(LevelDev L, const double* lvl, double* user,
                                                                const UserBox* ub, int ung) {
    const UserBox& U = ub[blockIdx.y];
    const int f = (int)blockIdx.z, d = f % 3;
    const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    const int n1 = U.hi[t1] - U.lo[t1] + 1, n2 = U.hi[t2] - U.lo[t2] + 1;
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= (int64_t)n1 * n2) return;
    int p[3];
    p[d] = f < 3 ? U.lo[d] - 1 : U.hi[d] + 1;
    p[t1] = U.lo[t1] + (int)(q % n1);
    p[t2] = U.lo[t2] + (int)(q / n1);
    int c[3] = {p[0], p[1], p[2]};  // the cell whose owner holds the value
    bool ghost = false;
    if (c[d] < 0 || c[d] >= L.n[d]) {
        if (L.periodic[d]) {
            c[d] = c[d] < 0 ? c[d] + L.n[d] : c[d] - L.n[d];
        } else {
            ghost = true;
        }
    }
    // the level box holding the user box's cell next to the ghost; its ghost
    // layer has the value of a physical or coarse/fine face
    int in[3] = {p[0], p[1], p[2]};
    in[d] = f < 3 ? U.lo[d] : U.hi[d];
    const int ob = L.nbox == 1 ? 0 : owner(L, in[0], in[1], in[2]);
    if (ob < 0) return;
    // a covered cell (neighbour box or periodic image) holds its own value
    const int o = ghost ? -1 : (L.nbox == 1 ? 0 : owner(L, c[0], c[1], c[2]));
    const int64_t ux = U.hi[0] - U.lo[0] + 1 + 2 * ung, uy = U.hi[1] - U.lo[1] + 1 + 2 * ung;
    const int64_t u = U.off + (p[0] - U.lo[0] + ung) + ux * ((p[1] - U.lo[1] + ung) + uy * (int64_t)(p[2] - U.lo[2] + ung));
    user[u] = o >= 0 ? lvl[cidx(L.box[o], c[0], c[1], c[2])] : lvl[cidx(L.box[ob], p[0], p[1], p[2])];
}

// Zero the edge and corner cells of the first ghost layer of every user box
// (cells one outside the box in two or three directions): item q of a box
// runs over its 4(nx+ny+nz) edge cells, then its 8 corners.
__global__ __launch_bounds__(kBlock) void k_zero_user_edges(double* user, const UserBox* ub, int ung) {
    const UserBox& U = ub[blockIdx.y];
    const int n[3] = {U.hi[0] - U.lo[0] + 1, U.hi[1] - U.lo[1] + 1, U.hi[2] - U.lo[2] + 1};
    int q = (int)(blockIdx.x * kBlock + threadIdx.x);
    int p[3];  // position relative to lo, -1 .. n
    int d = 0;
    for (; d < 3; ++d) {
        if (q < 4 * n[d]) break;
        q -= 4 * n[d];
    }
    if (d < 3) {
        const int e = q / n[d], t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
        p[d] = q % n[d];
        p[t1] = (e & 1) ? n[t1] : -1;
        p[t2] = (e & 2) ? n[t2] : -1;
    } else {
        if (q >= 8) return;
        for (int c = 0; c < 3; ++c) p[c] = ((q >> c) & 1) ? n[c] : -1;
    }
    const int64_t ux = n[0] + 2 * ung, uy = n[1] + 2 * ung;
    user[U.off + (p[0] + ung) + ux * ((p[1] + ung) + uy * (int64_t)(p[2] + ung))] = 0.0;
}

// Ghost values at physical domain faces (one layer) from a user FAB set with
// ung >= 1: the inhomogeneous Dirichlet data of setLevelBC.
template <class TL, class TU>
__global__ __launch_bounds__(kBlock) void k_copy_user_grown(LevelDev L, TL* lvl, const TU* user, const UserBox* ub, int ung) {
    const UserBox& U = ub[blockIdx.y];
    const int nx = U.hi[0] - U.lo[0] + 3, ny = U.hi[1] - U.lo[1] + 3, nz = U.hi[2] - U.lo[2] + 3;
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= (int64_t)nx * ny * nz) return;
    int p[3] = {U.lo[0] - 1 + (int)(q % nx), U.lo[1] - 1 + (int)((q / nx) % ny), U.lo[2] - 1 + (int)(q / ((int64_t)nx * ny))};
    int out = 0, in[3];
    for (int d = 0; d < 3; ++d) {
        in[d] = min(max(p[d], 0), L.n[d] - 1);
        if (in[d] != p[d]) ++out;
    }
    if (out != 1) return;  // only face ghosts outside the domain carry boundary values
    // the cell inside the domain next to the ghost; a patch level may not own it
    // (an edge ghost of a box on the patch boundary): then no level box has the ghost
    const int o = owner(L, in[0], in[1], in[2]);
    if (o < 0) return;
    const int64_t ux = U.hi[0] - U.lo[0] + 1 + 2 * ung, uy = U.hi[1] - U.lo[1] + 1 + 2 * ung;
    const int64_t u = U.off + (p[0] - U.lo[0] + ung) + ux * ((p[1] - U.lo[1] + ung) + uy * (int64_t)(p[2] - U.lo[2] + ung));
    lvl[cidx(L.box[o], p[0], p[1], p[2])] = static_cast<TL>(user[u]);
}

// Face values of direction d from a user face FAB set (nodal boxes, ghost ung).
template <class TL, class TU>
__global__ __launch_bounds__(kBlock) void k_copy_user_faces(LevelDev L, TL* lvl, const TU* user, const UserBox* ub, int ung, int d) {
    const UserBox& U = ub[blockIdx.y];
    const int nx = U.hi[0] - U.lo[0] + 1, ny = U.hi[1] - U.lo[1] + 1, nz = U.hi[2] - U.lo[2] + 1;
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= (int64_t)nx * ny * nz) return;
    int p[3] = {U.lo[0] + (int)(q % nx), U.lo[1] + (int)((q / nx) % ny), U.lo[2] + (int)(q / ((int64_t)nx * ny))};
    const int64_t ux = nx + 2 * ung, uy = ny + 2 * ung;
    const int64_t u = U.off + (p[0] - U.lo[0] + ung) + ux * ((p[1] - U.lo[1] + ung) + uy * (int64_t)(p[2] - U.lo[2] + ung));
    // Every level box holds both faces of its cells, so a face on a level box
    // boundary has two copies: the one of the box owning the cell on its high
    // side and the one of the box owning the cell on its low side.
    // A level that covers part of the domain (a patch with coarse/fine
    // boundary data) owns no cell beyond the patch: the face on the patch
    // boundary then has the one copy of the box inside.
    const TL v = static_cast<TL>(user[u]);
    int c[3] = {p[0], p[1], p[2]};
    if (c[d] < L.n[d]) {
        const int o = owner(L, c[0], c[1], c[2]);
        if (o >= 0) lvl[fidx(L.box[o], d, p[0], p[1], p[2])] = v;
    }
    c[d] = p[d] - 1;
    if (c[d] >= 0) {
        const int o = owner(L, c[0], c[1], c[2]);
        if (o >= 0) lvl[fidx(L.box[o], d, p[0], p[1], p[2])] = v;
    }
}

// Face values of direction d on the valid faces of each user face box, the
// loops of MLCellLinOp::compGrad (AMReX_MLCellLinOp.H:1392-1437) and
// MLCellABecLap::getFluxes (AMReX_MLCellABecLap.H:272-289):
//   grad / Poisson:  v = c * (x(p) - x(p - e_d))
//   ABecLaplacian:   v = ((-c) * b(p)) * (x(p) - x(p - e_d)),  c = b_scalar * dxinv
// then v *= post when scale != 0.  Both cells are read from the ghost-filled FAB
// of the level box owning the cell of the face inside the user box, as the
// reference reads them from the user box's own FAB: that cell is valid in the
// level box, so its neighbour along d is at worst in the box's ghost layer.
template <class T>
__global__ __launch_bounds__(kBlock) void k_face_values(LevelDev L, const T* x, const T* b, const UserBox* ub, int ung,
                                                      int d, T c, int scale, T post, T* user) {
    const UserBox& U = ub[blockIdx.y];
    const int nx = U.hi[0] - U.lo[0] + 1, ny = U.hi[1] - U.lo[1] + 1, nz = U.hi[2] - U.lo[2] + 1;
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= (int64_t)nx * ny * nz) return;
    const int p[3] = {U.lo[0] + (int)(q % nx), U.lo[1] + (int)((q / nx) % ny), U.lo[2] + (int)(q / ((int64_t)nx * ny))};
    int in[3] = {p[0], p[1], p[2]};
    if (p[d] > U.lo[d]) in[d] = p[d] - 1;
    const BoxDev& B = L.box[owner(L, in[0], in[1], in[2])];
    const int64_t hi = cidx(B, p[0], p[1], p[2]);
    const int64_t lo = hi - (d == 0 ? 1 : (d == 1 ? B.sy : B.sz));
    const T diff = x[hi] - x[lo];
    T v = b ? ((-c) * b[fidx(B, d, p[0], p[1], p[2])]) * diff : c * diff;
    if (scale) v *= post;
    const int64_t ux = nx + 2 * ung, uy = ny + 2 * ung;
    user[U.off + (p[0] - U.lo[0] + ung) + ux * ((p[1] - U.lo[1] + ung) + uy * (int64_t)(p[2] - U.lo[2] + ung))] = v;
}

// User FAB set (any ghost width, any index type) <-> global array.
template <class TU, class TG>
__global__ __launch_bounds__(kBlock) void k_user_global(const UserBox* ub, int ung, TU* user, TG* glob, int G0, int G1,
                                                      int G2, int gg, int rg, int to_user) {
    const UserBox& B = ub[blockIdx.y];
    int lo[3], hi[3];
    const int G[3] = {G0, G1, G2};
    for (int d = 0; d < 3; ++d) {
        lo[d] = max(B.lo[d] - rg, -gg);
        hi[d] = min(B.hi[d] + rg, G[d] - 1 + gg);
    }
    const int64_t nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (nx <= 0 || ny <= 0 || nz <= 0 || q >= nx * ny * nz) return;
    const int i = lo[0] + (int)(q % nx), j = lo[1] + (int)((q / nx) % ny), k = lo[2] + (int)(q / (nx * ny));
    const int64_t ux = B.hi[0] - B.lo[0] + 1 + 2 * ung, uy = B.hi[1] - B.lo[1] + 1 + 2 * ung;
    const int64_t u = B.off + (i - B.lo[0] + ung) + ux * ((j - B.lo[1] + ung) + uy * (int64_t)(k - B.lo[2] + ung));
    const int64_t g = (i + gg) + (int64_t)(G0 + 2 * gg) * ((j + gg) + (int64_t)(G1 + 2 * gg) * (k + gg));
    if (to_user) user[u] = static_cast<TU>(glob[g]);
    else glob[g] = static_cast<TG>(user[u]);
}

// ---------------------------------------------------------------------------
// BiCGStab bottom solve in a single work-group
// ---------------------------------------------------------------------------
// Threads of the work-group: NB = 64 (one wave, barriers nearly free) for the
// usual few-cell bottom levels, 1024 for larger ones.
constexpr int kBotWide = 1024, kBotNarrow = 64;

__device__ __forceinline__ bool flat_cell(const LevelDev& L, int64_t q, int& b, int& i, int& j, int& k) {
    if (q >= L.ncell) return false;
    b = 0;
    while (L.cell_start[b + 1] <= q) ++b;
    return cell_of(L.box[b], q - L.cell_start[b], i, j, k);
}

template <class T, int NB>
__device__ void blk_fill(const LevelDev& L, T* phi) {
    __syncthreads();
    for (int b = 0; b < L.nbox; ++b) {
        const BoxDev& B = L.box[b];
        for (int f = 0; f < 6; ++f) {
            const int d = f % 3;
            const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
            const int nq = (B.hi[t1] - B.lo[t1] + 1) * (B.hi[t2] - B.lo[t2] + 1);
            for (int q = threadIdx.x; q < nq; q += NB) ghost_cell<T>(L, B, f, q, phi, nullptr, 0);
        }
    }
    __syncthreads();
}

template <class T, int NB, class F>
__device__ __forceinline__ void blk_each(const LevelDev& L, F f) {
    for (int64_t q = threadIdx.x; q < L.ncell; q += NB) {
        int b, i, j, k;
        if (flat_cell(L, q, b, i, j, k)) f(L.box[b], cidx(L.box[b], i, j, k), i, j, k);
    }
}

template <class T, int NB>
__device__ T blk_dot(const LevelDev& L, const T* x, const T* y, double* lds) {
    // T-precision accumulation per thread, then a fixed-order tree
    T v = T(0);
    blk_each<T, NB>(L, [&](const BoxDev&, int64_t c, int, int, int) { v += x[c] * y[c]; });
    return static_cast<T>(block_reduce<NB>((double)v, 1, lds));
}
template <class T, int NB>
__device__ T blk_norminf(const LevelDev& L, const T* x, double* lds) {
    T v = T(0);
    blk_each<T, NB>(L, [&](const BoxDev&, int64_t c, int, int, int) { v = (T)nanmax((double)v, fabs((double)x[c])); });
    return static_cast<T>(block_reduce<NB>((double)v, 0, lds));
}
// out = A in (TR: A^T in on a dense level, no ghosts needed), left-Jacobi
// scaled when NORM
template <class T, int NB, bool NORM, bool TR>
__device__ void blk_apply(const LevelDev& L, const OpCoef<T>& op, T* out, T* in) {
    if constexpr (TR) {
        __syncthreads();
        auto x = [&](int ii, int jj, int kk) -> T { return in[lt::cell(L.box[0], ii, jj, kk)]; };
        blk_each<T, NB>(L, [&](const BoxDev& B, int64_t c, int i, int j, int k) {
            T y = lt::apply_dense_T<T>(L, B, op, x, i, j, k);
            if (NORM && op.kind == 0) y /= diag_cell(B, op, i, j, k, c);
            out[c] = y;
        });
    } else {
        blk_fill<T, NB>(L, in);
        blk_each<T, NB>(L, [&](const BoxDev& B, int64_t c, int i, int j, int k) {
            T y = apply_cell(B, op, in, i, j, k, c);
            if (NORM && op.kind == 0) y /= diag_cell(B, op, i, j, k, c);
            out[c] = y;
        });
    }
    __syncthreads();
}

// Krylov bottom solves in one work-group on the bottom rhs bb with a zero
// initial guess; sol is zero on entry, the work arrays are the kernel's.
// Both return the reference's code and leave the iteration count in *it.
//
// BiCGStab (MLCGSolverT::solve_bicgstab, AMReX_MLCGSolver.H:98-273): left
// Jacobi scaling (normalize) for ABecLaplacian, codes 1-4, 8, 9.
template <class T, int NB, bool TR>
__device__ int blk_bicgstab(const LevelDev& L, const OpCoef<T>& op, T* sol, const T* bb, T* p, T* r, T* rh, T* v, T* t,
                            T eps_rel, T eps_abs, int maxiter, double* lds, int* it) {
    // r = D^-1 b, rh = r
    blk_each<T, NB>(L, [&](const BoxDev& B, int64_t c, int i, int j, int k) {
        T x = bb[c];
        if (op.kind == 0) x /= diag_cell(B, op, i, j, k, c);
        r[c] = x;
        rh[c] = x;
    });
    __syncthreads();
    T rnorm = blk_norminf<T, NB>(L, r, lds);
    const T rnorm0 = rnorm;
    int ret = 0, iter = 1;
    T rho_1 = 0, alpha = 0, omega = 0;
    if (rnorm0 == T(0) || rnorm0 < eps_abs) {
        *it = 0;
        return 0;
    }
    for (; iter <= maxiter; ++iter) {
        const T rho = blk_dot<T, NB>(L, rh, r, lds);
        if (rho == T(0)) { ret = 1; break; }
        if (iter == 1) {
            blk_each<T, NB>(L, [&](const BoxDev&, int64_t c, int, int, int) { p[c] = r[c]; });
        } else {
            const T beta = (rho / rho_1) * (alpha / omega);
            blk_each<T, NB>(L, [&](const BoxDev&, int64_t c, int, int, int) {
                T pc = p[c] + (-omega) * v[c];
                p[c] = r[c] + beta * pc;
            });
        }
        blk_apply<T, NB, true, TR>(L, op, v, p);
        const T rhTv = blk_dot<T, NB>(L, rh, v, lds);
        if (rhTv != T(0.0)) alpha = rho / rhTv;
        else { ret = 2; break; }
        blk_each<T, NB>(L, [&](const BoxDev&, int64_t c, int, int, int) {
            sol[c] += alpha * p[c];
            r[c] += (-alpha) * v[c];
        });
        __syncthreads();
        rnorm = blk_norminf<T, NB>(L, r, lds);
        if (rnorm < eps_rel * rnorm0 || rnorm < eps_abs) break;
        blk_apply<T, NB, true, TR>(L, op, t, r);
        const T tt = blk_dot<T, NB>(L, t, t, lds);
        const T tr = blk_dot<T, NB>(L, t, r, lds);
        if (tt != T(0.0)) omega = tr / tt;
        else { ret = 3; break; }
        blk_each<T, NB>(L, [&](const BoxDev&, int64_t c, int, int, int) {
            sol[c] += omega * r[c];
            r[c] += (-omega) * t[c];
        });
        __syncthreads();
        rnorm = blk_norminf<T, NB>(L, r, lds);
        if (rnorm < eps_rel * rnorm0 || rnorm < eps_abs) break;
        if (omega == T(0)) { ret = 4; break; }
        rho_1 = rho;
    }
    if (ret == 0 && rnorm > eps_rel * rnorm0 && rnorm > eps_abs) ret = 8;
    if ((ret == 0 || ret == 8) && (rnorm < rnorm0)) {
        if (ret == 8) ret = 9;
    } else {
        for (int64_t q = threadIdx.x; q < L.nalloc; q += NB) sol[q] = T(0);
    }
    __syncthreads();
    *it = iter;
    return ret;
}

// CG (MLCGSolverT::solve_cg, AMReX_MLCGSolver.H:275-410): plain conjugate
// gradients on A, no scaling; codes 1 (rho or p.Ap zero), 8, 9.
template <class T, int NB, bool TR>
__device__ int blk_cg(const LevelDev& L, const OpCoef<T>& op, T* sol, const T* bb, T* p, T* r, T* q, T eps_rel,
                      T eps_abs, int maxiter, double* lds, int* it) {
    blk_each<T, NB>(L, [&](const BoxDev&, int64_t c, int, int, int) { r[c] = bb[c]; });
    __syncthreads();
    T rnorm = blk_norminf<T, NB>(L, r, lds);
    const T rnorm0 = rnorm;
    int ret = 0, iter = 1;
    T rho_1 = 0;
    if (rnorm0 == T(0) || rnorm0 < eps_abs) {
        *it = 0;
        return 0;
    }
    for (; iter <= maxiter; ++iter) {
        const T rho = blk_dot<T, NB>(L, r, r, lds);
        if (rho == T(0)) { ret = 1; break; }
        if (iter == 1) {
            blk_each<T, NB>(L, [&](const BoxDev&, int64_t c, int, int, int) { p[c] = r[c]; });
        } else {
            const T beta = rho / rho_1;
            blk_each<T, NB>(L, [&](const BoxDev&, int64_t c, int, int, int) { p[c] = r[c] + beta * p[c]; });
        }
        blk_apply<T, NB, false, TR>(L, op, q, p);
        const T pw = blk_dot<T, NB>(L, p, q, lds);
        T alpha;
        if (pw != T(0.0)) alpha = rho / pw;
        else { ret = 1; break; }
        blk_each<T, NB>(L, [&](const BoxDev&, int64_t c, int, int, int) {
            sol[c] += alpha * p[c];
            r[c] += (-alpha) * q[c];
        });
        __syncthreads();
        rnorm = blk_norminf<T, NB>(L, r, lds);
        if (rnorm < eps_rel * rnorm0 || rnorm < eps_abs) break;
        rho_1 = rho;
    }
    if (ret == 0 && rnorm > eps_rel * rnorm0 && rnorm > eps_abs) ret = 8;
    if ((ret == 0 || ret == 8) && (rnorm < rnorm0)) {
        if (ret == 8) ret = 9;
    } else {
        for (int64_t c = threadIdx.x; c < L.nalloc; c += NB) sol[c] = T(0);
    }
    __syncthreads();
    *it = iter;
    return ret;
}

// The Krylov bottom solve of MLMGT::actualBottomSolve (AMReX_MLMG.H:1529-1601)
// with the iteration loops on the device.  status[2] holds the bottom solver
// in force (amrex::BottomSolver codes: 2 bicgstab, 3 cg, 4 bicgcg, 5 cgbicg)
// and persists across calls: bicgcg and cgbicg retry a failed solve from zero
// with the other method and, when the retry succeeds, keep that method (the
// code becomes 3 or 2).  status[0] = return code, status[1] = iterations of
// every attempt.
template <class T, int NB, bool TR>
__global__ __launch_bounds__(NB) void k_bottom(LevelDev L, OpCoef<T> op, T* sol, const T* rhs_in, T* work,
                                               T eps_rel, T eps_abs, int maxiter, int singular, double volinv,
                                               int nsm_ok, int nsm_fail, int* status) {
    __shared__ double lds[NB / 64 + 1];
    const int64_t na = L.nalloc;
    T* p = work;
    T* r = work + na;
    T* rh = work + 2 * na;
    T* v = work + 3 * na;
    T* t = work + 4 * na;
    T* bb = work + 5 * na;
    int mode = status[2];
    __syncthreads();  // every thread has read the code before thread 0 may rewrite it
    auto clear = [&]() {
        for (int64_t q = threadIdx.x; q < 5 * na; q += NB) work[q] = T(0);
        for (int64_t q = threadIdx.x; q < na; q += NB) sol[q] = T(0);
        __syncthreads();
    };
    for (int64_t q = threadIdx.x; q < 6 * na; q += NB) work[q] = T(0);
    for (int64_t q = threadIdx.x; q < na; q += NB) sol[q] = T(0);
    __syncthreads();
    // bottom rhs, made solvable if singular (MLMG.H:1535-1545, makeSolvable)
    blk_each<T, NB>(L, [&](const BoxDev&, int64_t c, int, int, int) { bb[c] = rhs_in[c]; });
    __syncthreads();
    if (singular) {
        T s = T(0);
        blk_each<T, NB>(L, [&](const BoxDev&, int64_t c, int, int, int) { s += bb[c]; });
        const T offset = static_cast<T>(block_reduce<NB>((double)s, 1, lds)) * static_cast<T>(volinv);
        blk_each<T, NB>(L, [&](const BoxDev&, int64_t c, int, int, int) { bb[c] += -offset; });
        __syncthreads();
    }
    auto run = [&](bool use_cg, int* it) {
        return use_cg ? blk_cg<T, NB, TR>(L, op, sol, bb, p, r, v, eps_rel, eps_abs, maxiter, lds, it)
                      : blk_bicgstab<T, NB, TR>(L, op, sol, bb, p, r, rh, v, t, eps_rel, eps_abs, maxiter, lds, it);
    };
    bool use_cg = mode == 3 || mode == 5;
    int it = 0, iters = 0;
    int ret = run(use_cg, &it);
    iters = it;
    if (ret != 0 && (mode == 4 || mode == 5)) {
        use_cg = !use_cg;
        clear();
        ret = run(use_cg, &it);
        iters += it;
        if (ret == 0) mode = use_cg ? 3 : 2;
    }
    // MLMG::actualBottomSolve (MLMG.H:1594-1601): zero on failure, then smooth
    if (ret != 0 && ret != 9) {
        for (int64_t q = threadIdx.x; q < na; q += NB) sol[q] = T(0);
    }
    const int nsm = (ret == 0) ? nsm_ok : nsm_fail;
    for (int s = 0; s < nsm; ++s) {
        for (int rb = 0; rb < 2; ++rb) {
            if constexpr (TR) {
                __syncthreads();
                blk_each<T, NB>(L, [&](const BoxDev& B, int64_t, int i, int j, int k) {
                    if (((i + j + k + rb) & 1) == 0) lt::gsrb_T_cell(L, B, op, sol, rhs_in, i, j, k);
                });
            } else {
                blk_fill<T, NB>(L, sol);
                blk_each<T, NB>(L, [&](const BoxDev& B, int64_t, int i, int j, int k) {
                    if (((i + j + k + rb) & 1) == 0) gsrb_cell(B, op, sol, rhs_in, i, j, k);
                });
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        status[0] = ret;
        status[1] = iters;
        status[2] = mode;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
int64_t reduce_partials(const LevelDev& L, int) {
    return (int64_t)std::max(L.mtiles[1], 1) * L.nbox;
}

template <class T>
void fill_ghosts(const LevelDev& L, T* phi, const T* bcv, int inhomog, hipStream_t s) {
    // one thread per ghost cell of the largest face (the grid-stride loop covers any remainder)
    const unsigned gx = (unsigned)std::min<int64_t>(std::max<int64_t>((L.max_face + kBlock - 1) / kBlock, 1), 4096);
    hipLaunchKernelGGL(k_fill_ghosts<T>, dim3(gx, L.nbox, 6), dim3(kBlock), 0, s, L, phi, bcv, inhomog);
}
template <class T>
void gsrb(const LevelDev& L, int mbc, const OpCoef<T>& op, T* phi, const T* rhs, int rb, hipStream_t s) {
    (void)mbc;
    hipLaunchKernelGGL(k_gsrb<T>, tile_grid(L), dim3(kBlock), 0, s, L, op, phi, rhs, rb);
}
template <class T>
void gsrb_T(const LevelDev& L, const OpCoef<T>& op, T* phi, const T* rhs, int rb, hipStream_t s) {
    if (L.nbox != 1) throw std::runtime_error("gsrb_T: the transposed sweep needs a dense level");
    const int64_t nb = (int64_t)((L.n[0] + kTPX - 1) / kTPX) * ((L.n[1] + kTY - 1) / kTY) * ((L.n[2] + kTZ - 1) / kTZ);
    hipLaunchKernelGGL(k_gsrb_T<T>, dim3((unsigned)nb), dim3(kBlock), 0, s, L, op, phi, rhs, rb);
}
template <class T>
void resid_T(const LevelDev& L, const OpCoef<T>& op, T* out, const T* x, const T* b, hipStream_t s) {
    if (L.nbox != 1) throw std::runtime_error("resid_T: the transposed operator needs a dense level");
    dim3 g = tile_grid(L);
    g.y = 1;
    hipLaunchKernelGGL(k_resid_T<T>, g, dim3(kBlock), 0, s, L, op, out, x, b);
}
template <class T>
void apply(const LevelDev& L, int mbc, const OpCoef<T>& op, T* out, const T* x, const T* b, hipStream_t s) {
    (void)mbc;
    hipLaunchKernelGGL(k_apply<T>, tile_grid(L), dim3(kBlock), 0, s, L, op, out, x, b);
}
template <class T>
void normalize(const LevelDev& L, int mbc, const OpCoef<T>& op, T* x, hipStream_t s) {
    if (op.kind != 0) return;  // MLPoisson::normalize is a no-op in 3D
    (void)mbc;
    hipLaunchKernelGGL(k_normalize<T>, tile_grid(L), dim3(kBlock), 0, s, L, op, x);
}
template <class T>
void restrict_avg(const LevelDev& Lc, int mbc, const LevelDev& Lf, T* crse, const T* fine, int same, hipStream_t s) {
    (void)mbc;
    hipLaunchKernelGGL(k_restrict<T>, tile_grid(Lc), dim3(kBlock), 0, s, Lc, Lf, crse, fine, same);
}
template <class T>
void prolong_add(const LevelDev& Lf, int mbc, const LevelDev& Lc, T* fine, const T* crse, int same, hipStream_t s) {
    (void)mbc;
    hipLaunchKernelGGL(k_prolong_add<T>, tile_grid(Lf), dim3(kBlock), 0, s, Lf, Lc, fine, crse, same);
}
template <class T>
void fmg_interp(const LevelDev& Lf, const LevelDev& Lc, T* fine, const T* crse, int phys, hipStream_t s) {
    hipLaunchKernelGGL(k_fmg_interp<T>, tile_grid(Lf), dim3(kBlock), 0, s, Lf, Lc, fine, crse, phys);
}
template <class T>
void add_valid(const LevelDev& L, T* y, const T* x, hipStream_t s) {
    hipLaunchKernelGGL(k_add_valid<T>, tile_grid(L), dim3(kBlock), 0, s, L, y, x);
}
template <class T>
void avgdown_faces(const LevelDev& Lc, const LevelDev& Lf, T* bc, const T* bf, int d, int same, hipStream_t s) {
    const int64_t mx = Lc.mbf;
    hipLaunchKernelGGL(k_avgdown_faces<T>, dim3((unsigned)((mx + kBlock - 1) / kBlock), Lc.nbox), dim3(kBlock), 0, s,
                       Lc, Lf, bc, bf, d, same);
}
template <class T>
void set_val(T* x, int64_t n, T v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_val<T>, flat_grid(n), dim3(kBlock), 0, s, x, n, v);
}
template <class T>
void axpy(T* y, const T* x, T a, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_axpy<T>, flat_grid(n), dim3(kBlock), 0, s, y, x, a, n);
}
template <class TD, class TS>
void convert(TD* y, const TS* x, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_convert<TD, TS>), flat_grid(n), dim3(kBlock), 0, s, y, x, n);
}
template <class TD, class TS>
void add_convert(TD* y, const TS* x, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_add_convert<TD, TS>), flat_grid(n), dim3(kBlock), 0, s, y, x, n);
}
template <class T>
void add_scaled_dev_scalar(const LevelDev& L, int mbc, T* x, const double* sd, double factor, hipStream_t s) {
    (void)mbc;
    hipLaunchKernelGGL(k_add_scaled_scalar<T>, tile_grid(L), dim3(kBlock), 0, s, L, x, sd, factor);
}
template <class T>
void reduce(const LevelDev& L, int mbc, int op, const T* x, const T* y, double* partials, double* out, hipStream_t s) {
    (void)mbc;
    const dim3 g = tile_grid(L, 1);
    hipLaunchKernelGGL(k_reduce_partial<T>, g, dim3(kBlock), 0, s, L, op, x, y, partials);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(1024), 0, s, partials, (int64_t)g.x * g.y, op, out);
}
template <class TL, class TG>
void scatter_cells(const LevelDev& L, int mbc, TL* lvl, const TG* g, int grown, hipStream_t s) {
    (void)mbc;
    const int64_t mx = L.mbg;
    hipLaunchKernelGGL((k_scatter_cells<TL, TG>), dim3((unsigned)((mx + kBlock - 1) / kBlock), L.nbox), dim3(kBlock), 0, s,
                       L, lvl, g, grown, 1);
}
template <class TL, class TG>
void gather_cells(const LevelDev& L, int mbc, const TL* lvl, TG* g, int grown, hipStream_t s) {
    (void)mbc;
    const int64_t mx = L.mbg;
    hipLaunchKernelGGL((k_scatter_cells<TL, TG>), dim3((unsigned)((mx + kBlock - 1) / kBlock), L.nbox), dim3(kBlock), 0, s,
                       L, const_cast<TL*>(lvl), g, grown, 0);
}
template <class TL, class TG>
void scatter_faces(const LevelDev& L, TL* lvl, const TG* g, int d, hipStream_t s) {
    const int64_t mx = L.mbf;
    hipLaunchKernelGGL((k_scatter_faces<TL, TG>), dim3((unsigned)((mx + kBlock - 1) / kBlock), L.nbox), dim3(kBlock), 0, s,
                       L, lvl, const_cast<TG*>(g), d, 1);
}
template <class TL, class TG>
void gather_faces(const LevelDev& L, const TL* lvl, TG* g, int d, hipStream_t s) {
    const int64_t mx = L.mbf;
    hipLaunchKernelGGL((k_scatter_faces<TL, TG>), dim3((unsigned)((mx + kBlock - 1) / kBlock), L.nbox), dim3(kBlock), 0, s,
                       L, const_cast<TL*>(lvl), g, d, 0);
}
template <class TL, class TU>
void copy_user_to_level(const LevelDev& L, TL* lvl, const TU* user, const UserBox* ub, int nbox, int64_t mx, int ung, hipStream_t s) {
    hipLaunchKernelGGL((k_copy_user<TL, TU>), dim3((unsigned)((mx + kBlock - 1) / kBlock), nbox), dim3(kBlock), 0, s, L, lvl,
                       const_cast<TU*>(user), ub, ung, 1);
}
template <class TL, class TU>
void copy_level_to_user(const LevelDev& L, const TL* lvl, TU* user, const UserBox* ub, int nbox, int64_t mx, int ung, hipStream_t s) {
    hipLaunchKernelGGL((k_copy_user<TL, TU>), dim3((unsigned)((mx + kBlock - 1) / kBlock), nbox), dim3(kBlock), 0, s, L,
                       const_cast<TL*>(lvl), user, ub, ung, 0);
}
void face_ghosts_to_user(const LevelDev& L, const double* lvl, double* user, const UserBox* ub, int nbox,
                         int64_t max_face, int ung, hipStream_t s) {
    if (nbox <= 0 || max_face <= 0 || ung < 1) return;
    hipLaunchKernelGGL(k_face_ghosts_to_user, dim3((unsigned)((max_face + kBlock - 1) / kBlock), nbox, 6), dim3(kBlock),
                       0, s, L, lvl, user, ub, ung);
}
void zero_user_edge_ghosts(double* user, const UserBox* ub, int nbox, int64_t max_items, int ung, hipStream_t s) {
    if (nbox <= 0 || max_items <= 0 || ung < 1) return;
    hipLaunchKernelGGL(k_zero_user_edges, dim3((unsigned)((max_items + kBlock - 1) / kBlock), nbox), dim3(kBlock), 0, s,
                       user, ub, ung);
}
template <class TL, class TU>
void copy_user_grown_to_level(const LevelDev& L, TL* lvl, const TU* user, const UserBox* ub, int nbox, int64_t mx, int ung, hipStream_t s) {
    hipLaunchKernelGGL((k_copy_user_grown<TL, TU>), dim3((unsigned)((mx + kBlock - 1) / kBlock), nbox), dim3(kBlock), 0, s,
                       L, lvl, user, ub, ung);
}
template <class TL, class TU>
void copy_user_faces_to_level(const LevelDev& L, TL* lvl, const TU* user, const UserBox* ub, int nbox, int64_t mx, int ung, int d, hipStream_t s) {
    hipLaunchKernelGGL((k_copy_user_faces<TL, TU>), dim3((unsigned)((mx + kBlock - 1) / kBlock), nbox), dim3(kBlock), 0, s,
                       L, lvl, user, ub, ung, d);
}
template <class T>
void face_values(const LevelDev& L, const T* x, const T* b, const UserBox* ub, int nbox, int64_t mx, int ung, int d, T c,
                 int scale, T post, T* user, hipStream_t s) {
    hipLaunchKernelGGL(k_face_values<T>, dim3((unsigned)((mx + kBlock - 1) / kBlock), nbox), dim3(kBlock), 0, s, L, x, b,
                       ub, ung, d, c, scale, post, user);
}
template void face_values<double>(const LevelDev&, const double*, const double*, const UserBox*, int, int64_t, int, int,
                                  double, int, double, double*, hipStream_t);
template <class TU, class TG>
void user_global(const UserBox* ub_dev, int nbox, int64_t max_region, int ung, TU* user, TG* glob, const int* G, int gg,
                 int rg, int to_user, hipStream_t s) {
    hipLaunchKernelGGL((k_user_global<TU, TG>), dim3((unsigned)((max_region + kBlock - 1) / kBlock), nbox), dim3(kBlock), 0, s,
                       ub_dev, ung, user, glob, G[0], G[1], G[2], gg, rg, to_user);
}
template void user_global<double, double>(const UserBox*, int, int64_t, int, double*, double*, const int*, int, int, int, hipStream_t);
template void user_global<float, double>(const UserBox*, int, int64_t, int, float*, double*, const int*, int, int, int, hipStream_t);

template <class T>
void bottom_solve(const LevelDev& L, const OpCoef<T>& op, T* x, const T* b, T* work, T eps_rel, T eps_abs,
                  int maxiter, int singular, double volinv, int nsm_ok, int nsm_fail, int* status, int transpose,
                  hipStream_t s) {
    if (transpose && L.nbox != 1) throw std::runtime_error("bottom_solve: the transposed operator needs a dense level");
    auto go = [&](auto tr) {
        constexpr bool TR = decltype(tr)::value;
        if (L.nalloc <= 4096)
            hipLaunchKernelGGL((k_bottom<T, kBotNarrow, TR>), dim3(1), dim3(kBotNarrow), 0, s, L, op, x, b, work,
                               eps_rel, eps_abs, maxiter, singular, volinv, nsm_ok, nsm_fail, status);
        else
            hipLaunchKernelGGL((k_bottom<T, kBotWide, TR>), dim3(1), dim3(kBotWide), 0, s, L, op, x, b, work,
                               eps_rel, eps_abs, maxiter, singular, volinv, nsm_ok, nsm_fail, status);
    };
    if (transpose)
        go(std::true_type{});
    else
        go(std::false_type{});
}

template void set_val<int>(int*, int64_t, int, hipStream_t);

#define MGX_INST(T)                                                                                          \
    template void fill_ghosts<T>(const LevelDev&, T*, const T*, int, hipStream_t);                           \
    template void gsrb<T>(const LevelDev&, int, const OpCoef<T>&, T*, const T*, int, hipStream_t);           \
    template void apply<T>(const LevelDev&, int, const OpCoef<T>&, T*, const T*, const T*, hipStream_t);     \
    template void normalize<T>(const LevelDev&, int, const OpCoef<T>&, T*, hipStream_t);                     \
    template void restrict_avg<T>(const LevelDev&, int, const LevelDev&, T*, const T*, int, hipStream_t);    \
    template void prolong_add<T>(const LevelDev&, int, const LevelDev&, T*, const T*, int, hipStream_t);     \
    template void fmg_interp<T>(const LevelDev&, const LevelDev&, T*, const T*, int, hipStream_t);           \
    template void add_valid<T>(const LevelDev&, T*, const T*, hipStream_t);                                   \
    template void avgdown_faces<T>(const LevelDev&, const LevelDev&, T*, const T*, int, int, hipStream_t);   \
    template void set_val<T>(T*, int64_t, T, hipStream_t);                                                   \
    template void axpy<T>(T*, const T*, T, int64_t, hipStream_t);                                            \
    template void add_scaled_dev_scalar<T>(const LevelDev&, int, T*, const double*, double, hipStream_t);    \
    template void reduce<T>(const LevelDev&, int, int, const T*, const T*, double*, double*, hipStream_t);   \
    template void gsrb_T<T>(const LevelDev&, const OpCoef<T>&, T*, const T*, int, hipStream_t);            \
    template void resid_T<T>(const LevelDev&, const OpCoef<T>&, T*, const T*, const T*, hipStream_t);      \
    template void bottom_solve<T>(const LevelDev&, const OpCoef<T>&, T*, const T*, T*, T, T, int, int,    \
                                     double, int, int, int*, int, hipStream_t);
MGX_INST(double)
MGX_INST(float)

#define MGX_INST2(TA, TB)                                                                                    \
    template void convert<TA, TB>(TA*, const TB*, int64_t, hipStream_t);                                     \
    template void add_convert<TA, TB>(TA*, const TB*, int64_t, hipStream_t);                                 \
    template void scatter_cells<TA, TB>(const LevelDev&, int, TA*, const TB*, int, hipStream_t);             \
    template void gather_cells<TA, TB>(const LevelDev&, int, const TA*, TB*, int, hipStream_t);              \
    template void scatter_faces<TA, TB>(const LevelDev&, TA*, const TB*, int, hipStream_t);                  \
    template void gather_faces<TA, TB>(const LevelDev&, const TA*, TB*, int, hipStream_t);                   \
    template void copy_user_to_level<TA, TB>(const LevelDev&, TA*, const TB*, const UserBox*, int, int64_t,  \
                                             int, hipStream_t);                                              \
    template void copy_level_to_user<TA, TB>(const LevelDev&, const TA*, TB*, const UserBox*, int, int64_t,  \
                                             int, hipStream_t);                                              \
    template void copy_user_grown_to_level<TA, TB>(const LevelDev&, TA*, const TB*, const UserBox*, int,     \
                                                   int64_t, int, hipStream_t);                               \
    template void copy_user_faces_to_level<TA, TB>(const LevelDev&, TA*, const TB*, const UserBox*, int,     \
                                                   int64_t, int, int, hipStream_t);
MGX_INST2(double, double)
MGX_INST2(float, float)
MGX_INST2(double, float)
MGX_INST2(float, double)

}  // namespace mgx
