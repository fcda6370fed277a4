// The transposed level operator L^T on dense levels (one FAB covering the
// domain, box from 0), homogeneous boundary data: the fine residual of the
// adjoint solve and the adjoint's V-cycles (sweeps, residuals of corrections,
// the bottom solve).  No reference counterpart.
#pragma once
#include <hip/hip_runtime.h>

#include "mgx_device.H"

namespace mgx {
namespace lt {

__device__ __forceinline__ int64_t cell(const BoxDev& B, int i, int j, int k) {
    return B.off + (i + 1) + B.sy * (j + 1) + B.sz * (k + 1);
}

// Face value b_d at face index (i,j,k) (index along d runs 0..n_d).
template <class T>
__device__ __forceinline__ T bface(const BoxDev& B, const OpCoef<T>& op, int n0, int n1, int d, int i, int j, int k) {
    if (d == 0) return op.b[0][B.foff[0] + i + (int64_t)(n0 + 1) * (j + (int64_t)n1 * k)];
    if (d == 1) return op.b[1][B.foff[1] + i + (int64_t)n0 * (j + (int64_t)(n1 + 1) * k)];
    return op.b[2][B.foff[2] + i + (int64_t)n0 * (j + (int64_t)n1 * k)];
}

// (L^T x) at a real cell, homogeneous boundary data.  Row p of L is
// diag0 + sum over its six sides of w*(x_p - x_side) once each ghost is
// replaced by its closure (w = dh*b_face for ABecLaplacian, -dh for Poisson).
// In the transpose, a side inside the domain is unchanged (the face is shared);
// across a periodic wrap the off-diagonal takes w of the image cell's face on
// the far end; a physical face keeps only the closure's diagonal part
// w*(1 - c_1); and the Dirichlet weights c_m (m >= 2) of the boundary row move
// to the rows of the cells they read.
template <class T, class F>
__device__ __forceinline__ T apply_dense_T(const LevelDev& L, const BoxDev& B, const OpCoef<T>& op, F x, int i, int j,
                                           int k) {
    const int n0 = L.n[0], n1 = L.n[1];
    const int p[3] = {i, j, k};
    const T xc = x(i, j, k);
    T y = (op.kind == 0) ? op.alpha * op.a[cell(B, i, j, k)] * xc : T(0);
    // the cell's own faces: face index p (low side) and p + e_d (high side)
    int64_t of[3] = {0, 0, 0}, fs[3] = {1, n0, (int64_t)n0 * n1};
    if (op.kind == 0) {
        of[0] = B.foff[0] + i + (int64_t)(n0 + 1) * (j + (int64_t)n1 * k);
        of[1] = B.foff[1] + i + (int64_t)n0 * (j + (int64_t)(n1 + 1) * k);
        of[2] = B.foff[2] + i + (int64_t)n0 * (j + (int64_t)n1 * k);
    }
    for (int d = 0; d < 3; ++d) {
        const T dh = (op.kind == 0) ? op.dha[d] : op.dpo[d];
        auto wof = [&](const int* fq) { return (op.kind == 0) ? dh * bface(B, op, n0, n1, d, fq[0], fq[1], fq[2]) : -dh; };
        for (int side = 0; side < 2; ++side) {
            const T w = (op.kind == 0) ? dh * op.b[d][of[d] + side * fs[d]] : -dh;
            int q[3] = {p[0], p[1], p[2]};
            q[d] += side ? 1 : -1;
            if (q[d] >= 0 && q[d] < L.n[d]) {
                y += w * (xc - x(q[0], q[1], q[2]));
            } else if (L.periodic[d]) {
                q[d] = side ? 0 : L.n[d] - 1;
                int fo[3] = {p[0], p[1], p[2]};
                fo[d] = side ? 0 : L.n[d];
                y += w * xc - wof(fo) * x(q[0], q[1], q[2]);
            } else {
                const int f = d + 3 * side;
                const int bct = B.bct[f];
                const T c1 = bct == BC_DIRICHLET ? T(B.bcc[f][1])
                             : bct == BC_NEUMANN ? T(1) : bct == BC_REFLECT_ODD ? T(-1) : T(0);
                y += w * (T(1) - c1) * xc;
            }
        }
        if (L.periodic[d]) continue;
        for (int side = 0; side < 2; ++side) {
            const int f = d + 3 * side;
            if (B.bct[f] != BC_DIRICHLET) continue;
            const int t = side ? L.n[d] - 1 - p[d] : p[d];  // distance from the boundary row
            if (t < 1 || t > B.nxo[f] - 2) continue;
            int r[3] = {p[0], p[1], p[2]};
            r[d] = side ? L.n[d] - 1 : 0;
            int fq[3] = {r[0], r[1], r[2]};
            fq[d] += side;
            y += -wof(fq) * T(B.bcc[f][t + 1]) * x(r[0], r[1], r[2]);
        }
    }
    return y;
}

// The forward sweep's relaxation divisor at a cell (g_m_d of abec_gsrb,
// AMReX_MLABecLap_3D_K.H:210-264, and mlpoisson_gsrb, AMReX_MLPoisson_3D_K.H:
// 155-196), the divisor of the transposed sweep.
template <class T>
__device__ __forceinline__ T gmd(const LevelDev& L, const BoxDev& B, const OpCoef<T>& op, int i, int j, int k) {
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    const T cf0 = (i == 0 && B.bct[0] != BC_INTERIOR) ? T(B.cf[0]) : T(0.0);
    const T cf1 = (j == 0 && B.bct[1] != BC_INTERIOR) ? T(B.cf[1]) : T(0.0);
    const T cf2 = (k == 0 && B.bct[2] != BC_INTERIOR) ? T(B.cf[2]) : T(0.0);
    const T cf3 = (i == n0 - 1 && B.bct[3] != BC_INTERIOR) ? T(B.cf[3]) : T(0.0);
    const T cf4 = (j == n1 - 1 && B.bct[4] != BC_INTERIOR) ? T(B.cf[4]) : T(0.0);
    const T cf5 = (k == n2 - 1 && B.bct[5] != BC_INTERIOR) ? T(B.cf[5]) : T(0.0);
    if (op.kind == 0) {
        const T dhx = op.dhs[0], dhy = op.dhs[1], dhz = op.dhs[2];
        const T bxm = bface(B, op, n0, n1, 0, i, j, k), bxp = bface(B, op, n0, n1, 0, i + 1, j, k);
        const T bym = bface(B, op, n0, n1, 1, i, j, k), byp = bface(B, op, n0, n1, 1, i, j + 1, k);
        const T bzm = bface(B, op, n0, n1, 2, i, j, k), bzp = bface(B, op, n0, n1, 2, i, j, k + 1);
        const T gamma = op.alpha * op.a[cell(B, i, j, k)] + dhx * (bxm + bxp) + dhy * (bym + byp) + dhz * (bzm + bzp);
        return gamma - (dhx * (bxm * cf0 + bxp * cf3) + dhy * (bym * cf1 + byp * cf4) + dhz * (bzm * cf2 + bzp * cf5));
    }
    const T dhx = op.dpo[0], dhy = op.dpo[1], dhz = op.dpo[2];
    const T gamma = T(-2.) * (dhx + dhy + dhz);
    return gamma + dhx * (cf0 + cf3) + dhy * (cf1 + cf4) + dhz * (cf2 + cf5);
}

// One cell of the transposed sweep: x_p += omega / g_m_d (b_p - (L^T x)_p).
// With maxorder <= 3 a row of L^T reads cells of the other colour only.  The
// cell's own six face coefficients are loaded once for both (L^T x)_p and
// g_m_d; the arithmetic is apply_dense_T's and gmd's, in their order.
template <class T>
__device__ __forceinline__ void gsrb_T_cell(const LevelDev& L, const BoxDev& B, const OpCoef<T>& op, T* phi,
                                            const T* rhs, int i, int j, int k) {
    constexpr T omega = T(1.15);
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    const int p[3] = {i, j, k};
    const int64_t c = cell(B, i, j, k);
    const int64_t st[3] = {1, B.sy, B.sz};
    // cells are read at offsets from c along one direction
    auto x = [&](int ii, int jj, int kk) -> T {
        return phi[c + (ii - i) + st[1] * (jj - j) + st[2] * (kk - k)];
    };
    const bool abec = op.kind == 0;
    T bb[3][2];
    if (abec) {
        const int64_t ox = B.foff[0] + i + (int64_t)(n0 + 1) * (j + (int64_t)n1 * k);
        const int64_t oy = B.foff[1] + i + (int64_t)n0 * (j + (int64_t)(n1 + 1) * k);
        const int64_t oz = B.foff[2] + i + (int64_t)n0 * (j + (int64_t)n1 * k);
        bb[0][0] = op.b[0][ox];
        bb[0][1] = op.b[0][ox + 1];
        bb[1][0] = op.b[1][oy];
        bb[1][1] = op.b[1][oy + n0];
        bb[2][0] = op.b[2][oz];
        bb[2][1] = op.b[2][oz + (int64_t)n0 * n1];
    } else {
        for (int d = 0; d < 3; ++d) bb[d][0] = bb[d][1] = T(0);
    }
    const T av = abec ? op.a[c] : T(0);
    // (L^T x)_p, as apply_dense_T
    const T xc = phi[c];
    T y = abec ? op.alpha * av * xc : T(0);
    for (int d = 0; d < 3; ++d) {
        const T dh = abec ? op.dha[d] : op.dpo[d];
        auto wof = [&](const int* fq) { return abec ? dh * bface(B, op, n0, n1, d, fq[0], fq[1], fq[2]) : -dh; };
        for (int side = 0; side < 2; ++side) {
            const T w = abec ? dh * bb[d][side] : -dh;
            int q[3] = {p[0], p[1], p[2]};
            q[d] += side ? 1 : -1;
            if (q[d] >= 0 && q[d] < L.n[d]) {
                y += w * (xc - phi[c + (side ? st[d] : -st[d])]);
            } else if (L.periodic[d]) {
                q[d] = side ? 0 : L.n[d] - 1;
                int fo[3] = {p[0], p[1], p[2]};
                fo[d] = side ? 0 : L.n[d];
                y += w * xc - wof(fo) * x(q[0], q[1], q[2]);
            } else {
                const int f = d + 3 * side;
                const int bct = B.bct[f];
                const T c1 = bct == BC_DIRICHLET ? T(B.bcc[f][1])
                             : bct == BC_NEUMANN ? T(1) : bct == BC_REFLECT_ODD ? T(-1) : T(0);
                y += w * (T(1) - c1) * xc;
            }
        }
        if (L.periodic[d]) continue;
        for (int side = 0; side < 2; ++side) {
            const int f = d + 3 * side;
            if (B.bct[f] != BC_DIRICHLET) continue;
            const int t = side ? L.n[d] - 1 - p[d] : p[d];  // distance from the boundary row
            if (t < 1 || t > B.nxo[f] - 2) continue;
            int r[3] = {p[0], p[1], p[2]};
            r[d] = side ? L.n[d] - 1 : 0;
            int fq[3] = {r[0], r[1], r[2]};
            fq[d] += side;
            y += -wof(fq) * T(B.bcc[f][t + 1]) * x(r[0], r[1], r[2]);
        }
    }
    // g_m_d, as gmd
    const T cf0 = (i == 0 && B.bct[0] != BC_INTERIOR) ? T(B.cf[0]) : T(0.0);
    const T cf1 = (j == 0 && B.bct[1] != BC_INTERIOR) ? T(B.cf[1]) : T(0.0);
    const T cf2 = (k == 0 && B.bct[2] != BC_INTERIOR) ? T(B.cf[2]) : T(0.0);
    const T cf3 = (i == n0 - 1 && B.bct[3] != BC_INTERIOR) ? T(B.cf[3]) : T(0.0);
    const T cf4 = (j == n1 - 1 && B.bct[4] != BC_INTERIOR) ? T(B.cf[4]) : T(0.0);
    const T cf5 = (k == n2 - 1 && B.bct[5] != BC_INTERIOR) ? T(B.cf[5]) : T(0.0);
    T g_m_d;
    if (abec) {
        const T dhx = op.dhs[0], dhy = op.dhs[1], dhz = op.dhs[2];
        const T bxm = bb[0][0], bxp = bb[0][1], bym = bb[1][0], byp = bb[1][1], bzm = bb[2][0], bzp = bb[2][1];
        const T gamma = op.alpha * av + dhx * (bxm + bxp) + dhy * (bym + byp) + dhz * (bzm + bzp);
        g_m_d = gamma - (dhx * (bxm * cf0 + bxp * cf3) + dhy * (bym * cf1 + byp * cf4) + dhz * (bzm * cf2 + bzp * cf5));
    } else {
        const T dhx = op.dpo[0], dhy = op.dpo[1], dhz = op.dpo[2];
        const T gamma = T(-2.) * (dhx + dhy + dhz);
        g_m_d = gamma + dhx * (cf0 + cf3) + dhy * (cf1 + cf4) + dhz * (cf2 + cf5);
    }
    phi[c] = phi[c] + omega / g_m_d * (rhs[c] - y);
}

}  // namespace lt
}  // namespace mgx
