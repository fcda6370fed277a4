// Fused red+black GSRB sweep for dense MG levels (one FAB covering the
// domain), one pass over HBM per sweep.
//
// The reference applies one colour per launch with a ghost refresh between
// the colours (MLCellLinOpT::smooth, AMReX_MLCellLinOp.H:1205-1217; abec_gsrb,
// AMReX_MLABecLap_3D_K.H:210-264; mlpoisson_gsrb, AMReX_MLPoisson_3D_K.H:
// 155-196), which streams every array twice per sweep.  Here a work-group owns
// a TX x TY column of cells over a range of z planes and marches up in z with
// four planes of the solution in LDS:
//   step k:  load plane k+2 (old values) -> red update of plane k+1 on the tile
//            plus a one-cell ring -> black update of plane k on the tile ->
//            write plane k.
// Red values of the ring are recomputed redundantly from old data, so tiles
// never wait for each other; the sweep is out of place (reads pin, writes
// pout), which keeps every tile's inputs unmodified while other tiles run.
// Ghost values follow the reference exactly: before the red update of a cell
// they come from old values (physical closure or periodic image), before the
// black update from red-updated values.  Ghost cells beyond a physical face are
// closures evaluated in LDS just before they are read, from the two interior
// planes/columns along the normal, which hold the right old/red-updated values
// at the positions that read them (colours alternate along the normal).
#include <hip/hip_runtime.h>

#include "mgx_device.H"
#include "mgx_smoother.H"

namespace mgx {
namespace {

constexpr int kSmBlock = 256;

// Classification of a coordinate along one direction of the tile region:
// 0 real cell (after the periodic wrap), 1 ghost below the domain, 2 ghost
// above it, 3 not needed.
__device__ __forceinline__ int classify(int x, int n, int periodic, int& xw) {
    if (periodic) {
        xw = x % n;
        if (xw < 0) xw += n;
        return 0;
    }
    xw = x;
    if (x >= 0 && x < n) return 0;
    if (x == -1) return 1;
    if (x == n) return 2;
    return 3;
}

// Closure of a ghost cell from interior values v1 (first cell inside) and v2
// (second cell inside): mllinop_apply_bc_* (AMReX_MLLinOp_K.H:14-115), homogeneous.
template <class T>
__device__ __forceinline__ T closure(const BoxDev& B, int f, T v1, T v2) {
    const int bct = B.bct[f];
    if (bct == BC_NEUMANN) return v1;
    if (bct == BC_REFLECT_ODD) return -v1;
    const int NX = B.nxo[f];
    T tmp = T(0.0);
    if (NX > 1) tmp += v1 * T(B.bcc[f][1]);
    if (NX > 2) tmp += v2 * T(B.bcc[f][2]);
    return tmp;
}

template <class T, int TX, int TY>
__global__ __launch_bounds__(kSmBlock) void k_gsrb_sweep(LevelDev L, OpCoef<T> op, const T* __restrict__ pin,
                                                         T* __restrict__ pout, const T* __restrict__ rhs, int cz,
                                                         int zero_in, const T* __restrict__ crse, LevelDev Lc) {
    constexpr int RX = TX + 4, RY = TY + 4, RS = RX * RY;
    __shared__ T S[4][RS];
    const BoxDev& B = L.box[0];
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    const int p0 = L.periodic[0], p1 = L.periodic[1], p2 = L.periodic[2];
    const int ntx = (n0 + TX - 1) / TX, nty = (n1 + TY - 1) / TY;
    // XCD-aware order: consecutive work-groups (round-robin over the 8 XCDs)
    // take tiles that are far apart, so each XCD's L2 sees neighbouring tiles.
    const int nblk = gridDim.x;
    const int bid = (nblk % 8 == 0) ? (int)(blockIdx.x % 8) * (nblk / 8) + (int)(blockIdx.x / 8) : (int)blockIdx.x;
    const int tx = bid % ntx, ty = (bid / ntx) % nty, tz = bid / (ntx * nty);
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * cz, z1 = min(z0 + cz, n2);
    const int tid = threadIdx.x;
    const int64_t sy = B.sy, sz = B.sz;
    const int64_t fx1 = n0 + 1, fy1 = n1 + 1;

    auto slot = [](int z) { return (z + 4) & 3; };
    // global FAB index of a real cell (dense level: box lo = 0)
    auto gidx = [&](int x, int y, int z) -> int64_t { return B.off + (x + 1) + sy * (y + 1) + sz * (z + 1); };
    auto zclass = [&](int z, int& zw) { return classify(z, n2, p2, zw); };

    // Old value of a real cell, with the coarse correction added when this sweep
    // also performs the prolongation (fine += crse(i/2,j/2,k/2)).
    auto old_value = [&](int xw, int yw, int zw) -> T {
        T v = zero_in ? T(0) : pin[gidx(xw, yw, zw)];
        if (crse) {
            const BoxDev& C = Lc.box[0];
            v += crse[C.off + ((xw >> 1) + 1) + C.sy * ((yw >> 1) + 1) + C.sz * ((zw >> 1) + 1)];
        }
        return v;
    };

    auto load_plane = [&](int z) {
        int zw;
        if (zclass(z, zw) != 0) return;
        T* P = S[slot(z)];
        for (int p = tid; p < RS; p += kSmBlock) {
            const int lx = p % RX, ly = p / RX;
            int xw, yw;
            const int cx = classify(x0 - 2 + lx, n0, p0, xw);
            const int cy = classify(y0 - 2 + ly, n1, p1, yw);
            P[p] = (cx == 0 && cy == 0) ? old_value(xw, yw, zw) : T(0);
        }
    };

    // x/y ghosts of plane z from the plane's own interior values.
    auto refresh_xy = [&](int z) {
        int zw;
        if (zclass(z, zw) != 0) return;
        T* P = S[slot(z)];
        for (int p = tid; p < RS; p += kSmBlock) {
            const int lx = p % RX, ly = p / RX;
            int xw, yw;
            const int cx = classify(x0 - 2 + lx, n0, p0, xw);
            const int cy = classify(y0 - 2 + ly, n1, p1, yw);
            if (cx != 0 && cy == 0 && cx != 3) {
                const int s = cx == 1 ? 1 : -1;
                P[p] = closure(B, cx == 1 ? 0 : 3, P[p + s], P[p + 2 * s]);
            } else if (cy != 0 && cx == 0 && cy != 3) {
                const int s = cy == 1 ? RX : -RX;
                P[p] = closure(B, cy == 1 ? 1 : 4, P[p + s], P[p + 2 * s]);
            }
        }
    };

    // ghost plane below/above the domain from the two interior planes
    auto refresh_zghost = [&](int z) {
        int zw;
        const int c = zclass(z, zw);
        if (c != 1 && c != 2) return;
        const int s = c == 1 ? 1 : -1;
        const T* P1 = S[slot(z + s)];
        const T* P2 = S[slot(z + 2 * s)];
        T* G = S[slot(z)];
        for (int p = tid; p < RS; p += kSmBlock) {
            const int lx = p % RX, ly = p / RX;
            int xw, yw;
            const int cx = classify(x0 - 2 + lx, n0, p0, xw);
            const int cy = classify(y0 - 2 + ly, n1, p1, yw);
            if (cx == 0 && cy == 0) G[p] = closure(B, c == 1 ? 2 : 5, P1[p], n2 > 1 ? P2[p] : T(0));
        }
    };

    // One GSRB-SOR update of the cell at LDS position p of plane z (global xw, yw, zw).
    auto update = [&](int p, int z, int xw, int yw, int zw) {
        constexpr T omega = T(1.15);
        T* P = S[slot(z)];
        const T* Pm = S[slot(z - 1)];
        const T* Pp = S[slot(z + 1)];
        const int64_t c = gidx(xw, yw, zw);
        const T cf0 = (xw == 0 && B.bct[0] != BC_INTERIOR) ? T(B.cf[0]) : T(0.0);
        const T cf1 = (yw == 0 && B.bct[1] != BC_INTERIOR) ? T(B.cf[1]) : T(0.0);
        const T cf2 = (zw == 0 && B.bct[2] != BC_INTERIOR) ? T(B.cf[2]) : T(0.0);
        const T cf3 = (xw == n0 - 1 && B.bct[3] != BC_INTERIOR) ? T(B.cf[3]) : T(0.0);
        const T cf4 = (yw == n1 - 1 && B.bct[4] != BC_INTERIOR) ? T(B.cf[4]) : T(0.0);
        const T cf5 = (zw == n2 - 1 && B.bct[5] != BC_INTERIOR) ? T(B.cf[5]) : T(0.0);
        const T v = P[p];
        const T xm = P[p - 1], xp = P[p + 1], ym = P[p - RX], yp = P[p + RX], zm = Pm[p], zp = Pp[p];
        T res;
        T g_m_d;
        if (op.kind == 0) {
            const T dhx = op.dhs[0], dhy = op.dhs[1], dhz = op.dhs[2];
            const int64_t fxi = B.foff[0] + xw + fx1 * (yw + (int64_t)n1 * zw);
            const int64_t fyi = B.foff[1] + xw + (int64_t)n0 * (yw + fy1 * zw);
            const int64_t fzi = B.foff[2] + xw + (int64_t)n0 * (yw + (int64_t)n1 * zw);
            const T bxm = op.b[0][fxi], bxp = op.b[0][fxi + 1];
            const T bym = op.b[1][fyi], byp = op.b[1][fyi + n0];
            const T bzm = op.b[2][fzi], bzp = op.b[2][fzi + (int64_t)n0 * n1];
            const T gamma = op.alpha * op.a[c] + dhx * (bxm + bxp) + dhy * (bym + byp) + dhz * (bzm + bzp);
            g_m_d = gamma - (dhx * (bxm * cf0 + bxp * cf3) + dhy * (bym * cf1 + byp * cf4) + dhz * (bzm * cf2 + bzp * cf5));
            const T rho = dhx * (bxm * xm + bxp * xp) + dhy * (bym * ym + byp * yp) + dhz * (bzm * zm + bzp * zp);
            res = rhs[c] - (gamma * v - rho);
        } else {
            const T dhx = op.dpo[0], dhy = op.dpo[1], dhz = op.dpo[2];
            const T gamma = T(-2.) * (dhx + dhy + dhz);
            g_m_d = gamma + dhx * (cf0 + cf3) + dhy * (cf1 + cf4) + dhz * (cf2 + cf5);
            res = rhs[c] - gamma * v - dhx * (xm + xp) - dhy * (ym + yp) - dhz * (zm + zp);
        }
        P[p] = v + omega / g_m_d * res;
    };

    // Red update of plane z on the tile plus a one-cell ring (real cells only).
    auto red = [&](int z) {
        int zw;
        if (zclass(z, zw) != 0) return;
        constexpr int W = TX + 2, H = TY + 2, HW = W / 2;
        for (int q = tid; q < H * HW; q += kSmBlock) {
            const int r = q / HW, s2 = q % HW;
            const int y = y0 - 1 + r;
            const int par = (x0 - 1 + y + z) & 1;  // (x + y + z) even for red
            const int x = x0 - 1 + 2 * s2 + par;
            int xw, yw;
            if (classify(x, n0, p0, xw) != 0 || classify(y, n1, p1, yw) != 0) continue;
            update((x - x0 + 2) + RX * (r + 1), z, xw, yw, zw);
        }
    };

    // Black update of plane z on the cells the tile owns.
    auto black = [&](int z) {
        constexpr int HW = TX / 2;
        for (int q = tid; q < TY * HW; q += kSmBlock) {
            const int r = q / HW, s2 = q % HW;
            const int y = y0 + r;
            const int par = (x0 + y + z + 1) & 1;  // (x + y + z) odd for black
            const int x = x0 + 2 * s2 + par;
            if (x >= n0 || y >= n1) continue;
            update((x - x0 + 2) + RX * (r + 2), z, x, y, z);
        }
    };

    auto write_plane = [&](int z) {
        const T* P = S[slot(z)];
        for (int q = tid; q < TX * TY; q += kSmBlock) {
            const int x = x0 + q % TX, y = y0 + q / TX;
            if (x >= n0 || y >= n1) continue;
            pout[gidx(x, y, z)] = P[(x - x0 + 2) + RX * (y - y0 + 2)];
        }
    };

    // Prologue: old planes z0-2 .. z0+1, red updates of planes z0-1 and z0.
    for (int z = z0 - 2; z <= z0 + 1; ++z) load_plane(z);
    __syncthreads();
    for (int z = z0 - 2; z <= z0 + 1; ++z) refresh_xy(z);
    __syncthreads();
    refresh_zghost(z0 - 2);
    refresh_zghost(z0 - 1);
    refresh_zghost(z0 + 1);
    __syncthreads();
    red(z0 - 1);
    __syncthreads();
    red(z0);
    __syncthreads();
    refresh_xy(z0);
    __syncthreads();

    for (int k = z0; k < z1; ++k) {
        load_plane(k + 2);
        __syncthreads();
        refresh_xy(k + 2);
        refresh_zghost(k + 2);  // ghost plane above the domain: old closure for red(k+1)
        __syncthreads();
        red(k + 1);
        __syncthreads();
        refresh_xy(k + 1);        // red-updated ghosts of plane k+1 for its black update
        refresh_zghost(k - 1);    // ghost planes next to plane k: red-updated closure
        refresh_zghost(k + 1);
        __syncthreads();
        black(k);
        __syncthreads();
        write_plane(k);
    }
}

}  // namespace

template <class T>
void gsrb_sweep(const LevelDev& L, const OpCoef<T>& op, const T* pin, T* pout, const T* rhs, int zero_in,
                const T* crse, const LevelDev* Lc, hipStream_t s) {
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    const LevelDev lc = Lc ? *Lc : L;
    if (n0 >= 48) {
        constexpr int TX = 64, TY = 16;
        const int nt = ((n0 + TX - 1) / TX) * ((n1 + TY - 1) / TY);
        // enough z chunks for ~4 work-groups per CU
        // z chunks: several waves of work-groups per CU, at most 64 planes each
        int cz = n2;
        while (cz > 16 && (cz > 64 || (int64_t)nt * ((n2 + cz - 1) / cz) < 2048)) cz = (cz + 1) / 2;
        const int nz = (n2 + cz - 1) / cz;
        hipLaunchKernelGGL((k_gsrb_sweep<T, TX, TY>), dim3((unsigned)(nt * nz)), dim3(kSmBlock), 0, s, L, op, pin, pout, rhs,
                           cz, zero_in, crse, lc);
    } else {
        constexpr int TX = 16, TY = 16;
        const int nt = ((n0 + TX - 1) / TX) * ((n1 + TY - 1) / TY);
        int cz = n2;
        while (cz > 8 && (int64_t)nt * ((n2 + cz - 1) / cz) < 512) cz = (cz + 1) / 2;
        const int nz = (n2 + cz - 1) / cz;
        hipLaunchKernelGGL((k_gsrb_sweep<T, TX, TY>), dim3((unsigned)(nt * nz)), dim3(kSmBlock), 0, s, L, op, pin, pout, rhs,
                           cz, zero_in, crse, lc);
    }
}

template void gsrb_sweep<double>(const LevelDev&, const OpCoef<double>&, const double*, double*, const double*, int,
                                 const double*, const LevelDev*, hipStream_t);
template void gsrb_sweep<float>(const LevelDev&, const OpCoef<float>&, const float*, float*, const float*, int,
                                const float*, const LevelDev*, hipStream_t);

}  // namespace mgx
