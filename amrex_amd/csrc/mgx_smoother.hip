// Fused red+black GSRB sweep for dense MG levels (one FAB covering the
// domain), one pass over HBM per sweep.
//
// The reference applies one colour per launch with a ghost refresh between
// the colours (MLCellLinOpT::smooth, AMReX_MLCellLinOp.H:1205-1217; abec_gsrb,
// AMReX_MLABecLap_3D_K.H:210-264; mlpoisson_gsrb, AMReX_MLPoisson_3D_K.H:
// 155-196), which streams every array twice per sweep.  Here a work-group owns
// a tile of columns over a range of z planes and marches up in z: at step k it
// does the red update of plane k+1 and then the black update of plane k, so
// each plane's arrays are read from HBM once.  Red values of the one-cell ring
// around the tile are recomputed redundantly from old data, so tiles never
// wait for each other; the sweep is out of place (reads pin, writes pout),
// which keeps every tile's inputs unmodified while other tiles run.  Ghost
// values follow the reference exactly: before the red update of a cell they
// come from old values (physical closure or periodic image), before the black
// update from red-updated values.
#include <hip/hip_runtime.h>

#include <limits>
#include <stdexcept>

#include "mgx_device.H"
#include "mgx_smoother.H"

namespace mgx {
namespace {

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

// One GSRB-SOR update (abec_gsrb, AMReX_MLABecLap_3D_K.H:210-264;
// mlpoisson_gsrb, AMReX_MLPoisson_3D_K.H:155-196) from the cell's value v, its
// six neighbours (ghosts already closed), its coefficients and the undrrelxr
// weights of its boundary faces.
template <class T>
__device__ __forceinline__ T gsrb_value(const OpCoef<T>& op, T v, T xm, T xp, T ym, T yp, T zm, T zp, T acoef,
                                        T rhs, T bxm, T bxp, T bym, T byp, T bzm, T bzp, T cf0, T cf1, T cf2, T cf3,
                                        T cf4, T cf5) {
    constexpr T omega = T(1.15);
    if (op.kind == 0) {
        const T dhx = op.dhs[0], dhy = op.dhs[1], dhz = op.dhs[2];
        const T gamma = op.alpha * acoef + dhx * (bxm + bxp) + dhy * (bym + byp) + dhz * (bzm + bzp);
        const T g_m_d = gamma - (dhx * (bxm * cf0 + bxp * cf3) + dhy * (bym * cf1 + byp * cf4) + dhz * (bzm * cf2 + bzp * cf5));
        const T rho = dhx * (bxm * xm + bxp * xp) + dhy * (bym * ym + byp * yp) + dhz * (bzm * zm + bzp * zp);
        const T res = rhs - (gamma * v - rho);
        return v + omega / g_m_d * res;
    }
    const T dhx = op.dpo[0], dhy = op.dpo[1], dhz = op.dpo[2];
    const T gamma = T(-2.) * (dhx + dhy + dhz);
    const T g_m_d = gamma + dhx * (cf0 + cf3) + dhy * (cf1 + cf4) + dhz * (cf2 + cf5);
    const T res = rhs - gamma * v - dhx * (xm + xp) - dhy * (ym + yp) - dhz * (zm + zp);
    return v + omega / g_m_d * res;
}

// Coefficients of one cell update.
template <class T>
struct CellCoef {
    T a, rhs, bxm, bxp, bym, byp, bzm, bzp;
};

// Work-group: a TX x TY tile of columns over the z planes [z0, z1), one thread
// per x-adjacent pair of columns.  At step k the thread whose column has
// x + y + k odd does the red update of plane k+1 and, after the step's only
// barrier, the black update of plane k; the other column of the pair takes
// the next step.  Each thread keeps planes k-1 .. k+2 of its two columns in
// registers; LDS holds four planes of the tile with a two-cell halo for the
// in-plane neighbours.  Red values of the one-cell ring around the tile are
// recomputed from old data by ring threads.  Ghosts across physical faces are
// closed just before use from the cell and its inner neighbour, which hold the
// values the reference's ghost refresh would see (old before red, red-updated
// before black).
template <class T, int TX, int TY>
__global__ __launch_bounds__(TX / 2 * TY) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 3 : 4, 8))) void k_gsrb_sweep(LevelDev L, OpCoef<T> op, const T* __restrict__ pin,
                                                            T* __restrict__ pout, const T* __restrict__ rhs, int cz,
                                                            int zero_in, const T* __restrict__ crse, LevelDev Lc) {
    constexpr int NT = TX / 2 * TY;
    constexpr int RX = TX + 4, RY = TY + 4, RS = RX * RY;
    constexpr int HPT = (RS - TX * TY + NT - 1) / NT;  // halo positions per thread
    __shared__ T S[4][RS];
    const BoxDev& B = L.box[0];
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    const bool per0 = L.periodic[0] != 0, per1 = L.periodic[1] != 0, per2 = L.periodic[2] != 0;
    const int ntx = (n0 + TX - 1) / TX, nty = (n1 + TY - 1) / TY;
    // XCD-aware order: consecutive work-groups (round-robin over the 8 XCDs)
    // take tiles that are far apart, so each XCD's L2 sees neighbouring tiles.
    const int nblk = gridDim.x;
    const int bid = (nblk % 8 == 0) ? (int)(blockIdx.x % 8) * (nblk / 8) + (int)(blockIdx.x / 8) : (int)blockIdx.x;
    const int tx = bid % ntx, ty = (bid / ntx) % nty, tz = bid / (ntx * nty);
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * cz, z1 = min(z0 + cz, n2);
    const int ex = min(TX, n0 - x0), ey = min(TY, n1 - y0);  // real extent of the tile
    const int tid = threadIdx.x;
    const int px = tid % (TX / 2), py = tid / (TX / 2);
    const int xe = x0 + 2 * px, yc = y0 + py;
    const bool vrow = py < ey;
    const bool valid0 = vrow && 2 * px < ex, valid1 = vrow && 2 * px + 1 < ex;

    // 32-bit offsets from the FAB and face-array bases (the host checks they fit)
    const int sy = n0 + 2, sz = (n0 + 2) * (n1 + 2);
    const T* __restrict__ P0 = pin + B.off;
    T* __restrict__ Q0 = pout + B.off;
    const T* __restrict__ R0 = rhs + B.off;
    const T* __restrict__ A0 = op.kind == 0 ? op.a + B.off : nullptr;
    const T* __restrict__ BX = op.kind == 0 ? op.b[0] + B.foff[0] : nullptr;
    const T* __restrict__ BY = op.kind == 0 ? op.b[1] + B.foff[1] : nullptr;
    const T* __restrict__ BZ = op.kind == 0 ? op.b[2] + B.foff[2] : nullptr;
    auto cidx = [&](int x, int y, int z) -> int { return (x + 1) + sy * (y + 1) + sz * (z + 1); };

    // wrap a coordinate outside the domain along a periodic direction; false
    // if it is a ghost beyond a non-periodic face
    auto wrapc = [](int x, int n, bool per, int& xw) -> bool {
        xw = x;
        if (x >= 0 && x < n) return true;
        if (!per) return false;
        xw = x % n;
        if (xw < 0) xw += n;
        return true;
    };
    auto old_value = [&](int xw, int yw, int zw) -> T {
        T v = zero_in ? T(0) : P0[cidx(xw, yw, zw)];
        if (crse) {
            const BoxDev& C = Lc.box[0];
            v += crse[C.off + ((xw >> 1) + 1) + C.sy * ((yw >> 1) + 1) + C.sz * ((zw >> 1) + 1)];
        }
        return v;
    };
    auto load_coef = [&](int x, int y, int z, CellCoef<T>& c) {
        const int ci = cidx(x, y, z);
        c.rhs = R0[ci];
        if (op.kind == 0) {
            c.a = A0[ci];
            const int fx = x + (n0 + 1) * (y + n1 * z);
            const int fy = x + n0 * (y + (n1 + 1) * z);
            const int fz = x + n0 * (y + n1 * z);
            c.bxm = BX[fx];
            c.bxp = BX[fx + 1];
            c.bym = BY[fy];
            c.byp = BY[fy + n0];
            c.bzm = BZ[fz];
            c.bzp = BZ[fz + n0 * n1];
        } else {
            c.a = c.bxm = c.bxp = c.bym = c.byp = c.bzm = c.bzp = T(0);
        }
    };
    // update of real cell (x, y, z) with value v, in-plane neighbours from LDS
    // plane P at position p, z neighbours zm/zp (real or ghost), closures of
    // in-plane ghosts applied here
    auto update = [&](const T* P, int p, int x, int y, int z, T v, T zm, T zp, const CellCoef<T>& c) -> T {
        T xm = P[p - 1], xp = P[p + 1], ym = P[p - RX], yp = P[p + RX];
        T cf0 = T(0), cf1 = T(0), cf2 = T(0), cf3 = T(0), cf4 = T(0), cf5 = T(0);
        if (x == 0 && B.bct[0] != BC_INTERIOR) { xm = closure(B, 0, v, xp); cf0 = T(B.cf[0]); }
        if (x == n0 - 1 && B.bct[3] != BC_INTERIOR) { xp = closure(B, 3, v, xm); cf3 = T(B.cf[3]); }
        if (y == 0 && B.bct[1] != BC_INTERIOR) { ym = closure(B, 1, v, yp); cf1 = T(B.cf[1]); }
        if (y == n1 - 1 && B.bct[4] != BC_INTERIOR) { yp = closure(B, 4, v, ym); cf4 = T(B.cf[4]); }
        if (z == 0 && B.bct[2] != BC_INTERIOR) { zm = closure(B, 2, v, zp); cf2 = T(B.cf[2]); }
        if (z == n2 - 1 && B.bct[5] != BC_INTERIOR) { zp = closure(B, 5, v, zm); cf5 = T(B.cf[5]); }
        return gsrb_value<T>(op, v, xm, xp, ym, yp, zm, zp, c.a, c.rhs, c.bxm, c.bxp, c.bym, c.byp, c.bzm, c.bzp, cf0,
                             cf1, cf2, cf3, cf4, cf5);
    };

    // halo position h of the current tile (two-cell band around ex x ey)
    const int W = ex + 4;
    const int nh = 4 * W + 4 * ey;
    auto halo_pos = [&](int h, int& lx, int& ly) {
        if (h < 2 * W) { lx = h % W; ly = h / W; return; }
        h -= 2 * W;
        if (h < 4 * ey) { const int c = h & 3; ly = 2 + (h >> 2); lx = c < 2 ? c : ex + c; return; }
        h -= 4 * ey;
        lx = h % W;
        ly = ey + 2 + h / W;
    };
    // ring position of this thread (one-cell band around the tile), if any
    const int nring = 2 * (ex + 2) + 2 * ey;
    int rlx = 0, rly = 0;
    if (tid < nring) {
        int r = tid;
        if (r < ex + 2) { rlx = 1 + r; rly = 1; }
        else if ((r -= ex + 2) < ex + 2) { rlx = 1 + r; rly = ey + 2; }
        else if ((r -= ex + 2) < ey) { rlx = 1; rly = 2 + r; }
        else { r -= ey; rlx = ex + 2; rly = 2 + r; }
    }
    const int rgx = x0 - 2 + rlx, rgy = y0 - 2 + rly;
    int rxw = 0, ryw = 0;
    const bool ring_real = tid < nring && wrapc(rgx, n0, per0, rxw) && wrapc(rgy, n1, per1, ryw);
    const int rp = rlx + RX * rly;

    // column state: planes k-1 .. k+2 of the two columns
    T w0[4] = {T(0), T(0), T(0), T(0)}, w1[4] = {T(0), T(0), T(0), T(0)};
    const int pp0 = (2 * px + 2) + RX * (py + 2);  // LDS position of column 0

    // old values of plane z: the two columns, this thread's halo positions and
    // its ring position (ghost planes beyond a non-periodic face read as 0)
    auto load_plane = [&](int z, T& c0, T& c1, T* hv, T& rv) {
        int zw;
        c0 = c1 = rv = T(0);
        for (int q = 0; q < HPT; ++q) hv[q] = T(0);
        if (!wrapc(z, n2, per2, zw)) return;
        if (valid0) c0 = old_value(xe, yc, zw);
        if (valid1) c1 = old_value(xe + 1, yc, zw);
        for (int q = 0; q < HPT; ++q) {
            const int h = tid + q * NT;
            if (h < nh) {
                int lx, ly, xw, yw;
                halo_pos(h, lx, ly);
                if (wrapc(x0 - 2 + lx, n0, per0, xw) && wrapc(y0 - 2 + ly, n1, per1, yw)) hv[q] = old_value(xw, yw, zw);
            }
        }
        if (ring_real) rv = old_value(rxw, ryw, zw);
    };
    auto store_plane = [&](int z, T c0, T c1, const T* hv) {
        T* P = S[z & 3];
        if (valid0) P[pp0] = c0;
        if (valid1) P[pp0 + 1] = c1;
        for (int q = 0; q < HPT; ++q) {
            const int h = tid + q * NT;
            if (h < nh) {
                int lx, ly;
                halo_pos(h, lx, ly);
                P[lx + RX * ly] = hv[q];
            }
        }
    };
    // coefficients of the red cell step k updates (column a_k, plane k+1; a_k
    // is the column with x + y + k odd)
    auto red_coefs = [&](int k, CellCoef<T>& cr) {
        const int act = ((xe + yc + k) & 1) ? 0 : 1;
        int zr;
        if ((act == 0 ? valid0 : valid1) && wrapc(k + 1, n2, per2, zr)) load_coef(xe + act, yc, zr, cr);
    };

    // prologue: planes z0-2 and z0-1 in LDS and in column slots 1, 2; the
    // first step's plane z0 and coefficients in flight
    T hv[HPT], nhv[HPT];
    for (int q = 0; q < HPT; ++q) nhv[q] = T(0);
    T rz = T(0), nrz = T(0), pf0 = T(0), pf1 = T(0), npf0 = T(0), npf1 = T(0);
    load_plane(z0 - 2, w0[1], w1[1], hv, rz);
    store_plane(z0 - 2, w0[1], w1[1], hv);
    load_plane(z0 - 1, w0[2], w1[2], hv, rz);
    store_plane(z0 - 1, w0[2], w1[2], hv);
    CellCoef<T> cr{}, cb{}, nr{};
    load_plane(z0, pf0, pf1, hv, rz);
    red_coefs(z0 - 2, cr);
    __syncthreads();

    for (int k = z0 - 2; k < z1; ++k) {
        const int act = ((xe + yc + k) & 1) ? 0 : 1;
        const bool vact = act == 0 ? valid0 : valid1;
        const int xa = xe + act;
        int zr;
        const bool red_real = wrapc(k + 1, n2, per2, zr);
        const bool do_black = k >= z0;
        // ---- next step's plane k+3 and coefficients, in flight during this step
        if (k + 1 < z1) {
            load_plane(k + 3, npf0, npf1, nhv, nrz);
            red_coefs(k + 1, nr);
        }
        // ---- phase A: store plane k+2, red updates of plane k+1
        w0[3] = pf0;
        w1[3] = pf1;
        store_plane(k + 2, pf0, pf1, hv);
        if (vact && red_real) {
            T* P = S[(k + 1) & 3];
            const T v = act == 0 ? w0[2] : w1[2];
            const T zm = act == 0 ? w0[1] : w1[1];
            const T zp = act == 0 ? w0[3] : w1[3];
            const T r = update(P, pp0 + act, xa, yc, zr, v, zm, zp, cr);
            P[pp0 + act] = r;
            if (act == 0) w0[2] = r; else w1[2] = r;
        }
        if (ring_real && red_real && (((rgx + rgy + k + 1) & 1) == 0)) {
            CellCoef<T> cg;
            load_coef(rxw, ryw, zr, cg);
            T* P = S[(k + 1) & 3];
            const T v = P[rp];
            const T zm = S[k & 3][rp];
            P[rp] = update(P, rp, rxw, ryw, zr, v, zm, rz, cg);
        }
        // coefficients of the black cell, in flight across the barrier
        if (vact && do_black) load_coef(xa, yc, k, cb);
        __syncthreads();
        // ---- phase B: black update of plane k, output of plane k
        if (do_black) {
            if (vact) {
                const T* P = S[k & 3];
                const T v = act == 0 ? w0[1] : w1[1];
                const T zm = act == 0 ? w0[0] : w1[0];
                const T zp = act == 0 ? w0[2] : w1[2];
                const T b = update(P, pp0 + act, xa, yc, k, v, zm, zp, cb);
                if (act == 0) w0[1] = b; else w1[1] = b;
            }
            if (valid0) Q0[cidx(xe, yc, k)] = w0[1];
            if (valid1) Q0[cidx(xe + 1, yc, k)] = w1[1];
        }
        // ---- slide the column windows and the prefetched data
        w0[0] = w0[1]; w0[1] = w0[2]; w0[2] = w0[3];
        w1[0] = w1[1]; w1[1] = w1[2]; w1[2] = w1[3];
        pf0 = npf0;
        pf1 = npf1;
        rz = nrz;
        for (int q = 0; q < HPT; ++q) hv[q] = nhv[q];
        cr = nr;
    }
}

}  // namespace

template <class T, int TX, int TY>
static void launch_sweep(const LevelDev& L, const OpCoef<T>& op, const T* pin, T* pout, const T* rhs, int zero_in,
                         const T* crse, const LevelDev& lc, int min_cz, hipStream_t s) {
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    const int nt = ((n0 + TX - 1) / TX) * ((n1 + TY - 1) / TY);
    // z chunks of at most 64 planes, halved while the grid has fewer than
    // 2048 work-groups (each chunk re-reads two planes below it)
    int cz = n2;
    while (cz > min_cz && (cz > 64 || (int64_t)nt * ((n2 + cz - 1) / cz) < 2048)) cz = (cz + 1) / 2;
    const int nz = (n2 + cz - 1) / cz;
    hipLaunchKernelGGL((k_gsrb_sweep<T, TX, TY>), dim3((unsigned)(nt * nz)), dim3(TX / 2 * TY), 0, s, L, op, pin, pout,
                       rhs, cz, zero_in, crse, lc);
}

template <class T>
void gsrb_sweep(const LevelDev& L, const OpCoef<T>& op, const T* pin, T* pout, const T* rhs, int zero_in,
                const T* crse, const LevelDev* Lc, hipStream_t s) {
    // the kernel indexes the level's FAB and face arrays with 32-bit offsets
    const int64_t lim = std::numeric_limits<int>::max();
    if (L.nalloc >= lim || L.nface[0] >= lim || L.nface[1] >= lim || L.nface[2] >= lim)
        throw std::runtime_error("gsrb_sweep: level too large for 32-bit offsets");
    const LevelDev lc = Lc ? *Lc : L;
    if (L.n[0] >= 48)
        launch_sweep<T, 64, 8>(L, op, pin, pout, rhs, zero_in, crse, lc, 8, s);
    else
        launch_sweep<T, 16, 16>(L, op, pin, pout, rhs, zero_in, crse, lc, 4, s);
}

template void gsrb_sweep<double>(const LevelDev&, const OpCoef<double>&, const double*, double*, const double*, int,
                                 const double*, const LevelDev*, hipStream_t);
template void gsrb_sweep<float>(const LevelDev&, const OpCoef<float>&, const float*, float*, const float*, int,
                                const float*, const LevelDev*, hipStream_t);

}  // namespace mgx
