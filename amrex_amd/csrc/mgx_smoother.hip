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
//
// The march is bound by memory latency unless loads stay in flight across
// steps, so it is software-pipelined: every load of step k+1 is issued at the
// start of step k, from a valid address and without a branch around it, and
// nothing touches the loaded registers before step k+1 (the hardware's load
// counter is in order, so a use would wait for everything issued before it).
// Loads are raw buffer loads: a scalar resource per plane and a per-thread
// 32-bit byte offset fixed for the whole march.  Each step loads one plane of
// coefficients for both columns of a thread's pair (whole cache lines); the
// active column's feed the red update at once, the other column's are carried
// to the next step's black update.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>
#include <stdexcept>

#include "mgx_buffer.H"
#include "mgx_device.H"
#include "mgx_smoother.H"

namespace mgx {
namespace {

// Closure of the six faces as ghost = v2*k2 + v1*k1 (v1 the cell, v2 its
// inner neighbour; mllinop_apply_bc_*, AMReX_MLLinOp_K.H:14-115, homogeneous):
// Neumann k1 = 1; reflect-odd k1 = -1; Dirichlet of closure order NX k1 = c1
// (NX > 1), k2 = c2 (NX > 2).  cf = undrrelxr weight of the first interior
// cell.  For finite values each form rounds as the reference's closure does.
template <class T>
struct Closure {
    T k1[6], k2[6], cf[6];
};

template <class T>
__device__ __forceinline__ T close_ghost(const Closure<T>& C, int f, T v1, T v2) {
    return fma(v2, C.k2[f], v1 * C.k1[f]);
}

// Coefficients of one cell update.
template <class T>
struct CellCoef {
    T a, rhs, bxm, bxp, bym, byp, bzm, bzp;
};

// One GSRB-SOR update (abec_gsrb, AMReX_MLABecLap_3D_K.H:210-264;
// mlpoisson_gsrb, AMReX_MLPoisson_3D_K.H:155-196) of value v from its six
// neighbours.  bits names the cell's physical-face closures (1 x-lo, 2 y-lo,
// 4 z-lo, 8 x-hi, 16 y-hi, 32 z-hi), applied first; the closure code runs
// under one wave-uniform branch.  Without closures the undrrelxr terms vanish
// and g_m_d equals gamma bit for bit.
template <class T, int KIND>
__device__ __forceinline__ T gsrb_update(const Closure<T>& C, T alpha, const T* dh, int bits, T v, T xm, T xp, T ym,
                                         T yp, T zm, T zp, const CellCoef<T>& c) {
    constexpr T omega = T(1.15);
    T cf0 = T(0), cf1 = T(0), cf2 = T(0), cf3 = T(0), cf4 = T(0), cf5 = T(0);
    const bool edge = __builtin_amdgcn_ballot_w64(bits != 0) != 0;
    if (edge) {
        const bool b0 = bits & 1, b1 = bits & 2, b2 = bits & 4, b3 = bits & 8, b4 = bits & 16, b5 = bits & 32;
        xm = b0 ? close_ghost(C, 0, v, xp) : xm;
        cf0 = b0 ? C.cf[0] : T(0);
        xp = b3 ? close_ghost(C, 3, v, xm) : xp;
        cf3 = b3 ? C.cf[3] : T(0);
        ym = b1 ? close_ghost(C, 1, v, yp) : ym;
        cf1 = b1 ? C.cf[1] : T(0);
        yp = b4 ? close_ghost(C, 4, v, ym) : yp;
        cf4 = b4 ? C.cf[4] : T(0);
        zm = b2 ? close_ghost(C, 2, v, zp) : zm;
        cf2 = b2 ? C.cf[2] : T(0);
        zp = b5 ? close_ghost(C, 5, v, zm) : zp;
        cf5 = b5 ? C.cf[5] : T(0);
    }
    const T dhx = dh[0], dhy = dh[1], dhz = dh[2];
    if constexpr (KIND == 0) {
        const T gamma = alpha * c.a + dhx * (c.bxm + c.bxp) + dhy * (c.bym + c.byp) + dhz * (c.bzm + c.bzp);
        T g_m_d = gamma;
        if (edge)
            g_m_d = gamma - (dhx * (c.bxm * cf0 + c.bxp * cf3) + dhy * (c.bym * cf1 + c.byp * cf4) +
                             dhz * (c.bzm * cf2 + c.bzp * cf5));
        const T rho = dhx * (c.bxm * xm + c.bxp * xp) + dhy * (c.bym * ym + c.byp * yp) + dhz * (c.bzm * zm + c.bzp * zp);
        const T res = c.rhs - (gamma * v - rho);
        return v + omega / g_m_d * res;
    } else {
        const T gamma = T(-2.) * (dhx + dhy + dhz);
        T g_m_d = gamma;
        if (edge) g_m_d = gamma + dhx * (cf0 + cf3) + dhy * (cf1 + cf4) + dhz * (cf2 + cf5);
        const T res = c.rhs - gamma * v - dhx * (xm + xp) - dhy * (ym + yp) - dhz * (zm + zp);
        return v + omega / g_m_d * res;
    }
}

// In-plane byte offsets of one cell: FAB, x faces, y and z faces (the same
// in-plane layout x + n0*y), coarse FAB.
struct CellOff {
    uint32_t c, fx, fyz, cc;
};

// Work-group: a TX x TY tile of columns over the z planes [z0, z1), one thread
// per x-adjacent pair of columns; the rows of a wave share their parity, so
// the active column of a step is uniform over the wave.  At step k the active
// column (x + y + k odd) does the red update of plane k+1 and, after the
// step's only barrier, the black update of plane k.  LDS holds four planes of
// the tile with a two-cell halo (k-1 red-updated, k, k+1 red-updated in the
// step, k+2 old); the columns' own values are read back from it.  Halo
// positions are numbered ring-first (the one-cell band around the tile, whose
// red values are recomputed from old data), so the thread that recomputes a
// ring cell also loaded its old value of plane k+2.
// ZERO: pin is taken as zero; CRSE: the coarse correction is added to pin.
// VEC (the level's x extent is even, so both columns of every pair are real
// cells): each pair's two values of an array come in one load.
template <class T, int KIND, int TX, int TY, bool ZERO, bool CRSE, bool VEC>
// FP32 on 512-thread tiles: at most 128 VGPRs, so two work-groups share a CU
// (four waves per SIMD).
__global__ __launch_bounds__(TX / 2 * TY, (sizeof(T) == 4 && TX / 2 * TY == 512) ? 4 : 1) void k_gsrb_sweep(
    LevelDev L, OpCoef<T> op, const T* pin, T* pout,
                                                            const T* rhs, int cz, const T* crse, LevelDev Lc) {
    constexpr int NT = TX / 2 * TY;
    constexpr int PR = TX / 2;                   // thread pairs per row
    constexpr int RPW = PR >= 64 ? 1 : 64 / PR;  // rows per wave (a wave holds rows of one parity)
    constexpr int WPR = PR > 64 ? PR / 64 : 1;   // waves per row, for rows wider than a wave
    static_assert(PR > 64 ? PR % 64 == 0 : (64 % PR == 0 && TY % (2 * RPW) == 0), "tile shape");
    constexpr int RX = TX + 4, RS = RX * (TY + 4) + 1;  // + one slot for stores of lanes with nothing to store
    constexpr int DUMMY = RS - 1;
    // An LDS row holds its even columns, then its odd ones: the cells a wave
    // touches in one access (one parity, consecutive pairs) are consecutive
    // words, so stencil reads are free of bank conflicts.
    constexpr int H = RX / 2;
    static_assert(RX % 2 == 0, "split rows");
    auto lpos = [](int lx, int lyy) { return RX * lyy + (lx & 1) * H + (lx >> 1); };
    constexpr int HPT = (4 * TX + 4 * TY + 16 + NT - 1) / NT;  // halo positions per thread
    static_assert(2 * (TX + 2) + 2 * TY <= NT, "one ring cell per thread");
    constexpr uint32_t ST = sizeof(T);
    constexpr uint32_t NONE = 0xffffffffu;
    __shared__ T S[4][RS];
    __shared__ Closure<T> CL;

    const BoxDev& B = L.box[0];
    if (threadIdx.x < 6) {
        const int f = threadIdx.x;
        T k1 = T(0), k2 = T(0);
        if (B.bct[f] == BC_NEUMANN) {
            k1 = T(1);
        } else if (B.bct[f] == BC_REFLECT_ODD) {
            k1 = T(-1);
        } else {
            if (B.nxo[f] > 1) k1 = T(B.bcc[f][1]);
            if (B.nxo[f] > 2) k2 = T(B.bcc[f][2]);
        }
        CL.k1[f] = k1;
        CL.k2[f] = k2;
        CL.cf[f] = T(B.cf[f]);
    }
    const bool zlo_on = B.bct[2] != BC_INTERIOR, zhi_on = B.bct[5] != BC_INTERIOR;
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    const int per0 = L.periodic[0], per1 = L.periodic[1], per2 = L.periodic[2];
    const int sy = (int)B.sy;
    const char* gin = reinterpret_cast<const char*>(pin + B.off);
    char* gout = reinterpret_cast<char*>(pout + B.off);
    const char* grhs = reinterpret_cast<const char*>(rhs + B.off);
    const char* ga = KIND == 0 ? reinterpret_cast<const char*>(op.a + B.off) : nullptr;
    const char* gb[3];
    for (int d = 0; d < 3; ++d) gb[d] = KIND == 0 ? reinterpret_cast<const char*>(op.b[d] + B.foff[d]) : nullptr;
    const char* gc = nullptr;
    int csy = 0, csz = 0;
    if (CRSE) {
        const BoxDev& C = Lc.box[0];
        gc = reinterpret_cast<const char*>(crse + C.off);
        csy = (int)C.sy;
        csz = (int)C.sz;
    }
    const T alpha = op.alpha;
    T dh[3];
    for (int d = 0; d < 3; ++d) dh[d] = KIND == 0 ? op.dhs[d] : op.dpo[d];

    const int ntx = (n0 + TX - 1) / TX, nty = (n1 + TY - 1) / TY;
    // XCD-aware order: consecutive work-groups (round-robin over the 8 XCDs)
    // take tiles that are far apart, so each XCD's L2 sees neighbouring tiles.
    const int nblk = gridDim.x;
    const int bid = (nblk % 8 == 0) ? (int)(blockIdx.x % 8) * (nblk / 8) + (int)(blockIdx.x / 8) : (int)blockIdx.x;
    const int tx = bid % ntx, ty = (bid / ntx) % nty, tz = bid / (ntx * nty);
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * cz, z1 = min(z0 + cz, n2);
    const int ex = min(TX, n0 - x0), ey = min(TY, n1 - y0);  // real extent of the tile
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int px = PR > 64 ? lane + 64 * (wave % WPR) : lane % PR;
    const int ly = PR > 64 ? wave / WPR : (wave >> 1) * (2 * RPW) + (wave & 1) + 2 * (lane / PR);  // tile row
    const int xe = x0 + 2 * px, yc = y0 + ly;
    const int ypar = (y0 + (PR > 64 ? ly : wave)) & 1;  // parity of yc, uniform over the wave
    const bool vrow = ly < ey;
    const bool valid0 = vrow && 2 * px < ex, valid1 = vrow && 2 * px + 1 < ex;

    // wrap an in-plane coordinate at most two cells outside the domain; false
    // if it is a ghost beyond a non-periodic face
    auto wrapc = [](int x, int n, int per, int& xw) -> bool {
        xw = x;
        if (x >= 0 && x < n) return true;
        if (!per) return false;
        xw = x < 0 ? x + n : x - n;
        return true;
    };
    // plane to address for z (z within [-2 n2, 2 n2); clamped beyond a
    // non-periodic face); true if z is a real plane
    auto zaddr = [&](int z, int& zl) -> bool {
        if (z >= 0 && z < n2) {
            zl = z;
            return true;
        }
        if (per2) {
            zl = z < 0 ? z + n2 : z - n2;
            if (zl < 0) zl += n2;
            if (zl >= n2) zl -= n2;
            return true;
        }
        zl = min(max(z, 0), n2 - 1);
        return false;
    };
    auto offs = [&](int x, int y) -> CellOff {
        CellOff o;
        o.c = (uint32_t)((x + 1) + sy * (y + 1)) * ST;
        o.fx = (uint32_t)(x + (n0 + 1) * y) * ST;
        o.fyz = (uint32_t)(x + n0 * y) * ST;
        o.cc = (uint32_t)(((x >> 1) + 1) + csy * ((y >> 1) + 1)) * ST;
        return o;
    };
    auto face_bits = [&](int x, int y) -> int {
        int b = 0;
        if (x == 0 && B.bct[0] != BC_INTERIOR) b |= 1;
        if (y == 0 && B.bct[1] != BC_INTERIOR) b |= 2;
        if (x == n0 - 1 && B.bct[3] != BC_INTERIOR) b |= 8;
        if (y == n1 - 1 && B.bct[4] != BC_INTERIOR) b |= 16;
        return b;
    };
    auto zbits = [&](int z) -> int {
        int b = 0;
        if (z == 0 && zlo_on) b |= 4;
        if (z == n2 - 1 && zhi_on) b |= 32;
        return b;
    };

    // own columns: column 1 is column 0 + one cell in x; where column 1 is no
    // real cell its loads use column 0's offsets (d1 = 0), so every load stays
    // in bounds.  With VEC (n0 even, n0 >= 48) lanes without real columns take
    // the last pair (n0-2, n0-1), so the pair loads and the third x face
    // (x + 2 <= n0) stay inside their arrays too.
    const CellOff own = offs(min(xe, VEC ? n0 - 2 : n0 - 1), min(yc, n1 - 1));
    const uint32_t d1 = valid1 ? ST : 0u;
    const int bits0 = valid0 ? face_bits(xe, yc) : 0, bits1 = valid1 ? face_bits(xe + 1, yc) : 0;
    const int pp0 = lpos(2 * px + 2, ly + 2);  // LDS position of column 0; column 1 is at pp0 + H

    // halo band (two cells around ex x ey), ring first: LDS position and FAB /
    // coarse in-plane offsets (NONE: no position, or a ghost beyond a
    // non-periodic face)
    const int nring = 2 * (ex + 2) + 2 * ey, nband = 4 * ex + 4 * ey + 16;
    int hl[HPT];
    uint32_t ho[HPT], hc[HPT];
    int rgx = 0, rgy = 0, rpar_l = 0;
    for (int q = 0; q < HPT; ++q) {
        int h = tid + q * NT;
        hl[q] = -1;
        ho[q] = hc[q] = NONE;
        if (h >= nband) continue;
        int lx, lyy;
        if (h < nring) {
            if (h < ex + 2) {
                lx = 1 + h;
                lyy = 1;
            } else if ((h -= ex + 2) < ex + 2) {
                lx = 1 + h;
                lyy = ey + 2;
            } else if ((h -= ex + 2) < ey) {
                lx = 1;
                lyy = 2 + h;
            } else {
                h -= ey;
                lx = ex + 2;
                lyy = 2 + h;
            }
        } else {
            h -= nring;
            if (h < ex + 4) {
                lx = h;
                lyy = 0;
            } else if ((h -= ex + 4) < ex + 4) {
                lx = h;
                lyy = ey + 3;
            } else if ((h -= ex + 4) < ey + 2) {
                lx = 0;
                lyy = 1 + h;
            } else {
                h -= ey + 2;
                lx = ex + 3;
                lyy = 1 + h;
            }
        }
        hl[q] = lpos(lx, lyy);
        if (q == 0) {
            rpar_l = lx & 1;
            rgx = x0 - 2 + lx;
            rgy = y0 - 2 + lyy;
        }
        int xw, yw;
        if (wrapc(x0 - 2 + lx, n0, per0, xw) && wrapc(y0 - 2 + lyy, n1, per1, yw)) {
            const CellOff o = offs(xw, yw);
            ho[q] = o.c;
            hc[q] = o.cc;
        }
    }
    uint32_t hoa[HPT], hca[HPT];  // load offsets (a ghost position loads cell (-1, -1) and discards it)
    int shl[HPT];                 // LDS store positions
    for (int q = 0; q < HPT; ++q) {
        hoa[q] = ho[q] == NONE ? 0u : ho[q];
        hca[q] = hc[q] == NONE ? 0u : hc[q];
        shl[q] = hl[q] >= 0 ? hl[q] : DUMMY;
    }
    // ring cell of this thread (halo slot 0 when tid < nring); a ring
    // position that is no real cell addresses cell (0, 0), never used
    int rxw = 0, ryw = 0;
    const bool ring_real = tid < nring && wrapc(rgx, n0, per0, rxw) && wrapc(rgy, n1, per1, ryw);
    const CellOff ro = ring_real ? offs(rxw, ryw) : offs(0, 0);
    const int rbits = ring_real ? face_bits(rxw, ryw) : 0;
    const int rpar = (rgx + rgy) & 1;
    const int rp = max(hl[0], 0);
    // x neighbours of a cell of parity q in the split rows: left q ? -H : H-1, right q ? 1-H : H
    const int rl = rpar_l ? -H : H - 1, rr = rpar_l ? 1 - H : H;
    const int sp0 = valid0 ? pp0 : DUMMY, sp1 = valid1 ? pp0 + H : DUMMY;

    // plane strides in bytes and per-plane resources
    const uint32_t pl_c = (uint32_t)B.sz * ST;
    const uint32_t pl_fx = (uint32_t)(n0 + 1) * n1 * ST;
    const uint32_t pl_fy = (uint32_t)n0 * (n1 + 1) * ST;
    const uint32_t pl_fz = (uint32_t)n0 * n1 * ST;
    // a cell plane zl (-1 <= zl <= n2 exist in the FAB) and a face plane zl
    // with the next one (0 <= zl < n2; the z faces have n2+1 planes): every
    // offset used stays inside these bytes
    auto cell_rsrc = [&](const char* b, int zl) { return make_rsrc_n(b + (int64_t)pl_c * (zl + 1), pl_c); };
    auto face_rsrc = [&](int d, uint32_t pl, int zl) { return make_rsrc_n(gb[d] + (int64_t)pl * zl, 2 * pl); };

    // Raw loads of one step: old values of a plane (own columns, halo slots,
    // coarse terms), both columns' coefficients of the next plane, the ring
    // cell's coefficients of that plane.  Pair coefficients: cells xe, xe+1;
    // x faces xe .. xe+2; y faces of rows yc, yc+1; z faces of planes z, z+1.
    struct PairCoef {
        T rhs[2], a[2], bx[3], by[2], byu[2], bz[2], bzu[2];
    };
    struct Step {
        T p0, p1, hp[HPT], cc, hcc[HPT];
        PairCoef pc;
        CellCoef<T> cg;
    };
    auto load_vals = [&](int z, Step& d) {
        int zv;
        zaddr(z, zv);
        if constexpr (!ZERO) {
            const Rsrc P = cell_rsrc(gin, zv);
            if constexpr (VEC) {
                const Two<T> v = bld2<T>(P, own.c);
                d.p0 = v.x;
                d.p1 = v.y;
            } else {
                d.p0 = bld<T>(P, own.c);
                d.p1 = bld<T>(P, own.c + ST);
            }
            for (int q = 0; q < HPT; ++q) d.hp[q] = bld<T>(P, hoa[q]);
        }
        if constexpr (CRSE) {
            const Rsrc C = make_rsrc_n(gc + (int64_t)csz * ST * ((zv >> 1) + 1), (uint32_t)csz * ST);
            d.cc = bld<T>(C, own.cc);  // both columns share the coarse cell (xe even)
            for (int q = 0; q < HPT; ++q) d.hcc[q] = bld<T>(C, hca[q]);
        }
    };
    auto load_coef = [&](const CellOff& o, int zl, CellCoef<T>& c) {
        c.rhs = bld<T>(cell_rsrc(grhs, zl), o.c);
        if constexpr (KIND == 0) {
            c.a = bld<T>(cell_rsrc(ga, zl), o.c);
            const Rsrc FX = face_rsrc(0, pl_fx, zl);
            const Rsrc FY = face_rsrc(1, pl_fy, zl);
            const Rsrc FZ = face_rsrc(2, pl_fz, zl);
            c.bxm = bld<T>(FX, o.fx);
            c.bxp = bld<T>(FX, o.fx + ST);
            c.bym = bld<T>(FY, o.fyz);
            c.byp = bld<T>(FY, o.fyz + (uint32_t)n0 * ST);
            c.bzm = bld<T>(FZ, o.fyz);
            c.bzp = bld<T>(FZ, o.fyz + pl_fz);
        } else {
            c.a = c.bxm = c.bxp = c.bym = c.byp = c.bzm = c.bzp = T(0);
        }
    };
    auto load_pair = [&](int zl, PairCoef& c) {
        const Rsrc R = cell_rsrc(grhs, zl);
        if constexpr (VEC) {
            auto put = [](T* dst, Two<T> v) {
                dst[0] = v.x;
                dst[1] = v.y;
            };
            put(c.rhs, bld2<T>(R, own.c));
            if constexpr (KIND == 0) {
                const Rsrc A = cell_rsrc(ga, zl);
                const Rsrc FX = face_rsrc(0, pl_fx, zl);
                const Rsrc FY = face_rsrc(1, pl_fy, zl);
                const Rsrc FZ = face_rsrc(2, pl_fz, zl);
                const uint32_t up = (uint32_t)n0 * ST;
                put(c.a, bld2<T>(A, own.c));
                put(c.bx, bld2<T>(FX, own.fx));
                c.bx[2] = bld<T>(FX, own.fx + 2 * ST);
                put(c.by, bld2<T>(FY, own.fyz));
                put(c.byu, bld2<T>(FY, own.fyz + up));
                put(c.bz, bld2<T>(FZ, own.fyz));
                put(c.bzu, bld2<T>(FZ, own.fyz + pl_fz));
            }
            return;
        }
        c.rhs[0] = bld<T>(R, own.c);
        c.rhs[1] = bld<T>(R, own.c + d1);
        if constexpr (KIND == 0) {
            const Rsrc A = cell_rsrc(ga, zl);
            const Rsrc FX = face_rsrc(0, pl_fx, zl);
            const Rsrc FY = face_rsrc(1, pl_fy, zl);
            const Rsrc FZ = face_rsrc(2, pl_fz, zl);
            const uint32_t up = (uint32_t)n0 * ST;
            c.a[0] = bld<T>(A, own.c);
            c.a[1] = bld<T>(A, own.c + d1);
            c.bx[0] = bld<T>(FX, own.fx);
            c.bx[1] = bld<T>(FX, own.fx + ST);
            c.bx[2] = bld<T>(FX, own.fx + ST + d1);
            c.by[0] = bld<T>(FY, own.fyz);
            c.by[1] = bld<T>(FY, own.fyz + d1);
            c.byu[0] = bld<T>(FY, own.fyz + up);
            c.byu[1] = bld<T>(FY, own.fyz + up + d1);
            c.bz[0] = bld<T>(FZ, own.fyz);
            c.bz[1] = bld<T>(FZ, own.fyz + d1);
            c.bzu[0] = bld<T>(FZ, own.fyz + pl_fz);
            c.bzu[1] = bld<T>(FZ, own.fyz + pl_fz + d1);
        }
    };
    // step j's loads: old values of plane j+2, pair and ring coefficients of plane j+1
    auto load_step = [&](int j, Step& d) {
        int zr;
        zaddr(j + 1, zr);
        load_vals(j + 2, d);
        load_pair(zr, d.pc);
        load_coef(ro, zr, d.cg);
    };
    // coefficients of one column of a pair
    auto pick = [&](const PairCoef& c, int col) -> CellCoef<T> {
        CellCoef<T> r;
        const bool one = col != 0;
        r.rhs = one ? c.rhs[1] : c.rhs[0];
        if constexpr (KIND == 0) {
            r.a = one ? c.a[1] : c.a[0];
            r.bxm = one ? c.bx[1] : c.bx[0];
            r.bxp = one ? c.bx[2] : c.bx[1];
            r.bym = one ? c.by[1] : c.by[0];
            r.byp = one ? c.byu[1] : c.byu[0];
            r.bzm = one ? c.bz[1] : c.bz[0];
            r.bzp = one ? c.bzu[1] : c.bzu[0];
        } else {
            r.a = r.bxm = r.bxp = r.bym = r.byp = r.bzm = r.bzp = T(0);
        }
        return r;
    };
    // old value from raw loads (zero on a ghost plane or at a ghost position)
    auto oldv = [&](bool ok, T p, T c) -> T {
        if (!ok) return T(0);
        T v = ZERO ? T(0) : p;
        if constexpr (CRSE) v += c;
        return v;
    };
    // old values of plane z into LDS (lanes with nothing to store write the dummy slot)
    T hv[HPT];
    auto store_plane = [&](int z, const Step& d) {
        int zl;
        const bool real = zaddr(z, zl);
        T* P = S[z & 3];
        P[sp0] = oldv(real && valid0, d.p0, d.cc);
        P[sp1] = oldv(real && valid1, d.p1, d.cc);
        for (int q = 0; q < HPT; ++q) {
            hv[q] = oldv(real && ho[q] != NONE, d.hp[q], d.hcc[q]);
            P[shl[q]] = hv[q];
        }
    };

    // prologue: planes z0-2 and z0-1 in LDS, then step z0-2's loads
    Step sa, sb;
    load_vals(z0 - 2, sa);
    load_vals(z0 - 1, sb);
    store_plane(z0 - 2, sa);
    store_plane(z0 - 1, sb);
    load_step(z0 - 2, sa);
    __syncthreads();

    // One march step: consumes cur (loaded during the previous step) and
    // fills nxt; the loop is unrolled by two so the buffers swap roles
    // without copies (a copy would wait for the loads in flight).
    CellCoef<T> carry{};  // black-cell coefficients of this step (the other column of plane k)
    auto step = [&](int k, const Step& cur, Step& nxt) {
        const int act = ((ypar + k) & 1) ? 0 : 1;  // uniform over the wave
        const int abits = act == 0 ? bits0 : bits1;
        const int p = pp0 + act * H;
        const int al = act ? -H : H - 1, ar = act ? 1 - H : H;  // x neighbours of the active column
        int zr;
        const bool red_real = zaddr(k + 1, zr);
        const bool do_black = k >= z0 && k < z1;
        const bool ring_red = ring_real && red_real && ((rpar ^ ((k + 1) & 1)) == 0);
        // ---- the next step's loads, in flight during this step
        load_step(k + 1, nxt);
        // ---- phase A: store plane k+2, red updates of plane k+1 (tile and ring)
        store_plane(k + 2, cur);
        T* P1 = S[(k + 1) & 3];
        const int zb = zbits(zr);
        if (red_real) {
            // every lane computes; lanes without a real active column store to the dummy slot
            const T v = P1[p], zm = S[k & 3][p], zp = S[(k + 2) & 3][p];
            const T r = gsrb_update<T, KIND>(CL, alpha, dh, abits | zb, v, P1[p + al], P1[p + ar], P1[p - RX],
                                             P1[p + RX], zm, zp, pick(cur.pc, act));
            P1[(act == 0 ? valid0 : valid1) ? p : DUMMY] = r;
        }
        const CellCoef<T> next_carry = pick(cur.pc, 1 - act);
        if (ring_red) {
            const T v = P1[rp];
            const T zm = S[k & 3][rp];
            P1[rp] = gsrb_update<T, KIND>(CL, alpha, dh, rbits | zb, v, P1[rp + rl], P1[rp + rr], P1[rp - RX],
                                          P1[rp + RX], zm, hv[0], cur.cg);
        }
        __syncthreads();
        // ---- phase B: black update of plane k, output of plane k
        if (do_black) {
            const T* P0 = S[k & 3];
            const T v = P0[p], zm = S[(k - 1) & 3][p], zp = P1[p];
            const T bv = gsrb_update<T, KIND>(CL, alpha, dh, abits | zbits(k), v, P0[p + al], P0[p + ar], P0[p - RX],
                                              P0[p + RX], zm, zp, carry);
            const T other = P0[pp0 + (1 - act) * H];  // red-updated at step k-1
            char* Q = gout + (size_t)pl_c * (k + 1);
            if (valid0) *reinterpret_cast<T*>(Q + own.c) = act == 0 ? bv : other;
            if (valid1) *reinterpret_cast<T*>(Q + own.c + ST) = act == 0 ? other : bv;
        }
        carry = next_carry;
    };
    // an odd step count runs one extra step past z1 (no black update, no output)
    for (int k = z0 - 2; k < z1; k += 2) {
        step(k, sa, sb);
        step(k + 1, sb, sa);
    }
}

template <class T, int KIND, int TX, int TY, bool VEC>
void launch_sweep(const LevelDev& L, const OpCoef<T>& op, const T* pin, T* pout, const T* rhs, int zero_in,
                  const T* crse, const LevelDev& lc, int min_cz, hipStream_t s) {
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    const int nt = ((n0 + TX - 1) / TX) * ((n1 + TY - 1) / TY);
    int cz = n2;
    if (TX >= 128) {
        // large planes: a work-group fills a CU, so chunks of at most 512
        // planes, halved only while the grid has fewer work-groups than the
        // device has CUs (one round of work-groups, fewest re-read planes)
        const int cus = device_compute_units();
        while (cz > min_cz && (cz > 512 || (int64_t)nt * ((n2 + cz - 1) / cz) < cus)) cz = (cz + 1) / 2;
    } else {
        // z chunks of at most 64 planes, halved while the grid has fewer than
        // 2048 work-groups (each chunk re-reads two planes below it)
        while (cz > min_cz && (cz > 64 || (int64_t)nt * ((n2 + cz - 1) / cz) < 2048)) cz = (cz + 1) / 2;
    }
    const int nz = (n2 + cz - 1) / cz;
    const dim3 grid((unsigned)(nt * nz)), block(TX / 2 * TY);
    if (crse && zero_in)
        hipLaunchKernelGGL((k_gsrb_sweep<T, KIND, TX, TY, true, true, VEC>), grid, block, 0, s, L, op, pin, pout, rhs, cz,
                           crse, lc);
    else if (crse)
        hipLaunchKernelGGL((k_gsrb_sweep<T, KIND, TX, TY, false, true, VEC>), grid, block, 0, s, L, op, pin, pout, rhs, cz,
                           crse, lc);
    else if (zero_in)
        hipLaunchKernelGGL((k_gsrb_sweep<T, KIND, TX, TY, true, false, VEC>), grid, block, 0, s, L, op, pin, pout, rhs, cz,
                           crse, lc);
    else
        hipLaunchKernelGGL((k_gsrb_sweep<T, KIND, TX, TY, false, false, VEC>), grid, block, 0, s, L, op, pin, pout, rhs,
                           cz, crse, lc);
}

template <class T, int KIND>
void launch_kind(const LevelDev& L, const OpCoef<T>& op, const T* pin, T* pout, const T* rhs, int zero_in,
                 const T* crse, const LevelDev& lc, hipStream_t s) {
    if constexpr (sizeof(T) == 4) {
        // FP32 rows of 1 KB per array on the largest planes
        if (L.n[0] >= 512 && L.n[1] >= 256 && L.n[0] % 2 == 0) {
            launch_sweep<T, KIND, 256, 8, true>(L, op, pin, pout, rhs, zero_in, crse, lc, 8, s);
            return;
        }
    }
    if (L.n[0] >= 256 && L.n[1] >= 256 && L.n[0] % 2 == 0)
        launch_sweep<T, KIND, 128, 8, true>(L, op, pin, pout, rhs, zero_in, crse, lc, 8, s);
    else if (L.n[0] >= 48 && L.n[0] % 2 == 0)
        launch_sweep<T, KIND, 64, 8, true>(L, op, pin, pout, rhs, zero_in, crse, lc, 8, s);
    else if (L.n[0] >= 48)
        launch_sweep<T, KIND, 64, 8, false>(L, op, pin, pout, rhs, zero_in, crse, lc, 8, s);
    else
        launch_sweep<T, KIND, 16, 16, false>(L, op, pin, pout, rhs, zero_in, crse, lc, 4, s);
}

}  // namespace

template <class T>
void gsrb_sweep(const LevelDev& L, const OpCoef<T>& op, const T* pin, T* pout, const T* rhs, int zero_in,
                const T* crse, const LevelDev* Lc, hipStream_t s) {
    // in-plane byte offsets are 32-bit (planes are addressed through per-plane resources)
    const int64_t lim = std::numeric_limits<uint32_t>::max();
    const int64_t sz = (int64_t)(L.n[0] + 2) * (L.n[1] + 2);
    if (sz * (int64_t)sizeof(T) >= lim || (int64_t)(L.n[0] + 1) * (L.n[1] + 1) * (int64_t)sizeof(T) >= lim)
        throw std::runtime_error("gsrb_sweep: plane too large for 32-bit offsets");
    if (crse && !Lc) throw std::runtime_error("gsrb_sweep: coarse correction without its level");
    const LevelDev lc = Lc ? *Lc : L;
    if (op.kind == 0)
        launch_kind<T, 0>(L, op, pin, pout, rhs, zero_in, crse, lc, s);
    else
        launch_kind<T, 1>(L, op, pin, pout, rhs, zero_in, crse, lc, s);
}

template void gsrb_sweep<double>(const LevelDev&, const OpCoef<double>&, const double*, double*, const double*, int,
                                 const double*, const LevelDev*, hipStream_t);
template void gsrb_sweep<float>(const LevelDev&, const OpCoef<float>&, const float*, float*, const float*, int,
                                const float*, const LevelDev*, hipStream_t);

}  // namespace mgx
