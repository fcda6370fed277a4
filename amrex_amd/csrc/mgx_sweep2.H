// Two consecutive red+black GSRB sweeps (abec_gsrb, AMReX_MLABecLap_3D_K.H:
// 210-264; mlpoisson_gsrb, AMReX_MLPoisson_3D_K.H:155-196; each colour after a
// boundary refresh as MLCellLinOpT::smooth does, AMReX_MLCellLinOp.H:1205-1217)
// in one pass over HBM, for dense MG levels whose planes the 64x16 tiles cover
// exactly.  A V-cycle runs two sweeps back to back before and after the
// coarse correction (nu1 = nu2 = 2); fused, each array streams once per pair.
//
// A work-group owns a tile of columns over a range of z planes and marches up
// in z.  Step k runs four stages on four planes:
//   S1  red update of sweep 1 on plane k+3   (tile and the ring of depth 3)
//   S2  black update of sweep 1 on plane k+2 (tile and ring depth 2)
//   S3  red update of sweep 2 on plane k+1   (tile and ring depth 1)
//   S4  black update of sweep 2 on plane k   (tile), then plane k is written.
// Ring cells are recomputed from old data, so tiles never wait for each other.
// A cell is active in every stage of a step or in none ((x+y+k) odd), and each
// stage reads the current step's results only at its own (x, y) column, which
// the same thread computed: with a fixed column-to-thread assignment the only
// cross-thread dependences run from one step to the next, so one barrier per
// step suffices.  (The per-step body is therefore also a correct sequential
// program over the threads of a work-group.)
//
// Threads: one per x-adjacent pair of tile columns (TX/2 x TY), the first
// NRP of them also own an adjacent pair of ring cells (horizontal pairs in the
// rows above and below the tile, vertical pairs beside it); adjacent cells
// have opposite colours, so exactly one cell of each pair is active per step.
// LDS holds eight planes (z & 7) of the tile with a halo of four cells.
// Coefficients of a plane are loaded once, for both cells of each pair, and
// kept in registers for the four steps that use them; every global load of
// step k+1 is issued at the start of step k (raw buffer loads, scalar
// resource per plane), and the loop is unrolled by two so the two load
// buffers swap roles without copies.
#pragma once

#include <cstdint>

#include "mgx_buffer.H"

namespace mgx {
namespace sw2 {

#define MGX_HD __host__ __device__ __forceinline__

template <class T>
struct Closure {
    T k1[6], k2[6], cf[6];
};
template <class T>
struct CellCoef {
    T a, rhs, bxm, bxp, bym, byp, bzm, bzp;
};
// both cells of a tile pair (cells x, x+1 of row y): x faces x .. x+2
template <class T>
struct PairCoef {
    T rhs[2], a[2], bx[3], by[2], byu[2], bz[2], bzu[2];
};
template <class T>
struct RingCoef {
    CellCoef<T> c[2];
};

// Arguments of one launch: a dense level stored as one ghost-grown FAB.
template <class T>
struct Args {
    int n0, n1, n2;
    int per0, per1, per2;
    int face_on[6];          // physical face (closure applies): x-lo, y-lo, z-lo, x-hi, y-hi, z-hi
    Closure<T> cl;
    T alpha, dh[3];
    const T* pin;            // FAB base (offset applied)
    T* pout;
    const T* rhs;
    const T* a;
    const T* b[3];           // face arrays (offset applied)
    int sy;                  // cell row stride n0+2
    int64_t sz;              // cell plane stride
    const T* crse;           // coarse FAB base (CRSE)
    int csy;
    int64_t csz;
    int cz;                  // planes per z chunk
};

template <class T, int KIND, class Mem>
MGX_HD T update(const Closure<T>& C, T alpha, const T* dh, int bits, T v, T xm, T xp, T ym, T yp, T zm, T zp,
                const CellCoef<T>& c) {
    constexpr T omega = T(1.15);
    T cf0 = T(0), cf1 = T(0), cf2 = T(0), cf3 = T(0), cf4 = T(0), cf5 = T(0);
    const bool edge = Mem::any(bits != 0);
    if (edge) {
        const bool b0 = bits & 1, b1 = bits & 2, b2 = bits & 4, b3 = bits & 8, b4 = bits & 16, b5 = bits & 32;
        // ghost = c2 * inner + c1 * cell (mllinop_apply_bc_*, AMReX_MLLinOp_K.H:14-115)
        xm = b0 ? fma(xp, C.k2[0], v * C.k1[0]) : xm;
        cf0 = b0 ? C.cf[0] : T(0);
        xp = b3 ? fma(xm, C.k2[3], v * C.k1[3]) : xp;
        cf3 = b3 ? C.cf[3] : T(0);
        ym = b1 ? fma(yp, C.k2[1], v * C.k1[1]) : ym;
        cf1 = b1 ? C.cf[1] : T(0);
        yp = b4 ? fma(ym, C.k2[4], v * C.k1[4]) : yp;
        cf4 = b4 ? C.cf[4] : T(0);
        zm = b2 ? fma(zp, C.k2[2], v * C.k1[2]) : zm;
        cf2 = b2 ? C.cf[2] : T(0);
        zp = b5 ? fma(zm, C.k2[5], v * C.k1[5]) : zp;
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

// Geometry of a TX x TY tile with a halo of four cells in LDS.
template <int TX, int TY>
struct Geo {
    static constexpr int NT = TX / 2 * TY;           // threads
    static constexpr int RX = TX + 8, RY = TY + 8;   // LDS plane extents
    static constexpr int PS = RX * RY + 1;           // plane stride (+1: a slot for discarded stores)
    static constexpr int DUMMY = RX * RY;
    static constexpr int NHP = (TX + 6) / 2;         // horizontal ring pairs per ring row
    static constexpr int NRP = 6 * NHP + 3 * TY;     // ring pairs: 6 rows above/below, 6 columns of TY/2 beside
    static constexpr int N4 = 2 * RX + 2 * (RY - 2); // cells at distance 4
    static constexpr int NSLOT = 8;
    static constexpr int H = RX / 2;                 // an LDS row holds its even columns, then its odd ones
    static_assert(TX % 2 == 0 && TY % 2 == 0 && (TX / 2 * TY) % 64 == 0, "whole waves");
    static_assert(NRP <= NT && N4 <= NT, "one ring pair and one distance-4 cell per thread");
    // cells of one colour in a row are consecutive words: the active cells of
    // a wave's lanes hit distinct banks
    MGX_HD static int lpos(int lx, int ly) { return RX * ly + (lx & 1) * H + (lx >> 1); }
    // offsets of the x neighbours of a cell whose lx has parity q
    MGX_HD static int xl(int q) { return q ? -H : H - 1; }
    MGX_HD static int xr(int q) { return q ? 1 - H : H; }
    // Chebyshev distance of LDS cell (lx, ly) from the tile (lx in [4, TX+3], ly in [4, TY+3])
    MGX_HD static int dist(int lx, int ly) {
        int dx = lx < 4 ? 4 - lx : (lx > TX + 3 ? lx - (TX + 3) : 0);
        int dy = ly < 4 ? 4 - ly : (ly > TY + 3 ? ly - (TY + 3) : 0);
        return dx > dy ? dx : dy;
    }
    // ring pair q < NRP: LDS cells of its two members
    MGX_HD static void ring_pair(int q, int& lx0, int& ly0, int& lx1, int& ly1) {
        if (q < 6 * NHP) {
            const int row = q / NHP, j = q % NHP;
            ly0 = ly1 = row < 3 ? 1 + row : TY + 4 + (row - 3);
            lx0 = 1 + 2 * j;
            lx1 = lx0 + 1;
        } else {
            q -= 6 * NHP;
            const int col = q / (TY / 2), j = q % (TY / 2);
            lx0 = lx1 = col < 3 ? 1 + col : TX + 4 + (col - 3);
            ly0 = 4 + 2 * j;
            ly1 = ly0 + 1;
        }
    }
    // distance-4 cell q < N4
    MGX_HD static void ring4(int q, int& lx, int& ly) {
        if (q < RX) {
            lx = q;
            ly = 0;
        } else if ((q -= RX) < RX) {
            lx = q;
            ly = RY - 1;
        } else if ((q -= RX) < RY - 2) {
            lx = 0;
            ly = 1 + q;
        } else {
            q -= RY - 2;
            lx = RX - 1;
            ly = 1 + q;
        }
    }
};

// In-plane byte offsets of one cell: FAB, x faces, y and z faces (x + n0*y), coarse FAB.
struct CellOff {
    uint32_t c, fx, fyz, cc;
};

constexpr uint32_t kNone = 0xffffffffu;

// One thread of the work-group: its constants and the registers it keeps
// across steps.
template <class T, int KIND, int TX, int TY, bool ZERO, bool CRSE, class Mem>
struct Thread {
    using G = Geo<TX, TY>;
    using R = typename Mem::Rsrc;
    static constexpr uint32_t ST = sizeof(T);

    // ---- constants
    Args<T> A;
    int z0, z1;
    int x0, y0;
    int tid;
    int yc;                 // own row
    int pp;                 // LDS position of own column 0
    CellOff own;            // own pair (columns xe, xe+1)
    int obits[2];           // in-plane closure bits of the own columns
    bool has_ring;
    int rl[2];              // LDS positions of the ring cells
    int rq[2];              // parity of their lx
    int rd[2];              // their distances from the tile
    int rbits[2];
    bool rreal[2];
    CellOff ro[2];
    int rpar0;              // (x + y) & 1 of ring cell 0 (unwrapped)
    int hl[3];              // halo store positions: ring cells, distance-4 cell
    uint32_t ho[3], hc[3];  // their FAB / coarse in-plane offsets (kNone: ghost)
    uint32_t pl_c, pl_fx, pl_fy, pl_fz;

    // ---- registers across steps
    struct Loads {
        T p[2], h[3], cc, hcc[3];
        PairCoef<T> pc;
        RingCoef<T> rc;
    };
    Loads buf[2];
    PairCoef<T> pc1, pc2;   // planes k+2, k+1
    CellCoef<T> c3;         // plane k, active column
    RingCoef<T> rh1;        // ring coefficients of plane k+2
    CellCoef<T> r3;         // ring plane k+1, active cell

    MGX_HD bool wrapc(int x, int n, int per, int& xw) const {
        xw = x;
        if (x >= 0 && x < n) return true;
        if (!per) return false;
        xw = x < 0 ? x + n : x - n;
        return true;
    }
    // plane to address for z (clamped beyond a non-periodic face); true for a real plane
    MGX_HD bool zaddr(int z, int& zl) const {
        const int n2 = A.n2;
        if (z >= 0 && z < n2) {
            zl = z;
            return true;
        }
        if (A.per2) {
            zl = z;
            while (zl < 0) zl += n2;
            while (zl >= n2) zl -= n2;
            return true;
        }
        zl = z < 0 ? 0 : n2 - 1;
        return false;
    }
    MGX_HD CellOff offs(int x, int y) const {
        CellOff o;
        o.c = (uint32_t)((x + 1) + A.sy * (y + 1)) * ST;
        o.fx = (uint32_t)(x + (A.n0 + 1) * y) * ST;
        o.fyz = (uint32_t)(x + A.n0 * y) * ST;
        o.cc = (uint32_t)(((x >> 1) + 1) + A.csy * ((y >> 1) + 1)) * ST;
        return o;
    }
    MGX_HD int face_bits(int x, int y) const {
        int b = 0;
        if (x == 0 && A.face_on[0]) b |= 1;
        if (y == 0 && A.face_on[1]) b |= 2;
        if (x == A.n0 - 1 && A.face_on[3]) b |= 8;
        if (y == A.n1 - 1 && A.face_on[4]) b |= 16;
        return b;
    }
    MGX_HD int zbits(int z) const {
        int b = 0;
        if (z == 0 && A.face_on[2]) b |= 4;
        if (z == A.n2 - 1 && A.face_on[5]) b |= 32;
        return b;
    }

    MGX_HD void init(const Args<T>& args, int bid, int nblk, int t) {
        A = args;
        tid = t;
        const int n0 = A.n0, n1 = A.n1, n2 = A.n2;
        const int ntx = n0 / TX, nty = n1 / TY;
        // XCD-aware order: consecutive work-groups (round-robin over the 8 XCDs)
        // take tiles far apart, so each XCD's L2 sees neighbouring tiles
        const int b = (nblk % 8 == 0) ? (bid % 8) * (nblk / 8) + bid / 8 : bid;
        const int tx = b % ntx, ty = (b / ntx) % nty, tz = b / (ntx * nty);
        x0 = tx * TX;
        y0 = ty * TY;
        z0 = tz * A.cz;
        z1 = z0 + A.cz < n2 ? z0 + A.cz : n2;
        const int px = tid % (TX / 2), py = tid / (TX / 2);
        yc = y0 + py;
        const int xe = x0 + 2 * px;
        pp = G::lpos(4 + 2 * px, 4 + py);
        own = offs(xe, yc);
        obits[0] = face_bits(xe, yc);
        obits[1] = face_bits(xe + 1, yc);
        // ring pair and halo cells
        for (int q = 0; q < 3; ++q) {
            hl[q] = G::DUMMY;
            ho[q] = hc[q] = kNone;
        }
        has_ring = tid < G::NRP;
        rl[0] = rl[1] = G::DUMMY;
        rq[0] = rq[1] = 0;
        rd[0] = rd[1] = 99;
        rbits[0] = rbits[1] = 0;
        rreal[0] = rreal[1] = false;
        ro[0] = ro[1] = offs(0, 0);
        rpar0 = 0;
        auto halo = [&](int q, int lx, int ly) {
            hl[q] = G::lpos(lx, ly);
            int xw, yw;
            if (wrapc(x0 - 4 + lx, n0, A.per0, xw) && wrapc(y0 - 4 + ly, n1, A.per1, yw)) {
                const CellOff o = offs(xw, yw);
                ho[q] = o.c;
                hc[q] = o.cc;
            }
        };
        if (has_ring) {
            int lx[2], ly[2];
            G::ring_pair(tid, lx[0], ly[0], lx[1], ly[1]);
            for (int c = 0; c < 2; ++c) {
                rl[c] = G::lpos(lx[c], ly[c]);
                rq[c] = lx[c] & 1;
                rd[c] = G::dist(lx[c], ly[c]);
                int xw, yw;
                rreal[c] = wrapc(x0 - 4 + lx[c], n0, A.per0, xw) && wrapc(y0 - 4 + ly[c], n1, A.per1, yw);
                if (rreal[c]) {
                    ro[c] = offs(xw, yw);
                    rbits[c] = face_bits(xw, yw);
                }
                halo(c, lx[c], ly[c]);
            }
            rpar0 = ((x0 - 4 + lx[0]) + (y0 - 4 + ly[0])) & 1;
        }
        if (tid < G::N4) {
            int lx, ly;
            G::ring4(tid, lx, ly);
            halo(2, lx, ly);
        }
        pl_c = (uint32_t)(A.sz * ST);
        pl_fx = (uint32_t)(n0 + 1) * n1 * ST;
        pl_fy = (uint32_t)n0 * (n1 + 1) * ST;
        pl_fz = (uint32_t)n0 * n1 * ST;
    }

    // ---- loads
    MGX_HD R cell_rsrc(const T* base, int zl) const {
        return Mem::rsrc(reinterpret_cast<const char*>(base) + (int64_t)pl_c * (zl + 1), pl_c);
    }
    MGX_HD R face_rsrc(int d, uint32_t pl, int zl) const {
        return Mem::rsrc(reinterpret_cast<const char*>(A.b[d]) + (int64_t)pl * zl, 2 * pl);
    }
    // old values of plane z: own pair, halo cells, coarse terms
    MGX_HD void load_vals(int z, Loads& d) const {
        int zv;
        zaddr(z, zv);
        if constexpr (!ZERO) {
            const R P = cell_rsrc(A.pin, zv);
            const auto v = Mem::template ld2<T>(P, own.c);
            d.p[0] = v.x;
            d.p[1] = v.y;
            for (int q = 0; q < 3; ++q) d.h[q] = Mem::template ld<T>(P, ho[q] == kNone ? 0u : ho[q]);
        }
        if constexpr (CRSE) {
            const R C = Mem::rsrc(reinterpret_cast<const char*>(A.crse) + (int64_t)A.csz * ST * ((zv >> 1) + 1),
                                  (uint32_t)(A.csz * ST));
            d.cc = Mem::template ld<T>(C, own.cc);
            for (int q = 0; q < 3; ++q) d.hcc[q] = Mem::template ld<T>(C, hc[q] == kNone ? 0u : hc[q]);
        }
    }
    MGX_HD void load_cell(const CellOff& o, int zl, CellCoef<T>& c) const {
        c.rhs = Mem::template ld<T>(cell_rsrc(A.rhs, zl), o.c);
        if constexpr (KIND == 0) {
            c.a = Mem::template ld<T>(cell_rsrc(A.a, zl), o.c);
            const R FX = face_rsrc(0, pl_fx, zl), FY = face_rsrc(1, pl_fy, zl), FZ = face_rsrc(2, pl_fz, zl);
            c.bxm = Mem::template ld<T>(FX, o.fx);
            c.bxp = Mem::template ld<T>(FX, o.fx + ST);
            c.bym = Mem::template ld<T>(FY, o.fyz);
            c.byp = Mem::template ld<T>(FY, o.fyz + (uint32_t)A.n0 * ST);
            c.bzm = Mem::template ld<T>(FZ, o.fyz);
            c.bzp = Mem::template ld<T>(FZ, o.fyz + pl_fz);
        } else {
            c.a = c.bxm = c.bxp = c.bym = c.byp = c.bzm = c.bzp = T(0);
        }
    }
    MGX_HD void load_pair(int zl, PairCoef<T>& c) const {
        auto put = [](T* dst, Two<T> v) {
            dst[0] = v.x;
            dst[1] = v.y;
        };
        put(c.rhs, Mem::template ld2<T>(cell_rsrc(A.rhs, zl), own.c));
        if constexpr (KIND == 0) {
            const R FX = face_rsrc(0, pl_fx, zl), FY = face_rsrc(1, pl_fy, zl), FZ = face_rsrc(2, pl_fz, zl);
            const uint32_t up = (uint32_t)A.n0 * ST;
            put(c.a, Mem::template ld2<T>(cell_rsrc(A.a, zl), own.c));
            put(c.bx, Mem::template ld2<T>(FX, own.fx));
            c.bx[2] = Mem::template ld<T>(FX, own.fx + 2 * ST);
            put(c.by, Mem::template ld2<T>(FY, own.fyz));
            put(c.byu, Mem::template ld2<T>(FY, own.fyz + up));
            put(c.bz, Mem::template ld2<T>(FZ, own.fyz));
            put(c.bzu, Mem::template ld2<T>(FZ, own.fyz + pl_fz));
        }
    }
    // step j's loads: old values of plane j+4, coefficients of plane j+3
    MGX_HD void load_step(int j, Loads& d) const {
        int zc;
        zaddr(j + 3, zc);
        load_vals(j + 4, d);
        load_pair(zc, d.pc);
        load_cell(ro[0], zc, d.rc.c[0]);
        load_cell(ro[1], zc, d.rc.c[1]);
    }
    MGX_HD static CellCoef<T> pick(const PairCoef<T>& c, int col) {
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
    }
    MGX_HD static CellCoef<T> sel(bool one, const CellCoef<T>& a, const CellCoef<T>& b) {
        CellCoef<T> r;
        r.a = one ? b.a : a.a;
        r.rhs = one ? b.rhs : a.rhs;
        r.bxm = one ? b.bxm : a.bxm;
        r.bxp = one ? b.bxp : a.bxp;
        r.bym = one ? b.bym : a.bym;
        r.byp = one ? b.byp : a.byp;
        r.bzm = one ? b.bzm : a.bzm;
        r.bzp = one ? b.bzp : a.bzp;
        return r;
    }
    MGX_HD static T oldv(bool ok, T p, T c) {
        if (!ok) return T(0);
        T v = ZERO ? T(0) : p;
        if constexpr (CRSE) v += c;
        return v;
    }
    MGX_HD static T* slot(T* S, int z) { return S + (z & (G::NSLOT - 1)) * G::PS; }

    // old values of plane z into LDS
    MGX_HD void store_plane(T* S, int z, const Loads& d) const {
        int zl;
        const bool real = zaddr(z, zl);
        T* P = slot(S, z);
        P[pp] = oldv(real, d.p[0], d.cc);
        P[pp + G::H] = oldv(real, d.p[1], d.cc);
        for (int q = 0; q < 3; ++q) P[hl[q]] = oldv(real && ho[q] != kNone, d.h[q], d.hcc[q]);
    }

    // one GSRB update of LDS cell p (x-neighbour parity q) on plane z
    MGX_HD void cell(T* S, int z, int p, int q, int bits, const CellCoef<T>& c) const {
        T* P = slot(S, z);
        const T v = P[p];
        P[p] = update<T, KIND, Mem>(A.cl, A.alpha, A.dh, bits | zbits(z), v, P[p + G::xl(q)], P[p + G::xr(q)],
                                    P[p - G::RX], P[p + G::RX], slot(S, z - 1)[p], slot(S, z + 1)[p], c);
    }

    MGX_HD void prologue(T* S) {
        const int k = z0 - 6;
        load_vals(k + 2, buf[0]);
        load_vals(k + 3, buf[1]);
        store_plane(S, k + 2, buf[0]);
        store_plane(S, k + 3, buf[1]);
        load_step(k, buf[0]);
    }

    // step k: consumes buf[B] (loaded during the previous step), fills buf[B^1]
    template <int B>
    MGX_HD void step(T* S, int k) {
        Loads& cur = buf[B];
        Loads& nxt = buf[B ^ 1];
        load_step(k + 1, nxt);
        store_plane(S, k + 4, cur);
        const int act = ((yc + k) & 1) ? 0 : 1;  // the column with (x + y + k) odd
        const int p = act ? pp + G::H : pp;  // own column 0 has even lx
        const int ob = act ? obits[1] : obits[0];
        int zl;
        // S1: sweep-1 red on plane k+3
        if (zaddr(k + 3, zl)) cell(S, k + 3, p, act, ob, pick(cur.pc, act));
        const bool r1 = ((rpar0 + k) & 1) == 0;  // ring cell 1 active (cell 0 has (x + y + k) even)
        const int rp = r1 ? rl[1] : rl[0];
        const int rpq = r1 ? rq[1] : rq[0];
        const bool rok = has_ring && (r1 ? rreal[1] : rreal[0]);
        const int rb = r1 ? rbits[1] : rbits[0], rdist = r1 ? rd[1] : rd[0];
        if (rok && zaddr(k + 3, zl)) cell(S, k + 3, rp, rpq, rb, sel(r1, cur.rc.c[0], cur.rc.c[1]));
        // S2: sweep-1 black on plane k+2 (tile and ring depth 2)
        if (k >= z0 - 4 && zaddr(k + 2, zl)) {
            cell(S, k + 2, p, act, ob, pick(pc1, act));
            if (rok && rdist <= 2) cell(S, k + 2, rp, rpq, rb, sel(r1, rh1.c[0], rh1.c[1]));
        }
        // S3: sweep-2 red on plane k+1 (tile and ring depth 1)
        if (k >= z0 - 2 && zaddr(k + 1, zl)) {
            cell(S, k + 1, p, act, ob, pick(pc2, act));
            if (rok && rdist <= 1) cell(S, k + 1, rp, rpq, rb, r3);
        }
        // S4: sweep-2 black on plane k, and the output of plane k
        if (k >= z0 && k < z1) {
            cell(S, k, p, act, ob, c3);
            const T* P = slot(S, k);
            Mem::store2(reinterpret_cast<char*>(A.pout) + (int64_t)pl_c * (k + 1) + own.c, P[pp], P[pp + G::H]);
        }
        // rotate: next step's active column is the other one
        c3 = pick(pc2, act ^ 1);
        pc2 = pc1;
        pc1 = cur.pc;
        r3 = sel(!r1, rh1.c[0], rh1.c[1]);
        rh1 = cur.rc;
    }
};

#undef MGX_HD

}  // namespace sw2
}  // namespace mgx
