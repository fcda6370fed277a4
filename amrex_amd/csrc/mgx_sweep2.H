// Two consecutive red+black GSRB sweeps (abec_gsrb, AMReX_MLABecLap_3D_K.H:
// 210-264; mlpoisson_gsrb, AMReX_MLPoisson_3D_K.H:155-196; each colour after a
// boundary refresh as MLCellLinOpT::smooth does, AMReX_MLCellLinOp.H:1205-1217)
// in one pass over HBM, for dense MG levels whose planes the 64x16 tiles cover
// exactly.  A V-cycle runs two sweeps back to back before and after the
// coarse correction (nu1 = nu2 = 2); fused, each array streams once per pair.
//
// A work-group owns a tile of columns over a range of z planes and marches up
// in z.  Step k runs four stages on four planes:
//   S1  red update of sweep 1 on plane k+3   (tile and ring cells up to L1 distance 3)
//   S2  black update of sweep 1 on plane k+2 (tile and ring up to distance 2)
//   S3  red update of sweep 2 on plane k+1   (tile and ring up to distance 1)
//   S4  black update of sweep 2 on plane k   (tile), then plane k is written.
// Distances are dx + dy outside the tile: a stage reads only the face
// neighbours of what it updates, so each stage's ring is one smaller in L1
// than the previous one's and the ring's far corners are never needed.  Ring
// cells are recomputed from old data, so tiles never wait for each other.  A
// cell is active in every stage of a step or in none ((x+y+k) odd), and each
// stage reads the current step's results only at its own (x, y) column, which
// the same thread computed: with a fixed column-to-thread assignment the only
// cross-thread dependences run from one step to the next, so one barrier per
// step suffices.  (The per-step body is therefore also a correct sequential
// program over the threads of a work-group.)
//
// Threads: one per x-adjacent pair of tile columns; the two rows a wave covers
// have the same parity, so which column of its pairs is active is the same
// for every lane of a wave and the kernel is compiled once per parity with
// static register indexing.  Ring cells come in adjacent pairs (one of each
// colour), grouped in two colour classes of (3 TX + 3 TY + 6) / 2 pairs, each
// class on waves of its own matching parity.  The cells at L1 distance 4 hold
// old values only (one per thread).  LDS holds eight planes (z & 7) of the
// tile with a halo of four cells.
//
// Coefficients of a plane are loaded once, for both cells of each pair, and
// kept in registers for the four steps that use them; a cell's diagonal and
// omega / (diagonal - boundary term) are formed at its first update and
// reused by its second.  The loads of the tile's plane k+4 are issued at the
// start of step k, a step ahead of their use; the ring's coefficients at the
// start of the step that uses them, whose tile stages run first.  Raw buffer
// loads through whole-array resources with the plane as a scalar offset; the
// loop is unrolled by four so the plane buffers rotate without copies, and
// the closure table sits in LDS (it is read only on physical faces).
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
// Coefficients of one cell update: right-hand side, the diagonal g (holds a
// until formed), w = omega / (g - boundary term), the six face coefficients.
template <class T>
struct Cell {
    T rhs, g, w, bxm, bxp, bym, byp, bzm, bzp;
};
// both cells of a tile pair (cells x, x+1 of row y): x faces x .. x+2
template <class T>
struct PairCoef {
    T rhs[2], g[2], w[2], bx[3], by[2], byu[2], bz[2], bzu[2];
};
template <class T>
struct RingCoef {
    Cell<T> m[2];
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

constexpr double kOmega = 1.15;

// Diagonal of one update: g = alpha a + sum dh (b- + b+) (ABecLaplacian) or
// -2 sum dh (Poisson) and w = omega / (g - delta), delta the closure's
// boundary term on physical faces (bits: x-lo 1, y-lo 2, z-lo 4, x-hi 8,
// y-hi 16, z-hi 32).  g_int / w_int: the Poisson values of interior cells.
template <class T, int KIND, class Mem>
MGX_HD void diag(const Closure<T>& C, T alpha, const T* dh, T g_int, T w_int, int bits, Cell<T>& c) {
    constexpr T omega = T(kOmega);
    const T dhx = dh[0], dhy = dh[1], dhz = dh[2];
    const bool edge = Mem::any(bits != 0);
    if constexpr (KIND == 0) {
        const T gamma = alpha * c.g + dhx * (c.bxm + c.bxp) + dhy * (c.bym + c.byp) + dhz * (c.bzm + c.bzp);
        T g_m_d = gamma;
        if (edge) {
            const T cf0 = (bits & 1) ? C.cf[0] : T(0), cf1 = (bits & 2) ? C.cf[1] : T(0);
            const T cf2 = (bits & 4) ? C.cf[2] : T(0), cf3 = (bits & 8) ? C.cf[3] : T(0);
            const T cf4 = (bits & 16) ? C.cf[4] : T(0), cf5 = (bits & 32) ? C.cf[5] : T(0);
            g_m_d = gamma - (dhx * (c.bxm * cf0 + c.bxp * cf3) + dhy * (c.bym * cf1 + c.byp * cf4) +
                             dhz * (c.bzm * cf2 + c.bzp * cf5));
        }
        c.g = gamma;
        c.w = omega / g_m_d;
    } else {
        c.g = g_int;
        c.w = w_int;
        if (edge) {
            const T cf0 = (bits & 1) ? C.cf[0] : T(0), cf1 = (bits & 2) ? C.cf[1] : T(0);
            const T cf2 = (bits & 4) ? C.cf[2] : T(0), cf3 = (bits & 8) ? C.cf[3] : T(0);
            const T cf4 = (bits & 16) ? C.cf[4] : T(0), cf5 = (bits & 32) ? C.cf[5] : T(0);
            c.w = omega / (g_int + dhx * (cf0 + cf3) + dhy * (cf1 + cf4) + dhz * (cf2 + cf5));
        }
    }
}

// One GSRB update of value v from its six neighbours, with the diagonal of
// diag(); on physical faces the closure replaces the ghost neighbour:
// ghost = k2 * inner + k1 * cell (mllinop_apply_bc_*, AMReX_MLLinOp_K.H:14-115).
template <class T, int KIND, class Mem>
MGX_HD T relax(const Closure<T>& C, const T* dh, int bits, T v, T xm, T xp, T ym, T yp, T zm, T zp,
               const Cell<T>& c) {
    if (Mem::any(bits != 0)) {
        const bool b0 = bits & 1, b1 = bits & 2, b2 = bits & 4, b3 = bits & 8, b4 = bits & 16, b5 = bits & 32;
        xm = b0 ? fma(xp, C.k2[0], v * C.k1[0]) : xm;
        xp = b3 ? fma(xm, C.k2[3], v * C.k1[3]) : xp;
        ym = b1 ? fma(yp, C.k2[1], v * C.k1[1]) : ym;
        yp = b4 ? fma(ym, C.k2[4], v * C.k1[4]) : yp;
        zm = b2 ? fma(zp, C.k2[2], v * C.k1[2]) : zm;
        zp = b5 ? fma(zm, C.k2[5], v * C.k1[5]) : zp;
    }
    const T dhx = dh[0], dhy = dh[1], dhz = dh[2];
    if constexpr (KIND == 0) {
        const T rho = dhx * (c.bxm * xm + c.bxp * xp) + dhy * (c.bym * ym + c.byp * yp) + dhz * (c.bzm * zm + c.bzp * zp);
        const T res = c.rhs - (c.g * v - rho);
        return v + c.w * res;
    } else {
        const T res = c.rhs - c.g * v - dhx * (xm + xp) - dhy * (ym + yp) - dhz * (zm + zp);
        return v + c.w * res;
    }
}

// Geometry of a TX x TY tile with a halo of four cells in LDS.
template <int TX, int TY>
struct Geo {
    static constexpr int NTT = TX / 2 * TY;          // tile threads (waves 0 .. NTT/64 - 1)
    static constexpr int TPR = TX / 2;               // threads per tile row
    static constexpr int RPW = 64 / TPR;             // tile rows per wave
    static constexpr int RX = TX + 8, RY = TY + 8;   // LDS plane extents
    static constexpr int PS = RX * RY + 1;           // plane stride (+1: a slot for discarded stores)
    static constexpr int DUMMY = RX * RY;
    static constexpr int NRC = (3 * TX + 3 * TY + 6) / 2;  // ring pairs of one colour class
    static constexpr int WPC = (NRC + 63) / 64;            // waves of one class
    static constexpr int NT = NTT + 2 * WPC * 64;          // threads: tile waves, then ring waves
    static constexpr int N4 = 2 * TX + 2 * TY + 12;        // cells at L1 distance 4
    static constexpr int NSLOT = 8;
    static constexpr int H = RX / 2;                 // an LDS row holds its even columns, then its odd ones
    static_assert(TPR <= 64 && 64 % TPR == 0 && TY % (2 * RPW) == 0, "waves of whole rows of one parity");
    static_assert(NTT % 64 == 0, "whole tile waves");
    static_assert(N4 <= 2 * WPC * 64, "one distance-4 cell per ring thread");
    // cells of one colour in a row are consecutive words: the active cells of
    // a wave's lanes hit distinct banks
    MGX_HD static int lpos(int lx, int ly) { return RX * ly + (lx & 1) * H + (lx >> 1); }
    // offsets of the x neighbours of a cell whose lx has parity q
    MGX_HD static constexpr int xl(int q) { return q ? -H : H - 1; }
    MGX_HD static constexpr int xr(int q) { return q ? 1 - H : H; }
    // tile row of thread t: wave w covers rows of parity w & 1
    MGX_HD static int row(int t) {
        const int w = t >> 6, i = (t / TPR) % RPW;
        return 2 * RPW * (w >> 1) + (w & 1) + 2 * i;
    }
    // L1 distance of LDS cell (lx, ly) from the tile (lx in [4, TX+3], ly in [4, TY+3])
    MGX_HD static int dist(int lx, int ly) {
        const int dx = lx < 4 ? 4 - lx : (lx > TX + 3 ? lx - (TX + 3) : 0);
        const int dy = ly < 4 ? 4 - ly : (ly > TY + 3 ? ly - (TY + 3) : 0);
        return dx + dy;
    }
    // ring pair q < NRC of colour class c (the parity of lx + ly of member 0):
    // rows at dy = 1, 2, 3 above the tile (class 1) or below it (class 0),
    // then three columns of vertical pairs beside it (dx = 1, 3 left and 2
    // right for class 1; the other three for class 0)
    MGX_HD static void ring_pair(int c, int q, int& lx0, int& ly0, int& lx1, int& ly1) {
        for (int dy = 1; dy <= 3; ++dy) {
            const int np = TX / 2 + 3 - dy;
            if (q < np) {
                ly0 = ly1 = c ? 4 - dy : TY + 3 + dy;
                lx0 = 1 + dy + 2 * q;
                lx1 = lx0 + 1;
                return;
            }
            q -= np;
        }
        const int col = q / (TY / 2), j = q % (TY / 2);
        const int lx = c ? (col == 0 ? 3 : (col == 1 ? 1 : TX + 5)) : (col == 0 ? 2 : (col == 1 ? TX + 4 : TX + 6));
        lx0 = lx1 = lx;
        ly0 = 4 + 2 * j;
        ly1 = ly0 + 1;
    }
    // distance-4 cell q < N4: rows above and below, columns beside, and the
    // cells (dx, dy) = (1, 3), (2, 2), (3, 1) at the four corners
    MGX_HD static void ring4(int q, int& lx, int& ly) {
        if (q < 2 * TX) {
            lx = 4 + q % TX;
            ly = q < TX ? 0 : RY - 1;
            return;
        }
        q -= 2 * TX;
        if (q < 2 * TY) {
            lx = q < TY ? 0 : RX - 1;
            ly = 4 + q % TY;
            return;
        }
        q -= 2 * TY;
        const int corner = q / 3, i = q % 3;
        const int dx = 1 + i, dy = 3 - i;
        lx = (corner & 1) ? TX + 3 + dx : 4 - dx;
        ly = (corner & 2) ? TY + 3 + dy : 4 - dy;
    }
};

// In-plane byte offsets of one cell: FAB, x faces, y and z faces (x + n0*y), coarse FAB.
struct CellOff {
    uint32_t c, fx, fyz, cc;
};

constexpr uint32_t kNone = 0xffffffffu;

// What every thread of a work-group keeps: the launch's constants, the tile's
// place, whole-array resources and the helpers both roles use.
template <class T, int KIND, int TX, int TY, bool ZERO, bool CRSE, class Mem>
struct Base {
    using G = Geo<TX, TY>;
    using R = typename Mem::Rsrc;
    static constexpr uint32_t ST = sizeof(T);

    Args<T> A;
    T g_int, w_int;         // Poisson diagonal of interior cells
    int z0, z1;
    int x0, y0;
    int p0;                 // column 0 (ring member 0) is active in the first step iff p0 == 1
    uint32_t pl_c, pl_cc, pl_fx, pl_fy, pl_fz;
    const Closure<T>* cl;   // closure table (LDS in the kernel)
    R r_in, r_out, r_rhs, r_a, r_bx, r_by, r_bz, r_crse;  // whole-array resources

    MGX_HD bool wrapc(int x, int n, int per, int& xw) const {
        xw = x;
        if (x >= 0 && x < n) return true;
        if (!per) return false;
        xw = x < 0 ? x + n : x - n;
        return true;
    }
    // plane to address for z (clamped beyond a non-periodic face); true for a
    // real plane.  z lies in [-6, n2 + 8) and n2 >= 8: one wrap suffices.
    MGX_HD bool zaddr(int z, int& zl) const {
        const int n2 = A.n2;
        if (z >= 0 && z < n2) {
            zl = z;
            return true;
        }
        if (A.per2) {
            zl = z < 0 ? z + n2 : z - n2;
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
    // par: the x + y parity of column 0 (ring member 0)
    MGX_HD void init_base(const Args<T>& args, const Closure<T>* clp, int bid, int nblk, int par) {
        A = args;
        cl = clp;
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
        // a cell is active when x + y + k is odd
        p0 = (par + z0 - 6) & 1;
        {
            const T dhx = A.dh[0], dhy = A.dh[1], dhz = A.dh[2];
            g_int = T(-2.) * (dhx + dhy + dhz);
            w_int = T(kOmega) / g_int;
        }
        pl_c = (uint32_t)(A.sz * ST);
        pl_cc = (uint32_t)(A.csz * ST);
        pl_fx = (uint32_t)(n0 + 1) * n1 * ST;
        pl_fy = (uint32_t)n0 * (n1 + 1) * ST;
        pl_fz = (uint32_t)n0 * n1 * ST;
        auto rs = [](const void* p, uint64_t n) { return Mem::rsrc(reinterpret_cast<const char*>(p), (uint32_t)n); };
        r_in = rs(A.pin, (uint64_t)pl_c * (n2 + 2));
        r_out = rs(A.pout, (uint64_t)pl_c * (n2 + 2));
        r_rhs = rs(A.rhs, (uint64_t)pl_c * (n2 + 2));
        r_a = rs(KIND == 0 ? A.a : A.rhs, (uint64_t)pl_c * (n2 + 2));
        r_bx = rs(KIND == 0 ? A.b[0] : A.rhs, (uint64_t)pl_fx * n2);
        r_by = rs(KIND == 0 ? A.b[1] : A.rhs, (uint64_t)pl_fy * n2);
        r_bz = rs(KIND == 0 ? A.b[2] : A.rhs, (uint64_t)pl_fz * (n2 + 1));
        r_crse = rs(A.crse, (uint64_t)pl_cc * ((n2 >> 1) + 2));
    }

    // coefficients of one cell on plane zl
    MGX_HD void load_cell(const CellOff& o, int zl, Cell<T>& c) const {
        const uint32_t sc = pl_c * (uint32_t)(zl + 1);
        c.rhs = Mem::template ld<T>(r_rhs, o.c, sc);
        if constexpr (KIND == 0) {
            const uint32_t sz = pl_fz * (uint32_t)zl;
            c.g = Mem::template ld<T>(r_a, o.c, sc);
            c.bxm = Mem::template ld<T>(r_bx, o.fx, pl_fx * (uint32_t)zl);
            c.bxp = Mem::template ld<T>(r_bx, o.fx + ST, pl_fx * (uint32_t)zl);
            c.bym = Mem::template ld<T>(r_by, o.fyz, pl_fy * (uint32_t)zl);
            c.byp = Mem::template ld<T>(r_by, o.fyz + (uint32_t)A.n0 * ST, pl_fy * (uint32_t)zl);
            c.bzm = Mem::template ld<T>(r_bz, o.fyz, sz);
            c.bzp = Mem::template ld<T>(r_bz, o.fyz, sz + pl_fz);
        }
    }
    MGX_HD static T oldv(bool ok, T p, T c) {
        if (!ok) return T(0);
        T v = ZERO ? T(0) : p;
        if constexpr (CRSE) v += c;
        return v;
    }
    MGX_HD static T* slot(T* S, int z) { return S + (z & (G::NSLOT - 1)) * G::PS; }

    // a cell's value and its four in-plane neighbours on one plane
    struct Nb {
        T v, xm, xp, ym, yp;
    };
    MGX_HD static Nb nb(const T* P, int p, int pxl, int pxr) {
        return Nb{P[p], P[pxl], P[pxr], P[p - G::RX], P[p + G::RX]};
    }
    MGX_HD T relax_nb(int bits, const Nb& n, T zm, T zp, const Cell<T>& c) const {
        return relax<T, KIND, Mem>(*cl, A.dh, bits, n.v, n.xm, n.xp, n.ym, n.yp, zm, zp, c);
    }
    MGX_HD void diag_cell(int bits, Cell<T>& c) const {
        diag<T, KIND, Mem>(*cl, A.alpha, A.dh, g_int, w_int, bits, c);
    }
};

// A tile thread: a pair of tile columns (x, x+1 of one row).
//
// Step k, the S_-th of a group of four (k - z0 + 6 = S_ mod 4), column
// (ring member) C = PAR ? 0 : 1 active.  Every LDS value a step reads was
// written in an earlier step: the in-plane neighbours of an active cell are
// inactive cells, untouched this step, and its column's values on planes
// k-1 .. k+3 were last written two steps ago.  So a step reads them all at
// once, runs its column's stages in registers (each stage's result is the
// next one's upper z neighbour), and writes the results as they come.  The
// loads of plane k+4's coefficients and plane k+5's old values are issued a
// step ahead of their use.
template <class T, int KIND, int TX, int TY, bool ZERO, bool CRSE, class Mem>
struct TileThread : Base<T, KIND, TX, TY, ZERO, CRSE, Mem> {
    using B = Base<T, KIND, TX, TY, ZERO, CRSE, Mem>;
    using G = typename B::G;
    using Nb = typename B::Nb;
    using B::A;
    using B::z0;
    using B::z1;
    using B::pl_c;

    int pp;                 // LDS position of own column 0
    CellOff own;            // own pair (columns xe, xe+1)
    int obits[2];           // in-plane closure bits of the own columns
    struct Vals {
        T v[2], c;
    };
    Vals V[2];              // old values of plane k+4 (V[s & 1]) and k+5 (loading)
    PairCoef<T> Q[4];       // coefficients of plane j in Q[(j - z0 + 6) & 3]

    MGX_HD void init(const Args<T>& args, const Closure<T>* clp, int bid, int nblk, int t) {
        B::init_base(args, clp, bid, nblk, (t >> 6) & 1);
        const int px = t % G::TPR, py = G::row(t);
        const int yc = B::y0 + py, xe = B::x0 + 2 * px;
        pp = G::lpos(4 + 2 * px, 4 + py);
        own = B::offs(xe, yc);
        obits[0] = B::face_bits(xe, yc);
        obits[1] = B::face_bits(xe + 1, yc);
        for (int i = 0; i < 4; ++i) Q[i] = PairCoef<T>{};
        V[0] = V[1] = Vals{};
    }
    MGX_HD void load_vals(int z, Vals& d) const {
        int zv;
        B::zaddr(z, zv);
        if constexpr (!ZERO) {
            const auto v = Mem::template ld2<T>(B::r_in, own.c, pl_c * (uint32_t)(zv + 1));
            d.v[0] = v.x;
            d.v[1] = v.y;
        }
        if constexpr (CRSE) d.c = Mem::template ld<T>(B::r_crse, own.cc, B::pl_cc * (uint32_t)((zv >> 1) + 1));
    }
    MGX_HD void load_coefs(int z, PairCoef<T>& c) const {
        int zl;
        B::zaddr(z, zl);
        auto put = [](T* dst, Two<T> v) {
            dst[0] = v.x;
            dst[1] = v.y;
        };
        const uint32_t sc = pl_c * (uint32_t)(zl + 1);
        put(c.rhs, Mem::template ld2<T>(B::r_rhs, own.c, sc));
        if constexpr (KIND == 0) {
            const uint32_t sx = B::pl_fx * (uint32_t)zl, sy = B::pl_fy * (uint32_t)zl, sz = B::pl_fz * (uint32_t)zl;
            put(c.g, Mem::template ld2<T>(B::r_a, own.c, sc));
            put(c.bx, Mem::template ld2<T>(B::r_bx, own.fx, sx));
            c.bx[2] = Mem::template ld<T>(B::r_bx, own.fx + 2 * B::ST, sx);
            put(c.by, Mem::template ld2<T>(B::r_by, own.fyz, sy));
            put(c.byu, Mem::template ld2<T>(B::r_by, own.fyz + (uint32_t)A.n0 * B::ST, sy));
            put(c.bz, Mem::template ld2<T>(B::r_bz, own.fyz, sz));
            put(c.bzu, Mem::template ld2<T>(B::r_bz, own.fyz, sz + B::pl_fz));
        }
    }
    template <int C>
    MGX_HD static Cell<T> cell(const PairCoef<T>& q) {
        Cell<T> c;
        c.rhs = q.rhs[C];
        c.g = q.g[C];
        c.w = q.w[C];
        c.bxm = q.bx[C];
        c.bxp = q.bx[C + 1];
        c.bym = q.by[C];
        c.byp = q.byu[C];
        c.bzm = q.bz[C];
        c.bzp = q.bzu[C];
        return c;
    }
    // old values of plane z into LDS
    MGX_HD void store_plane(T* S, int z, const Vals& d, T& o0, T& o1) const {
        int zl;
        const bool real = B::zaddr(z, zl);
        T* P = B::slot(S, z);
        o0 = B::oldv(real, d.v[0], d.c);
        o1 = B::oldv(real, d.v[1], d.c);
        P[pp] = o0;
        P[pp + G::H] = o1;
    }
    MGX_HD void prologue(T* S) {
        const int k = z0 - 6;
        T o0, o1;
        load_vals(k + 2, V[0]);
        load_vals(k + 3, V[1]);
        store_plane(S, k + 2, V[0], o0, o1);
        store_plane(S, k + 3, V[1], o0, o1);
        load_vals(k + 4, V[0]);
        load_coefs(k + 3, Q[3]);
    }
    // S1 sweep-1 red on plane k+3 and S2 sweep-1 black on plane k+2 (first
    // updates: form the diagonal), S3 sweep-2 red on plane k+1 and S4 sweep-2
    // black on plane k (diagonals from S1, S2), then plane k is written.
    template <int S_, int PAR>
    MGX_HD void step(T* S, int k) {
        constexpr int C = PAR ? 0 : 1;
        constexpr int QK = S_ & 3, Q1 = (S_ + 1) & 3, Q2 = (S_ + 2) & 3, Q3 = (S_ + 3) & 3;
        constexpr int XL = G::xl(C), XR = G::xr(C);
        int zl;
        // S4's coefficients leave Q[QK] before plane k+4's loads land there
        const Cell<T> c4 = cell<C>(Q[QK]);
        load_coefs(k + 4, Q[QK]);
        load_vals(k + 5, V[(S_ + 1) & 1]);
        T* Pm = B::slot(S, k - 1);
        T* P0 = B::slot(S, k);
        T* P1 = B::slot(S, k + 1);
        T* P2 = B::slot(S, k + 2);
        T* P3 = B::slot(S, k + 3);
        const int p = pp + C * G::H;
        const Nb t3 = B::nb(P3, p, p + XL, p + XR), t2 = B::nb(P2, p, p + XL, p + XR);
        const Nb t1 = B::nb(P1, p, p + XL, p + XR), t0 = B::nb(P0, p, p + XL, p + XR);
        const T tm = Pm[p], other = P0[pp + (1 - C) * G::H];
        T o0, o1;
        store_plane(S, k + 4, V[S_ & 1], o0, o1);
        const bool on3 = B::zaddr(k + 3, zl);
        const bool on2 = k >= z0 - 4 && B::zaddr(k + 2, zl);
        const bool on1 = k >= z0 - 2 && B::zaddr(k + 1, zl);
        const bool on0 = k >= z0 && k < z1;
        T u3 = t3.v, u2 = t2.v, u1 = t1.v;
        if (on3) {
            const int bits = obits[C] | B::zbits(k + 3);
            Cell<T> c = cell<C>(Q[Q3]);
            B::diag_cell(bits, c);
            Q[Q3].g[C] = c.g;
            Q[Q3].w[C] = c.w;
            u3 = B::relax_nb(bits, t3, t2.v, C == 0 ? o0 : o1, c);
            P3[p] = u3;
        }
        if (on2) {
            const int bits = obits[C] | B::zbits(k + 2);
            Cell<T> c = cell<C>(Q[Q2]);
            B::diag_cell(bits, c);
            Q[Q2].g[C] = c.g;
            Q[Q2].w[C] = c.w;
            u2 = B::relax_nb(bits, t2, t1.v, u3, c);
            P2[p] = u2;
        }
        if (on1) {
            u1 = B::relax_nb(obits[C] | B::zbits(k + 1), t1, t0.v, u2, cell<C>(Q[Q1]));
            P1[p] = u1;
        }
        if (on0) {
            const T u0 = B::relax_nb(obits[C] | B::zbits(k), t0, tm, u1, c4);
            P0[p] = u0;
            Mem::template st2<T>(B::r_out, own.c, pl_c * (uint32_t)(k + 1), C == 0 ? u0 : other, C == 0 ? other : u0);
        }
    }
};

// A ring thread: a pair of adjacent ring cells (one of each colour) and, the
// first N4 of the ring threads, a cell at distance 4 whose old values it
// stores.  The stages S1 .. S3 on the active ring cell (ring of depth 3, 2, 1),
// the same way as a tile thread's.
template <class T, int KIND, int TX, int TY, bool ZERO, bool CRSE, class Mem>
struct RingThread : Base<T, KIND, TX, TY, ZERO, CRSE, Mem> {
    using B = Base<T, KIND, TX, TY, ZERO, CRSE, Mem>;
    using G = typename B::G;
    using Nb = typename B::Nb;
    using B::A;
    using B::z0;
    using B::pl_c;

    bool has_ring;
    int rl[2];              // LDS positions of the ring cells
    int rxl[2], rxr[2];     // and of their x neighbours
    int rlim[2];            // stages that update them: 4 - distance (0: none)
    int rbits[2];
    CellOff ro[2];
    int hl4;                // LDS position of the distance-4 cell
    uint32_t ho4, hc4;      // its FAB / coarse in-plane offsets (kNone: ghost)
    struct Vals {
        T v[3], c[3];       // both ring cells, the distance-4 cell; coarse terms
    };
    Vals V[2];
    RingCoef<T> RQ[4];      // coefficients of plane j in RQ[(j - z0 + 6) & 3]

    // r: index among the ring threads
    MGX_HD void init(const Args<T>& args, const Closure<T>* clp, int bid, int nblk, int r) {
        const int rw = r >> 6;
        B::init_base(args, clp, bid, nblk, rw & 1);
        const int n0 = A.n0, n1 = A.n1;
        const int q = (rw >> 1) * 64 + (r & 63);
        has_ring = q < G::NRC;
        hl4 = G::DUMMY;
        ho4 = hc4 = kNone;
        for (int m = 0; m < 2; ++m) {
            rl[m] = rxl[m] = rxr[m] = G::DUMMY;
            rlim[m] = 0;
            rbits[m] = 0;
            ro[m] = B::offs(0, 0);
        }
        if (has_ring) {
            int lx[2], ly[2];
            G::ring_pair(rw & 1, q, lx[0], ly[0], lx[1], ly[1]);
            for (int m = 0; m < 2; ++m) {
                rl[m] = G::lpos(lx[m], ly[m]);
                rxl[m] = rl[m] + G::xl(lx[m] & 1);
                rxr[m] = rl[m] + G::xr(lx[m] & 1);
                int xw, yw;
                if (B::wrapc(B::x0 - 4 + lx[m], n0, A.per0, xw) && B::wrapc(B::y0 - 4 + ly[m], n1, A.per1, yw)) {
                    ro[m] = B::offs(xw, yw);
                    rbits[m] = B::face_bits(xw, yw);
                    rlim[m] = 4 - G::dist(lx[m], ly[m]);
                }
            }
        }
        if (r < G::N4) {
            int lx, ly, xw, yw;
            G::ring4(r, lx, ly);
            hl4 = G::lpos(lx, ly);
            if (B::wrapc(B::x0 - 4 + lx, n0, A.per0, xw) && B::wrapc(B::y0 - 4 + ly, n1, A.per1, yw)) {
                const CellOff o = B::offs(xw, yw);
                ho4 = o.c;
                hc4 = o.cc;
            }
        }
        for (int i = 0; i < 4; ++i) RQ[i] = RingCoef<T>{};
        V[0] = V[1] = Vals{};
    }
    MGX_HD void load_vals(int z, Vals& d) const {
        int zv;
        B::zaddr(z, zv);
        if constexpr (!ZERO) {
            const uint32_t so = pl_c * (uint32_t)(zv + 1);
            d.v[0] = Mem::template ld<T>(B::r_in, ro[0].c, so);
            d.v[1] = Mem::template ld<T>(B::r_in, ro[1].c, so);
            d.v[2] = Mem::template ld<T>(B::r_in, ho4 == kNone ? 0u : ho4, so);
        }
        if constexpr (CRSE) {
            const uint32_t so = B::pl_cc * (uint32_t)((zv >> 1) + 1);
            d.c[0] = Mem::template ld<T>(B::r_crse, ro[0].cc, so);
            d.c[1] = Mem::template ld<T>(B::r_crse, ro[1].cc, so);
            d.c[2] = Mem::template ld<T>(B::r_crse, hc4 == kNone ? 0u : hc4, so);
        }
    }
    MGX_HD void load_ring(int z, RingCoef<T>& r) const {
        int zl;
        B::zaddr(z, zl);
        B::load_cell(ro[0], zl, r.m[0]);
        B::load_cell(ro[1], zl, r.m[1]);
    }
    MGX_HD void store_plane(T* S, int z, const Vals& d, T& o0, T& o1) const {
        int zl;
        const bool real = B::zaddr(z, zl);
        T* P = B::slot(S, z);
        o0 = B::oldv(real && rlim[0] > 0, d.v[0], d.c[0]);
        o1 = B::oldv(real && rlim[1] > 0, d.v[1], d.c[1]);
        P[rl[0]] = o0;
        P[rl[1]] = o1;
        P[hl4] = B::oldv(real && ho4 != kNone, d.v[2], d.c[2]);
    }
    MGX_HD void prologue(T* S) {
        const int k = z0 - 6;
        T o0, o1;
        load_vals(k + 2, V[0]);
        load_vals(k + 3, V[1]);
        store_plane(S, k + 2, V[0], o0, o1);
        store_plane(S, k + 3, V[1], o0, o1);
        load_vals(k + 4, V[0]);
        if (has_ring) load_ring(k + 3, RQ[3]);
    }
    template <int S_, int PAR>
    MGX_HD void step(T* S, int k) {
        constexpr int C = PAR ? 0 : 1;
        constexpr int QK = S_ & 3, Q1 = (S_ + 1) & 3, Q2 = (S_ + 2) & 3, Q3 = (S_ + 3) & 3;
        int zl;
        if (has_ring) load_ring(k + 4, RQ[QK]);
        load_vals(k + 5, V[(S_ + 1) & 1]);
        T* P0 = B::slot(S, k);
        T* P1 = B::slot(S, k + 1);
        T* P2 = B::slot(S, k + 2);
        T* P3 = B::slot(S, k + 3);
        const int q = rl[C];
        const Nb r3 = B::nb(P3, q, rxl[C], rxr[C]), r2 = B::nb(P2, q, rxl[C], rxr[C]);
        const Nb r1 = B::nb(P1, q, rxl[C], rxr[C]);
        const T r0 = P0[q];
        T o0, o1;
        store_plane(S, k + 4, V[S_ & 1], o0, o1);
        const bool on3 = B::zaddr(k + 3, zl);
        const bool on2 = k >= z0 - 4 && B::zaddr(k + 2, zl);
        const bool on1 = k >= z0 - 2 && B::zaddr(k + 1, zl);
        const int lim = rlim[C];
        T v3 = r3.v, v2 = r2.v;
        if (on3 && lim >= 1) {
            const int bits = rbits[C] | B::zbits(k + 3);
            Cell<T>& c = RQ[Q3].m[C];
            B::diag_cell(bits, c);
            v3 = B::relax_nb(bits, r3, r2.v, C == 0 ? o0 : o1, c);
            P3[q] = v3;
        }
        if (on2 && lim >= 2) {
            const int bits = rbits[C] | B::zbits(k + 2);
            Cell<T> c = RQ[Q2].m[C];
            B::diag_cell(bits, c);
            v2 = B::relax_nb(bits, r2, r1.v, v3, c);
            P2[q] = v2;
        }
        if (on1 && lim >= 3) P1[q] = B::relax_nb(rbits[C] | B::zbits(k + 1), r1, r0, v2, RQ[Q1].m[C]);
    }
};

#undef MGX_HD

}  // namespace sw2
}  // namespace mgx
