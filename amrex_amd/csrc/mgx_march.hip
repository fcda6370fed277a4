// Residual passes of dense MG levels as pipelined z-marches.
//
//   update_residual (forward): sol_out = sol + cor; r = rhs - L(sol_out) with
//       the level BC (inhomogeneous Dirichlet data when given); per-block max|r|
//       and sum r; r stored as the next V-cycle's right-hand side.  Covers
//       MLMGT::oneIter's LocalAdd (AMReX_MLMG.H:1264), computeResidual
//       (AMReX_MLMG.H:1652) and ResNormInf (AMReX_MLMG.H:1808).
//   resid_restrict: crse = average_down(res - L(cor)) with homogeneous BC:
//       computeResOfCorrection + restriction of mgVcycle (AMReX_MLMG.H:1330-1339)
//       without storing the fine residual of the correction.
//
// The layout is the fused sweep's: a work-group owns a 64x8 tile of columns
// over a range of z planes, one thread per x-adjacent pair of columns, and
// marches up in z with four planes of x (tile + one-cell halo) in LDS.  Every
// load of step k+1 (x of plane k+2, both columns' coefficients of plane k+1)
// is an unconditional raw buffer load issued at the start of step k and not
// touched before step k+1, so loads stay in flight across the step's barrier.
// Ghosts across physical faces are closed inline from LDS values (levels whose
// closures reach a third interior cell use the per-cell kernels instead).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>
#include <stdexcept>

#include "mgx_buffer.H"
#include "mgx_device.H"
#include "mgx_march.H"

namespace mgx {
namespace {

// A work-group owns a tile of MX x kMY columns (MX = 64, or 128 on large
// planes, where rows of 512+ bytes per array keep HBM accesses long), one
// thread per x pair: a wave holds 32 pairs of two rows, MX/64 waves side by
// side.
constexpr int kMY = 8;
template <int MX>
struct MTile {
    static constexpr int T = MX / 2 * kMY;   // threads per work-group
    static constexpr int R = MX + 2;         // LDS row length (tile + halo)
    static constexpr int S = R * (kMY + 2);  // LDS plane
    // An LDS row holds its even columns, then its odd ones, so the cells a
    // wave reads in one access (one column of consecutive pairs) are
    // consecutive words.
    static constexpr int H = R / 2;
    static_assert(R % 2 == 0, "split rows");
    static_assert(2 * (MX + 2) + 2 * kMY <= T, "one halo position per thread");
};
int march_width(const LevelDev& L) { return (L.n[0] >= 256 && L.n[1] >= 256) ? 128 : 64; }

// ghost = v2*k2 + v1*k1 (+ k0*boundary value), the closure of
// mllinop_apply_bc_* (AMReX_MLLinOp_K.H:14-115) for NX <= 3; k0 = 0 for
// Neumann and reflect-odd faces.  For finite values it rounds as that loop does.
template <class T>
struct ClosureA {
    T k1[6], k2[6], k0[6];
};

// L(x) at a cell: mlabeclap_adotx (AMReX_MLABecLap_3D_K.H:9-28) and
// mlpoisson_adotx (AMReX_MLPoisson_3D_K.H:9-16), written as the per-cell
// kernels write them.
template <class T, int KIND>
__device__ __forceinline__ T adotx(T alpha, const T* dh, T a, T bxm, T bxp, T bym, T byp, T bzm, T bzp, T xc, T xm,
                                   T xp, T ym, T yp, T zm, T zp) {
    if constexpr (KIND == 0) {
        return alpha * a * xc - dh[0] * (bxp * (xp - xc) - bxm * (xc - xm)) - dh[1] * (byp * (yp - xc) - bym * (xc - ym)) -
               dh[2] * (bzp * (zp - xc) - bzm * (xc - zm));
    } else {
        return dh[0] * (xm - T(2.0) * xc + xp) + dh[1] * (ym - T(2.0) * xc + yp) + dh[2] * (zm - T(2.0) * xc + zp);
    }
}

__device__ __forceinline__ double nanmax_m(double a, double b) { return (b > a || b != b) ? b : a; }

// MODE 0: update_residual (T = double, x = sol + cor with cor of type TC when
// COR, r of type TR).  MODE 1: resid_restrict (x = cor, rhs = res, T = TC = TR).
// VEC (n0 even: both columns of every pair are real cells): each pair's two
// values of an array come in one load.
template <int MODE, int KIND, class T, class TC, class TR, bool COR, bool VEC, int MX>
__global__ __launch_bounds__(MTile<MX>::T) void k_march(LevelDev L, OpCoef<T> op, const T* __restrict__ xin,
                                               const TC* __restrict__ cor, T* __restrict__ sol_out,
                                               const T* __restrict__ rhs, const double* __restrict__ bcv, int inhomog,
                                               TR* __restrict__ res, double* __restrict__ partials, LevelDev Lc,
                                               T* __restrict__ crse, int cz) {
    constexpr uint32_t ST = sizeof(T), SC = sizeof(TC);
    constexpr int kMX = MX, kMT = MTile<MX>::T, kMR = MTile<MX>::R, kMS = MTile<MX>::S, kMH = MTile<MX>::H;
    constexpr int WPR = MX / 64;  // waves side by side over a row pair
    auto mpos = [](int lx, int lyy) { return kMR * lyy + (lx & 1) * kMH + (lx >> 1); };
    __shared__ T S[4][kMS + 1];  // + one slot for stores of lanes with nothing to store
    __shared__ ClosureA<T> CL;
    __shared__ double red[2 * kMT / 64];
    constexpr int DUMMY = kMS;

    const BoxDev& B = L.box[0];
    if (threadIdx.x < 6) {
        const int f = threadIdx.x;
        T k1 = T(0), k2 = T(0), k0 = T(0);
        if (B.bct[f] == BC_NEUMANN) {
            k1 = T(1);
        } else if (B.bct[f] == BC_REFLECT_ODD) {
            k1 = T(-1);
        } else if (B.bct[f] == BC_DIRICHLET) {
            if (B.nxo[f] > 1) k1 = T(B.bcc[f][1]);
            if (B.nxo[f] > 2) k2 = T(B.bcc[f][2]);
            k0 = T(B.bcc[f][0]);
        }
        CL.k1[f] = k1;
        CL.k2[f] = k2;
        CL.k0[f] = k0;
    }
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    const int per0 = L.periodic[0], per1 = L.periodic[1], per2 = L.periodic[2];
    const bool zlo_on = B.bct[2] != BC_INTERIOR, zhi_on = B.bct[5] != BC_INTERIOR;
    const int sy = (int)B.sy;
    const int64_t sz = B.sz;
    const char* gx = reinterpret_cast<const char*>(xin + B.off);
    const char* gcor = COR ? reinterpret_cast<const char*>(cor + B.off) : nullptr;
    const char* grhs = reinterpret_cast<const char*>(rhs + B.off);
    const char* ga = KIND == 0 ? reinterpret_cast<const char*>(op.a + B.off) : nullptr;
    const char* gb[3];
    for (int d = 0; d < 3; ++d) gb[d] = KIND == 0 ? reinterpret_cast<const char*>(op.b[d] + B.foff[d]) : nullptr;
    const int64_t cbytes = (L.nalloc - B.off) * (int64_t)ST, cobytes = (L.nalloc - B.off) * (int64_t)SC;
    int64_t fbytes[3];
    for (int d = 0; d < 3; ++d) fbytes[d] = (L.nface[d] - B.foff[d]) * (int64_t)ST;
    const T alpha = op.alpha;
    T dh[3];
    for (int d = 0; d < 3; ++d) dh[d] = KIND == 0 ? op.dha[d] : op.dpo[d];

    const int ntx = (n0 + kMX - 1) / kMX, nty = (n1 + kMY - 1) / kMY;
    // XCD-aware order, as in the sweep
    const int nblk = gridDim.x;
    const int bid = (nblk % 8 == 0) ? (int)(blockIdx.x % 8) * (nblk / 8) + (int)(blockIdx.x / 8) : (int)blockIdx.x;
    const int tx = bid % ntx, ty = (bid / ntx) % nty, tz = bid / (ntx * nty);
    const int x0 = tx * kMX, y0 = ty * kMY, z0 = tz * cz, z1 = min(z0 + cz, n2);
    const int ex = min(kMX, n0 - x0), ey = min(kMY, n1 - y0);
    const int tid = threadIdx.x;
    const int mwave = tid >> 6, mlane = tid & 63;
    const int px = (mlane & 31) + 32 * (mwave % WPR), ly = 2 * (mwave / WPR) + (mlane >> 5);
    const int xe = x0 + 2 * px, yc = y0 + ly;
    const bool vrow = ly < ey;
    const bool valid0 = vrow && 2 * px < ex, valid1 = vrow && 2 * px + 1 < ex;
    const uint32_t d1 = valid1 ? 1u : 0u;  // column 1 falls back to column 0 where it is no real cell

    auto wrapc = [](int x, int n, int per, int& xw) -> bool {
        xw = x;
        if (x >= 0 && x < n) return true;
        if (!per) return false;
        xw = x < 0 ? x + n : x - n;
        return true;
    };
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
    // in-plane element offsets of a real cell (with VEC, of a real pair)
    const int ox = min(xe, VEC ? n0 - 2 : n0 - 1), oy = min(yc, n1 - 1);
    const uint32_t oc = (uint32_t)((ox + 1) + sy * (oy + 1));
    const uint32_t ofx = (uint32_t)(ox + (n0 + 1) * oy), ofyz = (uint32_t)(ox + n0 * oy);
    auto face_bits = [&](int x, int y) -> int {
        int b = 0;
        if (x == 0 && B.bct[0] != BC_INTERIOR) b |= 1;
        if (y == 0 && B.bct[1] != BC_INTERIOR) b |= 2;
        if (x == n0 - 1 && B.bct[3] != BC_INTERIOR) b |= 8;
        if (y == n1 - 1 && B.bct[4] != BC_INTERIOR) b |= 16;
        return b;
    };
    const int bits0 = valid0 ? face_bits(xe, yc) : 0, bits1 = valid1 ? face_bits(xe + 1, yc) : 0;
    const int pp0 = mpos(2 * px + 1, ly + 1), pp1 = mpos(2 * px + 2, ly + 1);  // LDS positions of the columns

    // halo ring (one cell around ex x ey): LDS position and in-plane offset
    const int nring = 2 * (ex + 2) + 2 * ey;
    int hls = DUMMY;
    uint32_t hoa = 0;
    bool hreal = false;
    if (tid < nring) {
        int h = tid, lx, lyy;
        if (h < ex + 2) {
            lx = h;
            lyy = 0;
        } else if ((h -= ex + 2) < ex + 2) {
            lx = h;
            lyy = ey + 1;
        } else if ((h -= ex + 2) < ey) {
            lx = 0;
            lyy = 1 + h;
        } else {
            h -= ey;
            lx = ex + 1;
            lyy = 1 + h;
        }
        hls = mpos(lx, lyy);
        int xw, yw;
        if (wrapc(x0 - 1 + lx, n0, per0, xw) && wrapc(y0 - 1 + lyy, n1, per1, yw)) {
            hoa = (uint32_t)((xw + 1) + sy * (yw + 1));
            hreal = true;
        }
    }
    const int sp0 = valid0 ? pp0 : DUMMY, sp1 = valid1 ? pp1 : DUMMY;

    const uint32_t pl_fx = (uint32_t)(n0 + 1) * n1, pl_fy = (uint32_t)n0 * (n1 + 1), pl_fz = (uint32_t)n0 * n1;
    auto cell_rsrc = [&](const char* b, int zl, uint32_t es, int64_t bytes) {
        const int64_t o = (int64_t)sz * es * (zl + 1);
        return make_rsrc(b + o, bytes - o);
    };
    auto face_rsrc = [&](int d, uint32_t pl, int zl) {
        const int64_t o = (int64_t)pl * ST * zl;
        return make_rsrc(gb[d] + o, fbytes[d] - o);
    };

    // raw loads of step j: x of plane j+1 (own columns, halo), both columns'
    // coefficients of plane j
    struct Step {
        T p0, p1, hp;
        TC c0, c1, hc;
        T rhs[2], a[2], bx[3], by[2], byu[2], bz[2], bzu[2];
    };
    auto load_vals = [&](int z, Step& d) {
        int zv;
        zaddr(z, zv);
        const Rsrc P = cell_rsrc(gx, zv, ST, cbytes);
        if constexpr (VEC) {
            const Two<T> v = bld2<T>(P, oc * ST);
            d.p0 = v.x;
            d.p1 = v.y;
        } else {
            d.p0 = bld<T>(P, oc * ST);
            d.p1 = bld<T>(P, (oc + d1) * ST);
        }
        d.hp = bld<T>(P, hoa * ST);
        if constexpr (COR) {
            const Rsrc C = cell_rsrc(gcor, zv, SC, cobytes);
            if constexpr (VEC) {
                const Two<TC> v = bld2<TC>(C, oc * SC);
                d.c0 = v.x;
                d.c1 = v.y;
            } else {
                d.c0 = bld<TC>(C, oc * SC);
                d.c1 = bld<TC>(C, (oc + d1) * SC);
            }
            d.hc = bld<TC>(C, hoa * SC);
        }
    };
    auto load_step = [&](int j, Step& d) {
        load_vals(j + 1, d);
        int zc;
        zaddr(j, zc);
        const Rsrc R = cell_rsrc(grhs, zc, ST, cbytes);
        if constexpr (VEC) {
            auto put = [](T* dst, Two<T> v) {
                dst[0] = v.x;
                dst[1] = v.y;
            };
            put(d.rhs, bld2<T>(R, oc * ST));
            if constexpr (KIND == 0) {
                const Rsrc A = cell_rsrc(ga, zc, ST, cbytes);
                const Rsrc FX = face_rsrc(0, pl_fx, zc);
                const Rsrc FY = face_rsrc(1, pl_fy, zc);
                const Rsrc FZ = face_rsrc(2, pl_fz, zc);
                put(d.a, bld2<T>(A, oc * ST));
                put(d.bx, bld2<T>(FX, ofx * ST));
                d.bx[2] = bld<T>(FX, (ofx + 2) * ST);
                put(d.by, bld2<T>(FY, ofyz * ST));
                put(d.byu, bld2<T>(FY, (ofyz + n0) * ST));
                put(d.bz, bld2<T>(FZ, ofyz * ST));
                put(d.bzu, bld2<T>(FZ, (ofyz + pl_fz) * ST));
            }
            return;
        }
        d.rhs[0] = bld<T>(R, oc * ST);
        d.rhs[1] = bld<T>(R, (oc + d1) * ST);
        if constexpr (KIND == 0) {
            const Rsrc A = cell_rsrc(ga, zc, ST, cbytes);
            const Rsrc FX = face_rsrc(0, pl_fx, zc);
            const Rsrc FY = face_rsrc(1, pl_fy, zc);
            const Rsrc FZ = face_rsrc(2, pl_fz, zc);
            d.a[0] = bld<T>(A, oc * ST);
            d.a[1] = bld<T>(A, (oc + d1) * ST);
            d.bx[0] = bld<T>(FX, ofx * ST);
            d.bx[1] = bld<T>(FX, (ofx + 1) * ST);
            d.bx[2] = bld<T>(FX, (ofx + 1 + d1) * ST);
            d.by[0] = bld<T>(FY, ofyz * ST);
            d.by[1] = bld<T>(FY, (ofyz + d1) * ST);
            d.byu[0] = bld<T>(FY, (ofyz + n0) * ST);
            d.byu[1] = bld<T>(FY, (ofyz + n0 + d1) * ST);
            d.bz[0] = bld<T>(FZ, ofyz * ST);
            d.bz[1] = bld<T>(FZ, (ofyz + d1) * ST);
            d.bzu[0] = bld<T>(FZ, (ofyz + pl_fz) * ST);
            d.bzu[1] = bld<T>(FZ, (ofyz + pl_fz + d1) * ST);
        }
    };
    auto xval = [&](T p, TC c) -> T {
        if constexpr (COR) return p + static_cast<T>(c);
        else return p;
    };
    auto store_plane = [&](int z, const Step& d) {
        int zl;
        const bool real = zaddr(z, zl);
        T* P = S[z & 3];
        P[sp0] = real ? xval(d.p0, d.c0) : T(0);
        P[sp1] = real ? xval(d.p1, d.c1) : T(0);
        P[hls] = (real && hreal) ? xval(d.hp, d.hc) : T(0);
    };
    // boundary value of the ghost cell (gi, gj, gk) (inhomogeneous Dirichlet data)
    auto bval = [&](int gi, int gj, int gk) -> T {
        return static_cast<T>(bcv[B.off + (gi + 1) + (int64_t)sy * (gj + 1) + sz * (gk + 1)]);
    };

    // prologue: planes z0-1 and z0 in LDS, then step z0's loads
    Step sa, sb;
    load_vals(z0 - 1, sa);
    load_vals(z0, sb);
    store_plane(z0 - 1, sa);
    store_plane(z0, sb);
    load_step(z0, sa);

    double vmax = 0.0, vsum = 0.0;
    T acc = T(0);
    auto step = [&](int k, const Step& cur, Step& nxt) {
        load_step(k + 1, nxt);
        store_plane(k + 1, cur);
        __syncthreads();
        if (k >= z1) return;
        const T* P = S[k & 3];
        const T* Pm = S[(k - 1) & 3];
        const T* Pp = S[(k + 1) & 3];
        int zb = 0;
        if (k == 0 && zlo_on) zb |= 4;
        if (k == n2 - 1 && zhi_on) zb |= 32;
        T r[2];
        for (int c = 0; c < 2; ++c) {
            // column 0 is odd in the tile (x neighbours at -H and 1-H), column 1 even (H-1 and H)
            const int p = c == 0 ? pp0 : pp1;
            const int xl = c == 0 ? -kMH : kMH - 1, xr = c == 0 ? 1 - kMH : kMH;
            const int bits = (c == 0 ? bits0 : bits1) | zb;
            const T xc = P[p];
            T xm = P[p + xl], xp = P[p + xr], ym = P[p - kMR], yp = P[p + kMR], zm = Pm[p], zp = Pp[p];
            if (__builtin_amdgcn_ballot_w64(bits != 0) != 0) {
                const bool b0 = bits & 1, b1 = bits & 2, b2 = bits & 4, b3 = bits & 8, b4 = bits & 16, b5 = bits & 32;
                const int x = xe + c;
                auto gh = [&](int f, T v1, T v2, int gi, int gj, int gk) -> T {
                    T g = fma(v2, CL.k2[f], v1 * CL.k1[f]);
                    if (MODE == 0 && inhomog) g = fma(bval(gi, gj, gk), CL.k0[f], g);
                    return g;
                };
                if (b0) xm = gh(0, xc, xp, x - 1, yc, k);
                if (b3) xp = gh(3, xc, xm, x + 1, yc, k);
                if (b1) ym = gh(1, xc, yp, x, yc - 1, k);
                if (b4) yp = gh(4, xc, ym, x, yc + 1, k);
                if (b2) zm = gh(2, xc, zp, x, yc, k - 1);
                if (b5) zp = gh(5, xc, zm, x, yc, k + 1);
            }
            T y;
            if constexpr (KIND == 0)
                y = adotx<T, 0>(alpha, dh, cur.a[c], cur.bx[c], cur.bx[c + 1], cur.by[c], cur.byu[c], cur.bz[c],
                                cur.bzu[c], xc, xm, xp, ym, yp, zm, zp);
            else
                y = adotx<T, 1>(alpha, dh, T(0), T(0), T(0), T(0), T(0), T(0), T(0), xc, xm, xp, ym, yp, zm, zp);
            r[c] = cur.rhs[c] + T(-1.0) * y;
        }
        if constexpr (MODE == 0) {
            const int64_t cell = B.off + (int64_t)oc + sz * (k + 1);
            if (valid0) {
                if (sol_out) sol_out[cell] = P[pp0];
                res[cell] = static_cast<TR>(r[0]);
                vmax = nanmax_m(vmax, fabs((double)r[0]));
                vsum += (double)r[0];
            }
            if (valid1) {
                if (sol_out) sol_out[cell + 1] = P[pp1];
                res[cell + 1] = static_cast<TR>(r[1]);
                vmax = nanmax_m(vmax, fabs((double)r[1]));
                vsum += (double)r[1];
            }
        } else {
            // children (x pair, this row, this plane) summed here, the row
            // pair by a shuffle, the plane pair over two steps
            const T s = (valid0 ? r[0] : T(0)) + (valid1 ? r[1] : T(0));
            if (((k - z0) & 1) == 0) {
                acc = s;
            } else {
                acc += s;
                const T t = acc + __shfl_down(acc, 32, 64);
                if (valid0 && (ly & 1) == 0) {
                    const BoxDev& C = Lc.box[0];
                    crse[C.off + ((xe >> 1) + 1) + C.sy * ((yc >> 1) + 1) + C.sz * ((k >> 1) + 1)] = (T(1.0) / T(8)) * t;
                }
            }
        }
    };
    for (int k = z0; k < z1; k += 2) {
        step(k, sa, sb);
        step(k + 1, sb, sa);
    }
    if constexpr (MODE == 0) {
        for (int o = 32; o > 0; o >>= 1) {
            vmax = nanmax_m(vmax, __shfl_xor(vmax, o, 64));
            vsum += __shfl_xor(vsum, o, 64);
        }
        const int lane = tid & 63, wid = tid >> 6;
        if (lane == 0) {
            red[2 * wid] = vmax;
            red[2 * wid + 1] = vsum;
        }
        __syncthreads();
        if (tid == 0) {
            for (int w = 1; w < kMT / 64; ++w) {
                vmax = nanmax_m(vmax, red[2 * w]);
                vsum += red[2 * w + 1];
            }
            partials[2 * (int64_t)blockIdx.x] = vmax;
            partials[2 * (int64_t)blockIdx.x + 1] = vsum;
        }
    }
}

// z chunk: at most 512 planes, halved while the grid has fewer than per_cu
// work-groups per CU (the kernel's occupancy: one round of work-groups, few
// re-read planes), at least 8 planes; even (a restriction pairs fine planes)
int march_chunk(const LevelDev& L, int* nblocks, int per_cu) {
    const int mx = march_width(L);
    const int nt = ((L.n[0] + mx - 1) / mx) * ((L.n[1] + kMY - 1) / kMY);
    const int64_t target = (int64_t)device_compute_units() * per_cu;
    int cz = std::min(512, L.n[2] + (L.n[2] & 1));
    while (cz > 8 && (int64_t)nt * ((L.n[2] + cz - 1) / cz) < target) cz /= 2;
    cz += cz & 1;
    *nblocks = nt * ((L.n[2] + cz - 1) / cz);
    return cz;
}
// work-groups per CU (512 threads each): the ABecLaplacian update (FP64
// arrays, about 190 VGPRs) and FP64 restriction fit one, the FP32
// restriction and the Poisson kernels two
inline int update_per_cu(int kind) { return kind == 0 ? 1 : 2; }
template <class T>
inline int restrict_per_cu(int kind) { return (sizeof(T) == 4 || kind != 0) ? 2 : 1; }

bool plane_fits(const LevelDev& L, size_t es) {
    const int64_t lim = std::numeric_limits<uint32_t>::max();
    return (int64_t)(L.n[0] + 2) * (L.n[1] + 2) * (int64_t)es < lim &&
           (int64_t)(L.n[0] + 1) * (L.n[1] + 1) * (int64_t)es < lim;
}

}  // namespace

bool march_ok(const LevelDev& L) {
    // one box covering the domain (its cells are the domain's)
    return L.nbox == 1 && L.ncell == (int64_t)L.n[0] * L.n[1] * L.n[2] && L.max_nxo <= 3 && L.n[0] >= 32 && L.n[1] >= 8 && L.n[2] >= 2 && plane_fits(L, 8);
}

int64_t march_partials(const LevelDev& L, int kind) {
    int nb = 0;
    march_chunk(L, &nb, update_per_cu(kind));
    return 2 * (int64_t)nb;
}

template <class TC, class TR, bool VEC, int MX>
void launch_update_residual(const LevelDev& L, const OpCoef<double>& op, const double* sol, const TC* cor,
                            double* sol_out, const double* rhs, const double* bcv, int inhomog, TR* res,
                            double* partials, hipStream_t s) {
    int nb = 0;
    const int cz = march_chunk(L, &nb, update_per_cu(op.kind));
    const dim3 grid((unsigned)nb), block(MTile<MX>::T);
    if (op.kind == 0) {
        if (cor)
            hipLaunchKernelGGL((k_march<0, 0, double, TC, TR, true, VEC, MX>), grid, block, 0, s, L, op, sol, cor, sol_out,
                               rhs, bcv, inhomog, res, partials, L, nullptr, cz);
        else
            hipLaunchKernelGGL((k_march<0, 0, double, TC, TR, false, VEC, MX>), grid, block, 0, s, L, op, sol, cor,
                               sol_out, rhs, bcv, inhomog, res, partials, L, nullptr, cz);
    } else {
        if (cor)
            hipLaunchKernelGGL((k_march<0, 1, double, TC, TR, true, VEC, MX>), grid, block, 0, s, L, op, sol, cor, sol_out,
                               rhs, bcv, inhomog, res, partials, L, nullptr, cz);
        else
            hipLaunchKernelGGL((k_march<0, 1, double, TC, TR, false, VEC, MX>), grid, block, 0, s, L, op, sol, cor,
                               sol_out, rhs, bcv, inhomog, res, partials, L, nullptr, cz);
    }
}

template <class TC, class TR>
void march_update_residual(const LevelDev& L, const OpCoef<double>& op, const double* sol, const TC* cor,
                           double* sol_out, const double* rhs, const double* bcv, int inhomog, TR* res,
                           double* partials, hipStream_t s) {
    if (L.n[0] % 2 == 0 && march_width(L) == 128)
        launch_update_residual<TC, TR, true, 128>(L, op, sol, cor, sol_out, rhs, bcv, inhomog, res, partials, s);
    else if (L.n[0] % 2 == 0)
        launch_update_residual<TC, TR, true, 64>(L, op, sol, cor, sol_out, rhs, bcv, inhomog, res, partials, s);
    else
        launch_update_residual<TC, TR, false, 64>(L, op, sol, cor, sol_out, rhs, bcv, inhomog, res, partials, s);
}

template <class T, int MX>
void launch_resid_restrict(const LevelDev& Lf, const OpCoef<T>& op, const T* cor, const T* res, const LevelDev& Lc,
                           T* crse, hipStream_t s) {
    int nb = 0;
    const int cz = march_chunk(Lf, &nb, restrict_per_cu<T>(op.kind));
    const dim3 grid((unsigned)nb), block(MTile<MX>::T);
    // the fine extents are even here, so every pair is two real cells
    if (op.kind == 0)
        hipLaunchKernelGGL((k_march<1, 0, T, T, T, false, true, MX>), grid, block, 0, s, Lf, op, cor,
                           static_cast<const T*>(nullptr), static_cast<T*>(nullptr), res, nullptr, 0,
                           static_cast<T*>(nullptr), nullptr, Lc, crse, cz);
    else
        hipLaunchKernelGGL((k_march<1, 1, T, T, T, false, true, MX>), grid, block, 0, s, Lf, op, cor,
                           static_cast<const T*>(nullptr), static_cast<T*>(nullptr), res, nullptr, 0,
                           static_cast<T*>(nullptr), nullptr, Lc, crse, cz);
}

template <class T>
void march_resid_restrict(const LevelDev& Lf, const OpCoef<T>& op, const T* cor, const T* res, const LevelDev& Lc,
                          T* crse, hipStream_t s) {
    if ((Lf.n[0] | Lf.n[1] | Lf.n[2]) & 1) throw std::runtime_error("march_resid_restrict: odd fine extent");
    if (march_width(Lf) == 128)
        launch_resid_restrict<T, 128>(Lf, op, cor, res, Lc, crse, s);
    else
        launch_resid_restrict<T, 64>(Lf, op, cor, res, Lc, crse, s);
}

template void march_update_residual<double, double>(const LevelDev&, const OpCoef<double>&, const double*,
                                                    const double*, double*, const double*, const double*, int,
                                                    double*, double*, hipStream_t);
template void march_update_residual<float, float>(const LevelDev&, const OpCoef<double>&, const double*,
                                                  const float*, double*, const double*, const double*, int, float*,
                                                  double*, hipStream_t);
template void march_update_residual<float, double>(const LevelDev&, const OpCoef<double>&, const double*,
                                                   const float*, double*, const double*, const double*, int,
                                                   double*, double*, hipStream_t);
template void march_resid_restrict<double>(const LevelDev&, const OpCoef<double>&, const double*, const double*,
                                           const LevelDev&, double*, hipStream_t);
template void march_resid_restrict<float>(const LevelDev&, const OpCoef<float>&, const float*, const float*,
                                          const LevelDev&, float*, hipStream_t);

}  // namespace mgx
