// ============================================================================
// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the AMReX 24.10 MLMG single-level cell-centred solve
// (MLMG + MLABecLaplacian / MLPoisson, GSRB smoother, BiCGStab bottom).
// Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
// may load this library, and only as the checker or the timed CPU baseline.
// The product (amrex_amd) never calls it.
//
// Reference files followed (paths relative to the AMReX source tree):
//   Src/LinearSolvers/MLMG/AMReX_MLMG.H             solve 356-541, prepareForSolve
//        937-1116, oneIter 1228-1293, mgVcycle 1308-1415, actualBottomSolve
//        1502-1610, makeSolvable 1846-1873
//   Src/LinearSolvers/MLMG/AMReX_MLLinOp.H          defineGrids 795-1030
//   Src/LinearSolvers/MLMG/AMReX_MLCellLinOp.H      applyBC 683-893, restriction
//        946-952, interpolation 955-999, smooth 1205-1217, residuals 1220-1270,
//        getSolvabilityOffset 1559-1593, computeVolInv 1944-2004
//   Src/LinearSolvers/MLMG/AMReX_MLLinOp_K.H        apply_bc 14-115,
//        comp_interp_coef0 329-406
//   Src/Boundary/AMReX_LOUtil_K.H                   poly_interp_coeff 24-37
//   Src/LinearSolvers/MLMG/AMReX_MLABecLap_3D_K.H   adotx 9-28, normalize 60-75,
//        abec_gsrb 210-264
//   Src/LinearSolvers/MLMG/AMReX_MLABecLaplacian.H  Fsmooth 858-1000 (dh=b/h^2),
//        averageDownCoeffs 612-660, update_singular_flags 716-773
//   Src/LinearSolvers/MLMG/AMReX_MLPoisson_3D_K.H   adotx 9-16, gsrb 155-196
//   Src/LinearSolvers/MLMG/AMReX_MLCGSolver.H       solve_bicgstab 98-273
//   Src/Base/AMReX_MultiFabUtil_3D_C.H              avgdown_faces 173-215,
//        avgdown 377-396
//   Src/Base/AMReX_BoxList.cpp                      maxSize 765-820
//
// Storage: every MG level is held as one ghost-grown global array over the
// level's domain.  The reference's single-level results do not depend on the
// box decomposition (colours use global indices, ghosts are refreshed before
// each colour, transfers are pointwise, interior box faces have mask 0), so
// this layout reproduces them; BoxArrays are still built (maxSize chopping,
// agglomeration) because the Dirichlet closure order depends on box length.
// Reductions (dot, sum) are summed in global (k,j,i) order, which can differ
// from the reference's per-box order in the last bit.
//
// Beyond the reference: precision mode 2 ("mixed") runs an FP32 V-cycle with
// an FP64 bottom solve inside an FP64 defect-correction loop.  It has no
// reference counterpart; it is checked against the FP64 solve.  The adjoint
// part (row_of .. grad_coeffs, Params::transpose) has no reference counterpart
// either: it writes the homogeneous operator as explicit matrix rows (checked
// against fapply), transposes them by scattering, solves A^T x = b by the same
// MLMG loop with the fine residual b - A^T x, and differentiates the operator
// with respect to its coefficients (checked against dense brute force).
// ============================================================================
#include <stdexcept>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

namespace orc {

enum { BC_INTERIOR = 0, BC_DIRICHLET = 101, BC_NEUMANN = 102, BC_REFLECT_ODD = 103 };

struct IBox {
    int lo[3], hi[3];
    int len(int d) const { return hi[d] - lo[d] + 1; }
    double npts() const { return (double)len(0) * len(1) * len(2); }
    bool operator==(const IBox& o) const {
        for (int d = 0; d < 3; ++d)
            if (lo[d] != o.lo[d] || hi[d] != o.hi[d]) return false;
        return true;
    }
};

inline int coarsen_i(int i, int r) { return (i < 0) ? -((-i - 1) / r) - 1 : i / r; }

inline IBox coarsen(const IBox& b, const int* r) {
    IBox c;
    for (int d = 0; d < 3; ++d) { c.lo[d] = coarsen_i(b.lo[d], r[d]); c.hi[d] = coarsen_i(b.hi[d], r[d]); }
    return c;
}
inline IBox refine(const IBox& b, const int* r) {
    IBox c;
    for (int d = 0; d < 3; ++d) { c.lo[d] = b.lo[d] * r[d]; c.hi[d] = (b.hi[d] + 1) * r[d] - 1; }
    return c;
}
// Box::coarsenable (AMReX_Box.H:745-755)
inline bool coarsenable(const IBox& b, const int* r, const int* minw) {
    for (int d = 0; d < 3; ++d) if (b.len(d) < r[d] * minw[d]) return false;
    return refine(coarsen(b, r), r) == b;
}

// BoxList::maxSize (AMReX_BoxList.cpp:765-820)
std::vector<IBox> max_size(const std::vector<IBox>& in, int chunk) {
    std::vector<IBox> out;
    for (const IBox& bx : in) {
        int ratio[3] = {1, 1, 1}, numblk[3] = {1, 1, 1}, extra[3] = {0, 0, 0}, sz[3];
        for (int d = 0; d < 3; ++d) {
            sz[d] = bx.len(d);
            if (bx.len(d) > chunk) {
                int bs = chunk, nlen = bx.len(d);
                while (bs % 2 == 0 && nlen % 2 == 0) { ratio[d] *= 2; bs /= 2; nlen /= 2; }
                numblk[d] = (nlen + bs - 1) / bs;
                sz[d] = nlen / numblk[d];
                extra[d] = nlen - sz[d] * numblk[d];
            }
        }
        if (numblk[0] == 1 && numblk[1] == 1 && numblk[2] == 1) { out.push_back(bx); continue; }
        auto piece = [&](int d, int b, int& lo, int& hi) {
            lo = (b < extra[d]) ? b * (sz[d] + 1) * ratio[d] : (b * sz[d] + extra[d]) * ratio[d];
            hi = (b < extra[d]) ? lo + (sz[d] + 1) * ratio[d] - 1 : lo + sz[d] * ratio[d] - 1;
            lo += bx.lo[d]; hi += bx.lo[d];
        };
        for (int k = 0; k < numblk[2]; ++k)
            for (int j = 0; j < numblk[1]; ++j)
                for (int i = 0; i < numblk[0]; ++i) {
                    IBox nb;
                    piece(0, i, nb.lo[0], nb.hi[0]);
                    piece(1, j, nb.lo[1], nb.hi[1]);
                    piece(2, k, nb.lo[2], nb.hi[2]);
                    out.push_back(nb);
                }
    }
    return out;
}

// poly_interp_coeff (AMReX_LOUtil_K.H:24-37): Lagrange weights at xInt.
template <class T>
void poly_interp_coeff(T xInt, const T* x, int N, T* c) {
    for (int j = 0; j < N; ++j) {
        T num = T(1.0), den = T(1.0);
        for (int i = 0; i < N; ++i) {
            if (i != j) { num *= xInt - x[i]; den *= x[j] - x[i]; }
        }
        c[j] = num / den;
    }
}

// ---------------------------------------------------------------------------
// Fields
// ---------------------------------------------------------------------------
template <class T>
struct Fld {  // cell field over [0,n) with one ghost layer
    int n[3] = {0, 0, 0};
    int64_t sy = 0, sz = 0;
    std::vector<T> v;
    void define(const int* nn) {
        for (int d = 0; d < 3; ++d) n[d] = nn[d];
        sy = n[0] + 2;
        sz = sy * (n[1] + 2);
        v.assign((size_t)sz * (n[2] + 2), T(0));
    }
    T& operator()(int i, int j, int k) { return v[(size_t)((i + 1) + sy * (j + 1) + sz * (k + 1))]; }
    const T& operator()(int i, int j, int k) const { return v[(size_t)((i + 1) + sy * (j + 1) + sz * (k + 1))]; }
    void set_all(T x) { std::fill(v.begin(), v.end(), x); }
};

template <class T>
struct FaceArr {  // face field of direction d, no ghosts
    int m[3] = {0, 0, 0};
    std::vector<T> v;
    void define(const int* n, int d) {
        for (int e = 0; e < 3; ++e) m[e] = n[e] + (e == d ? 1 : 0);
        v.assign((size_t)m[0] * m[1] * m[2], T(0));
    }
    T& operator()(int i, int j, int k) { return v[(size_t)(i + (int64_t)m[0] * (j + (int64_t)m[1] * k))]; }
    const T& operator()(int i, int j, int k) const { return v[(size_t)(i + (int64_t)m[0] * (j + (int64_t)m[1] * k))]; }
};

// ---------------------------------------------------------------------------
// Problem description
// ---------------------------------------------------------------------------
struct Params {
    int n[3];
    double prob_lo[3], prob_hi[3];
    int periodic[3];
    int bc_lo[3], bc_hi[3];
    int max_grid_size, agg_grid_size, do_agglomeration;
    int maxorder, op;  // op: 0 ABecLaplacian, 1 Poisson
    double alpha, beta;
    int precision;     // 0 f64, 1 f32 throughout, 2 mixed
    int nu1, nu2, max_iter, verbose;
    double tol_rel, tol_abs;
    int bottom_maxiter;
    double bottom_reltol;  // <= 0: reference default for the precision
    int fixed_iters;
    int max_coarsening_level;
    int transpose;  // solve A^T x = b (homogeneous boundary data)
    int max_fmg_iters;  // outer iterations that use the F-cycle (MLMG::setMaxFmgIter)
    // mixed precision: FP32 V-cycles per FP64 refinement step (1, 2; 0: 1 or 2
    // chosen per step from the measured contraction)
    int refine_vcycles;
    // Dirichlet boundary location per domain face (dir + 3*side), in level-0
    // cells outside the face: 0 on physical faces (MLMGBndry.H:124-129); a
    // coarse/fine face at 0.5*ratio*dx of level 0 (MLMGBndry.H:138-145) is
    // ratio/2.  Every MG level keeps the level-0 distance (MLCellLinOp.H:586-592).
    double bloc[6];
    // bottom solver, the codes of amrex::BottomSolver (AMReX_MLLinOp.H:30-32):
    // 0 Default (= bicgstab), 1 smoother, 2 bicgstab, 3 cg, 4 bicgcg, 5 cgbicg
    int bottom_solver;
};

// ---------------------------------------------------------------------------
// One MG level
// ---------------------------------------------------------------------------
template <class T>
struct Level {
    int n[3];
    double dx[3], dxinv[3];
    std::vector<IBox> ba;
    Fld<T> a;          // cell coefficient (ghosts unused)
    FaceArr<T> b[3];   // face coefficients
    int bct[6];        // per domain face (orientation dir + 3*side)
    std::vector<signed char> nxo[6];  // Dirichlet closure order per face cell
    T coef[6][5][4];   // Lagrange weights per face for each order NX
    bool has_a = false;
    bool agg = false;  // BoxArray made by agglomeration (MFIter-unsafe against the finer level)

    int tlen(int f, int which) const {  // tangential extents of face f
        const int d = f % 3;
        const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        return which == 0 ? n[t1] : n[t2];
    }
    // Undrrelxr value at the first interior cell (AMReX_MLLinOp_K.H:329-406);
    // only used where the face is a physical boundary (mask > 0).
    T cf(int f, int u, int w) const {
        switch (bct[f]) {
            case BC_DIRICHLET: return coef[f][(int)nxo[f][(size_t)u + (size_t)tlen(f, 0) * w]][1];
            case BC_NEUMANN: return T(1.0);
            case BC_REFLECT_ODD: return T(1.0);
            default: return T(0.0);
        }
    }
};

// Build the ghost-closure tables of a level (bcl = 0 on physical faces,
// MLMGBndry.H:124-139; bcl = bloc * dx0 on coarse/fine faces).  The first
// Lagrange node is -bcl * dxinv of the level (AMReX_MLLinOp_K.H:96-100).
// NX = min(blen+1, maxorder), AMReX_MLLinOp_K.H:96.
template <class T>
void setup_bc(Level<T>& L, const Params& P) {
    for (int f = 0; f < 6; ++f) {
        const int d = f % 3;
        const double dx0 = (P.prob_hi[d] - P.prob_lo[d]) / P.n[d];
        const T bcl = static_cast<T>(P.bloc[f] * dx0);
        for (int m = 0; m < 5; ++m) for (int q = 0; q < 4; ++q) L.coef[f][m][q] = T(0);
        for (int NX = 1; NX <= 4; ++NX) {
            T x[4] = {-bcl * T(L.dxinv[d]), T(0.5), T(1.5), T(2.5)};
            poly_interp_coeff(-T(0.5), x, NX, L.coef[f][NX]);
        }
    }
    for (int f = 0; f < 6; ++f) {
        const int d = f % 3, side = f / 3;
        L.bct[f] = P.periodic[d] ? BC_INTERIOR : (side == 0 ? P.bc_lo[d] : P.bc_hi[d]);
        const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        L.nxo[f].assign((size_t)L.n[t1] * L.n[t2], (signed char)std::min(3, P.maxorder));
        if (L.bct[f] != BC_DIRICHLET) continue;
        const int pos = side == 0 ? 0 : L.n[d] - 1;
        for (const IBox& bx : L.ba) {
            if ((side == 0 ? bx.lo[d] : bx.hi[d]) != pos) continue;
            const int NX = std::min(bx.len(d) + 1, P.maxorder);
            for (int w = bx.lo[t2]; w <= bx.hi[t2]; ++w)
                for (int u = bx.lo[t1]; u <= bx.hi[t1]; ++u)
                    L.nxo[f][(size_t)u + (size_t)L.n[t1] * w] = (signed char)NX;
        }
    }
}

// Fill the face ghosts of phi: periodic wrap (FillBoundary, cross stencil) and
// the physical closure (mllinop_apply_bc_*).  bcv, when given, is a ghost-grown
// array whose ghost cells hold inhomogeneous Dirichlet values.
template <class T>
void fill_ghosts(const Level<T>& L, Fld<T>& phi, const Fld<T>* bcv) {
    const int* n = L.n;
    for (int f = 0; f < 6; ++f) {
        const int d = f % 3, side = f / 3;
        const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        const int g = side == 0 ? -1 : n[d];        // ghost index along d
        const int s = side == 0 ? 1 : -1;           // step into the interior
        for (int w = 0; w < n[t2]; ++w) {
            for (int u = 0; u < n[t1]; ++u) {
                int idx[3];
                idx[d] = g; idx[t1] = u; idx[t2] = w;
                auto at = [&](int off) -> T& {
                    int q[3] = {idx[0], idx[1], idx[2]};
                    q[d] += off;
                    return phi(q[0], q[1], q[2]);
                };
                T& gh = at(0);
                switch (L.bct[f]) {
                    case BC_INTERIOR: {
                        int q[3] = {idx[0], idx[1], idx[2]};
                        q[d] = (side == 0) ? n[d] - 1 : 0;
                        gh = phi(q[0], q[1], q[2]);
                        break;
                    }
                    case BC_NEUMANN: gh = at(s); break;
                    case BC_REFLECT_ODD: gh = -at(s); break;
                    case BC_DIRICHLET: {
                        const int NX = L.nxo[f][(size_t)u + (size_t)n[t1] * w];
                        T tmp = T(0.0);
                        for (int m = 1; m < NX; ++m) tmp += at(m * s) * L.coef[f][NX][m];
                        gh = tmp;
                        if (bcv) gh += (*bcv)(idx[0], idx[1], idx[2]) * L.coef[f][NX][0];
                        break;
                    }
                    default: break;
                }
            }
        }
    }
}

template <class T>
struct Scal {  // per-level operator scalars, cast the way the reference casts them
    T alpha, beta;
    T dhs[3];   // smoother: b_scalar / (h*h)      (MLABecLaplacian.H Fsmooth)
    T dxi[3];   // apply:    RT(dxinv)              (mlabeclap_adotx)
    T dpo[3];   // Poisson:  RT(dxinv*dxinv)
};
template <class T>
Scal<T> scalars(const Level<T>& L, const Params& P) {
    Scal<T> s;
    s.alpha = T(P.alpha);
    s.beta = T(P.beta);
    for (int d = 0; d < 3; ++d) {
        s.dhs[d] = s.beta / static_cast<T>(L.dx[d] * L.dx[d]);
        s.dxi[d] = static_cast<T>(L.dxinv[d]);
        s.dpo[d] = static_cast<T>(L.dxinv[d] * L.dxinv[d]);
    }
    return s;
}

// One colour of GSRB (abec_gsrb, AMReX_MLABecLap_3D_K.H:210-264; mlpoisson_gsrb,
// AMReX_MLPoisson_3D_K.H:155-196).  Ghosts must be current.
template <class T>
void gsrb(const Level<T>& L, const Params& P, Fld<T>& phi, const Fld<T>& rhs, int rb) {
    const Scal<T> S = scalars(L, P);
    const int* n = L.n;
    constexpr T omega = T(1.15);
    const bool per[3] = {L.bct[0] == BC_INTERIOR, L.bct[1] == BC_INTERIOR, L.bct[2] == BC_INTERIOR};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k) {
        for (int j = 0; j < n[1]; ++j) {
            for (int i = 0; i < n[0]; ++i) {
                if ((i + j + k + rb) % 2 != 0) continue;
                const T cf0 = (i == 0 && !per[0]) ? L.cf(0, j, k) : T(0.0);
                const T cf1 = (j == 0 && !per[1]) ? L.cf(1, i, k) : T(0.0);
                const T cf2 = (k == 0 && !per[2]) ? L.cf(2, i, j) : T(0.0);
                const T cf3 = (i == n[0] - 1 && !per[0]) ? L.cf(3, j, k) : T(0.0);
                const T cf4 = (j == n[1] - 1 && !per[1]) ? L.cf(4, i, k) : T(0.0);
                const T cf5 = (k == n[2] - 1 && !per[2]) ? L.cf(5, i, j) : T(0.0);
                if (P.op == 0) {
                    const T dhx = S.dhs[0], dhy = S.dhs[1], dhz = S.dhs[2];
                    const T bxm = L.b[0](i, j, k), bxp = L.b[0](i + 1, j, k);
                    const T bym = L.b[1](i, j, k), byp = L.b[1](i, j + 1, k);
                    const T bzm = L.b[2](i, j, k), bzp = L.b[2](i, j, k + 1);
                    const T gamma = S.alpha * L.a(i, j, k)
                        + dhx * (bxm + bxp)
                        + dhy * (bym + byp)
                        + dhz * (bzm + bzp);
                    const T g_m_d = gamma
                        - (dhx * (bxm * cf0 + bxp * cf3)
                         + dhy * (bym * cf1 + byp * cf4)
                         + dhz * (bzm * cf2 + bzp * cf5));
                    const T rho = dhx * (bxm * phi(i - 1, j, k) + bxp * phi(i + 1, j, k))
                                + dhy * (bym * phi(i, j - 1, k) + byp * phi(i, j + 1, k))
                                + dhz * (bzm * phi(i, j, k - 1) + bzp * phi(i, j, k + 1));
                    const T res = rhs(i, j, k) - (gamma * phi(i, j, k) - rho);
                    phi(i, j, k) = phi(i, j, k) + omega / g_m_d * res;
                } else {
                    const T dhx = S.dpo[0], dhy = S.dpo[1], dhz = S.dpo[2];
                    const T gamma = T(-2.) * (dhx + dhy + dhz);
                    const T g_m_d = gamma + dhx * (cf0 + cf3) + dhy * (cf1 + cf4) + dhz * (cf2 + cf5);
                    const T res = rhs(i, j, k) - gamma * phi(i, j, k)
                        - dhx * (phi(i - 1, j, k) + phi(i + 1, j, k))
                        - dhy * (phi(i, j - 1, k) + phi(i, j + 1, k))
                        - dhz * (phi(i, j, k - 1) + phi(i, j, k + 1));
                    phi(i, j, k) = phi(i, j, k) + omega / g_m_d * res;
                }
            }
        }
    }
}

// The adjoint's V-cycles smooth, take correction residuals with and bottom-
// solve the transposed operator where its red-black sweep is well defined
// (maxorder <= 3, gsrb_T below); otherwise they use the forward operator and
// only the outer residual is b - A^T x.
inline bool transposed_cycle(const Params& P) { return P.transpose && P.maxorder <= 3; }
template <class T>
void gsrb_T(const Level<T>& L, const Params& P, Fld<T>& phi, const Fld<T>& rhs, int rb);

// linop.smooth (AMReX_MLCellLinOp.H:1205-1217): applyBC then colour, twice.
// The adjoint cycle's transposed sweep needs no ghosts.
template <class T>
void smooth(const Level<T>& L, const Params& P, Fld<T>& phi, const Fld<T>& rhs) {
    if (transposed_cycle(P)) {
        for (int rb = 0; rb < 2; ++rb) gsrb_T(L, P, phi, rhs, rb);
        return;
    }
    for (int rb = 0; rb < 2; ++rb) {
        fill_ghosts<T>(L, phi, nullptr);
        gsrb(L, P, phi, rhs, rb);
    }
}

// Fapply on filled ghosts (mlabeclap_adotx / mlpoisson_adotx).
template <class T>
void fapply(const Level<T>& L, const Params& P, Fld<T>& out, const Fld<T>& x) {
    const Scal<T> S = scalars(L, P);
    const int* n = L.n;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k) {
        for (int j = 0; j < n[1]; ++j) {
            for (int i = 0; i < n[0]; ++i) {
                if (P.op == 0) {
                    const T dhx = S.beta * S.dxi[0] * S.dxi[0];
                    const T dhy = S.beta * S.dxi[1] * S.dxi[1];
                    const T dhz = S.beta * S.dxi[2] * S.dxi[2];
                    const T xc = x(i, j, k);
                    out(i, j, k) = S.alpha * L.a(i, j, k) * xc
                        - dhx * (L.b[0](i + 1, j, k) * (x(i + 1, j, k) - xc)
                               - L.b[0](i, j, k) * (xc - x(i - 1, j, k)))
                        - dhy * (L.b[1](i, j + 1, k) * (x(i, j + 1, k) - xc)
                               - L.b[1](i, j, k) * (xc - x(i, j - 1, k)))
                        - dhz * (L.b[2](i, j, k + 1) * (x(i, j, k + 1) - xc)
                               - L.b[2](i, j, k) * (xc - x(i, j, k - 1)));
                } else {
                    const T dhx = S.dpo[0], dhy = S.dpo[1], dhz = S.dpo[2];
                    const T xc = x(i, j, k);
                    out(i, j, k) = dhx * (x(i - 1, j, k) - T(2.0) * xc + x(i + 1, j, k))
                                 + dhy * (x(i, j - 1, k) - T(2.0) * xc + x(i, j + 1, k))
                                 + dhz * (x(i, j, k - 1) - T(2.0) * xc + x(i, j, k + 1));
                }
            }
        }
    }
}

// mlabeclap_normalize (AMReX_MLABecLap_3D_K.H:60-75); MLPoisson: no-op in 3D.
template <class T>
void normalize(const Level<T>& L, const Params& P, Fld<T>& x) {
    if (P.op != 0) return;
    const Scal<T> S = scalars(L, P);
    const T dhx = S.beta * S.dxi[0] * S.dxi[0];
    const T dhy = S.beta * S.dxi[1] * S.dxi[1];
    const T dhz = S.beta * S.dxi[2] * S.dxi[2];
    const int* n = L.n;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i)
                x(i, j, k) /= S.alpha * L.a(i, j, k)
                    + dhx * (L.b[0](i, j, k) + L.b[0](i + 1, j, k))
                    + dhy * (L.b[1](i, j, k) + L.b[1](i, j + 1, k))
                    + dhz * (L.b[2](i, j, k) + L.b[2](i, j, k + 1));
}

// r = b - L(x) (apply + Xpay with -1).
template <class T>
void residual(const Level<T>& L, const Params& P, Fld<T>& r, Fld<T>& x, const Fld<T>& b,
              const Fld<T>* bcv) {
    fill_ghosts<T>(L, x, bcv);
    fapply(L, P, r, x);
    const int* n = L.n;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) r(i, j, k) = b(i, j, k) + T(-1.0) * r(i, j, k);
}

// ---------------------------------------------------------------------------
// Adjoint support: explicit rows of the homogeneous operator and their transpose
// ---------------------------------------------------------------------------
template <class T>
struct Row {
    int cell[1 + 6 * 4][3];
    T val[1 + 6 * 4];
    int cnt = 0;
    void add(const int* q, T v) {
        for (int d = 0; d < 3; ++d) cell[cnt][d] = q[d];
        val[cnt++] = v;
    }
};

// Row (i,j,k) of the level operator with homogeneous boundary data: the
// coefficient of every cell whose value the output at (i,j,k) reads once each
// ghost is replaced by its closure (periodic image, Neumann / reflect-odd
// mirror, Dirichlet Lagrange combination as in fill_ghosts).  Both operators
// are written as diag0 + sum over the six sides of w * (x_p - x_side), with
// w = dh*b_face for ABecLaplacian and w = -dh for Poisson.
template <class T>
void row_of(const Level<T>& L, const Params& P, int i, int j, int k, Row<T>& R) {
    const Scal<T> S = scalars(L, P);
    const int* n = L.n;
    const int p[3] = {i, j, k};
    R.cnt = 0;
    R.add(p, T(0));
    T diag = (P.op == 0) ? S.alpha * L.a(i, j, k) : T(0);
    for (int d = 0; d < 3; ++d) {
        const T dh = (P.op == 0) ? S.beta * S.dxi[d] * S.dxi[d] : S.dpo[d];
        for (int side = 0; side < 2; ++side) {
            T w;
            if (P.op == 0) {
                int fq[3] = {i, j, k};
                fq[d] += side;
                w = dh * L.b[d](fq[0], fq[1], fq[2]);
            } else {
                w = -dh;
            }
            diag += w;
            int q[3] = {i, j, k};
            q[d] += side == 0 ? -1 : 1;
            if (q[d] >= 0 && q[d] < n[d]) {
                R.add(q, -w);
                continue;
            }
            const int f = d + 3 * side;
            switch (L.bct[f]) {
                case BC_INTERIOR:
                    q[d] = side == 0 ? n[d] - 1 : 0;
                    R.add(q, -w);
                    break;
                case BC_NEUMANN: diag -= w; break;      // ghost = x_p
                case BC_REFLECT_ODD: diag += w; break;  // ghost = -x_p
                case BC_DIRICHLET: {
                    const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
                    const int NX = L.nxo[f][(size_t)p[t1] + (size_t)n[t1] * p[t2]];
                    for (int m = 1; m < NX; ++m) {
                        int r[3] = {i, j, k};
                        r[d] = side == 0 ? m - 1 : n[d] - m;
                        R.add(r, -w * L.coef[f][NX][m]);
                    }
                    break;
                }
                default: break;
            }
        }
    }
    R.val[0] = diag;
}

// out = A x from the explicit rows (homogeneous boundary data).
template <class T>
void apply_rows(const Level<T>& L, const Params& P, Fld<T>& out, const Fld<T>& x) {
    const int* n = L.n;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k) {
        Row<T> R;
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                row_of(L, P, i, j, k, R);
                T y = T(0);
                for (int e = 0; e < R.cnt; ++e) y += R.val[e] * x(R.cell[e][0], R.cell[e][1], R.cell[e][2]);
                out(i, j, k) = y;
            }
    }
}

// out = A^T x: every row p scatters A(p,q) x_p into out_q.  A row of plane k
// writes only planes k-1..k+1, except the rows of the first and last planes
// (periodic wrap, Dirichlet closure along z); those run first, then the other
// planes in three passes of planes k = 1 + c (mod 3), planes of one pass in
// parallel.
template <class T>
void apply_transpose(const Level<T>& L, const Params& P, Fld<T>& out, const Fld<T>& x) {
    const int* n = L.n;
    out.set_all(T(0));
    auto plane = [&](int k) {
        Row<T> R;
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                row_of(L, P, i, j, k, R);
                const T xp = x(i, j, k);
                for (int e = 0; e < R.cnt; ++e) out(R.cell[e][0], R.cell[e][1], R.cell[e][2]) += R.val[e] * xp;
            }
    };
    plane(0);
    if (n[2] > 1) plane(n[2] - 1);
    for (int c = 0; c < 3; ++c) {
#pragma omp parallel for schedule(static)
        for (int k = 1 + c; k < n[2] - 1; k += 3) plane(k);
    }
}

// r = b - A^T x.
template <class T>
void residual_transpose(const Level<T>& L, const Params& P, Fld<T>& r, const Fld<T>& x, const Fld<T>& b) {
    apply_transpose(L, P, r, x);
    each_valid(r, [&](int i, int j, int k) { r(i, j, k) = b(i, j, k) + T(-1.0) * r(i, j, k); });
}

// (A^T x) at (i,j,k) with homogeneous boundary data, gathered from column
// (i,j,k) of the operator whose rows row_of writes: a face inside the domain
// is shared (w (x_p - x_q)); across a periodic wrap the off-diagonal takes w
// of the image cell's face on the far end; a physical face keeps only the
// closure's diagonal part w (1 - c_1); and a cell at distance t = 1 .. NX-2
// from a Dirichlet face picks up the boundary row's weight -w c_{t+1} on that
// row's cell.  The accumulation order is the device's (apply_dense_T).
template <class T>
T apply_T_at(const Level<T>& L, const Params& P, const Scal<T>& S, const Fld<T>& x, int i, int j, int k) {
    const int* n = L.n;
    const int p[3] = {i, j, k};
    const T xc = x(i, j, k);
    T y = (P.op == 0) ? S.alpha * L.a(i, j, k) * xc : T(0);
    for (int d = 0; d < 3; ++d) {
        const T dh = (P.op == 0) ? S.beta * S.dxi[d] * S.dxi[d] : S.dpo[d];
        auto wof = [&](const int* fq) { return (P.op == 0) ? dh * L.b[d](fq[0], fq[1], fq[2]) : -dh; };
        const bool per = L.bct[d] == BC_INTERIOR;
        const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        const size_t fidx = (size_t)p[t1] + (size_t)n[t1] * p[t2];
        for (int side = 0; side < 2; ++side) {
            int fq[3] = {i, j, k};
            fq[d] += side;
            const T w = wof(fq);
            int q[3] = {i, j, k};
            q[d] += side ? 1 : -1;
            if (q[d] >= 0 && q[d] < n[d]) {
                y += w * (xc - x(q[0], q[1], q[2]));
            } else if (per) {
                q[d] = side ? 0 : n[d] - 1;
                int fo[3] = {i, j, k};
                fo[d] = side ? 0 : n[d];
                y += w * xc - wof(fo) * x(q[0], q[1], q[2]);
            } else {
                const int f = d + 3 * side;
                const T c1 = L.bct[f] == BC_DIRICHLET ? L.coef[f][(int)L.nxo[f][fidx]][1]
                             : L.bct[f] == BC_NEUMANN ? T(1) : L.bct[f] == BC_REFLECT_ODD ? T(-1) : T(0);
                y += w * (T(1) - c1) * xc;
            }
        }
        if (per) continue;
        for (int side = 0; side < 2; ++side) {
            const int f = d + 3 * side;
            if (L.bct[f] != BC_DIRICHLET) continue;
            const int NX = L.nxo[f][fidx];
            const int t = side ? n[d] - 1 - p[d] : p[d];
            if (t < 1 || t > NX - 2) continue;
            int r[3] = {i, j, k};
            r[d] = side ? n[d] - 1 : 0;
            int fq[3] = {r[0], r[1], r[2]};
            fq[d] += side;
            y += -wof(fq) * L.coef[f][NX][t + 1] * x(r[0], r[1], r[2]);
        }
    }
    return y;
}

// One colour of the transposed sweep (no reference counterpart):
// x_p += omega / g_m_d (b_p - (A^T x)_p) on cells with (i+j+k+rb) even, g_m_d
// the forward sweep's relaxation divisor (abec_gsrb / mlpoisson_gsrb).  With
// maxorder <= 3 a row of A^T reads only cells of the other colour (the moved
// Dirichlet weight c_2 couples the second cell to the boundary cell), so the
// colour's cells update independently; where A is symmetric this is the
// forward sweep up to rounding.
template <class T>
void gsrb_T(const Level<T>& L, const Params& P, Fld<T>& phi, const Fld<T>& rhs, int rb) {
    const Scal<T> S = scalars(L, P);
    const int* n = L.n;
    constexpr T omega = T(1.15);
    const bool per[3] = {L.bct[0] == BC_INTERIOR, L.bct[1] == BC_INTERIOR, L.bct[2] == BC_INTERIOR};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k) {
        for (int j = 0; j < n[1]; ++j) {
            for (int i = 0; i < n[0]; ++i) {
                if ((i + j + k + rb) % 2 != 0) continue;
                const T cf0 = (i == 0 && !per[0]) ? L.cf(0, j, k) : T(0.0);
                const T cf1 = (j == 0 && !per[1]) ? L.cf(1, i, k) : T(0.0);
                const T cf2 = (k == 0 && !per[2]) ? L.cf(2, i, j) : T(0.0);
                const T cf3 = (i == n[0] - 1 && !per[0]) ? L.cf(3, j, k) : T(0.0);
                const T cf4 = (j == n[1] - 1 && !per[1]) ? L.cf(4, i, k) : T(0.0);
                const T cf5 = (k == n[2] - 1 && !per[2]) ? L.cf(5, i, j) : T(0.0);
                T g_m_d;
                if (P.op == 0) {
                    const T dhx = S.dhs[0], dhy = S.dhs[1], dhz = S.dhs[2];
                    const T bxm = L.b[0](i, j, k), bxp = L.b[0](i + 1, j, k);
                    const T bym = L.b[1](i, j, k), byp = L.b[1](i, j + 1, k);
                    const T bzm = L.b[2](i, j, k), bzp = L.b[2](i, j, k + 1);
                    const T gamma = S.alpha * L.a(i, j, k) + dhx * (bxm + bxp) + dhy * (bym + byp) + dhz * (bzm + bzp);
                    g_m_d = gamma - (dhx * (bxm * cf0 + bxp * cf3) + dhy * (bym * cf1 + byp * cf4) +
                                     dhz * (bzm * cf2 + bzp * cf5));
                } else {
                    const T dhx = S.dpo[0], dhy = S.dpo[1], dhz = S.dpo[2];
                    const T gamma = T(-2.) * (dhx + dhy + dhz);
                    g_m_d = gamma + dhx * (cf0 + cf3) + dhy * (cf1 + cf4) + dhz * (cf2 + cf5);
                }
                const T y = apply_T_at(L, P, S, phi, i, j, k);
                phi(i, j, k) = phi(i, j, k) + omega / g_m_d * (rhs(i, j, k) - y);
            }
        }
    }
}

// Residual of a correction (homogeneous boundary data) and the homogeneous
// operator, transposed in the adjoint's cycles.
template <class T>
void resid_cor(const Level<T>& L, const Params& P, Fld<T>& r, Fld<T>& x, const Fld<T>& b) {
    if (transposed_cycle(P)) residual_transpose(L, P, r, x, b);
    else residual<T>(L, P, r, x, b, nullptr);
}
template <class T>
void apply_hom(const Level<T>& L, const Params& P, Fld<T>& out, Fld<T>& x) {
    if (transposed_cycle(P)) {
        apply_transpose(L, P, out, x);
    } else {
        fill_ghosts<T>(L, x, nullptr);
        fapply(L, P, out, x);
    }
}

// Face values of a solution whose ghosts are filled with the level's
// solution boundary data (applyBC Inhomogeneous/Solution before either call):
//   mode 0, MLCellLinOpT::compGrad (AMReX_MLCellLinOp.H:1393-1442):
//       g = dxinv_d * (phi_i - phi_{i-1});
//   mode 1, MLCellABecLapT::getFluxes (AMReX_MLCellABecLap.H:272-291) over
//       FFlux: ABecLaplacian (AMReX_MLABecLaplacian.H:1147-1216,
//       mlabeclap_flux_x, AMReX_MLABecLap_3D_K.H:79-96)
//       f = -(b_scalar*dxinv_d) * b_face * (phi_i - phi_{i-1}), Poisson
//       (AMReX_MLPoisson.H:857-1050, mlpoisson_flux_x, AMReX_MLPoisson_3D_K.H:
//       36-50) f = dxinv_d * (phi_i - phi_{i-1}); then times 1/b_scalar
//       where that is not 1 (b_scalar = -1 for Poisson, AMReX_MLPoisson.H:69).
// The inhomogeneous-Neumann and Robin additions are outside the supported BCs.
template <class T>
void face_values(const Level<T>& L, const Params& P, const Fld<T>& phi, int mode, FaceArr<T>* out) {
    const T bscalar = P.op == 0 ? T(P.beta) : T(-1.0);
    const T betainv = T(1.0) / bscalar;
    for (int d = 0; d < 3; ++d) {
        FaceArr<T>& f = out[d];
        const T dxi = static_cast<T>(L.dxinv[d]);
        const T fac = bscalar * static_cast<T>(L.dxinv[d]);
        for (int k = 0; k < f.m[2]; ++k)
            for (int j = 0; j < f.m[1]; ++j)
                for (int i = 0; i < f.m[0]; ++i) {
                    const int im = i - (d == 0), jm = j - (d == 1), km = k - (d == 2);
                    const T diff = phi(i, j, k) - phi(im, jm, km);
                    T v;
                    if (mode == 0) {
                        v = dxi * diff;
                    } else {
                        v = P.op == 0 ? -fac * L.b[d](i, j, k) * diff : dxi * diff;
                        if (betainv != T(1.0)) v *= betainv;
                    }
                    f(i, j, k) = v;
                }
    }
}

// Gradient of J(phi) through the solve A(a, b) phi = rhs, given the adjoint
// lam = A^{-T} dJ/dphi: dJ/dc = -lam . d(A phi)/dc for each coefficient c.
// Row p of A phi holds alpha*a_p*phi_p + sum_sides w*(phi_p - phi_side) with
// w = beta*dxinv^2*b_face and phi_side the neighbour, periodic image or the
// boundary ghost (inhomogeneous Dirichlet data included), so
//   dJ/da_p    = -lam_p * alpha * phi_p
//   dJ/db_face += -lam_p * beta*dxinv^2 * (phi_p - phi_side)  for each cell p touching the face.
// phi must have its ghosts filled (fill_ghosts with the level's boundary data).
template <class T>
void grad_coeffs(const Level<T>& L, const Params& P, const Fld<T>& lam, const Fld<T>& phi, Fld<T>& ga,
                 FaceArr<T>* gb) {
    const Scal<T> S = scalars(L, P);
    const int* n = L.n;
    ga.set_all(T(0));
    for (int d = 0; d < 3; ++d) std::fill(gb[d].v.begin(), gb[d].v.end(), T(0));
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const T lp = lam(i, j, k), pp = phi(i, j, k);
                ga(i, j, k) = -lp * S.alpha * pp;
                for (int d = 0; d < 3; ++d) {
                    const T dh = S.beta * S.dxi[d] * S.dxi[d];
                    for (int side = 0; side < 2; ++side) {
                        int q[3] = {i, j, k};
                        q[d] += side == 0 ? -1 : 1;
                        int fq[3] = {i, j, k};
                        fq[d] += side;
                        gb[d](fq[0], fq[1], fq[2]) += -lp * dh * (pp - phi(q[0], q[1], q[2]));
                    }
                }
            }
}

template <class T>
T norminf(const Fld<T>& x) {
    T m = T(0);
    for (int k = 0; k < x.n[2]; ++k)
        for (int j = 0; j < x.n[1]; ++j)
            for (int i = 0; i < x.n[0]; ++i) m = std::max(m, std::abs(x(i, j, k)));
    return m;
}
template <class T>
T dot(const Fld<T>& x, const Fld<T>& y) {
    T s = T(0);
    for (int k = 0; k < x.n[2]; ++k)
        for (int j = 0; j < x.n[1]; ++j)
            for (int i = 0; i < x.n[0]; ++i) s += x(i, j, k) * y(i, j, k);
    return s;
}
template <class T>
T sum(const Fld<T>& x) {
    T s = T(0);
    for (int k = 0; k < x.n[2]; ++k)
        for (int j = 0; j < x.n[1]; ++j)
            for (int i = 0; i < x.n[0]; ++i) s += x(i, j, k);
    return s;
}
template <class T, class F>
void each_valid(Fld<T>& x, F f) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < x.n[2]; ++k)
        for (int j = 0; j < x.n[1]; ++j)
            for (int i = 0; i < x.n[0]; ++i) f(i, j, k);
}

// average_down (amrex_avgdown, AMReX_MultiFabUtil_3D_C.H:377-396), ratio 2.
template <class T>
void restrict_avg(Fld<T>& crse, const Fld<T>& fine) {
    const T volfrac = T(1.0) / T(8);
    each_valid(crse, [&](int i, int j, int k) {
        T c = 0;
        for (int kr = 0; kr < 2; ++kr)
            for (int jr = 0; jr < 2; ++jr)
                for (int ir = 0; ir < 2; ++ir) c += fine(2 * i + ir, 2 * j + jr, 2 * k + kr);
        crse(i, j, k) = volfrac * c;
    });
}
// MG prolongation (MLCellLinOp.H:955-999): fine += crse(i/2,j/2,k/2).
template <class T>
void prolong_add(Fld<T>& fine, const Fld<T>& crse) {
    each_valid(fine, [&](int i, int j, int k) { fine(i, j, k) += crse(i / 2, j / 2, k / 2); });
}

// ---------------------------------------------------------------------------
// Hierarchy (MLLinOpT::defineGrids, AMReX_MLLinOp.H:857-1030, one AMR level)
// ---------------------------------------------------------------------------
struct LevelGeom {
    int n[3];
    std::vector<IBox> ba;
    bool agg = false;  // a new BoxArray from agglomeration (not a coarsening of level 0's)
};

std::vector<LevelGeom> define_grids(const Params& P) {
    std::vector<LevelGeom> out;
    IBox dom{{0, 0, 0}, {P.n[0] - 1, P.n[1] - 1, P.n[2] - 1}};
    std::vector<IBox> ba0 = max_size({dom}, P.max_grid_size);
    LevelGeom L0;
    for (int d = 0; d < 3; ++d) L0.n[d] = P.n[d];
    L0.ba = ba0;
    out.push_back(L0);
    const int mgr[3] = {2, 2, 2}, minw[3] = {2, 2, 2};
    const int agg = P.agg_grid_size > 0 ? P.agg_grid_size : 8;  // CPU default (MLLinOp.H:66-72)
    const bool aggable = ba0.size() > 1 && P.do_agglomeration;  // domain is covered
    if (aggable) {
        IBox dbx = dom, bbx = dom;
        const double nbxs = (double)ba0.size();
        const double thr = (double)agg * agg * agg;
        std::vector<IBox> domainboxes{dbx};
        std::vector<int> agg_flag{0};
        std::vector<int> acc{1};
        for (int lev = 0; lev < P.max_coarsening_level; ++lev) {
            bool ok = true;
            for (int d = 0; d < 3; ++d) {
                int rd[3] = {1, 1, 1};
                rd[d] = 2;
                if (!(coarsenable(dbx, rd, minw) && coarsenable(bbx, rd, minw))) ok = false;
            }
            if (!ok) break;
            acc.push_back(acc.back() * 2);
            dbx = coarsen(dbx, mgr);
            bbx = coarsen(bbx, mgr);
            domainboxes.push_back(dbx);
            agg_flag.push_back((bbx.npts() / nbxs) < 0.999 * thr);
        }
        bool agged = false;
        std::vector<IBox> cur = ba0;
        for (size_t lev = 1; lev < domainboxes.size(); ++lev) {
            int r[3] = {acc[lev], acc[lev], acc[lev]};
            bool ba_ok = true;
            for (const IBox& b : ba0) if (!coarsenable(b, r, minw)) ba_ok = false;
            LevelGeom G;
            for (int d = 0; d < 3; ++d) G.n[d] = domainboxes[lev].len(d);
            if (!agged && !agg_flag[lev] && ba_ok) {
                for (const IBox& b : ba0) G.ba.push_back(coarsen(b, r));
            } else {
                bool prev_ok = true;
                const int one[3] = {1, 1, 1};
                for (const IBox& b : out.back().ba) if (!coarsenable(b, mgr, one)) prev_ok = false;
                if (!prev_ok) break;
                G.ba = max_size({domainboxes[lev]}, agg);
                G.agg = true;
                agged = true;
            }
            out.push_back(G);
        }
    } else {
        // No agglomeration: coarsen the BoxArray while every box and the domain
        // stay coarsenable (the consolidation branch on one rank, MLLinOp.H:1028-1102).
        std::vector<IBox> cur = ba0;
        IBox dbx = dom;
        for (int lev = 0; lev < P.max_coarsening_level; ++lev) {
            bool ok = coarsenable(dbx, mgr, minw);
            for (const IBox& b : cur) if (!coarsenable(b, mgr, minw)) ok = false;
            if (!ok) break;
            dbx = coarsen(dbx, mgr);
            std::vector<IBox> nb;
            for (const IBox& b : cur) nb.push_back(coarsen(b, mgr));
            cur = nb;
            LevelGeom G;
            for (int d = 0; d < 3; ++d) G.n[d] = dbx.len(d);
            G.ba = cur;
            out.push_back(G);
        }
    }
    return out;
}

// ---------------------------------------------------------------------------
// Operator: levels with coefficients
// ---------------------------------------------------------------------------
template <class T>
struct Operator {
    Params P;
    std::vector<Level<T>> lev;
    bool singular = false;
    double volinv0 = 0, volinv_bot = 0;
};

template <class T>
void build_operator(Operator<T>& op, const Params& P, const double* a, const double* bx,
                    const double* by, const double* bz) {
    op.P = P;
    std::vector<LevelGeom> G = define_grids(P);
    op.lev.resize(G.size());
    for (size_t l = 0; l < G.size(); ++l) {
        Level<T>& L = op.lev[l];
        for (int d = 0; d < 3; ++d) {
            L.n[d] = G[l].n[d];
            L.dx[d] = (P.prob_hi[d] - P.prob_lo[d]) / (double)L.n[d];
            L.dxinv[d] = 1.0 / L.dx[d];
        }
        L.ba = G[l].ba;
        L.agg = G[l].agg;
        L.a.define(L.n);
        for (int d = 0; d < 3; ++d) L.b[d].define(L.n, d);
        setup_bc(L, P);
    }
    Level<T>& L0 = op.lev[0];
    const int* n = L0.n;
    if (P.op == 0) {
        for (int k = 0; k < n[2]; ++k)
            for (int j = 0; j < n[1]; ++j)
                for (int i = 0; i < n[0]; ++i)
                    L0.a(i, j, k) = T(a ? a[i + (int64_t)n[0] * (j + (int64_t)n[1] * k)] : 1.0);
        const double* src[3] = {bx, by, bz};
        for (int d = 0; d < 3; ++d)
            for (size_t p = 0; p < L0.b[d].v.size(); ++p) L0.b[d].v[p] = T(src[d] ? src[d][p] : 1.0);
    }
    // averageDownCoeffsSameAmrLevel (MLABecLaplacian.H:624-660)
    for (size_t l = 1; l < op.lev.size(); ++l) {
        Level<T>& C = op.lev[l];
        const Level<T>& F = op.lev[l - 1];
        if (P.op != 0) continue;
        if (P.alpha == 0.0) C.a.set_all(T(0));
        else restrict_avg(C.a, F.a);
        for (int d = 0; d < 3; ++d) {
            const T facInv = T(1.0) / T(4);
            const int* m = C.b[d].m;
#pragma omp parallel for schedule(static)
            for (int k = 0; k < m[2]; ++k)
                for (int j = 0; j < m[1]; ++j)
                    for (int i = 0; i < m[0]; ++i) {
                        T c = T(0.);
                        const int ii = 2 * i, jj = 2 * j, kk = 2 * k;
                        if (d == 0) {
                            for (int kr = 0; kr < 2; ++kr) for (int jr = 0; jr < 2; ++jr) c += F.b[0](ii, jj + jr, kk + kr);
                        } else if (d == 1) {
                            for (int kr = 0; kr < 2; ++kr) for (int ir = 0; ir < 2; ++ir) c += F.b[1](ii + ir, jj, kk + kr);
                        } else {
                            for (int jr = 0; jr < 2; ++jr) for (int ir = 0; ir < 2; ++ir) c += F.b[2](ii + ir, jj + jr, kk);
                        }
                        C.b[d](i, j, k) = c * facInv;
                    }
        }
    }
    // update_singular_flags (MLABecLaplacian.H:716-745), MLPoisson.H:140-160
    bool dirichlet = false;
    for (int d = 0; d < 3; ++d)
        if (!P.periodic[d] && (P.bc_lo[d] == BC_DIRICHLET || P.bc_hi[d] == BC_DIRICHLET)) dirichlet = true;
    op.singular = false;
    if (!dirichlet) {
        if (P.op == 1 || P.alpha == 0.0) op.singular = true;
        else {
            T asum = sum(L0.a), amax = norminf(L0.a);
            op.singular = std::abs(asum) <= amax * T(1.e-12);
        }
    }
    // computeVolInv (MLCellLinOp.H:1944-1980)
    auto npts = [](const Level<T>& L) { return (double)L.n[0] * L.n[1] * L.n[2]; };
    op.volinv0 = 1.0 / npts(op.lev.front());
    op.volinv_bot = 1.0 / npts(op.lev.back());
}

// ---------------------------------------------------------------------------
// BiCGStab bottom solve (MLCGSolverT::solve_bicgstab, AMReX_MLCGSolver.H:98-273)
// ---------------------------------------------------------------------------
template <class T>
int bicgstab(const Level<T>& L, const Params& P, Fld<T>& sol, const Fld<T>& rhs, T eps_rel,
             T eps_abs, int maxiter, int* niter_out) {
    Fld<T> p, r, rh, v, t;
    p.define(L.n); r.define(L.n); rh.define(L.n); v.define(L.n); t.define(L.n);
    each_valid(r, [&](int i, int j, int k) { r(i, j, k) = rhs(i, j, k); });
    normalize(L, P, r);
    rh.v = r.v;
    T rnorm = norminf(r);
    const T rnorm0 = rnorm;
    int ret = 0, iter = 1;
    T rho_1 = 0, alpha = 0, omega = 0;
    if (rnorm0 == 0 || rnorm0 < eps_abs) { *niter_out = 0; return ret; }
    for (; iter <= maxiter; ++iter) {
        const T rho = dot(rh, r);
        if (rho == 0) { ret = 1; break; }
        if (iter == 1) {
            each_valid(p, [&](int i, int j, int k) { p(i, j, k) = r(i, j, k); });
        } else {
            const T beta = (rho / rho_1) * (alpha / omega);
            each_valid(p, [&](int i, int j, int k) { p(i, j, k) += -omega * v(i, j, k); });
            each_valid(p, [&](int i, int j, int k) { p(i, j, k) = r(i, j, k) + beta * p(i, j, k); });
        }
        apply_hom(L, P, v, p);
        normalize(L, P, v);
        const T rhTv = dot(rh, v);
        if (rhTv != T(0.0)) alpha = rho / rhTv;
        else { ret = 2; break; }
        each_valid(sol, [&](int i, int j, int k) { sol(i, j, k) += alpha * p(i, j, k); });
        each_valid(r, [&](int i, int j, int k) { r(i, j, k) += -alpha * v(i, j, k); });
        rnorm = norminf(r);
        if (rnorm < eps_rel * rnorm0 || rnorm < eps_abs) break;
        apply_hom(L, P, t, r);
        normalize(L, P, t);
        const T tt = dot(t, t), tr = dot(t, r);
        if (tt != T(0.0)) omega = tr / tt;
        else { ret = 3; break; }
        each_valid(sol, [&](int i, int j, int k) { sol(i, j, k) += omega * r(i, j, k); });
        each_valid(r, [&](int i, int j, int k) { r(i, j, k) += -omega * t(i, j, k); });
        rnorm = norminf(r);
        if (rnorm < eps_rel * rnorm0 || rnorm < eps_abs) break;
        if (omega == 0) { ret = 4; break; }
        rho_1 = rho;
    }
    if (ret == 0 && rnorm > eps_rel * rnorm0 && rnorm > eps_abs) ret = 8;
    if ((ret == 0 || ret == 8) && (rnorm < rnorm0)) {
        if (ret == 8) ret = 9;
    } else {
        sol.set_all(T(0.0));
    }
    *niter_out = iter;
    return ret;
}

// CG bottom solve (MLCGSolverT::solve_cg, AMReX_MLCGSolver.H:275-410) with a
// zero initial guess: plain conjugate gradients on A (no Jacobi scaling),
// return codes 1 (rho or p.Ap zero), 8 (not converged), 9 (not converged
// but reduced: the iterate is kept).
template <class T>
int cg(const Level<T>& L, const Params& P, Fld<T>& sol, const Fld<T>& rhs, T eps_rel, T eps_abs, int maxiter,
       int* niter_out) {
    Fld<T> p, r, q;
    p.define(L.n); r.define(L.n); q.define(L.n);
    each_valid(r, [&](int i, int j, int k) { r(i, j, k) = rhs(i, j, k); });
    T rnorm = norminf(r);
    const T rnorm0 = rnorm;
    T rho_1 = 0;
    int ret = 0, iter = 1;
    if (rnorm0 == 0 || rnorm0 < eps_abs) { *niter_out = 0; return ret; }
    for (; iter <= maxiter; ++iter) {
        const T rho = dot(r, r);
        if (rho == 0) { ret = 1; break; }
        if (iter == 1) {
            each_valid(p, [&](int i, int j, int k) { p(i, j, k) = r(i, j, k); });
        } else {
            const T beta = rho / rho_1;
            each_valid(p, [&](int i, int j, int k) { p(i, j, k) = r(i, j, k) + beta * p(i, j, k); });
        }
        apply_hom(L, P, q, p);
        const T pw = dot(p, q);
        T alpha;
        if (pw != T(0.0)) alpha = rho / pw;
        else { ret = 1; break; }
        each_valid(sol, [&](int i, int j, int k) { sol(i, j, k) += alpha * p(i, j, k); });
        each_valid(r, [&](int i, int j, int k) { r(i, j, k) += -alpha * q(i, j, k); });
        rnorm = norminf(r);
        if (rnorm < eps_rel * rnorm0 || rnorm < eps_abs) break;
        rho_1 = rho;
    }
    if (ret == 0 && rnorm > eps_rel * rnorm0 && rnorm > eps_abs) ret = 8;
    if ((ret == 0 || ret == 8) && (rnorm < rnorm0)) {
        if (ret == 8) ret = 9;
    } else {
        sol.set_all(T(0.0));
    }
    *niter_out = iter;
    return ret;
}

// The Krylov part of MLMGT::actualBottomSolve (AMReX_MLMG.H:1567-1601): CG
// first for cg/cgbicg, BiCGStab first otherwise; bicgcg and cgbicg retry a
// failed solve from zero with the other method and, when the retry succeeds,
// keep that method for every later bottom solve (mode is updated).  iters:
// the iterations of every attempt.  x is zero on return when the last
// attempt failed with a code other than 9.
template <class T>
int bottom_krylov(const Level<T>& L, const Params& P, Fld<T>& x, const Fld<T>& b, T eps_rel, T eps_abs, int maxiter,
                  int& mode, int* iters) {
    auto run = [&](bool use_cg, int* it) {
        return use_cg ? cg(L, P, x, b, eps_rel, eps_abs, maxiter, it) : bicgstab(L, P, x, b, eps_rel, eps_abs, maxiter, it);
    };
    bool use_cg = mode == 3 || mode == 5;
    int it = 0;
    int ret = run(use_cg, &it);
    *iters = it;
    if (ret != 0 && (mode == 4 || mode == 5)) {
        use_cg = !use_cg;
        x.set_all(T(0));
        ret = run(use_cg, &it);
        *iters += it;
        if (ret == 0) mode = use_cg ? 3 : 2;
    }
    if (ret != 0 && ret != 9) x.set_all(T(0));
    return ret;
}

template <class T>
void make_solvable(Fld<T>& f, double volinv) {
    const T offset = sum(f) * static_cast<T>(volinv);
    each_valid(f, [&](int i, int j, int k) { f(i, j, k) += -offset; });
}

// ---------------------------------------------------------------------------
// MLMG driver (single AMR level)
// ---------------------------------------------------------------------------
struct Stats {
    int niters;
    int nlevels;
    int status;  // 0 converged, 1 failed to converge, 2 blew up
    int bottom_iters_total;
    double init_resnorm, rhsnorm, final_resnorm, max_norm;
    double res_history[256];
    double solve_seconds;
    double setup_seconds;  // operator build, copies, initial residual
    double iter_seconds;   // the iteration loop
    int nvcycles;          // V- and F-cycles run (more than niters with refine_vcycles != 1)
    int bottom_solver_final;  // bottom solver after the solve (bicgcg/cgbicg switch for good)
};

// Per-level work arrays of a V-cycle in precision V (hold: cor_hold of the F-cycle).
template <class V>
struct Work {
    std::vector<Fld<V>> res, rescor, cor, hold;
    int bottom_mode = 2;  // the bottom solver in force (Params::bottom_solver codes)
};
// The bottom solver a solve starts with (Default means BiCGStab for these operators).
// (hypre and petsc, codes 6 and 7, are refused at the C entry points)
inline int bottom_mode_of(const Params& P) { return P.bottom_solver == 0 ? 2 : P.bottom_solver; }

template <class V>
void make_work(Work<V>& w, const std::vector<Level<V>>& lev) {
    const size_t nl = lev.size();
    w.res.resize(nl); w.rescor.resize(nl); w.cor.resize(nl); w.hold.resize(nl);
    for (size_t l = 0; l < nl; ++l) {
        w.res[l].define(lev[l].n); w.rescor[l].define(lev[l].n); w.cor[l].define(lev[l].n);
        w.hold[l].define(lev[l].n);
    }
}

// Bottom solve on level nl-1 with precision B coefficients; cor/res in V.
template <class V, class B>
int bottom_solve(const Operator<B>& opb, const Operator<V>& opv, Work<V>& w, bool singular,
                 int* iters) {
    const Params& P = opv.P;
    const size_t lb = opv.lev.size() - 1;
    const Level<B>& L = opb.lev[lb];
    Fld<B> x, b;
    x.define(L.n); b.define(L.n);
    each_valid(b, [&](int i, int j, int k) { b(i, j, k) = static_cast<B>(w.res[lb](i, j, k)); });
    if (singular) make_solvable(b, opb.volinv_bot);
    if (w.bottom_mode == 1) {
        // BottomSolver::smoother (AMReX_MLMG.H:1521-1527): nuf sweeps from zero
        // on the bottom rhs as it stands
        Fld<V>& cor = w.cor[lb];
        cor.set_all(V(0));
        for (int s = 0; s < 8 /* nuf */; ++s) smooth(opv.lev[lb], P, cor, w.res[lb]);
        *iters = 0;
        return 0;
    }
    const double reltol = P.bottom_reltol > 0 ? P.bottom_reltol
                          : (std::is_same<B, double>::value ? 1.e-4 : 1.e-3);
    const int ret = bottom_krylov(L, opb.P, x, b, static_cast<B>(reltol), static_cast<B>(-1.0), P.bottom_maxiter,
                                  w.bottom_mode, iters);
    Fld<V>& cor = w.cor[lb];
    cor.set_all(V(0));
    each_valid(cor, [&](int i, int j, int k) { cor(i, j, k) = static_cast<V>(x(i, j, k)); });
    const int nsm = (ret == 0) ? 0 /* nub */ : 8 /* nuf */;
    for (int s = 0; s < nsm; ++s) smooth(opv.lev[lb], P, cor, w.res[lb]);
    return ret;
}

// mgVcycle (AMReX_MLMG.H:1308-1415) on amrlev 0 from MG level top; input
// res[top], output cor[top].  warm: the top level's smoothing starts from
// cor[top] as it stands (the second V-cycle of a mixed-precision refinement
// step) instead of zero.
template <class V, class B>
void vcycle(const Operator<V>& opv, const Operator<B>& opb, Work<V>& w, bool singular, int* bot_iters,
            size_t top = 0, bool warm = false) {
    const Params& P = opv.P;
    const size_t nb = opv.lev.size() - 1;
    for (size_t l = top; l < nb; ++l) {
        if (!(warm && l == top)) w.cor[l].set_all(V(0.0));
        for (int s = 0; s < P.nu1; ++s) smooth(opv.lev[l], P, w.cor[l], w.res[l]);
        resid_cor<V>(opv.lev[l], P, w.rescor[l], w.cor[l], w.res[l]);
        restrict_avg(w.res[l + 1], w.rescor[l]);
    }
    int it = 0;
    if (nb == 0) {
        w.cor[0].set_all(V(0));
    }
    bottom_solve(opb, opv, w, singular, &it);
    *bot_iters += it;
    for (size_t l = nb; l-- > top;) {
        prolong_add(w.cor[l], w.cor[l + 1]);
        for (int s = 0; s < P.nu2; ++s) smooth(opv.lev[l], P, w.cor[l], w.res[l]);
    }
}

// FillBoundary(periodicity) over every ghost cell (the crse.FillBoundary of
// interpAssign, AMReX_MLCellLinOp.H:1002-1016): a ghost cell whose periodic
// image is a valid cell takes its value; ghost cells outside a non-periodic
// face keep theirs.
template <class T>
void fill_periodic_all(const Level<T>& L, Fld<T>& phi) {
    const int* n = L.n;
    for (int k = -1; k <= n[2]; ++k)
        for (int j = -1; j <= n[1]; ++j)
            for (int i = -1; i <= n[0]; ++i) {
                int p[3] = {i, j, k};
                bool ghost = false, image = true;
                for (int d = 0; d < 3; ++d) {
                    if (p[d] >= 0 && p[d] < n[d]) continue;
                    ghost = true;
                    if (L.bct[d] != BC_INTERIOR) { image = false; break; }
                    p[d] = p[d] < 0 ? p[d] + n[d] : p[d] - n[d];
                }
                if (ghost && image) phi(i, j, k) = phi(p[0], p[1], p[2]);
            }
}

// mlmg_lin_cc_interp_r2 (AMReX_MLMG_3D_K.H:9-38): fine valid cells = trilinear
// interpolation of crse (ratio 2), crse ghosts as they stand.
template <class T>
void interp_assign(Fld<T>& fine, const Fld<T>& crse) {
    each_valid(fine, [&](int i, int j, int k) {
        const int ic = i / 2, jc = j / 2, kc = k / 2;
        const int io = 2 * (i - 2 * ic) - 1, jo = 2 * (j - 2 * jc) - 1, ko = 2 * (k - 2 * kc) - 1;
        fine(i, j, k) = T(0.421875) * crse(ic, jc, kc) + T(0.140625) * crse(ic + io, jc, kc) +
                        T(0.140625) * crse(ic, jc + jo, kc) + T(0.140625) * crse(ic, jc, kc + ko) +
                        T(0.046875) * crse(ic, jc + jo, kc + ko) + T(0.046875) * crse(ic + io, jc, kc + ko) +
                        T(0.046875) * crse(ic + io, jc + jo, kc) + T(0.015625) * crse(ic + io, jc + jo, kc + ko);
    });
}

// Coarse data the F-cycle interpolation reads (MLCellLinOp::interpAssign,
// AMReX_MLCellLinOp.H:1002-1026).  When the two MG levels share their
// BoxArray lineage (MFIter-safe), the coarse correction itself after
// FillBoundary: ghosts outside a physical face keep what the last boundary fill
// wrote there (the closure before the last smoothing colour of that level's
// cycle, computed without the held correction; 0 at the bottom level, and 0 on
// edges and corners).  Otherwise (the coarse level was agglomerated) a zeroed
// copy that receives valid and periodic-image values only: 0 outside physical
// faces.
template <class T>
void fmg_coarse_view(const Level<T>& Lc, Fld<T>& crse, Fld<T>& tmp) {
    if (!Lc.agg) {
        fill_periodic_all(Lc, crse);
        return;
    }
    tmp.define(Lc.n);
    tmp.set_all(T(0));
    each_valid(tmp, [&](int i, int j, int k) { tmp(i, j, k) = crse(i, j, k); });
    fill_periodic_all(Lc, tmp);
}

// mgFcycle (AMReX_MLMG.H:1422-1457) on amrlev 0: input res[0], output cor[0].
template <class V, class B>
void fcycle(const Operator<V>& opv, const Operator<B>& opb, Work<V>& w, bool singular, int* bot_iters) {
    const size_t nb = opv.lev.size() - 1;
    for (size_t l = 1; l <= nb; ++l) restrict_avg(w.res[l], w.res[l - 1]);  // avgDownResMG
    int it = 0;
    bottom_solve(opb, opv, w, singular, &it);
    *bot_iters += it;
    for (size_t l = nb; l-- > 0;) {
        const Level<V>& Lc = opv.lev[l + 1];
        Fld<V> tmp;
        fmg_coarse_view(Lc, w.cor[l + 1], tmp);
        interp_assign(w.cor[l], Lc.agg ? tmp : w.cor[l + 1]);                   // interpCorrection
        resid_cor<V>(opv.lev[l], opv.P, w.rescor[l], w.cor[l], w.res[l]);  // computeResOfCorrection
        each_valid(w.res[l], [&](int i, int j, int k) { w.res[l](i, j, k) = w.rescor[l](i, j, k); });
        std::swap(w.cor[l], w.hold[l]);
        vcycle(opv, opb, w, singular, bot_iters, l);
        each_valid(w.cor[l], [&](int i, int j, int k) { w.cor[l](i, j, k) += w.hold[l](i, j, k); });
    }
}

// V-cycles of the next refinement step of the mixed form: refine_vcycles if
// 1 or 2; for 0, one if one V-cycle at the last measured contraction rho is
// predicted to reach the target, else two.  The other forms run one.
inline int refine_cycles(const Params& P, double fine, double rho, double target) {
    if (P.precision != 2) return 1;
    if (P.refine_vcycles == 1 || P.refine_vcycles == 2) return P.refine_vcycles;
    return fine * rho <= target ? 1 : 2;
}

// MLMGT::solve (AMReX_MLMG.H:356-541) for precision modes 0 (T=double) and
// 1 (T=float); mode 2 runs an FP32 V-cycle inside an FP64 outer loop.
template <class T, class V>
int solve(const Params& P, const double* a, const double* bx, const double* by, const double* bz,
          const double* rhs_in, const double* bcv_in, double* sol_io, Stats* st) {
    auto t0 = std::chrono::steady_clock::now();
    Operator<T> op;
    build_operator(op, P, a, bx, by, bz);
    Operator<V> opv_store;
    const Operator<V>* opv;
    if constexpr (std::is_same<T, V>::value) {
        opv = &op;
    } else {
        build_operator(opv_store, P, a, bx, by, bz);
        opv = &opv_store;
    }
    const Level<T>& L0 = op.lev[0];
    const int* n = L0.n;
    Fld<T> sol, rhs, res;
    sol.define(n); rhs.define(n); res.define(n);
    std::unique_ptr<Fld<T>> bcv;
    if (bcv_in) {
        bcv.reset(new Fld<T>);
        bcv->define(n);
        for (size_t p = 0; p < bcv->v.size(); ++p) bcv->v[p] = T(bcv_in[p]);
    }
    auto gidx = [&](int i, int j, int k) { return i + (int64_t)n[0] * (j + (int64_t)n[1] * k); };
    each_valid(sol, [&](int i, int j, int k) { sol(i, j, k) = T(sol_io[gidx(i, j, k)]); });
    each_valid(rhs, [&](int i, int j, int k) { rhs(i, j, k) = T(rhs_in[gidx(i, j, k)]); });
    if (op.singular) make_solvable(rhs, op.volinv0);
    Work<V> w;
    make_work(w, opv->lev);
    w.bottom_mode = bottom_mode_of(P);
    st->nlevels = (int)op.lev.size();
    st->bottom_iters_total = 0;
    st->status = 0;

    // fine residual: rhs - A sol, or rhs - A^T sol for the adjoint (homogeneous)
    auto fine_residual = [&]() {
        if (P.transpose) residual_transpose(L0, P, res, sol, rhs);
        else residual(L0, P, res, sol, rhs, bcv.get());
    };
    fine_residual();
    const T resnorm0 = norminf(res), rhsnorm0 = norminf(rhs);
    st->init_resnorm = resnorm0;
    st->rhsnorm = rhsnorm0;
    const T max_norm = (rhsnorm0 >= resnorm0) ? rhsnorm0 : resnorm0;
    st->max_norm = max_norm;
    const T res_target = std::max(T(P.tol_abs), std::max(T(P.tol_rel), T(1.e-16)) * max_norm);
    int iters = 0;
    st->nvcycles = 0;
    T fine = resnorm0;
    double rho = 0.1;  // contraction per V-cycle assumed before the first measured step
    auto t1 = std::chrono::steady_clock::now();
    st->setup_seconds = std::chrono::duration<double>(t1 - t0).count();
    if (resnorm0 > res_target) {
        bool converged = false;
        const int niters = P.fixed_iters ? P.fixed_iters : P.max_iter;
        for (int iter = 0; iter < niters; ++iter) {
            if (op.singular) make_solvable(res, op.volinv0);
            each_valid(w.res[0], [&](int i, int j, int k) { w.res[0](i, j, k) = static_cast<V>(res(i, j, k)); });
            const bool fmg = iter < P.max_fmg_iters;
            const int m = fmg ? 1 : refine_cycles(P, (double)fine, rho, (double)res_target);
            if (fmg) fcycle(*opv, op, w, op.singular, &st->bottom_iters_total);
            else vcycle(*opv, op, w, op.singular, &st->bottom_iters_total);
            if (m == 2) vcycle(*opv, op, w, op.singular, &st->bottom_iters_total, 0, true);
            st->nvcycles += m;
            each_valid(sol, [&](int i, int j, int k) { sol(i, j, k) += static_cast<T>(w.cor[0](i, j, k)); });
            const T before = fine;
            fine_residual();
            fine = norminf(res);
            if (before > 0 && fine > 0) rho = std::pow((double)fine / (double)before, 1.0 / m);
            iters = iter + 1;
            if (iter < 256) st->res_history[iter] = fine;
            if (P.verbose >= 2)
                std::printf("MLMG(oracle): Iteration %3d Fine resid/max_norm = %.6e\n", iters, (double)(fine / max_norm));
            if (fine <= res_target) { converged = true; break; }
            if (fine > T(1.e20) * max_norm) { st->status = 2; break; }
        }
        if (!converged && !P.fixed_iters && st->status == 0) st->status = 1;
    }
    st->iter_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
    st->niters = iters;
    st->bottom_solver_final = w.bottom_mode;
    st->final_resnorm = fine;
    each_valid(sol, [&](int i, int j, int k) { sol_io[gidx(i, j, k)] = (double)sol(i, j, k); });
    st->solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return st->status;
}

// ---------------------------------------------------------------------------
// GMRES with the MLMG preconditioner (GMRES<V,M>, AMReX_GMRES.H:44-413, and
// GMRESMLMGT, AMReX_GMRES_MLMG.H:111-231), FP64, single AMR level.
// ---------------------------------------------------------------------------
struct GmresOpts {
    int use_precond, precond_niters, restart, maxiter;
};
struct GmresOut {
    int iters, status;
    double resnorm, rnorm0;
    double res_hist[512];  // |grs| after each iteration
    int nhist;
};

inline void gmres_mlmg(const Params& P, const double* a, const double* bx, const double* by, const double* bz,
                       const double* rhs_in, const double* bcv_in, double* sol_io, const GmresOpts& o, GmresOut* out) {
    using Vec = Fld<double>;
    Operator<double> op;
    build_operator(op, P, a, bx, by, bz);
    const Level<double>& L0 = op.lev[0];
    const int* n = L0.n;
    Work<double> w;
    make_work(w, op.lev);
    w.bottom_mode = bottom_mode_of(P);
    std::unique_ptr<Vec> bcv;
    if (bcv_in) {
        bcv.reset(new Vec);
        bcv->define(n);
        for (size_t p = 0; p < bcv->v.size(); ++p) bcv->v[p] = bcv_in[p];
    }
    auto gidx = [&](int i, int j, int k) { return i + (int64_t)n[0] * (j + (int64_t)n[1] * k); };
    auto mk = [&]() { Vec v; v.define(n); return v; };
    // vector operations of GMRESMLMGT over the valid cells (AMReX_GMRES_MLMG.H:120-178)
    auto dot = [&](const Vec& x, const Vec& y) {
        double sum = 0.0;
        for (int k = 0; k < n[2]; ++k)
            for (int j = 0; j < n[1]; ++j)
                for (int i = 0; i < n[0]; ++i) sum += x(i, j, k) * y(i, j, k);
        return sum;
    };
    auto norm2 = [&](const Vec& x) { return std::sqrt(dot(x, x)); };
    auto scale = [&](Vec& x, double f) { each_valid(x, [&](int i, int j, int k) { x(i, j, k) *= f; }); };
    auto assign = [&](Vec& x, const Vec& y) { each_valid(x, [&](int i, int j, int k) { x(i, j, k) = y(i, j, k); }); };
    auto increment = [&](Vec& x, const Vec& y, double f) {
        each_valid(x, [&](int i, int j, int k) { x(i, j, k) += f * y(i, j, k); });
    };
    auto apply_h = [&](Vec& lhs, Vec& rhs) {  // homogeneous BC, correction state
        fill_ghosts<double>(L0, rhs, nullptr);
        fapply(L0, P, lhs, rhs);
    };
    int bot = 0;
    auto precond = [&](Vec& lhs, const Vec& rhs) {
        if (!o.use_precond) { assign(lhs, rhs); return; }
        for (int ic = 0; ic < o.precond_niters; ++ic) {
            if (ic == 0) {
                assign(w.res[0], rhs);
            } else {
                residual<double>(L0, P, w.rescor[0], w.cor[0], w.res[0], nullptr);  // computeResOfCorrection
                assign(w.res[0], w.rescor[0]);
            }
            vcycle(op, op, w, op.singular, &bot);
            if (ic == 0) assign(lhs, w.cor[0]);
            else increment(lhs, w.cor[0], 1.0);
        }
    };
    Vec sol = mk(), rhs = mk();
    each_valid(sol, [&](int i, int j, int k) { sol(i, j, k) = sol_io[gidx(i, j, k)]; });
    each_valid(rhs, [&](int i, int j, int k) { rhs(i, j, k) = rhs_in[gidx(i, j, k)]; });
    // GMRESMLMGT::solve without the property of zero: res = L(sol) - rhs (MLMG::apply,
    // inhomogeneous BC), GMRES on L(cor) = res, sol -= cor
    Vec res = mk(), cor = mk();
    fill_ghosts<double>(L0, sol, bcv.get());
    fapply(L0, P, res, sol);
    increment(res, rhs, -1.0);
    // GMRES<V,M>::solve
    const int rs = o.restart;
    std::vector<double> hh((size_t)(rs + 2) * (rs + 1)), hes(hh.size()), grs(rs + 2), cc(rs + 1), ss(rs + 1);
    auto H = [&](std::vector<double>& t, int i, int j) -> double& { return t[(size_t)i * (rs + 1) + j]; };
    std::vector<Vec> vv;
    vv.reserve(rs + 1);
    vv.push_back(mk());
    Vec tmp_rhs = mk(), tmp_lhs = mk();
    assign(vv[0], res);
    cor.set_all(0.0);
    int its = 0, status = -1;
    double rnorm0 = 0.0, mres = 0.0;
    out->nhist = 0;
    auto converged = [&](double r0, double r) { return (r < r0 * P.tol_rel) || (r < P.tol_abs); };
    auto cycle = [&]() {
        mres = norm2(vv[0]);
        grs[0] = mres;
        if (mres == 0.0) { status = 0; return; }
        scale(vv[0], 1.0 / mres);
        if (its == 0) rnorm0 = mres;
        status = converged(rnorm0, mres) ? 0 : -1;
        int it = 0;
        while (it < rs && its < o.maxiter) {
            if (status == 0) break;
            while ((int)vv.size() < it + 2) vv.push_back(mk());
            precond(tmp_lhs, vv[(size_t)it]);
            apply_h(vv[(size_t)it + 1], tmp_lhs);
            // classical Gram-Schmidt, twice (gram_schmidt_orthogonalization)
            Vec& v1 = vv[(size_t)it + 1];
            for (int j = 0; j <= it; ++j) { H(hh, j, it) = 0.0; H(hes, j, it) = 0.0; }
            std::vector<double> lhh((size_t)it + 1);
            for (int cnt = 0; cnt < 2; ++cnt) {
                for (int j = 0; j <= it; ++j) lhh[(size_t)j] = dot(v1, vv[(size_t)j]);
                for (int j = 0; j <= it; ++j) {
                    increment(v1, vv[(size_t)j], -lhh[(size_t)j]);
                    H(hh, j, it) += lhh[(size_t)j];
                    H(hes, j, it) -= lhh[(size_t)j];
                }
            }
            const double tt = norm2(v1);
            const bool happyend = tt < 1.e-99;
            if (!happyend) scale(v1, 1.0 / tt);
            H(hh, it + 1, it) = tt;
            H(hes, it + 1, it) = tt;
            // update_hessenberg: Givens rotations
            for (int j = 1; j <= it; ++j) {
                const double t0 = H(hh, j - 1, it);
                H(hh, j - 1, it) = cc[(size_t)j - 1] * t0 + ss[(size_t)j - 1] * H(hh, j, it);
                H(hh, j, it) = cc[(size_t)j - 1] * H(hh, j, it) - ss[(size_t)j - 1] * t0;
            }
            if (!happyend) {
                const double t1 = std::sqrt(H(hh, it, it) * H(hh, it, it) + H(hh, it + 1, it) * H(hh, it + 1, it));
                cc[(size_t)it] = H(hh, it, it) / t1;
                ss[(size_t)it] = H(hh, it + 1, it) / t1;
                grs[(size_t)it + 1] = -(ss[(size_t)it] * grs[(size_t)it]);
                grs[(size_t)it] = cc[(size_t)it] * grs[(size_t)it];
                H(hh, it, it) = cc[(size_t)it] * H(hh, it, it) + ss[(size_t)it] * H(hh, it + 1, it);
                mres = std::abs(grs[(size_t)it + 1]);
            } else {
                mres = 0.0;
            }
            ++it;
            ++its;
            if (out->nhist < 512) out->res_hist[out->nhist++] = mres;
            status = converged(rnorm0, mres) ? 0 : -1;
            if (happyend) break;
        }
        // build_solution(xx, it - 1)
        const int top = it - 1;
        if (top < 0) return;
        if (H(hh, top, top) != 0.0) grs[(size_t)top] /= H(hh, top, top);
        else grs[(size_t)top] = 0.0;
        for (int ii = 1; ii <= top; ++ii) {
            const int k = top - ii;
            double t2 = grs[(size_t)k];
            for (int j = k + 1; j <= top; ++j) t2 -= H(hh, k, j) * grs[(size_t)j];
            grs[(size_t)k] = t2 / H(hh, k, k);
        }
        tmp_rhs.set_all(0.0);
        for (int ii = 0; ii <= top; ++ii) increment(tmp_rhs, vv[(size_t)ii], grs[(size_t)ii]);
        precond(tmp_lhs, tmp_rhs);
        increment(cor, tmp_lhs, 1.0);
    };
    cycle();
    while (status == -1 && its < o.maxiter) {
        // compute_residual: vv0 = res - L(cor)
        assign(tmp_lhs, cor);
        apply_h(tmp_rhs, tmp_lhs);
        each_valid(vv[0], [&](int i, int j, int k) { vv[0](i, j, k) = 1.0 * res(i, j, k) + (-1.0) * tmp_rhs(i, j, k); });
        cycle();
    }
    if (status == -1 && its >= o.maxiter) status = 1;
    increment(sol, cor, -1.0);
    each_valid(sol, [&](int i, int j, int k) { sol_io[gidx(i, j, k)] = sol(i, j, k); });
    out->iters = its;
    out->status = status;
    out->resnorm = mres;
    out->rnorm0 = rnorm0;
}

template <class T>
void load_field(Fld<T>& f, const double* src, bool grown) {
    if (grown) {
        for (size_t p = 0; p < f.v.size(); ++p) f.v[p] = T(src[p]);
    } else {
        const int* n = f.n;
        each_valid(f, [&](int i, int j, int k) { f(i, j, k) = T(src[i + (int64_t)n[0] * (j + (int64_t)n[1] * k)]); });
    }
}
template <class T>
void store_field(const Fld<T>& f, double* dst, bool grown) {
    if (grown) {
        for (size_t p = 0; p < f.v.size(); ++p) dst[p] = (double)f.v[p];
    } else {
        const int* n = f.n;
        for (int k = 0; k < n[2]; ++k)
            for (int j = 0; j < n[1]; ++j)
                for (int i = 0; i < n[0]; ++i) dst[i + (int64_t)n[0] * (j + (int64_t)n[1] * k)] = (double)f(i, j, k);
    }
}

template <class T>
int level_op(const Params& P, int which, const double* a, const double* bx, const double* by,
             const double* bz, const double* bcv_in, const double* in1, double* io, double* out,
             int count, int level) {
    Operator<T> op;
    build_operator(op, P, a, bx, by, bz);
    if (level < 0 || level >= (int)op.lev.size()) return -1;
    const Level<T>& L = op.lev[level];
    Fld<T> x, y, bcv;
    x.define(L.n); y.define(L.n);
    if (bcv_in) { bcv.define(L.n); load_field(bcv, bcv_in, true); }
    switch (which) {
        case 0: {  // smooth x count times against rhs in1 (homogeneous BC); io grown
            load_field(x, io, true);
            load_field(y, in1, false);
            for (int s = 0; s < count; ++s) smooth(L, P, x, y);
            store_field(x, io, true);
            break;
        }
        case 1: {  // out = L(io) with BC (inhomogeneous if bcv given); io grown, ghosts filled
            load_field(x, io, true);
            fill_ghosts<T>(L, x, bcv_in ? &bcv : nullptr);
            fapply(L, P, y, x);
            store_field(x, io, true);
            store_field(y, out, false);
            break;
        }
        case 2: {  // out = in1 - L(io)
            Fld<T> b;
            b.define(L.n);
            load_field(b, in1, false);
            load_field(x, io, true);
            residual(L, P, y, x, b, bcv_in ? &bcv : nullptr);
            store_field(x, io, true);
            store_field(y, out, false);
            break;
        }
        case 3: {  // one colour (count = redblack) after filling ghosts (transposed: gsrb_T)
            load_field(x, io, true);
            load_field(y, in1, false);
            if (transposed_cycle(P)) {
                gsrb_T(L, P, x, y, count);
                store_field(x, io, true);
                break;
            }
            fill_ghosts<T>(L, x, nullptr);
            gsrb(L, P, x, y, count);
            store_field(x, io, true);
            break;
        }
        case 4: {  // out = coefficients of level: a then bx, by, bz concatenated
            size_t off = 0;
            const int* n = L.n;
            for (int k = 0; k < n[2]; ++k)
                for (int j = 0; j < n[1]; ++j)
                    for (int i = 0; i < n[0]; ++i) out[off++] = (double)L.a(i, j, k);
            for (int d = 0; d < 3; ++d)
                for (T v : L.b[d].v) out[off++] = (double)v;
            break;
        }
        case 5: {  // BiCGStab on this level: io (valid, n^3) = solution of L x = in1
            Fld<T> b;
            b.define(L.n);
            load_field(b, in1, false);
            int it = 0;
            const int ret = bicgstab(L, P, x, b, T(P.bottom_reltol > 0 ? P.bottom_reltol : 1e-4), T(-1.0),
                                     P.bottom_maxiter, &it);
            store_field(x, io, false);
            if (out) { out[0] = ret; out[1] = it; }
            break;
        }
        case 13: {  // the bottom Krylov solve of actualBottomSolve on this level, solver
                    // P.bottom_solver (cg, bicgstab, bicgcg, cgbicg): io (valid) = solution of
                    // L x = in1; out = {return code, iterations of all attempts, solver after}
            if (P.bottom_solver < 2 || P.bottom_solver > 5) return -1;
            Fld<T> b;
            b.define(L.n);
            load_field(b, in1, false);
            int it = 0, mode = P.bottom_solver;
            const int ret = bottom_krylov(L, P, x, b, T(P.bottom_reltol > 0 ? P.bottom_reltol : 1e-4), T(-1.0),
                                          P.bottom_maxiter, mode, &it);
            store_field(x, io, false);
            out[0] = ret;
            out[1] = it;
            out[2] = mode;
            break;
        }
        case 10: {  // F-cycle interpolation to level-1: io = this level's correction (grown, ghosts
                    // as given); count 1: through fmg_coarse_view, 0: ghosts as given; out = valid
                    // cells of level-1 by mlmg_lin_cc_interp_r2; io = the data the interpolation read
            if (level < 1) return -1;
            load_field(x, io, true);
            Fld<T> tmp;
            Fld<T>* src = &x;
            if (count & 1) {
                fmg_coarse_view(L, x, tmp);
                if (L.agg) src = &tmp;
            }
            Fld<T> f;
            f.define(op.lev[(size_t)level - 1].n);
            interp_assign(f, *src);
            store_field(*src, io, true);
            store_field(f, out, false);
            break;
        }
        case 6: {  // out = A^T io (homogeneous BC); io grown, valid part used
            load_field(x, io, true);
            apply_transpose(L, P, y, x);
            store_field(y, out, false);
            break;
        }
        case 7: {  // gradient: in1 = lam (valid), io = phi (grown; ghosts filled here, bcv if given);
                   // out = [dJ/da (n^3), dJ/dbx, dJ/dby, dJ/dbz]
            Fld<T> lam, ga;
            lam.define(L.n);
            ga.define(L.n);
            load_field(lam, in1, false);
            load_field(x, io, true);
            fill_ghosts<T>(L, x, bcv_in ? &bcv : nullptr);
            FaceArr<T> gb[3];
            for (int d = 0; d < 3; ++d) gb[d].define(L.n, d);
            grad_coeffs(L, P, lam, x, ga, gb);
            store_field(ga, out, false);
            size_t off = (size_t)L.n[0] * L.n[1] * L.n[2];
            for (int d = 0; d < 3; ++d)
                for (T v : gb[d].v) out[off++] = (double)v;
            break;
        }
        case 9: {  // face values of the solution io (grown; ghosts filled here, bcv if given):
                   // count = 0 gradient, 1 flux; out = [x faces, y faces, z faces]
            load_field(x, io, true);
            fill_ghosts<T>(L, x, bcv_in ? &bcv : nullptr);
            FaceArr<T> fv[3];
            for (int d = 0; d < 3; ++d) fv[d].define(L.n, d);
            face_values(L, P, x, count, fv);
            size_t off = 0;
            for (int d = 0; d < 3; ++d)
                for (T v : fv[d].v) out[off++] = (double)v;
            break;
        }
        case 8: {  // out = A io from the explicit rows (homogeneous BC); io grown, valid part used
            load_field(x, io, true);
            apply_rows(L, P, y, x);
            store_field(y, out, false);
            break;
        }
        default: return -2;
    }
    return 0;
}


// ---------------------------------------------------------------------------
// Two-level AMR composite solve: MLPoisson on a periodic coarse level (amrlev
// 0) and fine boxes of refinement ratio 2 (amrlev 1) anywhere in the periodic
// domain (a ghost cell beyond a domain face is the periodic image).  Restates MLMGT::solve/oneIter for two AMR levels (AMReX_MLMG.H:
// 356-541, 1228-1293, 1652-1747), the coarse/fine boundary of
// MLCellLinOpT::updateSolBC/updateCorBC + applyBC (AMReX_MLCellLinOp.H:
// 652-893) with InterpBndryData's o3 interpolation (AMReX_InterpBndryData.H:
// 161-267, AMReX_InterpBndryData_3D_K.H:23-118), MLCellLinOpT::reflux
// (AMReX_MLCellLinOp.H:1273-1344) with YAFluxRegister's CrseAdd/FineAdd/
// Reflux (AMReX_YAFluxRegister.H:343-540, AMReX_YAFluxRegister_3D_K.H), the
// AMR correction interpolation mlmg_lin_cc_interp_r2 (AMReX_MLMG_3D_K.H:9-38)
// and average_down between the levels.  The fine AMR level has one MG level
// (ratio 2), so its "V-cycle" is nu1 smoothing sweeps.
// ---------------------------------------------------------------------------
template <class T>
struct BoxFld {  // one fine box with one ghost layer
    IBox b;
    int m[3] = {0, 0, 0};
    std::vector<T> v;
    void define(const IBox& bx) {
        b = bx;
        for (int d = 0; d < 3; ++d) m[d] = bx.len(d) + 2;
        v.assign((size_t)m[0] * m[1] * m[2], T(0));
    }
    T& operator()(int i, int j, int k) {
        return v[(size_t)((i - b.lo[0] + 1) + (int64_t)m[0] * ((j - b.lo[1] + 1) + (int64_t)m[1] * (k - b.lo[2] + 1)))];
    }
    const T& operator()(int i, int j, int k) const {
        return v[(size_t)((i - b.lo[0] + 1) + (int64_t)m[0] * ((j - b.lo[1] + 1) + (int64_t)m[1] * (k - b.lo[2] + 1)))];
    }
};
template <class T>
using FineFld = std::vector<BoxFld<T>>;

struct FineGeom {
    int n[3];                      // fine domain
    double dxinv[3];
    std::vector<IBox> boxes;
    IBox region;                   // owner-map extent: bounding box of the boxes grown by 3
    std::vector<int> owner;        // fine box covering a cell of region, or -1
    std::vector<std::array<int, 6>> nxo;  // closure order NX per box face
    double coef[5][4];             // c/f Lagrange weights by NX (bcl = dx: node -1)
    // physical domain faces (a fine box may touch them): per face (dir +
    // 3*side) the boundary type, and the Dirichlet closure weights by NX with
    // the boundary value at the face (bcl = 0: node 0, MLMGBndry.H:124-139)
    int periodic[3] = {1, 1, 1};
    int bct[6] = {BC_INTERIOR, BC_INTERIOR, BC_INTERIOR, BC_INTERIOR, BC_INTERIOR, BC_INTERIOR};
    double pcoef[5][4];

    bool inside(int i, int j, int k) const {  // not beyond a non-periodic domain face
        const int p[3] = {i, j, k};
        for (int d = 0; d < 3; ++d)
            if (!periodic[d] && (p[d] < 0 || p[d] >= n[d])) return false;
        return true;
    }
    // periodic image of a fine position inside the fine domain
    void wrap(int& i, int& j, int& k) const {
        i = ((i % n[0]) + n[0]) % n[0];
        j = ((j % n[1]) + n[1]) % n[1];
        k = ((k % n[2]) + n[2]) % n[2];
    }
    // fine box covering a position (its periodic image), -1 when none does, -2
    // beyond a non-periodic domain face
    int owner_at(int i, int j, int k) const {
        if (!inside(i, j, k)) return -2;
        wrap(i, j, k);
        if (i < region.lo[0] || i > region.hi[0] || j < region.lo[1] || j > region.hi[1] || k < region.lo[2] ||
            k > region.hi[2])
            return -1;
        const int64_t mx = region.len(0), my = region.len(1);
        return owner[(size_t)((i - region.lo[0]) + mx * ((j - region.lo[1]) + my * (int64_t)(k - region.lo[2])))];
    }
    // BndryData mask of a fine position: 0 covered by a fine box, 1 not covered,
    // 2 outside the domain
    int mask(int i, int j, int k) const {
        const int o = owner_at(i, j, k);
        return o >= 0 ? 0 : (o == -1 ? 1 : 2);
    }
    // undrrelxr of a box face cell whose ghost lies beyond a physical face
    // (AMReX_MLLinOp_K.H:329-406), closure order NX
    double phys_cf(int f, int NX) const {
        switch (bct[f]) {
            case BC_DIRICHLET: return pcoef[NX][1];
            case BC_NEUMANN: return 1.0;
            case BC_REFLECT_ODD: return 1.0;
            default: return 0.0;
        }
    }

    // MLABecLaplacian on AMR level 1 (abec false: MLPoisson): a per box (valid
    // cells, x fastest) and b per box and direction (the box's faces, x fastest)
    bool abec = false;
    double alpha = 0, beta = 0;
    std::vector<std::vector<double>> a;
    std::vector<std::array<std::vector<double>, 3>> b;
    double av(size_t q, int i, int j, int k) const {
        const IBox& x = boxes[q];
        return a[q][(size_t)((i - x.lo[0]) + (int64_t)x.len(0) * ((j - x.lo[1]) + (int64_t)x.len(1) * (k - x.lo[2])))];
    }
    // b along d at face (i,j,k) of box q (face i lies between cells i-1 and i along d)
    double bv(size_t q, int d, int i, int j, int k) const {
        const IBox& x = boxes[q];
        const int m0 = x.len(0) + (d == 0), m1 = x.len(1) + (d == 1);
        return b[q][(size_t)d][(size_t)((i - x.lo[0]) + (int64_t)m0 * ((j - x.lo[1]) + (int64_t)m1 * (k - x.lo[2])))];
    }
};

inline void build_fine_geom(FineGeom& G, const Params& P, const std::vector<IBox>& boxes) {
    const int* nc = P.n;
    const int maxorder = P.maxorder;
    for (int d = 0; d < 3; ++d) {
        G.n[d] = 2 * nc[d];
        G.dxinv[d] = (double)G.n[d];  // unit cube
        G.periodic[d] = P.periodic[d];
    }
    for (int f = 0; f < 6; ++f) {
        const int d = f % 3;
        G.bct[f] = P.periodic[d] ? BC_INTERIOR : (f < 3 ? P.bc_lo[d] : P.bc_hi[d]);
    }
    G.boxes = boxes;
    for (int d = 0; d < 3; ++d) { G.region.lo[d] = 1 << 30; G.region.hi[d] = -(1 << 30); }
    for (const IBox& b : boxes)
        for (int d = 0; d < 3; ++d) {  // owner positions are looked up wrapped into the domain
            G.region.lo[d] = std::min(G.region.lo[d], std::max(b.lo[d] - 3, 0));
            G.region.hi[d] = std::max(G.region.hi[d], std::min(b.hi[d] + 3, G.n[d] - 1));
        }
    if (boxes.empty())
        for (int d = 0; d < 3; ++d) { G.region.lo[d] = 0; G.region.hi[d] = -1; }
    const int64_t mx = std::max(0, G.region.len(0)), my = std::max(0, G.region.len(1)), mz = std::max(0, G.region.len(2));
    G.owner.assign((size_t)(mx * my * mz), -1);
    for (size_t q = 0; q < boxes.size(); ++q) {
        const IBox& b = boxes[q];
        for (int k = b.lo[2]; k <= b.hi[2]; ++k)
            for (int j = b.lo[1]; j <= b.hi[1]; ++j)
                for (int i = b.lo[0]; i <= b.hi[0]; ++i)
                    G.owner[(size_t)((i - G.region.lo[0]) + mx * ((j - G.region.lo[1]) + my * (int64_t)(k - G.region.lo[2])))] = (int)q;
    }
    G.nxo.resize(boxes.size());
    for (size_t q = 0; q < boxes.size(); ++q)
        for (int f = 0; f < 6; ++f) G.nxo[q][f] = std::min(boxes[q].len(f % 3) + 1, maxorder);
    for (int m = 0; m < 5; ++m) for (int c = 0; c < 4; ++c) G.coef[m][c] = 0.0;
    for (int NX = 1; NX <= 4; ++NX) {
        // bcl = 0.5 * ratio * dx = dx at a coarse/fine face (MLMGBndry.H:145)
        const double x[4] = {-1.0, 0.5, 1.5, 2.5};
        poly_interp_coeff(-0.5, x, NX, G.coef[NX]);
        const double xp[4] = {-0.0, 0.5, 1.5, 2.5};
        poly_interp_coeff(-0.5, xp, NX, G.pcoef[NX]);
    }
}

template <class T>
void fine_define(FineFld<T>& f, const FineGeom& G) {
    f.resize(G.boxes.size());
    for (size_t q = 0; q < G.boxes.size(); ++q) f[q].define(G.boxes[q]);
}

// Visit the face ghost cells of box q: fn(face, i, j, k) with (i,j,k) the ghost.
template <class F>
void each_face_ghost(const IBox& b, F fn) {
    for (int f = 0; f < 6; ++f) {
        const int d = f % 3, side = f / 3;
        const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
        for (int w = b.lo[t2]; w <= b.hi[t2]; ++w)
            for (int u = b.lo[t1]; u <= b.hi[t1]; ++u) {
                int p[3];
                p[d] = side == 0 ? b.lo[d] - 1 : b.hi[d] + 1;
                p[t1] = u;
                p[t2] = w;
                fn(f, p[0], p[1], p[2]);
            }
    }
}

// Ghost refresh of the fine level: face ghosts covered by another fine box get
// its value (FillBoundary, cross stencil); the others (mask 1) the Dirichlet
// closure through the coarse/fine boundary value bcv (null: homogeneous);
// ghosts beyond a physical face (mask 2) the face's closure (mllinop_apply_bc_*)
// with the Dirichlet value in bcv.
template <class T>
void fine_fill(const FineGeom& G, FineFld<T>& phi, const FineFld<T>* bcv) {
    for (size_t q = 0; q < G.boxes.size(); ++q) {
        BoxFld<T>& P = phi[q];
        each_face_ghost(G.boxes[q], [&](int f, int i, int j, int k) {
            const int o = G.owner_at(i, j, k);
            if (o >= 0) {
                int wi = i, wj = j, wk = k;
                G.wrap(wi, wj, wk);
                P(i, j, k) = phi[(size_t)o](wi, wj, wk);
                return;
            }
            const int d = f % 3, s = f < 3 ? 1 : -1;
            const int NX = G.nxo[q][f];
            int c1[3] = {i, j, k};
            c1[d] += s;
            if (o == -2 && G.bct[f] == BC_NEUMANN) { P(i, j, k) = P(c1[0], c1[1], c1[2]); return; }
            if (o == -2 && G.bct[f] == BC_REFLECT_ODD) { P(i, j, k) = -P(c1[0], c1[1], c1[2]); return; }
            const double* w = o == -2 ? G.pcoef[NX] : G.coef[NX];
            T tmp = T(0.0);
            for (int m = 1; m < NX; ++m) {
                int c[3] = {i, j, k};
                c[d] += m * s;
                tmp += P(c[0], c[1], c[2]) * T(w[m]);
            }
            if (bcv) tmp += (*bcv)[q](i, j, k) * T(w[0]);
            P(i, j, k) = tmp;
        });
    }
}

// One colour of mlpoisson_gsrb (abec_gsrb for an ABecLaplacian) on the fine
// level; undrrelxr = c_1 where the face ghost is not covered (mask > 0), 0
// where it is.
template <class T>
void fine_gsrb(const FineGeom& G, FineFld<T>& phi, const FineFld<T>& rhs, int rb) {
    const T dh[3] = {T(G.dxinv[0] * G.dxinv[0]), T(G.dxinv[1] * G.dxinv[1]), T(G.dxinv[2] * G.dxinv[2])};
    // b_scalar / (h*h), the smoother's scale (MLABecLaplacian Fsmooth)
    const T dhs[3] = {T(G.beta) / T((1.0 / G.dxinv[0]) * (1.0 / G.dxinv[0])),
                      T(G.beta) / T((1.0 / G.dxinv[1]) * (1.0 / G.dxinv[1])),
                      T(G.beta) / T((1.0 / G.dxinv[2]) * (1.0 / G.dxinv[2]))};
    constexpr T omega = T(1.15);
    for (size_t q = 0; q < G.boxes.size(); ++q) {
        const IBox& b = G.boxes[q];
        BoxFld<T>& P = phi[q];
        const BoxFld<T>& R = rhs[q];
        for (int k = b.lo[2]; k <= b.hi[2]; ++k)
            for (int j = b.lo[1]; j <= b.hi[1]; ++j)
                for (int i = b.lo[0]; i <= b.hi[0]; ++i) {
                    if (((i + j + k + rb) & 1) != 0) continue;
                    const int p[3] = {i, j, k};
                    T cf[6];
                    for (int f = 0; f < 6; ++f) {
                        const int d = f % 3;
                        const bool at = f < 3 ? p[d] == b.lo[d] : p[d] == b.hi[d];
                        int g[3] = {i, j, k};
                        g[d] += f < 3 ? -1 : 1;
                        const int mk = at ? G.mask(g[0], g[1], g[2]) : 0;
                        cf[f] = mk == 1 ? T(G.coef[G.nxo[q][f]][1])
                                        : (mk == 2 ? T(G.phys_cf(f, G.nxo[q][f])) : T(0.0));
                    }
                    if (G.abec) {
                        // abec_gsrb (AMReX_MLABecLap_3D_K.H:210-264) with the box's coefficients
                        const T bxm = T(G.bv(q, 0, i, j, k)), bxp = T(G.bv(q, 0, i + 1, j, k));
                        const T bym = T(G.bv(q, 1, i, j, k)), byp = T(G.bv(q, 1, i, j + 1, k));
                        const T bzm = T(G.bv(q, 2, i, j, k)), bzp = T(G.bv(q, 2, i, j, k + 1));
                        const T gamma = T(G.alpha) * T(G.av(q, i, j, k)) + dhs[0] * (bxm + bxp) + dhs[1] * (bym + byp) +
                                        dhs[2] * (bzm + bzp);
                        const T g_m_d = gamma - (dhs[0] * (bxm * cf[0] + bxp * cf[3]) + dhs[1] * (bym * cf[1] + byp * cf[4]) +
                                                 dhs[2] * (bzm * cf[2] + bzp * cf[5]));
                        const T rho = dhs[0] * (bxm * P(i - 1, j, k) + bxp * P(i + 1, j, k)) +
                                      dhs[1] * (bym * P(i, j - 1, k) + byp * P(i, j + 1, k)) +
                                      dhs[2] * (bzm * P(i, j, k - 1) + bzp * P(i, j, k + 1));
                        const T res = R(i, j, k) - (gamma * P(i, j, k) - rho);
                        P(i, j, k) = P(i, j, k) + omega / g_m_d * res;
                        continue;
                    }
                    const T gamma = T(-2.) * (dh[0] + dh[1] + dh[2]);
                    const T g_m_d = gamma + dh[0] * (cf[0] + cf[3]) + dh[1] * (cf[1] + cf[4]) + dh[2] * (cf[2] + cf[5]);
                    const T res = R(i, j, k) - gamma * P(i, j, k) - dh[0] * (P(i - 1, j, k) + P(i + 1, j, k))
                                - dh[1] * (P(i, j - 1, k) + P(i, j + 1, k)) - dh[2] * (P(i, j, k - 1) + P(i, j, k + 1));
                    P(i, j, k) = P(i, j, k) + omega / g_m_d * res;
                }
    }
}

// r = b - L(x) on the fine level, ghosts of x current (mlpoisson_adotx or
// mlabeclap_adotx + Xpay).
template <class T>
void fine_residual(const FineGeom& G, FineFld<T>& r, const FineFld<T>& x, const FineFld<T>& b) {
    const T dh[3] = {T(G.dxinv[0] * G.dxinv[0]), T(G.dxinv[1] * G.dxinv[1]), T(G.dxinv[2] * G.dxinv[2])};
    const T dha[3] = {T(G.beta) * T(G.dxinv[0]) * T(G.dxinv[0]), T(G.beta) * T(G.dxinv[1]) * T(G.dxinv[1]),
                      T(G.beta) * T(G.dxinv[2]) * T(G.dxinv[2])};
    for (size_t q = 0; q < G.boxes.size(); ++q) {
        const IBox& bx = G.boxes[q];
        const BoxFld<T>& X = x[q];
        for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
            for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
                for (int i = bx.lo[0]; i <= bx.hi[0]; ++i) {
                    const T xc = X(i, j, k);
                    T y;
                    if (G.abec) {
                        // mlabeclap_adotx (AMReX_MLABecLap_3D_K.H:9-28)
                        y = T(G.alpha) * T(G.av(q, i, j, k)) * xc
                          - dha[0] * (T(G.bv(q, 0, i + 1, j, k)) * (X(i + 1, j, k) - xc) - T(G.bv(q, 0, i, j, k)) * (xc - X(i - 1, j, k)))
                          - dha[1] * (T(G.bv(q, 1, i, j + 1, k)) * (X(i, j + 1, k) - xc) - T(G.bv(q, 1, i, j, k)) * (xc - X(i, j - 1, k)))
                          - dha[2] * (T(G.bv(q, 2, i, j, k + 1)) * (X(i, j, k + 1) - xc) - T(G.bv(q, 2, i, j, k)) * (xc - X(i, j, k - 1)));
                    } else {
                        y = dh[0] * (X(i - 1, j, k) - T(2.0) * xc + X(i + 1, j, k))
                          + dh[1] * (X(i, j - 1, k) - T(2.0) * xc + X(i, j + 1, k))
                          + dh[2] * (X(i, j, k - 1) - T(2.0) * xc + X(i, j, k + 1));
                    }
                    r[q](i, j, k) = b[q](i, j, k) + T(-1.0) * y;
                }
    }
}

// Coarse value with periodic wrap (the BndryRegister / ParallelCopy of coarse
// data around the fine boxes uses the coarse periodicity).
template <class T>
inline T cwrap(const Fld<T>& c, int i, int j, int k) {
    const int* n = c.n;
    i = ((i % n[0]) + n[0]) % n[0];
    j = ((j % n[1]) + n[1]) % n[1];
    k = ((k % n[2]) + n[2]) % n[2];
    return c(i, j, k);
}
inline int coarsen2(int i) { return i >= 0 ? i / 2 : -((-i + 1) / 2); }
// Coarse value as the correction interpolation and the coarse/fine boundary
// interpolation see it: the periodic image, 0 beyond a non-periodic face
// (interpCorrection copies valid cells and periodic images into a zeroed
// array, AMReX_MLMG.H:1719-1750).
template <class T>
inline T cval0(const FineGeom& G, const Fld<T>& c, int i, int j, int k) {
    const int p[3] = {i, j, k};
    for (int d = 0; d < 3; ++d)
        if (!G.periodic[d] && (p[d] < 0 || p[d] >= c.n[d])) return T(0);
    return cwrap(c, i, j, k);
}

// Coarse/fine boundary values of every face ghost of the fine level from the
// coarse field (interpbndrydata_{x,y,z}_o3 with the fine-level masks).
template <class T>
void fine_bndry_values(const FineGeom& G, const Fld<T>& crse, FineFld<T>& bcv, const FineFld<T>* phys = nullptr) {
    for (size_t q = 0; q < G.boxes.size(); ++q) {
        each_face_ghost(G.boxes[q], [&](int f, int i, int j, int k) {
            const int o = G.owner_at(i, j, k);
            if (o >= 0) { bcv[q](i, j, k) = T(0); return; }
            if (o == -2) { bcv[q](i, j, k) = phys ? (*phys)[q](i, j, k) : T(0); return; }
            const int d = f % 3;
            const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;  // tangential dims in the kernel's order
            const int p[3] = {i, j, k};
            const int cc[3] = {coarsen2(i), coarsen2(j), coarsen2(k)};
            auto C = [&](int a1, int a2) {  // coarse value at tangential offsets (a1 along t1, a2 along t2)
                int c[3] = {cc[0], cc[1], cc[2]};
                c[t1] += a1;
                c[t2] += a2;
                return cval0(G, crse, c[0], c[1], c[2]);
            };
            auto M = [&](int o1, int o2) {  // mask at fine tangential offsets
                int c[3] = {p[0], p[1], p[2]};
                c[t1] += o1;
                c[t2] += o2;
                return G.mask(c[0], c[1], c[2]);
            };
            const T c0 = C(0, 0);
            // direction t1
            int lo = (M(-2, 0) == 1) ? -1 : 0, hi = (M(2, 0) == 1) ? 1 : 0;
            T fac = (hi == lo + 1) ? T(1.0) : T(0.5);
            const T d1 = fac * (C(hi, 0) - C(lo, 0));
            const T d11 = (hi == lo + 2) ? T(0.5) * (C(1, 0) - T(2.) * c0 + C(-1, 0)) : T(0.);
            // direction t2
            lo = (M(0, -2) == 1) ? -1 : 0;
            hi = (M(0, 2) == 1) ? 1 : 0;
            fac = (hi == lo + 1) ? T(1.0) : T(0.5);
            const T d2 = fac * (C(0, hi) - C(0, lo));
            const T d22 = (hi == lo + 2) ? T(0.5) * (C(0, 1) - T(2.) * c0 + C(0, -1)) : T(0.);
            const T d12 = (M(-2, -2) == 1 && M(2, -2) == 1 && M(-2, 2) == 1 && M(2, 2) == 1)
                              ? T(0.25) * (C(1, 1) - C(-1, 1) + C(-1, -1) - C(1, -1))
                              : T(0.0);
            const T y1 = -T(0.5) + (p[t1] - cc[t1] * 2 + T(0.5)) / 2;
            const T y2 = -T(0.5) + (p[t2] - cc[t2] * 2 + T(0.5)) / 2;
            bcv[q](i, j, k) = c0 + y1 * d1 + (y1 * y1) * d11 + y2 * d2 + (y2 * y2) * d22 + y1 * y2 * d12;
        });
    }
}

// crse (covered cells) = mean of the 8 fine children (average_down, ratio 2).
template <class T>
void fine_avgdown(const FineGeom& G, Fld<T>& crse, const FineFld<T>& fine) {
    for (size_t q = 0; q < G.boxes.size(); ++q) {
        const IBox& b = G.boxes[q];
        for (int K = b.lo[2] / 2; K <= b.hi[2] / 2; ++K)
            for (int J = b.lo[1] / 2; J <= b.hi[1] / 2; ++J)
                for (int I = b.lo[0] / 2; I <= b.hi[0] / 2; ++I) {
                    T c = 0;
                    for (int kr = 0; kr < 2; ++kr)
                        for (int jr = 0; jr < 2; ++jr)
                            for (int ir = 0; ir < 2; ++ir) c += fine[q](2 * I + ir, 2 * J + jr, 2 * K + kr);
                    crse(I, J, K) = (T(1.0) / T(8)) * c;
                }
    }
}

// fine = trilinear interpolation of crse (mlmg_lin_cc_interp_r2), assigned.
template <class T>
void fine_interp(const FineGeom& G, FineFld<T>& fine, const Fld<T>& crse) {
    for (size_t q = 0; q < G.boxes.size(); ++q) {
        const IBox& b = G.boxes[q];
        for (int k = b.lo[2]; k <= b.hi[2]; ++k) {
            const int kc = k / 2, ko = 2 * (k - kc * 2) - 1;
            for (int j = b.lo[1]; j <= b.hi[1]; ++j) {
                const int jc = j / 2, jo = 2 * (j - jc * 2) - 1;
                for (int i = b.lo[0]; i <= b.hi[0]; ++i) {
                    const int ic = i / 2, io = 2 * (i - ic * 2) - 1;
                    fine[q](i, j, k) = T(0.421875) * cval0(G, crse, ic, jc, kc)
                        + T(0.140625) * cval0(G, crse, ic + io, jc, kc)
                        + T(0.140625) * cval0(G, crse, ic, jc + jo, kc)
                        + T(0.140625) * cval0(G, crse, ic, jc, kc + ko)
                        + T(0.046875) * cval0(G, crse, ic, jc + jo, kc + ko)
                        + T(0.046875) * cval0(G, crse, ic + io, jc, kc + ko)
                        + T(0.046875) * cval0(G, crse, ic + io, jc + jo, kc)
                        + T(0.015625) * cval0(G, crse, ic + io, jc + jo, kc + ko);
                }
            }
        }
    }
}

// Coarse cells of the coarse/fine interface: fine_cell (covered), crse_fine
// boundary (within one cell, diagonals included, of a covered cell), else 0.
enum { FR_CRSE = 0, FR_BOUNDARY = 1, FR_FINE = 2 };
inline std::vector<signed char> coarse_flags(const FineGeom& G, const int* nc) {
    std::vector<signed char> fl((size_t)nc[0] * nc[1] * nc[2], FR_CRSE);
    signed char none = FR_CRSE;
    auto at = [&](int i, int j, int k) -> signed char& {
        const int p[3] = {i, j, k};
        for (int d = 0; d < 3; ++d)
            if (!G.periodic[d] && (p[d] < 0 || p[d] >= nc[d])) {
                none = FR_CRSE;  // beyond a physical face: no coarse cell
                return none;
            }
        i = ((i % nc[0]) + nc[0]) % nc[0];
        j = ((j % nc[1]) + nc[1]) % nc[1];
        k = ((k % nc[2]) + nc[2]) % nc[2];
        return fl[(size_t)(i + (int64_t)nc[0] * (j + (int64_t)nc[1] * k))];
    };
    for (const IBox& b : G.boxes)
        for (int K = b.lo[2] / 2 - 1; K <= b.hi[2] / 2 + 1; ++K)
            for (int J = b.lo[1] / 2 - 1; J <= b.hi[1] / 2 + 1; ++J)
                for (int I = b.lo[0] / 2 - 1; I <= b.hi[0] / 2 + 1; ++I)
                    if (at(I, J, K) == FR_CRSE) at(I, J, K) = FR_BOUNDARY;
    for (const IBox& b : G.boxes)
        for (int K = b.lo[2] / 2; K <= b.hi[2] / 2; ++K)
            for (int J = b.lo[1] / 2; J <= b.hi[1] / 2; ++J)
                for (int I = b.lo[0] / 2; I <= b.hi[0] / 2; ++I) at(I, J, K) = FR_FINE;
    return fl;
}

// Reflux of the coarse residual: at a coarse cell next to the fine region the
// coarse flux through each coarse/fine face is replaced by the mean of the four
// fine fluxes (Poisson flux dxinv*(phi_i - phi_{i-1}), MLPoisson FFlux).
// res += (F_crse - <F_fine>)/dx_c on the low side of the fine region and
// -(F_crse - <F_fine>)/dx_c on its high side.  crse and fine need current ghosts.
// For an ABecLaplacian both fluxes are FFlux's -b_scalar*dxinv*b_face*diff
// (mlabeclap_flux_*, AMReX_MLABecLap_3D_K.H:79-96) with the coarse level's
// face coefficients (L0) and the fine box's.
template <class T>
void reflux(const FineGeom& G, const int* nc, const std::vector<signed char>& fl, Fld<T>& res,
            const Fld<T>& crse, const FineFld<T>& fine, const Level<T>* L0 = nullptr) {
    const T dxc[3] = {T(nc[0]), T(nc[1]), T(nc[2])};  // coarse dxinv
    const T fscale = G.abec ? T(-G.beta) : T(1);
    auto flag = [&](int i, int j, int k) -> signed char {
        const int p[3] = {i, j, k};
        for (int d = 0; d < 3; ++d)
            if (!G.periodic[d] && (p[d] < 0 || p[d] >= nc[d])) return FR_CRSE;
        i = ((i % nc[0]) + nc[0]) % nc[0];
        j = ((j % nc[1]) + nc[1]) % nc[1];
        k = ((k % nc[2]) + nc[2]) % nc[2];
        return fl[(size_t)(i + (int64_t)nc[0] * (j + (int64_t)nc[1] * k))];
    };
    for (int k = 0; k < nc[2]; ++k)
        for (int j = 0; j < nc[1]; ++j)
            for (int i = 0; i < nc[0]; ++i) {
                if (flag(i, j, k) != FR_BOUNDARY) continue;
                const int c[3] = {i, j, k};
                T d = 0;
                for (int dir = 0; dir < 3; ++dir) {
                    const int t1 = (dir == 0) ? 1 : 0, t2 = (dir == 2) ? 1 : 2;
                    for (int side = 0; side < 2; ++side) {  // side 1: fine region above (c + e_dir)
                        int nb[3] = {c[0], c[1], c[2]};
                        nb[dir] += side ? 1 : -1;
                        if (flag(nb[0], nb[1], nb[2]) != FR_FINE) continue;
                        // coarse flux through the face between c and nb
                        int hi[3] = {c[0], c[1], c[2]};
                        if (side) hi[dir] += 1;
                        int lo[3] = {hi[0], hi[1], hi[2]};
                        lo[dir] -= 1;
                        const T bc = G.abec ? L0->b[dir](hi[0], hi[1], hi[2]) : T(1);
                        const T fc = fscale * dxc[dir] * bc * (crse(hi[0], hi[1], hi[2]) - crse(lo[0], lo[1], lo[2]));
                        // the four fine faces of that coarse face (fine index 2*hi along dir)
                        const int fi = 2 * hi[dir];
                        const int cf_cell = side ? fi : fi - 1;  // fine cell inside the fine region
                        int pfine[3];
                        pfine[dir] = cf_cell;
                        pfine[t1] = 2 * c[t1];
                        pfine[t2] = 2 * c[t2];
                        const int q = G.owner_at(pfine[0], pfine[1], pfine[2]);
                        // the coarse/fine face in the frame of box q: its low face when the fine
                        // region is above c, else its high face (2*hi[dir] unless the face is
                        // the periodic image across a domain face)
                        const int fq = side ? G.boxes[(size_t)q].lo[dir] : G.boxes[(size_t)q].hi[dir] + 1;
                        T fsum = 0;
                        for (int a2 = 0; a2 < 2; ++a2)
                            for (int a1 = 0; a1 < 2; ++a1) {
                                int h2[3], l2[3];
                                h2[dir] = fq;
                                l2[dir] = fq - 1;
                                h2[t1] = l2[t1] = 2 * c[t1] + a1;
                                h2[t2] = l2[t2] = 2 * c[t2] + a2;
                                const T bf = G.abec ? T(G.bv((size_t)q, dir, h2[0], h2[1], h2[2])) : T(1);
                                fsum += fscale * T(G.dxinv[dir]) * bf *
                                        (fine[(size_t)q](h2[0], h2[1], h2[2]) - fine[(size_t)q](l2[0], l2[1], l2[2]));
                            }
                        // CrseAdd adds +-dtdx_c*F_crse, FineAdd -+dt/(dx_f*8)*sum F_fine
                        // (AMReX_YAFluxRegister_3D_K.H), dt = 1
                        const T dtdx_f = T(G.dxinv[dir]) / T(8);
                        if (side) d += dxc[dir] * fc - dtdx_f * fsum;
                        else d += -dxc[dir] * fc + dtdx_f * fsum;
                    }
                }
                res(i, j, k) += d;
            }
}


// Fine arrays of the C interface: the boxes' valid cells, x fastest, box after box.
template <class T>
void fine_load(FineFld<T>& f, const double* src) {
    size_t off = 0;
    for (BoxFld<T>& B : f) {
        for (int k = B.b.lo[2]; k <= B.b.hi[2]; ++k)
            for (int j = B.b.lo[1]; j <= B.b.hi[1]; ++j)
                for (int i = B.b.lo[0]; i <= B.b.hi[0]; ++i) B(i, j, k) = T(src[off++]);
    }
}
template <class T>
void fine_store(const FineFld<T>& f, double* dst) {
    size_t off = 0;
    for (const BoxFld<T>& B : f) {
        for (int k = B.b.lo[2]; k <= B.b.hi[2]; ++k)
            for (int j = B.b.lo[1]; j <= B.b.hi[1]; ++j)
                for (int i = B.b.lo[0]; i <= B.b.hi[0]; ++i) dst[off++] = (double)B(i, j, k);
    }
}
template <class T, class F>
void fine_each(FineFld<T>& f, F fn) {
    for (size_t q = 0; q < f.size(); ++q) {
        BoxFld<T>& B = f[q];
        for (int k = B.b.lo[2]; k <= B.b.hi[2]; ++k)
            for (int j = B.b.lo[1]; j <= B.b.hi[1]; ++j)
                for (int i = B.b.lo[0]; i <= B.b.hi[0]; ++i) fn(q, i, j, k);
    }
}
template <class T>
T fine_norminf(const FineFld<T>& f) {
    T m = 0;
    for (const BoxFld<T>& B : f)
        for (int k = B.b.lo[2]; k <= B.b.hi[2]; ++k)
            for (int j = B.b.lo[1]; j <= B.b.hi[1]; ++j)
                for (int i = B.b.lo[0]; i <= B.b.hi[0]; ++i) m = std::max(m, std::abs(B(i, j, k)));
    return m;
}
// max |x| over coarse cells not covered by the fine level (m_norm_fine_mask)
template <class T>
T masked_norminf(const Fld<T>& x, const std::vector<signed char>& fl) {
    T m = 0;
    const int* n = x.n;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i)
                if (fl[(size_t)(i + (int64_t)n[0] * (j + (int64_t)n[1] * k))] != FR_FINE) m = std::max(m, std::abs(x(i, j, k)));
    return m;
}

// Coefficients of the next two-level ABecLaplacian call (orc_amr_set_coeffs):
// the coarse level's a and face b (global arrays), the fine boxes' a (valid
// cells, box after box) and b (per direction, per box, the box's faces).
struct AmrCoefIn {
    const double *a0 = nullptr, *bx0 = nullptr, *by0 = nullptr, *bz0 = nullptr, *a1 = nullptr, *b1 = nullptr;
};
inline AmrCoefIn& amr_coef_in() {
    static AmrCoefIn c;
    return c;
}
// The fine level's boundary data of the next two-level call (orc_amr_set_fine_bc,
// setLevelBC(1, data)): a ghost-grown (2n+2)^3 array over the fine domain whose
// face ghosts hold the Dirichlet values of the physical faces; null: homogeneous.
inline const double*& amr_fine_bc_in() {
    static const double* g = nullptr;
    return g;
}

struct AmrState {
    Params P0;
    Operator<double> op;
    // coarse coefficients after averageDownCoeffsToCoarseAmrLevel (ABecLaplacian)
    std::vector<double> ca, cb[3];
    FineGeom G;
    std::vector<signed char> fl;
    Fld<double> sol0, rhs0, res0, cor0;
    FineFld<double> sol1, rhs1, res1, cor1, rescor1, bsol, bcor;
    // the coarse level's boundary data (setLevelBC(0, ...)): Dirichlet values in
    // the face ghosts; absent: homogeneous
    Fld<double> bcv0;
    bool has_bcv0 = false;
    const Fld<double>* bcv() const { return has_bcv0 ? &bcv0 : nullptr; }
    // the fine level's Dirichlet values beyond physical faces (in the fine boxes'
    // ghost cells there); absent: homogeneous
    FineFld<double> pbc;
    bool has_pbc = false;
    const FineFld<double>* phys() const { return has_pbc ? &pbc : nullptr; }
};

// MLABecLaplacian over two AMR levels: the fine boxes' coefficients, and the
// coarse level's with the fine ones averaged onto the cells and faces the
// fine level covers, averageDownCoeffsToCoarseAmrLevel (AMReX_MLABecLaplacian.H:
// 694-712): a as the mean of the 8 children when alpha != 0 (average_down), b
// on every coarse face of a coarsened fine box (its boundary faces included)
// as the mean of the 4 fine faces (average_down_faces), written to the
// periodic images of a face on a periodic domain face as well.
inline void amr_abec_coeffs(AmrState& S, const Params& P0) {
    const AmrCoefIn& C = amr_coef_in();
    if (!C.a0 || !C.bx0 || !C.by0 || !C.bz0 || !C.a1 || !C.b1)
        throw std::runtime_error("two-level ABecLaplacian: coefficients missing");
    FineGeom& G = S.G;
    const int* n = P0.n;
    G.abec = true;
    G.alpha = P0.alpha;
    G.beta = P0.beta;
    const size_t nb = G.boxes.size();
    G.a.assign(nb, {});
    G.b.assign(nb, {});
    size_t off = 0;
    for (size_t q = 0; q < nb; ++q) {
        const IBox& x = G.boxes[q];
        const size_t m = (size_t)x.len(0) * x.len(1) * x.len(2);
        G.a[q].assign(C.a1 + off, C.a1 + off + m);
        off += m;
    }
    off = 0;
    for (int d = 0; d < 3; ++d)
        for (size_t q = 0; q < nb; ++q) {
            const IBox& x = G.boxes[q];
            const size_t m = (size_t)(x.len(0) + (d == 0)) * (x.len(1) + (d == 1)) * (x.len(2) + (d == 2));
            G.b[q][(size_t)d].assign(C.b1 + off, C.b1 + off + m);
            off += m;
        }
    const size_t ncell = (size_t)n[0] * n[1] * n[2];
    S.ca.assign(C.a0, C.a0 + ncell);
    const double* bin[3] = {C.bx0, C.by0, C.bz0};
    size_t nf[3];
    for (int d = 0; d < 3; ++d) {
        nf[d] = (size_t)(n[0] + (d == 0)) * (n[1] + (d == 1)) * (n[2] + (d == 2));
        S.cb[d].assign(bin[d], bin[d] + nf[d]);
    }
    auto cidx = [&](int i, int j, int k) { return (size_t)(i + (int64_t)n[0] * (j + (int64_t)n[1] * k)); };
    auto fidx = [&](int d, int i, int j, int k) {
        return (size_t)(i + (int64_t)(n[0] + (d == 0)) * (j + (int64_t)(n[1] + (d == 1)) * k));
    };
    for (size_t q = 0; q < nb; ++q) {
        const IBox& x = G.boxes[q];
        int lo[3], hi[3];
        for (int d = 0; d < 3; ++d) { lo[d] = x.lo[d] / 2; hi[d] = x.hi[d] / 2; }
        if (P0.alpha != 0.0)
            for (int K = lo[2]; K <= hi[2]; ++K)
                for (int J = lo[1]; J <= hi[1]; ++J)
                    for (int I = lo[0]; I <= hi[0]; ++I) {
                        double v = 0;
                        for (int c = 0; c < 8; ++c) v += G.av(q, 2 * I + (c & 1), 2 * J + ((c >> 1) & 1), 2 * K + (c >> 2));
                        S.ca[cidx(I, J, K)] = v / 8.0;
                    }
        for (int d = 0; d < 3; ++d) {
            const int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
            const bool per = P0.periodic[d] != 0;
            int F[3];
            for (F[d] = lo[d]; F[d] <= hi[d] + 1; ++F[d])
                for (F[t2] = lo[t2]; F[t2] <= hi[t2]; ++F[t2])
                    for (F[t1] = lo[t1]; F[t1] <= hi[t1]; ++F[t1]) {
                        double v = 0;
                        for (int a2 = 0; a2 < 2; ++a2)
                            for (int a1 = 0; a1 < 2; ++a1) {
                                int f[3];
                                f[d] = 2 * F[d];
                                f[t1] = 2 * F[t1] + a1;
                                f[t2] = 2 * F[t2] + a2;
                                v += G.bv(q, d, f[0], f[1], f[2]);
                            }
                        v /= 4.0;
                        S.cb[d][fidx(d, F[0], F[1], F[2])] = v;
                        if (per && (F[d] == 0 || F[d] == n[d])) {
                            int g[3] = {F[0], F[1], F[2]};
                            g[d] = F[d] == 0 ? n[d] : 0;
                            S.cb[d][fidx(d, g[0], g[1], g[2])] = v;
                        }
                    }
        }
    }
}

inline void amr_setup(AmrState& S, const Params& P0, int nbox, const int* boxes, const double* bcv0 = nullptr) {
    S.P0 = P0;
    if (bcv0) {
        S.bcv0.define(P0.n);
        for (size_t p = 0; p < S.bcv0.v.size(); ++p) S.bcv0.v[p] = bcv0[p];
        S.has_bcv0 = true;
    }
    std::vector<IBox> bl((size_t)nbox);
    for (int q = 0; q < nbox; ++q)
        for (int d = 0; d < 3; ++d) { bl[(size_t)q].lo[d] = boxes[6 * q + d]; bl[(size_t)q].hi[d] = boxes[6 * q + 3 + d]; }
    build_fine_geom(S.G, P0, bl);
    if (P0.op == 0) amr_abec_coeffs(S, P0);
    build_operator(S.op, P0, S.ca.empty() ? nullptr : S.ca.data(), S.cb[0].empty() ? nullptr : S.cb[0].data(),
                   S.cb[1].empty() ? nullptr : S.cb[1].data(), S.cb[2].empty() ? nullptr : S.cb[2].data());
    S.fl = coarse_flags(S.G, P0.n);
    if (const double* g = amr_fine_bc_in()) {
        fine_define(S.pbc, S.G);
        const int* nf = S.G.n;
        for (size_t q = 0; q < S.G.boxes.size(); ++q)
            each_face_ghost(S.G.boxes[q], [&](int, int i, int j, int k) {
                if (S.G.owner_at(i, j, k) != -2) return;
                S.pbc[q](i, j, k) = g[(i + 1) + (int64_t)(nf[0] + 2) * ((j + 1) + (int64_t)(nf[1] + 2) * (k + 1))];
            });
        S.has_pbc = true;
    }
    S.sol0.define(P0.n); S.rhs0.define(P0.n); S.res0.define(P0.n); S.cor0.define(P0.n);
    for (FineFld<double>* f : {&S.sol1, &S.rhs1, &S.res1, &S.cor1, &S.rescor1, &S.bsol, &S.bcor}) fine_define(*f, S.G);
}

// Fine residual with the coarse/fine boundary from sol0 (computeResidual(1):
// updateSolBC + solutionResidual); leaves bsol and the ghosts of sol1 current.
inline void amr_fine_residual(AmrState& S) {
    fine_bndry_values(S.G, S.sol0, S.bsol, S.phys());
    fine_fill(S.G, S.sol1, &S.bsol);
    fine_residual(S.G, S.res1, S.sol1, S.rhs1);
}
// Coarse residual plus reflux against the fine solution (its ghosts from bsol).
inline void amr_coarse_residual(AmrState& S) {
    const Level<double>& L0 = S.op.lev[0];
    residual<double>(L0, S.P0, S.res0, S.sol0, S.rhs0, S.bcv());
    fine_fill(S.G, S.sol1, &S.bsol);
    reflux(S.G, S.P0.n, S.fl, S.res0, S.sol0, S.sol1, &S.op.lev[0]);
}

// nu smoothing sweeps of cor1 against res1, homogeneous coarse/fine boundary
// (the mixed form, V = float: the sweeps run in FP32 on an FP32 copy of res1,
// and cor1 takes the FP32 result; cor1 is zero on entry)
template <class V>
inline void amr_fine_smooth(AmrState& S, int nu) {
    if constexpr (std::is_same<V, double>::value) {
        for (int s = 0; s < nu; ++s)
            for (int rb = 0; rb < 2; ++rb) {
                fine_fill<double>(S.G, S.cor1, nullptr);
                fine_gsrb(S.G, S.cor1, S.res1, rb);
            }
    } else {
        FineFld<V> c, r;
        fine_define(c, S.G);
        fine_define(r, S.G);
        fine_each(r, [&](size_t q, int i, int j, int k) { r[q](i, j, k) = static_cast<V>(S.res1[q](i, j, k)); });
        for (int s = 0; s < nu; ++s)
            for (int rb = 0; rb < 2; ++rb) {
                fine_fill<V>(S.G, c, nullptr);
                fine_gsrb(S.G, c, r, rb);
            }
        fine_each(S.cor1, [&](size_t q, int i, int j, int k) { S.cor1[q](i, j, k) = static_cast<double>(c[q](i, j, k)); });
    }
}

template <class V>
int amr_solve(const Params& P0, int nbox, const int* boxes, const double* rhs0_in, const double* rhs1_in,
              double* sol0_io, double* sol1_io, Stats* st, const double* bcv0 = nullptr) {
    auto t0 = std::chrono::steady_clock::now();
    AmrState S;
    amr_setup(S, P0, nbox, boxes, bcv0);
    Operator<V> opv_store;
    const Operator<V>* opv;
    if constexpr (std::is_same<V, double>::value) {
        opv = &S.op;
    } else {
        build_operator(opv_store, P0, S.ca.empty() ? nullptr : S.ca.data(), S.cb[0].empty() ? nullptr : S.cb[0].data(),
                       S.cb[1].empty() ? nullptr : S.cb[1].data(), S.cb[2].empty() ? nullptr : S.cb[2].data());
        opv = &opv_store;
    }
    const int* n = P0.n;
    auto gidx = [&](int i, int j, int k) { return i + (int64_t)n[0] * (j + (int64_t)n[1] * k); };
    each_valid(S.sol0, [&](int i, int j, int k) { S.sol0(i, j, k) = sol0_io[gidx(i, j, k)]; });
    each_valid(S.rhs0, [&](int i, int j, int k) { S.rhs0(i, j, k) = rhs0_in[gidx(i, j, k)]; });
    fine_load(S.sol1, sol1_io);
    fine_load(S.rhs1, rhs1_in);
    // prepareForSolve: averageDownSolutionRHS, makeSolvable over both levels
    fine_avgdown(S.G, S.sol0, S.sol1);
    fine_avgdown(S.G, S.rhs0, S.rhs1);
    if (S.op.singular) {
        const double offset = sum(S.rhs0) * S.op.volinv0;
        each_valid(S.rhs0, [&](int i, int j, int k) { S.rhs0(i, j, k) += -offset; });
        fine_each(S.rhs1, [&](size_t q, int i, int j, int k) { S.rhs1[q](i, j, k) += -offset; });
    }
    Work<V> w;
    make_work(w, opv->lev);
    w.bottom_mode = bottom_mode_of(P0);
    st->nlevels = (int)S.op.lev.size();
    st->status = 0;
    st->bottom_iters_total = 0;
    // computeMLResidual(1) and the composite norms
    amr_fine_residual(S);
    amr_coarse_residual(S);
    const double resnorm0 = std::max(masked_norminf(S.res0, S.fl), fine_norminf(S.res1));
    const double rhsnorm0 = std::max(masked_norminf(S.rhs0, S.fl), fine_norminf(S.rhs1));
    st->init_resnorm = resnorm0;
    st->rhsnorm = rhsnorm0;
    const double max_norm = rhsnorm0 >= resnorm0 ? rhsnorm0 : resnorm0;
    st->max_norm = max_norm;
    const double target = std::max(P0.tol_abs, std::max(P0.tol_rel, 1.e-16) * max_norm);
    auto t1 = std::chrono::steady_clock::now();
    st->setup_seconds = std::chrono::duration<double>(t1 - t0).count();
    double composite = resnorm0;
    int iters = 0;
    if (resnorm0 > target) {
        bool converged = false;
        const int niters = P0.fixed_iters ? P0.fixed_iters : P0.max_iter;
        for (int iter = 0; iter < niters; ++iter) {
            // ---- oneIter, AMR level 1 down: miniCycle, sol += cor, composite residual
            for (BoxFld<double>& B : S.cor1) std::fill(B.v.begin(), B.v.end(), 0.0);
            amr_fine_smooth<V>(S, P0.nu1);
            fine_each(S.sol1, [&](size_t q, int i, int j, int k) { S.sol1[q](i, j, k) += S.cor1[q](i, j, k); });
            residual<double>(S.op.lev[0], P0, S.res0, S.sol0, S.rhs0, S.bcv());
            fine_fill<double>(S.G, S.cor1, nullptr);
            fine_residual(S.G, S.rescor1, S.cor1, S.res1);
            std::swap(S.res1, S.rescor1);
            fine_fill(S.G, S.sol1, &S.bsol);
            reflux(S.G, n, S.fl, S.res0, S.sol0, S.sol1, &S.op.lev[0]);
            fine_avgdown(S.G, S.res0, S.res1);
            // ---- coarsest AMR level: makeSolvable, V-cycle, sol += cor
            if (S.op.singular) make_solvable(S.res0, S.op.volinv0);
            each_valid(w.res[0], [&](int i, int j, int k) { w.res[0](i, j, k) = static_cast<V>(S.res0(i, j, k)); });
            if (iter < P0.max_fmg_iters) fcycle(*opv, S.op, w, S.op.singular, &st->bottom_iters_total);
            else vcycle(*opv, S.op, w, S.op.singular, &st->bottom_iters_total);
            each_valid(S.cor0, [&](int i, int j, int k) { S.cor0(i, j, k) = static_cast<double>(w.cor[0](i, j, k)); });
            each_valid(S.sol0, [&](int i, int j, int k) { S.sol0(i, j, k) += S.cor0(i, j, k); });
            // ---- AMR level 1 up: interpolated correction, residual with the
            // coarse correction as boundary data, miniCycle
            fine_interp(S.G, S.cor1, S.cor0);
            fine_each(S.sol1, [&](size_t q, int i, int j, int k) { S.sol1[q](i, j, k) += S.cor1[q](i, j, k); });
            fine_bndry_values(S.G, S.cor0, S.bcor);
            fine_fill(S.G, S.cor1, &S.bcor);
            fine_residual(S.G, S.rescor1, S.cor1, S.res1);
            std::swap(S.res1, S.rescor1);
            for (BoxFld<double>& B : S.cor1) std::fill(B.v.begin(), B.v.end(), 0.0);
            amr_fine_smooth<V>(S, P0.nu1);
            fine_each(S.sol1, [&](size_t q, int i, int j, int k) { S.sol1[q](i, j, k) += S.cor1[q](i, j, k); });
            // averageDownAndSync
            fine_avgdown(S.G, S.sol0, S.sol1);
            // ---- convergence: fine residual, then the coarse one when the fine passes
            amr_fine_residual(S);
            const double fine = fine_norminf(S.res1);
            iters = iter + 1;
            if (iter < 256) st->res_history[iter] = fine;
            composite = fine;
            if (fine <= target) {
                amr_coarse_residual(S);
                const double crse = masked_norminf(S.res0, S.fl);
                composite = std::max(fine, crse);
                if (crse <= target) { converged = true; break; }
            }
            if (P0.verbose >= 2)
                std::printf("MLMG(oracle amr): Iteration %3d Fine resid/max_norm = %.6e\n", iters, fine / max_norm);
            if (composite > 1.e20 * max_norm || std::isnan(composite)) { st->status = 2; break; }
        }
        if (!converged && !P0.fixed_iters && st->status == 0) st->status = 1;
    }
    st->iter_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
    st->niters = iters;
    st->bottom_solver_final = w.bottom_mode;
    st->nvcycles = iters;
    st->final_resnorm = composite;
    each_valid(S.sol0, [&](int i, int j, int k) { sol0_io[gidx(i, j, k)] = S.sol0(i, j, k); });
    fine_store(S.sol1, sol1_io);
    st->solve_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return st->status;
}

// Composite residual of given (sol0, sol1) against (rhs0, rhs1), no solvability
// fix: res1 with the coarse/fine boundary from sol0, res0 refluxed; covered
// coarse cells keep the coarse operator's residual.
inline void amr_residual(const Params& P0, int nbox, const int* boxes, const double* rhs0, const double* rhs1,
                         const double* sol0, const double* sol1, double* res0, double* res1,
                         const double* bcv0 = nullptr) {
    AmrState S;
    amr_setup(S, P0, nbox, boxes, bcv0);
    const int* n = P0.n;
    auto gidx = [&](int i, int j, int k) { return i + (int64_t)n[0] * (j + (int64_t)n[1] * k); };
    each_valid(S.sol0, [&](int i, int j, int k) { S.sol0(i, j, k) = sol0[gidx(i, j, k)]; });
    each_valid(S.rhs0, [&](int i, int j, int k) { S.rhs0(i, j, k) = rhs0[gidx(i, j, k)]; });
    fine_load(S.sol1, sol1);
    fine_load(S.rhs1, rhs1);
    amr_fine_residual(S);
    amr_coarse_residual(S);
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) res0[gidx(i, j, k)] = S.res0(i, j, k);
    fine_store(S.res1, res1);
}

// Face gradients (mode 0) or fluxes (mode 1) of both levels of the two-level
// Poisson hierarchy, MLMG::getGradSolution / getFluxes (AMReX_MLMG.H:556-660):
// the coarse level as a single level with its own (periodic) boundary; each
// fine box from its own ghost-filled FAB, coarse/fine ghosts by the Dirichlet
// closure through the boundary values interpolated from sol0 (compGrad's
// applyBC with m_bndry_sol after updateSolBC, AMReX_MLCellLinOp.H:1392-1437).
// Fine face value dxinv * diff, times 1/b_scalar = -1 for the flux
// (AMReX_MLCellABecLap.H:272-289).  out0 = [x, y, z] coarse face arrays;
// out1 = per direction, per box, the box's faces (x fastest).
inline void amr_faces(const Params& P0, int nbox, const int* boxes, const double* sol0, const double* sol1, int mode,
                      double* out0, double* out1, const double* bcv0 = nullptr) {
    AmrState S;
    amr_setup(S, P0, nbox, boxes, bcv0);
    const int* n = P0.n;
    each_valid(S.sol0, [&](int i, int j, int k) { S.sol0(i, j, k) = sol0[i + (int64_t)n[0] * (j + (int64_t)n[1] * k)]; });
    fine_load(S.sol1, sol1);
    const Level<double>& L0 = S.op.lev[0];
    fill_ghosts<double>(L0, S.sol0, S.bcv());
    FaceArr<double> fv[3];
    for (int d = 0; d < 3; ++d) fv[d].define(L0.n, d);
    face_values(L0, S.P0, S.sol0, mode, fv);
    size_t off = 0;
    for (int d = 0; d < 3; ++d)
        for (double v : fv[d].v) out0[off++] = v;
    fine_bndry_values(S.G, S.sol0, S.bsol, S.phys());
    fine_fill(S.G, S.sol1, &S.bsol);
    off = 0;
    for (int d = 0; d < 3; ++d)
        for (size_t q = 0; q < S.G.boxes.size(); ++q) {
            const IBox& b = S.G.boxes[q];
            const BoxFld<double>& F = S.sol1[q];
            for (int k = b.lo[2]; k <= b.hi[2] + (d == 2); ++k)
                for (int j = b.lo[1]; j <= b.hi[1] + (d == 1); ++j)
                    for (int i = b.lo[0]; i <= b.hi[0] + (d == 0); ++i) {
                        double v = S.G.dxinv[d] * (F(i, j, k) - F(i - (d == 0), j - (d == 1), k - (d == 2)));
                        if (mode == 1) v *= S.G.abec ? -S.G.bv(q, d, i, j, k) : -1.0;
                        out1[off++] = v;
                    }
        }
}

// nsweeps of the composite cycle's fine smoothing (homogeneous coarse/fine
// boundary) of cor1 against res1, from cor1 as given.
inline void amr_fine_sweeps(const Params& P0, int nbox, const int* boxes, const double* res1, double* cor1,
                            int nsweeps) {
    AmrState S;
    amr_setup(S, P0, nbox, boxes);
    fine_load(S.res1, res1);
    fine_load(S.cor1, cor1);
    amr_fine_smooth<double>(S, nsweeps);
    fine_store(S.cor1, cor1);
}

// fine = lin_cc_interp_r2(crse)
inline void amr_interp(const Params& P0, int nbox, const int* boxes, const double* crse, double* fine) {
    AmrState S;
    amr_setup(S, P0, nbox, boxes);
    const int* n = P0.n;
    each_valid(S.cor0, [&](int i, int j, int k) { S.cor0(i, j, k) = crse[i + (int64_t)n[0] * (j + (int64_t)n[1] * k)]; });
    fine_interp(S.G, S.cor1, S.cor0);
    fine_store(S.cor1, fine);
}

}  // namespace orc

extern "C" {

typedef orc::Params orc_params;
typedef orc::Stats orc_stats;

int orc_params_size(void) { return (int)sizeof(orc::Params); }
int orc_stats_size(void) { return (int)sizeof(orc::Stats); }

// Full single-level MLMG solve.  Arrays use the global layouts of the
// workload generator; bcv (nullable) is ghost-grown (n+2)^3 with Dirichlet
// values in its face ghosts; sol is (n)^3, initial guess in, solution out.
int orc_solve(const orc_params* P, const double* a, const double* bx, const double* by,
              const double* bz, const double* rhs, const double* bcv, double* sol, orc_stats* st) {
    std::memset(st, 0, sizeof(*st));
    if (P->bottom_solver < 0 || P->bottom_solver > 5) return -1;
    switch (P->precision) {
        case 0: return orc::solve<double, double>(*P, a, bx, by, bz, rhs, bcv, sol, st);
        case 1: return orc::solve<float, float>(*P, a, bx, by, bz, rhs, bcv, sol, st);
        case 2: return orc::solve<double, float>(*P, a, bx, by, bz, rhs, bcv, sol, st);
        default: return -1;
    }
}

// Single operations on MG level `level` (see level_op for `which`).
int orc_level_op(const orc_params* P, int which, const double* a, const double* bx,
                 const double* by, const double* bz, const double* bcv, const double* in1,
                 double* io, double* out, int count, int level) {
    if (P->precision == 1)
        return orc::level_op<float>(*P, which, a, bx, by, bz, bcv, in1, io, out, count, level);
    return orc::level_op<double>(*P, which, a, bx, by, bz, bcv, in1, io, out, count, level);
}

// MG hierarchy: fills nlev, dims[3*l], nbox[l] and boxes (6 ints each, all
// levels concatenated) up to max_boxes.
int orc_hierarchy(const orc_params* P, int* nlev, int* dims, int* nbox, int* boxes, int max_boxes) {
    std::vector<orc::LevelGeom> G = orc::define_grids(*P);
    *nlev = (int)G.size();
    int off = 0;
    for (size_t l = 0; l < G.size(); ++l) {
        for (int d = 0; d < 3; ++d) dims[3 * l + d] = G[l].n[d];
        nbox[l] = (int)G[l].ba.size();
        for (const orc::IBox& b : G[l].ba) {
            if (off < max_boxes) {
                for (int d = 0; d < 3; ++d) { boxes[6 * off + d] = b.lo[d]; boxes[6 * off + 3 + d] = b.hi[d]; }
            }
            ++off;
        }
    }
    return off;
}

// Two-level AMR composite Poisson solve (coarse level: P, periodic; fine boxes
// of ratio 2: 6 ints each).  Fine arrays: valid cells of each box, x fastest,
// box after box.  precision 0: FP64; 2: FP32 coarse V-cycle.  bcv0 (nullable):
// the coarse level's boundary data, ghost-grown (n+2)^3 as for orc_solve.
int orc_amr_solve(const orc_params* P, int nbox, const int* boxes, const double* rhs0, const double* rhs1,
                  double* sol0, double* sol1, orc_stats* st, const double* bcv0) {
    std::memset(st, 0, sizeof(*st));
    if (P->bottom_solver < 0 || P->bottom_solver > 5) return -1;
    try {
        if (P->precision == 2) return orc::amr_solve<float>(*P, nbox, boxes, rhs0, rhs1, sol0, sol1, st, bcv0);
        return orc::amr_solve<double>(*P, nbox, boxes, rhs0, rhs1, sol0, sol1, st, bcv0);
    } catch (const std::exception&) {
        return -3;
    }
}
int orc_amr_residual(const orc_params* P, int nbox, const int* boxes, const double* rhs0, const double* rhs1,
                     const double* sol0, const double* sol1, double* res0, double* res1, const double* bcv0) {
    try {
        orc::amr_residual(*P, nbox, boxes, rhs0, rhs1, sol0, sol1, res0, res1, bcv0);
    } catch (const std::exception&) {
        return -3;
    }
    return 0;
}
int orc_amr_faces(const orc_params* P, int nbox, const int* boxes, const double* sol0, const double* sol1, int mode,
                  double* out0, double* out1, const double* bcv0) {
    try {
        orc::amr_faces(*P, nbox, boxes, sol0, sol1, mode, out0, out1, bcv0);
    } catch (const std::exception&) {
        return -3;
    }
    return 0;
}
typedef orc::GmresOut orc_gmres_out;
int orc_gmres_out_size(void) { return (int)sizeof(orc::GmresOut); }
// GMRESMLMG solve (FP64): opts = {use_precond, precond_niters, restart, maxiter};
// P supplies the operator, the MLMG settings of the preconditioner and tol_rel/tol_abs.
int orc_gmres(const orc_params* P, const double* a, const double* bx, const double* by, const double* bz,
              const double* rhs, const double* bcv, double* sol, const int* opts, orc_gmres_out* out) {
    std::memset(out, 0, sizeof(*out));
    orc::GmresOpts o{opts[0], opts[1], opts[2], opts[3]};
    orc::gmres_mlmg(*P, a, bx, by, bz, rhs, bcv, sol, o, out);
    return out->status;
}
// Coefficients read by the next two-level calls with op = 0 (ABecLaplacian);
// all null: none.  The arrays must stay alive for those calls.
int orc_amr_set_coeffs(const double* a0, const double* bx0, const double* by0, const double* bz0, const double* a1,
                       const double* b1) {
    orc::AmrCoefIn& c = orc::amr_coef_in();
    c.a0 = a0;
    c.bx0 = bx0;
    c.by0 = by0;
    c.bz0 = bz0;
    c.a1 = a1;
    c.b1 = b1;
    return 0;
}
// Boundary data of the fine level for the next two-level calls (see
// amr_fine_bc_in); null: homogeneous.  The array must stay alive for those calls.
int orc_amr_set_fine_bc(const double* g) {
    orc::amr_fine_bc_in() = g;
    return 0;
}
int orc_amr_fine_sweeps(const orc_params* P, int nbox, const int* boxes, const double* res1, double* cor1,
                        int nsweeps) {
    try {
        orc::amr_fine_sweeps(*P, nbox, boxes, res1, cor1, nsweeps);
    } catch (const std::exception&) {
        return -3;
    }
    return 0;
}
int orc_amr_interp(const orc_params* P, int nbox, const int* boxes, const double* crse, double* fine) {
    try {
        orc::amr_interp(*P, nbox, boxes, crse, fine);
    } catch (const std::exception&) {
        return -3;
    }
    return 0;
}

// Lagrange extrapolation weights (for the closure tests).
void orc_poly_interp_coeff(double xInt, const double* x, int N, double* c) {
    orc::poly_interp_coeff(xInt, x, N, c);
}

}  // extern "C"
