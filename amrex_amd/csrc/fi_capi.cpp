// C ABI of the library: the reference's FFI entry points for the MLMG path
// (Src/F_Interfaces/{Base,LinearSolvers}/*_fi.cpp) over this library's
// objects, plus the mgx_* extensions declared in include/amrex_mlmg.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>

#include "../../include/amrex_mlmg.h"
#include "amrex/AMReX_GMRES_MLMG.H"
#include "amrex/AMReX_MLMG_Core.H"
#include "amrex/AMReX_PlotFileUtil.H"
#include "mgx_impl.H"
#include "mgx_fused.H"
#include "mgx_kernels.H"
#include "mgx_mfreduce.H"
#include "mgx_probe.H"
#include "mgx_smoother.H"

using namespace amrex;

namespace {

thread_local std::string g_err;
thread_local int g_status = 0;
bool g_abort = false;

void set_error(int code, const std::string& msg) {
    g_status = code;
    g_err = msg;
    std::fprintf(stderr, "amrex_mlmg: %s\n", msg.c_str());
    if (g_abort) std::abort();
}

template <class F>
int guard(F f) {
    try {
        f();
        return 0;
    } catch (const MLMG::error& e) {
        set_error(2, e.what());
        return 2;
    } catch (const std::exception& e) {
        set_error(1, e.what());
        return 1;
    }
}

// Face MultiFabs in level-major order, x, y, z per level (AMReX_multigrid_fi.cpp:27-51).
static Vector<std::array<MultiFab*, 3>> face_sets(MLMG* m, void* a[]) {
    Vector<std::array<MultiFab*, 3>> v((size_t)m->numAMRLevels());
    int q = 0;
    for (auto& f : v)
        for (auto& p : f) p = static_cast<MultiFab*>(a[q++]);
    return v;
}

constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();

template <class T>
T* buf(amrex::Vector<std::unique_ptr<DeviceBuffer>>& keep, int64_t n) {
    keep.emplace_back(new DeviceBuffer(sizeof(T) * (size_t)std::max<int64_t>(n, 1)));
    return static_cast<T*>(keep.back()->ptr);
}

template <class T>
void level_op_t(mgx::LinOpImpl& lp, int l, int which, const double* in1, double* io, double* out, int count) {
    using namespace mgx;
    if (!lp.prepared) lp.prepare();
    if (l < 0 || l >= (int)lp.lev.size()) throw std::runtime_error("level_op: bad MG level");
    const auto& L = lp.lev[(size_t)l].plan;
    hipStream_t s = Gpu::gpuStream();
    amrex::Vector<std::unique_ptr<DeviceBuffer>> keep;
    T* x = buf<T>(keep, L.nalloc);
    T* y = buf<T>(keep, L.nalloc);
    set_val<T>(x, L.nalloc, T(0), s);
    set_val<T>(y, L.nalloc, T(0), s);
    const OpCoef<T> op = lp.coef<T>(l);
    const int mbc = (int)L.mbc;
    const T* bcv = nullptr;
    if (l == 0 && lp.bcv) {
        T* b = buf<T>(keep, L.nalloc);
        convert<T, double>(b, static_cast<const double*>(lp.bcv->ptr), L.nalloc, s);
        bcv = b;
    }
    auto smooth_once = [&](T* phi, const T* rhs) {
        for (int rb = 0; rb < 2; ++rb) {
            fill_ghosts<T>(L.dev, phi, nullptr, 0, s);
            gsrb<T>(L.dev, mbc, op, phi, rhs, rb, s);
        }
    };
    switch (which) {
        case 0:
            scatter_cells<T, double>(L.dev, mbc, x, io, 1, s);
            scatter_cells<T, double>(L.dev, mbc, y, in1, 0, s);
            for (int c = 0; c < count; ++c) smooth_once(x, y);
            gather_cells<T, double>(L.dev, mbc, x, io, 1, s);
            break;
        case 1:
        case 2: {
            scatter_cells<T, double>(L.dev, mbc, x, io, 1, s);
            T* b = nullptr;
            if (which == 2) {
                b = buf<T>(keep, L.nalloc);
                set_val<T>(b, L.nalloc, T(0), s);
                scatter_cells<T, double>(L.dev, mbc, b, in1, 0, s);
            }
            fill_ghosts<T>(L.dev, x, bcv, bcv ? 1 : 0, s);
            apply<T>(L.dev, mbc, op, y, x, b, s);
            gather_cells<T, double>(L.dev, mbc, x, io, 1, s);
            gather_cells<T, double>(L.dev, mbc, y, out, 0, s);
            break;
        }
        case 3:
            scatter_cells<T, double>(L.dev, mbc, x, io, 1, s);
            scatter_cells<T, double>(L.dev, mbc, y, in1, 0, s);
            fill_ghosts<T>(L.dev, x, nullptr, 0, s);
            gsrb<T>(L.dev, mbc, op, x, y, count, s);
            gather_cells<T, double>(L.dev, mbc, x, io, 1, s);
            break;
        case 15:
            // one colour (count = redblack) of the adjoint cycle's transposed sweep (dense levels)
            if (lp.maxorder > 3) throw std::runtime_error("level_op: the transposed sweep needs maxorder <= 3");
            scatter_cells<T, double>(L.dev, mbc, x, io, 1, s);
            scatter_cells<T, double>(L.dev, mbc, y, in1, 0, s);
            gsrb_T<T>(L.dev, op, x, y, count, s);
            gather_cells<T, double>(L.dev, mbc, x, io, 1, s);
            break;
        case 16: {
            // out = in1 - L^T(io), homogeneous BC (dense levels)
            scatter_cells<T, double>(L.dev, mbc, x, io, 1, s);
            T* b = buf<T>(keep, L.nalloc);
            set_val<T>(b, L.nalloc, T(0), s);
            scatter_cells<T, double>(L.dev, mbc, b, in1, 0, s);
            resid_T<T>(L.dev, op, y, x, b, s);
            gather_cells<T, double>(L.dev, mbc, y, out, 0, s);
            break;
        }
        case 4: {
            gather_cells<T, double>(L.dev, mbc, op.a, out, 0, s);
            int64_t off = L.dev.ncell;
            for (int d = 0; d < 3; ++d) {
                const int64_t nf = (int64_t)(L.dev.n[0] + (d == 0)) * (L.dev.n[1] + (d == 1)) * (L.dev.n[2] + (d == 2));
                gather_faces<T, double>(L.dev, op.b[d], out + off, d, s);
                off += nf;
            }
            break;
        }
        case 5:
        case 13:
        case 17: {
            // 5: BiCGStab, 200 iterations; 13: the Krylov bottom solve with
            // count = solver code (2 bicgstab, 3 cg, 4 bicgcg, 5 cgbicg) + 16 *
            // iteration limit, out = {code, iterations, solver after}; 17: 13 on
            // the transposed operator (dense levels)
            const int code = which == 5 ? 2 : (count & 15);
            const int maxit = which == 5 ? 200 : (count >> 4);
            if (code < 2 || code > 5 || maxit < 1) throw std::runtime_error("level_op: bad bottom solver or limit");
            scatter_cells<T, double>(L.dev, mbc, y, in1, 0, s);
            T* work = buf<T>(keep, 6 * L.nalloc);
            int* st = buf<int>(keep, 4);
            set_val<int>(st + 2, 1, code, s);
            const T reltol = std::is_same<T, double>::value ? T(1.e-4) : T(1.e-3);
            const double volinv = 1.0 / (double)L.geom.Domain().numPts();
            // 5 applies the 8 smoothing sweeps of a failed bottom solve (nuf) as
            // the MLMG bottom does; 13 returns the Krylov iterate itself
            bottom_solve<T>(L.dev, op, x, y, work, reltol, T(-1.0), maxit, (lp.singular && lp.enforce_solvable) ? 1 : 0,
                            volinv, 0, which == 5 ? 8 : 0, st, which == 17 ? 1 : 0, s);
            gather_cells<T, double>(L.dev, mbc, x, io, 0, s);
            int h[3];
            MGX_HIP_CHECK(hipMemcpyAsync(h, st, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
            MGX_HIP_CHECK(hipStreamSynchronize(s));
            const double hd[3] = {(double)h[0], (double)h[1], (double)h[2]};
            MGX_HIP_CHECK(hipMemcpy(out, hd, (which == 5 ? 2 : 3) * sizeof(double), hipMemcpyHostToDevice));
            break;
        }
        case 6: {
            if (l + 1 >= (int)lp.lev.size()) throw std::runtime_error("level_op: no coarser level");
            const auto& Lc = lp.lev[(size_t)l + 1].plan;
            T* c = buf<T>(keep, Lc.nalloc);
            scatter_cells<T, double>(L.dev, mbc, x, in1, 0, s);
            restrict_avg<T>(Lc.dev, (int)Lc.mbc, L.dev, c, x, lp.same_layout[(size_t)l + 1], s);
            gather_cells<T, double>(Lc.dev, (int)Lc.mbc, c, out, 0, s);
            break;
        }
        case 7: {
            if (l + 1 >= (int)lp.lev.size()) throw std::runtime_error("level_op: no coarser level");
            const auto& Lc = lp.lev[(size_t)l + 1].plan;
            T* c = buf<T>(keep, Lc.nalloc);
            scatter_cells<T, double>(L.dev, mbc, x, io, 0, s);
            scatter_cells<T, double>(Lc.dev, (int)Lc.mbc, c, in1, 0, s);
            prolong_add<T>(L.dev, mbc, Lc.dev, x, c, lp.same_layout[(size_t)l + 1], s);
            gather_cells<T, double>(L.dev, mbc, x, io, 0, s);
            break;
        }
        case 8: {  // `count` fused out-of-place sweeps (dense levels)
            if (!lp.dense || L.dev.nbox != 1) throw std::runtime_error("level_op: fused sweep needs dense storage");
            scatter_cells<T, double>(L.dev, mbc, x, io, 1, s);
            scatter_cells<T, double>(L.dev, mbc, y, in1, 0, s);
            T* z = buf<T>(keep, L.nalloc);
            set_val<T>(z, L.nalloc, T(0), s);
            T* a = x;
            T* b = z;
            for (int c = 0; c < count; ++c) {
                gsrb_sweep<T>(L.dev, op, a, b, y, 0, nullptr, nullptr, s);
                std::swap(a, b);
            }
            gather_cells<T, double>(L.dev, mbc, a, io, 1, s);
            break;
        }
        case 14: {  // `count` two-sweep launches (dense levels the two-sweep march covers); io grown
            if (!lp.dense || L.dev.nbox != 1 || !gsrb_sweep2_fits(L.dev))
                throw std::runtime_error("level_op: the two-sweep launch needs a dense level of whole 64x16 tiles");
            scatter_cells<T, double>(L.dev, mbc, x, io, 1, s);
            scatter_cells<T, double>(L.dev, mbc, y, in1, 0, s);
            T* z = buf<T>(keep, L.nalloc);
            set_val<T>(z, L.nalloc, T(0), s);
            T* a = x;
            T* b = z;
            for (int c = 0; c < count; ++c) {
                gsrb_sweep2<T>(L.dev, L.hbox[0], op, a, b, y, 0, nullptr, nullptr, s);
                std::swap(a, b);
            }
            gather_cells<T, double>(L.dev, mbc, a, io, 1, s);
            break;
        }
        case 9: {  // fused residual of level 0: out = in1 - L(io) (level BC), out2 = (max, sum);
                   // count = 1: out = in1 - L^T(io), homogeneous BC
            if (!lp.dense || l != 0) throw std::runtime_error("level_op: fused residual needs dense level 0");
            if constexpr (std::is_same<T, double>::value) {
                double* b = buf<double>(keep, L.nalloc);
                double* sc = buf<double>(keep, 2 + update_residual_partials(L.dev));
                set_val<double>(b, L.nalloc, 0.0, s);
                scatter_cells<double, double>(L.dev, mbc, x, io, 1, s);
                scatter_cells<double, double>(L.dev, mbc, b, in1, 0, s);
                update_residual<double, double>(L.dev, lp.coef<double>(0), x, static_cast<const double*>(nullptr),
                                                nullptr, b, static_cast<const double*>(lp.bcv ? lp.bcv->ptr : nullptr),
                                                lp.bcv ? 1 : 0, y, sc + 2, sc, count ? 1 : 0, s);
                gather_cells<double, double>(L.dev, mbc, y, out, 0, s);
                MGX_HIP_CHECK(hipMemcpyAsync(out + L.dev.ncell, sc, 2 * sizeof(double), hipMemcpyDeviceToDevice, s));
            } else {
                throw std::runtime_error("level_op 9 is FP64");
            }
            break;
        }
        case 10: {  // fused residual + restriction: out (level l+1) = avgdown(in1 - L(io)), homogeneous BC
            if (!lp.dense || l + 1 >= (int)lp.lev.size()) throw std::runtime_error("level_op: needs dense levels");
            const auto& Lc = lp.lev[(size_t)l + 1].plan;
            T* c = buf<T>(keep, Lc.nalloc);
            scatter_cells<T, double>(L.dev, mbc, x, io, 1, s);
            scatter_cells<T, double>(L.dev, mbc, y, in1, 0, s);
            resid_restrict<T>(L.dev, op, x, y, Lc.dev, c, s);
            gather_cells<T, double>(Lc.dev, (int)Lc.mbc, c, out, 0, s);
            break;
        }
        default:
            throw std::runtime_error("level_op: unknown operation");
    }
    MGX_HIP_CHECK(hipStreamSynchronize(s));
}

bool fused_ok(const mgx::LevelDev& L) {
    for (int d = 0; d < 3; ++d)
        if (L.periodic[d] && (L.n[d] % 2) != 0) return false;
    return true;
}

LPInfo make_info(int metric_term, int agglomeration, int consolidation, int max_coarsening_level) {
    LPInfo info;
    if (metric_term >= 0) info.setMetricTerm(metric_term);
    if (agglomeration >= 0) info.setAgglomeration(agglomeration);
    if (consolidation >= 0) info.setConsolidation(consolidation);
    info.setMaxCoarseningLevel(max_coarsening_level);
    return info;
}

}  // namespace

extern "C" {

// ---- BoxArray ---------------------------------------------------------------
void amrex_fi_new_boxarray(void** ba, int lo[3], int hi[3]) {
    guard([&] { *ba = new BoxArray(Box(IntVect(lo), IntVect(hi))); });
}
void amrex_fi_new_boxarray_from_bxfarr(void** ba, const int* bxs, int nsides, int ndims, int nbxs) {
    guard([&] {
        if (nsides != 2 || ndims < 3) throw std::runtime_error("new_boxarray_from_bxfarr: expects (2, 3) boxes");
        std::vector<Box> bl;
        for (int i = 0; i < nbxs; ++i) {
            const int* b = bxs + (size_t)2 * ndims * i;
            bl.emplace_back(IntVect(b[0], b[2], b[4]), IntVect(b[1], b[3], b[5]));
        }
        *ba = new BoxArray(std::move(bl));
    });
}
void amrex_fi_delete_boxarray(void* ba) { delete static_cast<BoxArray*>(ba); }
void amrex_fi_clone_boxarray(void** bao, const void* bai) {
    guard([&] {
        delete static_cast<BoxArray*>(*bao);
        *bao = new BoxArray(*static_cast<const BoxArray*>(bai));
    });
}
void amrex_fi_boxarray_maxsize(void* ba, int sz[3]) {
    guard([&] { static_cast<BoxArray*>(ba)->maxSize(IntVect(sz)); });
}
long long amrex_fi_boxarray_nboxes(const void* ba) { return static_cast<const BoxArray*>(ba)->size(); }
void amrex_fi_boxarray_get_box(const void* ba, int i, int lo[3], int hi[3]) {
    const Box b = (*static_cast<const BoxArray*>(ba))[i];
    for (int d = 0; d < 3; ++d) { lo[d] = b.smallEnd(d); hi[d] = b.bigEnd(d); }
}
long long amrex_fi_boxarray_numpts(const void* ba) { return static_cast<const BoxArray*>(ba)->numPts(); }

// ---- Geometry / DistributionMapping -----------------------------------------
void amrex_fi_new_geometry(void** geom, int lo[3], int hi[3]) {
    guard([&] { *geom = new Geometry(Box(IntVect(lo), IntVect(hi))); });
}
void mgx_new_geometry(void** geom, const int lo[3], const int hi[3], const double plo[3], const double phi[3],
                      const int per[3]) {
    guard([&] {
        RealBox rb({plo[0], plo[1], plo[2]}, {phi[0], phi[1], phi[2]});
        *geom = new Geometry(Box(IntVect(lo), IntVect(hi)), &rb, 0, per);
    });
}
void amrex_fi_delete_geometry(void* geom) { delete static_cast<Geometry*>(geom); }
void amrex_fi_geometry_get_intdomain(const void* geom, int lo[3], int hi[3]) {
    const Box& b = static_cast<const Geometry*>(geom)->Domain();
    for (int d = 0; d < 3; ++d) { lo[d] = b.smallEnd(d); hi[d] = b.bigEnd(d); }
}
void amrex_fi_new_distromap(void** dm, const void* ba) {
    guard([&] { *dm = new DistributionMapping(*static_cast<const BoxArray*>(ba)); });
}
void amrex_fi_delete_distromap(void* dm) { delete static_cast<DistributionMapping*>(dm); }

// ---- MultiFab ----------------------------------------------------------------
void amrex_fi_new_multifab(void** mf, const void** ba, const void** dm, int nc, const int* ng, const int* nodal) {
    guard([&] {
        auto* m = new MultiFab(convert(*static_cast<const BoxArray*>(*ba), IntVect(nodal)),
                               *static_cast<const DistributionMapping*>(*dm), nc, ng[0]);
        *mf = m;
        *ba = &m->boxArray();
        *dm = &m->DistributionMap();
    });
}
void amrex_fi_delete_multifab(void* mf) { delete static_cast<MultiFab*>(mf); }
int amrex_fi_multifab_ncomp(const void* mf) { return static_cast<const MultiFab*>(mf)->nComp(); }
void amrex_fi_multifab_ngrow(const void* mf, int* ngv) {
    for (int d = 0; d < 3; ++d) ngv[d] = static_cast<const MultiFab*>(mf)->nGrow();
}
void amrex_fi_multifab_dataptr_int(void* mf, int igrd, double** dp, int lo[3], int hi[3]) {
    guard([&] {
        auto* m = static_cast<MultiFab*>(mf);
        *dp = m->fabPtr(igrd);
        const Box b = m->fabbox(igrd);
        for (int d = 0; d < 3; ++d) { lo[d] = b.smallEnd(d); hi[d] = b.bigEnd(d); }
    });
}
double amrex_fi_multifab_sum(const void* mf, int comp) {
    double r = kNaN;
    guard([&] { r = static_cast<const MultiFab*>(mf)->sum(comp); });
    return r;
}
double amrex_fi_multifab_norm0(const void* mf, int comp) {
    double r = kNaN;
    guard([&] { r = static_cast<const MultiFab*>(mf)->norminf(comp); });
    return r;
}
int mgx_multifab_sum_norm0_noalloc(const void* mf, int comp, int ncomp, double* out_dev) {
    return guard([&] {
        if (!mf || !out_dev) throw std::runtime_error("mgx_multifab_sum_norm0_noalloc: null MultiFab or output");
        if (ncomp < 1 || ncomp > mgx::kMfReduceMaxComp)
            throw std::runtime_error("mgx_multifab_sum_norm0_noalloc: 1 to " + std::to_string(mgx::kMfReduceMaxComp) +
                                     " components per call, got " + std::to_string(ncomp));
        static_cast<const MultiFab*>(mf)->reduce(comp, ncomp, mgx::mf_reduce_scratch(), out_dev);
    });
}
// components [ic, ic+nc); ghost cells of those components are set too
void amrex_fi_multifab_setval(void* mf, double val, int ic, int nc, const int*) {
    guard([&] { static_cast<MultiFab*>(mf)->setVal(val, ic, nc); });
}
void* mgx_multifab_data(void* mf, long long* nelem) {
    auto* m = static_cast<MultiFab*>(mf);
    if (nelem) *nelem = m->numElements();
    return m->dataPtr();
}
int mgx_multifab_copy_global(void* mf, double* glob, const int dom_n[3], int gg, int rg, int to_mf) {
    return guard([&] {
        auto* m = static_cast<MultiFab*>(mf);
        std::vector<mgx::UserBox> ub((size_t)m->size());
        int64_t mx = 0;
        for (Long i = 0; i < m->size(); ++i) {
            const Box b = m->boxArray()[i];
            for (int d = 0; d < 3; ++d) { ub[(size_t)i].lo[d] = b.smallEnd(d); ub[(size_t)i].hi[d] = b.bigEnd(d); }
            ub[(size_t)i].off = m->offsets()[(size_t)i];
            mx = std::max<int64_t>(mx, (int64_t)(b.length(0) + 2 * rg) * (b.length(1) + 2 * rg) * (b.length(2) + 2 * rg));
        }
        if (rg > m->nGrow()) throw std::runtime_error("copy_global: region wider than the ghost layer");
        DeviceBuffer d(sizeof(mgx::UserBox) * ub.size());
        MGX_HIP_CHECK(hipMemcpy(d.ptr, ub.data(), sizeof(mgx::UserBox) * ub.size(), hipMemcpyHostToDevice));
        hipStream_t s = Gpu::gpuStream();
        mgx::user_global<double, double>(static_cast<const mgx::UserBox*>(d.ptr), (int)ub.size(), mx, m->nGrow(),
                                         m->dataPtr(), glob, dom_n, gg, rg, to_mf, s);
        MGX_HIP_CHECK(hipStreamSynchronize(s));
    });
}

int mgx_multifab_host(void* mf, double* host, long long n, int to_mf) {
    return guard([&] {
        auto* m = static_cast<MultiFab*>(mf);
        if (n != (long long)m->numElements()) throw std::runtime_error("mgx_multifab_host: size mismatch");
        MGX_HIP_CHECK(hipStreamSynchronize(Gpu::gpuStream()));
        if (to_mf) MGX_HIP_CHECK(hipMemcpy(m->dataPtr(), host, sizeof(double) * (size_t)n, hipMemcpyHostToDevice));
        else MGX_HIP_CHECK(hipMemcpy(host, m->dataPtr(), sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
    });
}

// ---- On-disk interchange: Src/F_Interfaces/Base/AMReX_vismf_fi.cpp, AMReX_plotfile_fi.cpp
void amrex_fi_write_multifab(const void* mf, const char* name) {
    guard([&] { VisMF::Write(*static_cast<const MultiFab*>(mf), std::string(name)); });
}
int mgx_write_multifab_set_ghost(void* mf, const char* name, int set_ghost) {
    return guard([&] {
        VisMF::Write(*static_cast<const MultiFab*>(mf), std::string(name), VisMF::NFiles, set_ghost != 0);
    });
}
void amrex_fi_read_multifab(void* mf, const char* name) {
    guard([&] { VisMF::Read(*static_cast<MultiFab*>(mf), std::string(name)); });
}
void amrex_fi_write_plotfile(const char* name, int nlevs, const void* mf[], const char* varname[], const void* geom[],
                             double time, const int level_steps[], const int ref_ratio[]) {
    guard([&] {
        Vector<const MultiFab*> mfs;
        for (int l = 0; l < nlevs; ++l) mfs.push_back(static_cast<const MultiFab*>(mf[l]));
        Vector<std::string> vn;
        for (int c = 0; c < mfs[0]->nComp(); ++c) vn.push_back(varname[c]);
        Vector<Geometry> g;
        for (int l = 0; l < nlevs; ++l) g.push_back(*static_cast<const Geometry*>(geom[l]));
        Vector<int> ls(level_steps, level_steps + nlevs);
        Vector<IntVect> rr;
        for (int l = 0; l + 1 < nlevs; ++l) rr.push_back(IntVect(ref_ratio[l]));
        WriteMultiLevelPlotfile(name, nlevs, mfs, vn, g, time, ls, rr);
    });
}
int mgx_read_multifab_new(void** mf, const char* name) {
    return guard([&] {
        std::unique_ptr<MultiFab> m(new MultiFab);
        VisMF::Read(*m, std::string(name));
        *mf = m.release();
    });
}
int mgx_multifab_info(const void* mf, int* nbox, int* ncomp, int* ngrow, int nodal[3]) {
    return guard([&] {
        auto* m = static_cast<const MultiFab*>(mf);
        *nbox = (int)m->size();
        *ncomp = m->nComp();
        *ngrow = m->nGrow();
        const IntVect t = m->size() > 0 ? m->boxArray()[0].ixType() : IntVect(0);
        for (int d = 0; d < 3; ++d) nodal[d] = t[d];
    });
}
int mgx_multifab_box(const void* mf, int i, int lo[3], int hi[3]) {
    return guard([&] {
        const Box b = static_cast<const MultiFab*>(mf)->boxArray()[i];
        for (int d = 0; d < 3; ++d) { lo[d] = b.smallEnd(d); hi[d] = b.bigEnd(d); }
    });
}
int mgx_multifab_copy(void* dst, const void* src, int srccomp, int dstcomp, int numcomp, int nghost) {
    return guard([&] {
        Copy(*static_cast<MultiFab*>(dst), *static_cast<const MultiFab*>(src), srccomp, dstcomp, numcomp, nghost);
    });
}
int mgx_new_plotfile_data(void** pf, const char* name) {
    return guard([&] { *pf = new PlotFileData(std::string(name)); });
}
void mgx_delete_plotfile_data(void* pf) { delete static_cast<PlotFileData*>(pf); }
int mgx_plotfile_data_info(void* pf, int* ncomp, int* finest_level, double* time, int* coord) {
    return guard([&] {
        auto* p = static_cast<PlotFileData*>(pf);
        *ncomp = p->nComp();
        *finest_level = p->finestLevel();
        *time = p->time();
        *coord = p->coordSys();
    });
}
int mgx_plotfile_data_varname(void* pf, int i, char* buf, int len) {
    return guard([&] {
        const std::string& v = static_cast<PlotFileData*>(pf)->varNames().at((size_t)i);
        std::snprintf(buf, (size_t)len, "%s", v.c_str());
    });
}
int mgx_plotfile_data_level(void* pf, int level, int dom_lo[3], int dom_hi[3], double prob_lo[3], double prob_hi[3],
                            double dx[3], int* ref_ratio, int* level_step) {
    return guard([&] {
        auto* p = static_cast<PlotFileData*>(pf);
        if (level < 0 || level > p->finestLevel()) throw std::runtime_error("PlotFileData: bad level");
        const Box d = p->probDomain(level);
        const auto lo = p->probLo(), hi = p->probHi(), h = p->cellSize(level);
        for (int k = 0; k < 3; ++k) {
            dom_lo[k] = d.smallEnd(k);
            dom_hi[k] = d.bigEnd(k);
            prob_lo[k] = lo[(size_t)k];
            prob_hi[k] = hi[(size_t)k];
            dx[k] = h[(size_t)k];
        }
        *ref_ratio = p->refRatio(level);
        *level_step = p->levelStep(level);
    });
}
int mgx_plotfile_data_get(void* pf, int level, const char* varname, void** mf) {
    return guard([&] {
        auto* p = static_cast<PlotFileData*>(pf);
        if (level < 0 || level > p->finestLevel()) throw std::runtime_error("PlotFileData: bad level");
        std::unique_ptr<MultiFab> m(new MultiFab(varname ? p->get(level, std::string(varname)) : p->get(level)));
        *mf = m.release();
    });
}
int mgx_multigrid_checkpoint(void* mlmg, void* a_sol[], void* a_rhs[], double tol_rel, double tol_abs,
                             const char* file_name) {
    return guard([&] {
        auto* m = static_cast<MLMG*>(mlmg);
        Vector<MultiFab*> sol;
        Vector<MultiFab const*> rhs;
        for (int l = 0; l < m->numAMRLevels(); ++l) {
            sol.push_back(static_cast<MultiFab*>(a_sol[l]));
            rhs.push_back(static_cast<const MultiFab*>(a_rhs[l]));
        }
        m->checkPoint(sol, rhs, tol_rel, tol_abs, file_name);
    });
}

// ---- Linear operators ----------------------------------------------------------
void amrex_fi_new_abeclaplacian(void** linop, int nlevels, const void* geom[], const void* ba[], const void* dm[],
                                int metric_term, int agglomeration, int consolidation, int max_coarsening_level) {
    guard([&] {
        Vector<Geometry> g;
        Vector<BoxArray> b;
        Vector<DistributionMapping> d;
        for (int i = 0; i < nlevels; ++i) {
            g.push_back(*static_cast<const Geometry*>(geom[i]));
            b.push_back(*static_cast<const BoxArray*>(ba[i]));
            d.push_back(*static_cast<const DistributionMapping*>(dm[i]));
        }
        *linop = static_cast<MLLinOp*>(new MLABecLaplacian(g, b, d, make_info(metric_term, agglomeration, consolidation, max_coarsening_level)));
    });
}
void amrex_fi_new_poisson(void** linop, int nlevels, const void* geom[], const void* ba[], const void* dm[],
                          int metric_term, int agglomeration, int consolidation, int max_coarsening_level) {
    guard([&] {
        Vector<Geometry> g;
        Vector<BoxArray> b;
        Vector<DistributionMapping> d;
        for (int i = 0; i < nlevels; ++i) {
            g.push_back(*static_cast<const Geometry*>(geom[i]));
            b.push_back(*static_cast<const BoxArray*>(ba[i]));
            d.push_back(*static_cast<const DistributionMapping*>(dm[i]));
        }
        *linop = static_cast<MLLinOp*>(new MLPoisson(g, b, d, make_info(metric_term, agglomeration, consolidation, max_coarsening_level)));
    });
}
void mgx_new_abeclaplacian(void** linop, const void* geom, const void* ba, int agg, int agglomeration, int maxc) {
    guard([&] {
        LPInfo info = make_info(-1, agglomeration, -1, maxc);
        if (agg > 0) info.setAgglomerationGridSize(agg);
        const BoxArray& b = *static_cast<const BoxArray*>(ba);
        *linop = static_cast<MLLinOp*>(new MLABecLaplacian({*static_cast<const Geometry*>(geom)}, {b}, {DistributionMapping(b)}, info));
    });
}
void mgx_new_poisson(void** linop, const void* geom, const void* ba, int agg, int agglomeration, int maxc) {
    guard([&] {
        LPInfo info = make_info(-1, agglomeration, -1, maxc);
        if (agg > 0) info.setAgglomerationGridSize(agg);
        const BoxArray& b = *static_cast<const BoxArray*>(ba);
        *linop = static_cast<MLLinOp*>(new MLPoisson({*static_cast<const Geometry*>(geom)}, {b}, {DistributionMapping(b)}, info));
    });
}
void amrex_fi_abeclap_set_scalars(void* linop, double a, double b) {
    guard([&] { dynamic_cast<MLABecLaplacian&>(*static_cast<MLLinOp*>(linop)).setScalars(a, b); });
}
void amrex_fi_abeclap_set_acoeffs(void* linop, int amrlev, const void* alpha) {
    guard([&] {
        dynamic_cast<MLABecLaplacian&>(*static_cast<MLLinOp*>(linop)).setACoeffs(amrlev, *static_cast<const MultiFab*>(alpha));
    });
}
void amrex_fi_abeclap_set_bcoeffs(void* linop, int amrlev, const void* beta[]) {
    guard([&] {
        std::array<MultiFab const*, 3> b{{static_cast<const MultiFab*>(beta[0]), static_cast<const MultiFab*>(beta[1]),
                                          static_cast<const MultiFab*>(beta[2])}};
        dynamic_cast<MLABecLaplacian&>(*static_cast<MLLinOp*>(linop)).setBCoeffs(amrlev, b);
    });
}
void amrex_fi_delete_linop(void* linop) { delete static_cast<MLLinOp*>(linop); }
void amrex_fi_linop_set_maxorder(void* linop, int ord) {
    guard([&] { static_cast<MLLinOp*>(linop)->setMaxOrder(ord); });
}
void amrex_fi_linop_set_domain_bc(void* linop, const int* ilobc, const int* ihibc) {
    guard([&] {
        std::array<LinOpBCType, 3> lo, hi;
        for (int d = 0; d < 3; ++d) { lo[d] = static_cast<LinOpBCType>(ilobc[d]); hi[d] = static_cast<LinOpBCType>(ihibc[d]); }
        static_cast<MLLinOp*>(linop)->setDomainBC(lo, hi);
    });
}
void amrex_fi_linop_set_coarse_fine_bc(void* linop, const void* crse, int ratio) {
    guard([&] {
        if (!linop) throw std::runtime_error("setCoarseFineBC: null operator");
        static_cast<MLLinOp*>(linop)->setCoarseFineBC(static_cast<const MultiFab*>(crse), ratio);
    });
}
int mgx_linop_set_coarse_fine_bc_type(void* linop, const void* crse, int ratio, int bc_type) {
    return guard([&] {
        if (!linop) throw std::runtime_error("setCoarseFineBC: null operator");
        static_cast<MLLinOp*>(linop)->setCoarseFineBC(static_cast<const MultiFab*>(crse), ratio,
                                                      static_cast<LinOpBCType>(bc_type));
    });
}
void amrex_fi_linop_set_level_bc(void* linop, int amrlev, const void* levelbcdata) {
    guard([&] { static_cast<MLLinOp*>(linop)->setLevelBC(amrlev, static_cast<const MultiFab*>(levelbcdata)); });
}
int mgx_linop_num_mg_levels(void* linop) { return static_cast<MLLinOp*>(linop)->NMGLevels(0); }
int mgx_linop_level_info(void* linop, int mglev, int n[3], int* nboxes, int* boxes, int max_boxes) {
    return guard([&] {
        auto* lp = static_cast<MLLinOp*>(linop);
        const Geometry& g = lp->Geom(0, mglev);
        const BoxArray& ba = lp->boxArray(0, mglev);
        for (int d = 0; d < 3; ++d) n[d] = g.Domain().length(d);
        *nboxes = (int)ba.size();
        if (boxes)
            for (int i = 0; i < std::min<int>((int)ba.size(), max_boxes); ++i)
                for (int d = 0; d < 3; ++d) { boxes[6 * i + d] = ba[i].smallEnd(d); boxes[6 * i + 3 + d] = ba[i].bigEnd(d); }
    });
}
int mgx_linop_level_op(void* linop, int mglev, int which, int precision, const double* in1, double* io, double* out,
                       int count) {
    return guard([&] {
        auto& lp = static_cast<MLLinOp*>(linop)->impl();
        if (precision == 0) level_op_t<double>(lp, mglev, which, in1, io, out, count);
        else level_op_t<float>(lp, mglev, which, in1, io, out, count);
    });
}

// Times nsweeps smoothing sweeps of MG level mglev with HIP events on the
// library stream: the fused single-launch sweep on dense levels, else ghost
// fill + red + ghost fill + black.  out_ms[0] = per sweep, out_ms[1] = per GSRB
// launch, out_ms[2] = per ghost-fill launch (0 for the fused sweep).
int mgx_hbm_probe(int mode, const void* src, void* dst, long long nbytes, int reps, double* out_ms) {
    return guard([&] {
        if (reps < 1 || !out_ms) throw std::runtime_error("mgx_hbm_probe: reps >= 1 and an output expected");
        hipStream_t s = Gpu::gpuStream();
        amrex::Vector<std::unique_ptr<DeviceBuffer>> keep;
        double* sink = buf<double>(keep, 1);
        hipEvent_t ev[2];
        for (auto& e : ev) MGX_HIP_CHECK(hipEventCreate(&e));
        mgx::hbm_probe(mode, src, dst, nbytes, sink, s);  // warm-up
        MGX_HIP_CHECK(hipEventRecord(ev[0], s));
        for (int r = 0; r < reps; ++r) mgx::hbm_probe(mode, src, dst, nbytes, sink, s);
        MGX_HIP_CHECK(hipEventRecord(ev[1], s));
        MGX_HIP_CHECK(hipEventSynchronize(ev[1]));
        float t = 0;
        MGX_HIP_CHECK(hipEventElapsedTime(&t, ev[0], ev[1]));
        *out_ms = (double)t / reps;
        for (auto& e : ev) (void)hipEventDestroy(e);
    });
}

int mgx_linop_time_smoother(void* linop, int mglev, int precision, int nsweeps, double* out_ms) {
    return guard([&] {
        auto& lp = static_cast<MLLinOp*>(linop)->impl();
        if (!lp.prepared) lp.prepare();
        const auto& L = lp.lev.at((size_t)mglev).plan;
        hipStream_t s = Gpu::gpuStream();
        auto run = [&](auto tag) {
            using T = decltype(tag);
            amrex::Vector<std::unique_ptr<DeviceBuffer>> keep;
            T* phi = buf<T>(keep, L.nalloc);
            T* rhs = buf<T>(keep, L.nalloc);
            mgx::set_val<T>(phi, L.nalloc, T(0.5), s);
            mgx::set_val<T>(rhs, L.nalloc, T(1.0), s);
            const mgx::OpCoef<T> op = lp.coef<T>(mglev);
            hipEvent_t ev[5];
            for (auto& e : ev) MGX_HIP_CHECK(hipEventCreate(&e));
            double tsweep = 0, tg = 0, tf = 0;
            const bool fused = lp.dense && L.dev.nbox == 1 && fused_ok(L.dev);
            if (fused) {
                T* alt = buf<T>(keep, L.nalloc);
                mgx::set_val<T>(alt, L.nalloc, T(0), s);
                for (int it = -1; it < nsweeps; ++it) {
                    MGX_HIP_CHECK(hipEventRecord(ev[0], s));
                    mgx::gsrb_sweep<T>(L.dev, op, phi, alt, rhs, 0, nullptr, nullptr, s);
                    MGX_HIP_CHECK(hipEventRecord(ev[1], s));
                    MGX_HIP_CHECK(hipEventSynchronize(ev[1]));
                    float t;
                    MGX_HIP_CHECK(hipEventElapsedTime(&t, ev[0], ev[1]));
                    if (it >= 0) tsweep += t;
                    std::swap(phi, alt);
                }
                for (auto& e : ev) (void)hipEventDestroy(e);
                out_ms[0] = tsweep / nsweeps;
                out_ms[1] = out_ms[0];
                out_ms[2] = 0.0;
                return;
            }
            for (int it = -1; it < nsweeps; ++it) {  // it = -1 warms up
                MGX_HIP_CHECK(hipEventRecord(ev[0], s));
                mgx::fill_ghosts<T>(L.dev, phi, nullptr, 0, s);
                MGX_HIP_CHECK(hipEventRecord(ev[1], s));
                mgx::gsrb<T>(L.dev, (int)L.mbc, op, phi, rhs, 0, s);
                MGX_HIP_CHECK(hipEventRecord(ev[2], s));
                mgx::fill_ghosts<T>(L.dev, phi, nullptr, 0, s);
                MGX_HIP_CHECK(hipEventRecord(ev[3], s));
                mgx::gsrb<T>(L.dev, (int)L.mbc, op, phi, rhs, 1, s);
                MGX_HIP_CHECK(hipEventRecord(ev[4], s));
                MGX_HIP_CHECK(hipEventSynchronize(ev[4]));
                float t[4];
                for (int q = 0; q < 4; ++q) MGX_HIP_CHECK(hipEventElapsedTime(&t[q], ev[q], ev[q + 1]));
                if (it >= 0) {
                    tsweep += t[0] + t[1] + t[2] + t[3];
                    tf += t[0] + t[2];
                    tg += t[1] + t[3];
                }
            }
            for (auto& e : ev) (void)hipEventDestroy(e);
            out_ms[0] = tsweep / nsweeps;
            out_ms[1] = tg / (2.0 * nsweeps);
            out_ms[2] = tf / (2.0 * nsweeps);
        };
        if (precision == 0) run(double{});
        else run(float{});
    });
}

int mgx_linop_time_sweep_pairs(void* linop, int mglev, int precision, int npairs, double* out_ms) {
    return guard([&] {
        auto& lp = static_cast<MLLinOp*>(linop)->impl();
        if (!lp.prepared) lp.prepare();
        const auto& L = lp.lev.at((size_t)mglev).plan;
        if (!lp.dense || L.dev.nbox != 1 || !mgx::gsrb_sweep2_fits(L.dev))
            throw std::runtime_error("time_sweep_pairs: the two-sweep launch needs a dense level of whole 64x16 tiles");
        if (npairs < 1) throw std::runtime_error("time_sweep_pairs: npairs must be positive");
        hipStream_t s = Gpu::gpuStream();
        auto run = [&](auto tag) {
            using T = decltype(tag);
            amrex::Vector<std::unique_ptr<DeviceBuffer>> keep;
            T* phi = buf<T>(keep, L.nalloc);
            T* rhs = buf<T>(keep, L.nalloc);
            T* alt = buf<T>(keep, L.nalloc);
            mgx::set_val<T>(phi, L.nalloc, T(0.5), s);
            mgx::set_val<T>(rhs, L.nalloc, T(1.0), s);
            mgx::set_val<T>(alt, L.nalloc, T(0), s);
            const mgx::OpCoef<T> op = lp.coef<T>(mglev);
            hipEvent_t ev[2];
            for (auto& e : ev) MGX_HIP_CHECK(hipEventCreate(&e));
            double tot = 0;
            for (int it = -1; it < npairs; ++it) {  // it = -1 warms up
                MGX_HIP_CHECK(hipEventRecord(ev[0], s));
                mgx::gsrb_sweep2<T>(L.dev, L.hbox[0], op, phi, alt, rhs, 0, nullptr, nullptr, s);
                MGX_HIP_CHECK(hipEventRecord(ev[1], s));
                MGX_HIP_CHECK(hipEventSynchronize(ev[1]));
                float t;
                MGX_HIP_CHECK(hipEventElapsedTime(&t, ev[0], ev[1]));
                if (it >= 0) tot += t;
                std::swap(phi, alt);
            }
            for (auto& e : ev) (void)hipEventDestroy(e);
            *out_ms = tot / npairs;
        };
        if (precision == 0) run(double{});
        else run(float{});
    });
}
int mgx_set_pair_sweeps(int on) {
    const int was = mgx::pair_sweeps_enabled() ? 1 : 0;
    mgx::set_pair_sweeps(on != 0);
    return was;
}

// ---- MLMG ------------------------------------------------------------------------
void amrex_fi_new_multigrid(void** mlmg, void* lp) {
    guard([&] { *mlmg = new MLMG(*static_cast<MLLinOp*>(lp)); });
}
void amrex_fi_delete_multigrid(void* mlmg) { delete static_cast<MLMG*>(mlmg); }
double amrex_fi_multigrid_solve(void* mlmg, void* a_sol[], void* a_rhs[], double tol_rel, double tol_abs) {
    double r = kNaN;
    guard([&] {
        auto* m = static_cast<MLMG*>(mlmg);
        const int n = m->numAMRLevels();
        Vector<MultiFab*> sol;
        Vector<MultiFab const*> rhs;
        for (int l = 0; l < n; ++l) {
            sol.push_back(static_cast<MultiFab*>(a_sol[l]));
            rhs.push_back(static_cast<const MultiFab*>(a_rhs[l]));
        }
        r = m->solve(sol, rhs, tol_rel, tol_abs);
    });
    return r;
}
double mgx_multigrid_solve_adjoint(void* mlmg, void* a_sol[], void* a_rhs[], double tol_rel, double tol_abs) {
    double r = kNaN;
    guard([&] {
        auto* m = static_cast<MLMG*>(mlmg);
        r = m->solveAdjoint(Vector<MultiFab*>{static_cast<MultiFab*>(a_sol[0])},
                            Vector<MultiFab const*>{static_cast<const MultiFab*>(a_rhs[0])}, tol_rel, tol_abs);
    });
    return r;
}
int mgx_new_gmres_mlmg(void** gmres, void* mlmg) {
    return guard([&] { *gmres = new GMRESMLMG(*static_cast<MLMG*>(mlmg)); });
}
void mgx_delete_gmres_mlmg(void* gmres) { delete static_cast<GMRESMLMG*>(gmres); }
double mgx_gmres_mlmg_solve(void* gmres, void* sol, const void* rhs, double tol_rel, double tol_abs) {
    double r = kNaN;
    guard([&] {
        auto* g = static_cast<GMRESMLMG*>(gmres);
        g->solve(*static_cast<MultiFab*>(sol), *static_cast<const MultiFab*>(rhs), tol_rel, tol_abs);
        r = g->getResidualNorm();
    });
    return r;
}
void mgx_gmres_mlmg_set_verbose(void* gmres, int v) { static_cast<GMRESMLMG*>(gmres)->setVerbose(v); }
void mgx_gmres_mlmg_set_max_iters(void* gmres, int n) { static_cast<GMRESMLMG*>(gmres)->setMaxIters(n); }
void mgx_gmres_mlmg_set_restart_length(void* gmres, int n) { static_cast<GMRESMLMG*>(gmres)->setRestartLength(n); }
void mgx_gmres_mlmg_set_precond_num_iters(void* gmres, int n) {
    static_cast<GMRESMLMG*>(gmres)->setPrecondNumIters(n);
}
int mgx_gmres_mlmg_use_precond(void* gmres, int flag) {
    return static_cast<GMRESMLMG*>(gmres)->usePrecond(flag != 0) ? 1 : 0;
}
void mgx_gmres_mlmg_set_property_of_zero(void* gmres, int flag) {
    static_cast<GMRESMLMG*>(gmres)->setPropertyOfZero(flag != 0);
}
int mgx_gmres_mlmg_stats(void* gmres, int* niters, int* status, double* resnorm, double* history, int max_history) {
    auto* g = static_cast<GMRESMLMG*>(gmres);
    *niters = g->getNumIters();
    *status = g->getStatus();
    *resnorm = g->getResidualNorm();
    const auto& h = g->getResidualHistory();
    const int n = std::min<int>((int)h.size(), max_history);
    for (int i = 0; i < n; ++i) history[i] = h[(size_t)i];
    return (int)h.size();
}
int mgx_multigrid_get_fluxes_sol(void* mlmg, void* a_fluxes[], void* a_sol[]) {
    return guard([&] {
        auto* m = static_cast<MLMG*>(mlmg);
        Vector<MultiFab*> sol;
        for (int l = 0; l < m->numAMRLevels(); ++l) sol.push_back(static_cast<MultiFab*>(a_sol[l]));
        m->getFluxes(face_sets(m, a_fluxes), sol);
    });
}
int mgx_multigrid_get_cell_fluxes(void* mlmg, void* a_fluxes[], void* a_sol[]) {
    return guard([&] {
        auto* m = static_cast<MLMG*>(mlmg);
        Vector<MultiFab*> flux, sol;
        for (int l = 0; l < m->numAMRLevels(); ++l) {
            flux.push_back(static_cast<MultiFab*>(a_fluxes[l]));
            if (a_sol) sol.push_back(static_cast<MultiFab*>(a_sol[l]));
        }
        if (a_sol) m->getFluxes(flux, sol);
        else m->getFluxes(flux);
    });
}
int mgx_abeclap_coeff_gradient(void* linop, const void* lam, const void* phi, void* grad_a, void* grad_b[3],
                               int accumulate) {
    return guard([&] {
        auto* lp = dynamic_cast<MLABecLaplacian*>(static_cast<MLLinOp*>(linop));
        if (!lp) throw std::runtime_error("mgx_abeclap_coeff_gradient: not an MLABecLaplacian");
        std::array<MultiFab*, 3> gb{{nullptr, nullptr, nullptr}};
        if (grad_b)
            for (int d = 0; d < 3; ++d) gb[d] = static_cast<MultiFab*>(grad_b[d]);
        lp->compCoeffGradient(0, *static_cast<const MultiFab*>(lam), *static_cast<const MultiFab*>(phi),
                              static_cast<MultiFab*>(grad_a), gb, accumulate != 0);
    });
}
void amrex_fi_multigrid_get_grad_solution(void* mlmg, void* a_grad_sol[]) {
    guard([&] {
        auto* m = static_cast<MLMG*>(mlmg);
        m->getGradSolution(face_sets(m, a_grad_sol));
    });
}
void amrex_fi_multigrid_get_fluxes(void* mlmg, void* a_fluxes[]) {
    guard([&] {
        auto* m = static_cast<MLMG*>(mlmg);
        m->getFluxes(face_sets(m, a_fluxes));
    });
}
void amrex_fi_multigrid_comp_residual(void* mlmg, void* a_res[], void* a_sol[], void* a_rhs[]) {
    guard([&] {
        static_cast<MLMG*>(mlmg)->compResidual(Vector<MultiFab*>{static_cast<MultiFab*>(a_res[0])},
                                               Vector<MultiFab*>{static_cast<MultiFab*>(a_sol[0])},
                                               Vector<MultiFab const*>{static_cast<const MultiFab*>(a_rhs[0])});
    });
}
void amrex_fi_multigrid_set_verbose(void* m, int v) { static_cast<MLMG*>(m)->setVerbose(v); }
void amrex_fi_multigrid_set_max_iter(void* m, int n) { static_cast<MLMG*>(m)->setMaxIter(n); }
void amrex_fi_multigrid_set_max_fmg_iter(void* m, int n) { static_cast<MLMG*>(m)->setMaxFmgIter(n); }
void amrex_fi_multigrid_set_fixed_iter(void* m, int n) { static_cast<MLMG*>(m)->setFixedIter(n); }
void amrex_fi_multigrid_set_bottom_solver(void* m, int s) {
    // same codes as AMReX_multigrid_fi.cpp:82-97
    BottomSolver b;
    switch (s) {
        case 0: b = BottomSolver::smoother; break;
        case 1: b = BottomSolver::bicgstab; break;
        case 2: b = BottomSolver::cg; break;
        case 3: b = BottomSolver::hypre; break;
        case 4: b = BottomSolver::petsc; break;
        default: set_error(1, "amrex_fi_multigrid_set_bottom_solver: unknown bottom solver"); return;
    }
    static_cast<MLMG*>(m)->setBottomSolver(b);
}
int mgx_multigrid_set_bottom_solver_code(void* m, int code) {
    return guard([&] {
        if (code < 0 || code > 7) throw std::runtime_error("setBottomSolver: unknown amrex::BottomSolver code");
        static_cast<MLMG*>(m)->setBottomSolver(static_cast<BottomSolver>(code));
    });
}
int mgx_multigrid_get_bottom_solver_code(void* m) { return static_cast<int>(static_cast<MLMG*>(m)->getBottomSolver()); }
void amrex_fi_multigrid_set_bottom_verbose(void* m, int n) { static_cast<MLMG*>(m)->setBottomVerbose(n); }
void amrex_fi_multigrid_set_always_use_bnorm(void* m, int f) { static_cast<MLMG*>(m)->setAlwaysUseBNorm(f); }
void amrex_fi_multigrid_set_final_fill_bc(void* m, int f) { static_cast<MLMG*>(m)->setFinalFillBC(f); }
int mgx_multigrid_set_precision(void* m, int mixed) {
    static_cast<MLMG*>(m)->setPrecision(mixed ? MLMG::Precision::Mixed : MLMG::Precision::Double);
    return 0;
}
int mgx_multigrid_set_refinement_vcycles(void* m, int n) {
    return guard([&] { static_cast<MLMG*>(m)->setRefinementVcycles(n); });
}
int mgx_multigrid_set_use_graph(void* m, int use) {
    static_cast<MLMG*>(m)->setUseGraph(use != 0);
    return 0;
}
int mgx_multigrid_set_bottom_tolerance(void* m, double reltol, double abstol) {
    static_cast<MLMG*>(m)->setBottomTolerance(reltol);
    static_cast<MLMG*>(m)->setBottomToleranceAbs(abstol);
    return 0;
}
int mgx_multigrid_set_bottom_max_iter(void* m, int n) {
    static_cast<MLMG*>(m)->setBottomMaxIter(n);
    return 0;
}
int mgx_multigrid_set_pre_post_smooth(void* m, int nu1, int nu2) {
    static_cast<MLMG*>(m)->setPreSmooth(nu1);
    static_cast<MLMG*>(m)->setPostSmooth(nu2);
    return 0;
}
int mgx_multigrid_get_stats(void* mlmg, mgx_solve_stats* st) {
    auto* m = static_cast<MLMG*>(mlmg);
    std::memset(st, 0, sizeof(*st));
    st->niters = m->getNumIters();
    st->nlevels = m->impl().lp.lev.empty() ? 0 : (int)m->impl().lp.lev.size();
    st->last_bottom_status = m->impl().last_bottom_status;
    st->last_bottom_iters = m->getNumCGIters().empty() ? 0 : m->getNumCGIters().back();
    st->init_resnorm = m->getInitResidual();
    st->rhsnorm = m->getInitRHS();
    st->final_resnorm = m->getFinalResidual();
    st->solve_seconds = m->getTimers()[0];
    st->iter_seconds = m->getTimers()[1];
    const auto& h = m->getResidualHistory();
    for (size_t i = 0; i < h.size() && i < 256; ++i) st->res_history[i] = h[i];
    st->nvcycles = m->getNumVcycles();
    return 0;
}

// ---- error handling, stream ------------------------------------------------------
int mgx_last_error(char* b, int len) {
    if (b && len > 0) {
        std::strncpy(b, g_err.c_str(), (size_t)len - 1);
        b[len - 1] = 0;
    }
    return g_status;
}
void mgx_clear_error(void) {
    g_status = 0;
    g_err.clear();
}
void mgx_set_abort_on_error(int flag) { g_abort = flag != 0; }
int mgx_set_stream(void* s) {
    return guard([&] {
        Gpu::setStream(static_cast<hipStream_t>(s));
    });
}
void* mgx_get_stream(void) { return (void*)Gpu::gpuStream(); }
int mgx_synchronize(void) {
    return guard([&] { Gpu::streamSynchronize(); });
}
int mgx_abi_version(void) { return 1; }
int mgx_set_dense_storage(int dense) {
    mgx::set_dense_default(dense != 0);
    return 0;
}

}  // extern "C"
