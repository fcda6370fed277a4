// Host side of the MI355X MLMG: base types, MG hierarchy, level plans and
// the solve orchestration.  All numerical work is issued to the HIP kernels
// of mgx_kernels.hip on one stream; the V-cycle of an outer iteration is
// captured once into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <numeric>

#include "amrex/AMReX_MLMG_Core.H"
#include "mgx_fused.H"
#include "mgx_impl.H"
#include "mgx_kernels.H"
#include "mgx_mfreduce.H"

namespace amrex {

// ---------------------------------------------------------------------------
// Box / BoxArray / Geometry
// ---------------------------------------------------------------------------
static inline int coarsen_i(int i, int r) { return (i < 0) ? -((-i - 1) / r) - 1 : i / r; }

Box& Box::coarsen(int r) {
    for (int d = 0; d < 3; ++d) {
        m_lo[d] = coarsen_i(m_lo[d], r);
        if (m_nodal[d]) m_hi[d] = coarsen_i(m_hi[d] - 1, r) + 1;  // nodal: coarsen the enclosed cells
        else m_hi[d] = coarsen_i(m_hi[d], r);
    }
    return *this;
}
Box& Box::refine(int r) {
    for (int d = 0; d < 3; ++d) {
        m_lo[d] *= r;
        m_hi[d] = m_nodal[d] ? m_hi[d] * r : (m_hi[d] + 1) * r - 1;
    }
    return *this;
}
// Box::coarsenable (AMReX_Box.H:745-755)
bool Box::coarsenable(int r, int min_width) const {
    for (int d = 0; d < 3; ++d) if (length(d) < r * min_width) return false;
    Box t = *this;
    t.coarsen(r).refine(r);
    return t == *this;
}
Box Box::operator&(const Box& o) const {
    Box r = *this;
    for (int d = 0; d < 3; ++d) {
        r.m_lo[d] = std::max(m_lo[d], o.m_lo[d]);
        r.m_hi[d] = std::min(m_hi[d], o.m_hi[d]);
    }
    return r;
}
Box coarsen(const Box& b, int r) { Box c = b; return c.coarsen(r); }
Box refine(const Box& b, int r) { Box c = b; return c.refine(r); }

// Chopping rule of BoxList::maxSize (AMReX_BoxList.cpp:765-820): pieces as
// even as possible, preferring power-of-two alignment.
BoxArray& BoxArray::maxSize(const IntVect& chunk) {
    std::vector<Box> out;
    for (const Box& bx : m_boxes) {
        int ratio[3] = {1, 1, 1}, numblk[3] = {1, 1, 1}, extra[3] = {0, 0, 0}, sz[3];
        for (int d = 0; d < 3; ++d) {
            sz[d] = bx.length(d);
            if (bx.length(d) > chunk[d]) {
                int bs = chunk[d], nlen = bx.length(d);
                while (bs % 2 == 0 && nlen % 2 == 0) { ratio[d] *= 2; bs /= 2; nlen /= 2; }
                numblk[d] = (nlen + bs - 1) / bs;
                sz[d] = nlen / numblk[d];
                extra[d] = nlen - sz[d] * numblk[d];
            }
        }
        auto piece = [&](int d, int b, int& lo, int& hi) {
            lo = (b < extra[d]) ? b * (sz[d] + 1) * ratio[d] : (b * sz[d] + extra[d]) * ratio[d];
            hi = (b < extra[d]) ? lo + (sz[d] + 1) * ratio[d] - 1 : lo + sz[d] * ratio[d] - 1;
            lo += bx.smallEnd(d);
            hi += bx.smallEnd(d);
        };
        if (numblk[0] * numblk[1] * numblk[2] == 1) { out.push_back(bx); continue; }
        for (int k = 0; k < numblk[2]; ++k)
            for (int j = 0; j < numblk[1]; ++j)
                for (int i = 0; i < numblk[0]; ++i) {
                    IntVect lo, hi;
                    piece(0, i, lo[0], hi[0]);
                    piece(1, j, lo[1], hi[1]);
                    piece(2, k, lo[2], hi[2]);
                    out.emplace_back(lo, hi, bx.ixType());
                }
    }
    m_boxes = std::move(out);
    return *this;
}
Long BoxArray::numPts() const {
    Long n = 0;
    for (const Box& b : m_boxes) n += b.numPts();
    return n;
}
Box BoxArray::minimalBox() const {
    Box r = m_boxes.at(0);
    IntVect lo = r.smallEnd(), hi = r.bigEnd();
    for (const Box& b : m_boxes)
        for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], b.smallEnd(d)); hi[d] = std::max(hi[d], b.bigEnd(d)); }
    return Box(lo, hi, r.ixType());
}
BoxArray& BoxArray::coarsen(int r) {
    for (Box& b : m_boxes) b.coarsen(r);
    return *this;
}
bool BoxArray::coarsenable(int r, int min_width) const {
    for (const Box& b : m_boxes) if (!b.coarsenable(r, min_width)) return false;
    return true;
}
BoxArray convert(const BoxArray& ba, const IntVect& nodal) {
    std::vector<Box> v;
    for (Box b : ba.boxes()) {
        b.enclosedCells();
        for (int d = 0; d < 3; ++d) if (nodal[d]) b.surroundingNodes(d);
        v.push_back(b);
    }
    return BoxArray(std::move(v));
}

Geometry::Geometry(const Box& dom, const RealBox* rb, int coord, const int* is_per) : m_domain(dom), m_coord(coord) {
    if (rb) m_rb = *rb;
    for (int d = 0; d < 3; ++d) {
        m_per[d] = is_per ? is_per[d] : 0;
        m_dx[d] = m_rb.length(d) / (double)m_domain.length(d);
        m_dxinv[d] = 1.0 / m_dx[d];
    }
}

namespace Gpu {
static hipStream_t g_own = nullptr;
static hipStream_t g_user = nullptr;
hipStream_t gpuStream() {
    if (g_user) return g_user;
    if (!g_own) {
        // a non-blocking stream of default priority, so V-cycles can be captured
        MGX_HIP_CHECK(hipStreamCreateWithFlags(&g_own, hipStreamNonBlocking));
    }
    return g_own;
}
// nullptr selects the library's own stream again
void setStream(hipStream_t s) { g_user = s; }
void streamSynchronize() { MGX_HIP_CHECK(hipStreamSynchronize(gpuStream())); }
}  // namespace Gpu

DeviceBuffer::DeviceBuffer(size_t nbytes) : bytes(nbytes) {
    if (nbytes) MGX_HIP_CHECK(hipMalloc(&ptr, nbytes));
}
DeviceBuffer::~DeviceBuffer() {
    if (ptr) (void)hipFree(ptr);
}

template <class T>
void FabArray<T>::define(const BoxArray& ba, const DistributionMapping& dm, int ncomp, int ngrow) {
    if (ncomp < 1) throw std::runtime_error("FabArray: ncomp must be positive");
    m_ba = ba;
    m_dm = dm;
    m_ncomp = ncomp;
    m_ngrow = ngrow;
    m_off.resize((size_t)ba.size());
    int64_t o = 0;
    for (Long i = 0; i < ba.size(); ++i) {
        m_off[(size_t)i] = o;
        o += fabbox(i).numPts() * ncomp;
    }
    m_total = o;
    m_buf = std::make_shared<DeviceBuffer>(sizeof(T) * (size_t)std::max<int64_t>(o, 1));
    m_doff = std::make_shared<DeviceBuffer>(sizeof(int64_t) * std::max<size_t>(m_off.size(), 1));
    MGX_HIP_CHECK(hipMemcpy(m_doff->ptr, m_off.data(), sizeof(int64_t) * m_off.size(), hipMemcpyHostToDevice));
    std::vector<mgx::UserBox> ub((size_t)ba.size());
    m_maxcells = m_maxgrown = 0;
    for (Long i = 0; i < ba.size(); ++i) {
        const Box b = ba[i];
        for (int d = 0; d < 3; ++d) { ub[(size_t)i].lo[d] = b.smallEnd(d); ub[(size_t)i].hi[d] = b.bigEnd(d); }
        ub[(size_t)i].off = m_off[(size_t)i];
        m_maxcells = std::max<int64_t>(m_maxcells, b.numPts());
        m_maxgrown = std::max<int64_t>(m_maxgrown, (int64_t)(b.length(0) + 2) * (b.length(1) + 2) * (b.length(2) + 2));
    }
    m_dub = std::make_shared<DeviceBuffer>(sizeof(mgx::UserBox) * std::max<size_t>(ub.size(), 1));
    MGX_HIP_CHECK(hipMemcpy(m_dub->ptr, ub.data(), sizeof(mgx::UserBox) * ub.size(), hipMemcpyHostToDevice));
    setVal(T(0));
}
template <class T>
void FabArray<T>::setVal(T v) {
    mgx::set_val<T>(dataPtr(), m_total, v, Gpu::gpuStream());
}
template <class T>
void FabArray<T>::copyToHost(std::vector<T>& out) const {
    out.resize((size_t)m_total);
    Gpu::streamSynchronize();
    MGX_HIP_CHECK(hipMemcpy(out.data(), dataPtr(), sizeof(T) * (size_t)m_total, hipMemcpyDeviceToHost));
}
template <class T>
void FabArray<T>::copyFromHost(const std::vector<T>& in) {
    if ((int64_t)in.size() != m_total) throw std::runtime_error("FabArray::copyFromHost: size mismatch");
    Gpu::streamSynchronize();
    MGX_HIP_CHECK(hipMemcpy(dataPtr(), in.data(), sizeof(T) * (size_t)m_total, hipMemcpyHostToDevice));
}
template <class T>
void FabArray<T>::setVal(T v, int comp, int ncomp) {
    if (comp < 0 || ncomp < 1 || comp + ncomp > m_ncomp) throw std::runtime_error("FabArray::setVal: bad component range");
    for (Long b = 0; b < m_ba.size(); ++b)
        mgx::set_val<T>(dataPtr() + compOffset(b, comp), fabbox(b).numPts() * ncomp, v, Gpu::gpuStream());
}
namespace {
// Library-owned buffers of the MultiFab reductions, one set per device.
struct MfReduceBufs {
    std::unique_ptr<DeviceBuffer> part, out;
};
MfReduceBufs& mf_reduce_bufs() {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static MfReduceBufs bufs[kMaxDevices];
    int dev = 0;
    MGX_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("FabArray reduction: device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    MfReduceBufs& b = bufs[dev];
    if (!b.part) {
        b.part.reset(new DeviceBuffer(sizeof(double) * (size_t)mgx::kMfReducePartials));
        b.out.reset(new DeviceBuffer(sizeof(double) * 2 * mgx::kMfReduceMaxComp));
    }
    return b;
}
}  // namespace

template <class T>
void FabArray<T>::reduce(int comp, int ncomp, double* partials_dev, double* out_dev) const {
    if (comp < 0 || ncomp < 1 || comp + ncomp > m_ncomp)
        throw std::runtime_error("FabArray reduction: bad component range");
    mgx::mf_reduce<T>(dataPtr(), static_cast<const mgx::UserBox*>(devBoxes()), (int)m_ba.size(), m_ngrow,
                      m_ba.numPts(), comp, ncomp, partials_dev, out_dev, Gpu::gpuStream());
}
template <class T>
void FabArray<T>::sum_max(int comp, double out[2]) const {
    MfReduceBufs& b = mf_reduce_bufs();
    reduce(comp, 1, static_cast<double*>(b.part->ptr), static_cast<double*>(b.out->ptr));
    MGX_HIP_CHECK(hipMemcpyAsync(out, b.out->ptr, 2 * sizeof(double), hipMemcpyDeviceToHost, Gpu::gpuStream()));
    Gpu::streamSynchronize();
}
template <class T>
T FabArray<T>::norminf(int comp) const {
    double r[2];
    sum_max(comp, r);
    return (T)r[1];
}
template <class T>
T FabArray<T>::sum(int comp) const {
    double r[2];
    sum_max(comp, r);
    return (T)r[0];
}
template class FabArray<double>;
template class FabArray<float>;

}  // namespace amrex

namespace mgx {

using amrex::Box;
using amrex::BoxArray;
using amrex::Geometry;
using amrex::IntVect;

// ---------------------------------------------------------------------------
// Level plans
// ---------------------------------------------------------------------------
void LevelPlan::build(const Geometry& g, const BoxArray& b) {
    geom = g;
    ba = b;
    const Box dom = g.Domain();
    for (int d = 0; d < 3; ++d)
        if (dom.smallEnd(d) != 0) throw std::runtime_error("mgx: the domain must start at index 0");
    const int nb = (int)ba.size();
    hbox.assign((size_t)nb, BoxDev{});
    int64_t off = 0, foff[3] = {0, 0, 0};
    cell_start.assign((size_t)nb + 1, 0);
    mbc = mbg = mbf = 0;
    mtiles[0] = mtiles[1] = 0;
    for (int i = 0; i < nb; ++i) {
        const Box bx = ba[i];
        BoxDev& B = hbox[(size_t)i];
        for (int d = 0; d < 3; ++d) { B.lo[d] = bx.smallEnd(d); B.hi[d] = bx.bigEnd(d); }
        const int64_t nx = bx.length(0), ny = bx.length(1), nz = bx.length(2);
        B.off = off;
        B.sy = nx + 2;
        B.sz = (nx + 2) * (ny + 2);
        off += (nx + 2) * (ny + 2) * (nz + 2);
        for (int d = 0; d < 3; ++d) {
            B.foff[d] = foff[d];
            const int64_t nf = (nx + (d == 0)) * (ny + (d == 1)) * (nz + (d == 2));
            foff[d] += nf;
            mbf = std::max(mbf, nf);
        }
        for (int f = 0; f < 6; ++f) { B.bct[f] = BC_INTERIOR; B.nxo[f] = 0; B.cf[f] = 0.0; for (double& c : B.bcc[f]) c = 0.0; }
        cell_start[(size_t)i + 1] = cell_start[(size_t)i] + nx * ny * nz;
        mbc = std::max(mbc, nx * ny * nz);
        mtiles[0] = std::max(mtiles[0], tiles_of((int)nx, (int)ny, (int)nz, kTileZ));
        mtiles[1] = std::max(mtiles[1], tiles_of((int)nx, (int)ny, (int)nz, kTileZRed));
        mbg = std::max(mbg, (nx + 2) * (ny + 2) * (nz + 2));
    }
    nalloc = off;
    for (int d = 0; d < 3; ++d) nface[d] = foff[d];
    // Ownership lookup: slabs bounded by every box edge in each direction.
    for (int d = 0; d < 3; ++d) {
        std::vector<int> cuts{0, dom.length(d)};
        for (int i = 0; i < nb; ++i) { cuts.push_back(hbox[(size_t)i].lo[d]); cuts.push_back(hbox[(size_t)i].hi[d] + 1); }
        std::sort(cuts.begin(), cuts.end());
        cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
        nslab[d] = (int)cuts.size() - 1;
        slab_of[d].assign((size_t)dom.length(d), 0);
        for (int s = 0; s < nslab[d]; ++s)
            for (int c = cuts[(size_t)s]; c < cuts[(size_t)s + 1]; ++c) slab_of[d][(size_t)c] = s;
        slab_lo[d].assign(cuts.begin(), cuts.end());
    }
    box_of.assign((size_t)nslab[0] * nslab[1] * nslab[2], -1);
    for (int i = 0; i < nb; ++i) {
        const BoxDev& B = hbox[(size_t)i];
        for (int s2 = slab_of[2][(size_t)B.lo[2]]; s2 <= slab_of[2][(size_t)B.hi[2]]; ++s2)
            for (int s1 = slab_of[1][(size_t)B.lo[1]]; s1 <= slab_of[1][(size_t)B.hi[1]]; ++s1)
                for (int s0 = slab_of[0][(size_t)B.lo[0]]; s0 <= slab_of[0][(size_t)B.hi[0]]; ++s0) {
                    int& o = box_of[(size_t)s0 + (size_t)nslab[0] * ((size_t)s1 + (size_t)nslab[1] * s2)];
                    if (o != -1) throw std::runtime_error("mgx: overlapping boxes in a BoxArray");
                    o = i;
                }
    }
    covered = std::find(box_of.begin(), box_of.end(), -1) == box_of.end();
    d_slab.clear();
    for (int d = 0; d < 3; ++d) {
        d_slab.emplace_back(new amrex::DeviceBuffer(sizeof(int) * slab_of[d].size()));
        MGX_HIP_CHECK(hipMemcpy(d_slab.back()->ptr, slab_of[d].data(), sizeof(int) * slab_of[d].size(), hipMemcpyHostToDevice));
    }
    d_boxof.reset(new amrex::DeviceBuffer(sizeof(int) * box_of.size()));
    MGX_HIP_CHECK(hipMemcpy(d_boxof->ptr, box_of.data(), sizeof(int) * box_of.size(), hipMemcpyHostToDevice));
    d_cellstart.reset(new amrex::DeviceBuffer(sizeof(int64_t) * cell_start.size()));
    MGX_HIP_CHECK(hipMemcpy(d_cellstart->ptr, cell_start.data(), sizeof(int64_t) * cell_start.size(), hipMemcpyHostToDevice));
    upload_boxes();
}

void LevelPlan::upload_boxes() {
    d_box.reset(new amrex::DeviceBuffer(sizeof(BoxDev) * std::max<size_t>(hbox.size(), 1)));
    MGX_HIP_CHECK(hipMemcpy(d_box->ptr, hbox.data(), sizeof(BoxDev) * hbox.size(), hipMemcpyHostToDevice));
    dev = LevelDev{};
    dev.box = static_cast<const BoxDev*>(d_box->ptr);
    dev.nbox = (int)hbox.size();
    for (int d = 0; d < 3; ++d) {
        dev.n[d] = geom.Domain().length(d);
        dev.periodic[d] = geom.isPeriodic(d) ? 1 : 0;
        dev.slab_of[d] = static_cast<const int*>(d_slab[(size_t)d]->ptr);
        dev.nslab[d] = nslab[d];
        dev.nface[d] = nface[d];
    }
    dev.box_of = static_cast<const int*>(d_boxof->ptr);
    dev.cell_start = static_cast<const int64_t*>(d_cellstart->ptr);
    dev.ncell = cell_start.back();
    dev.nalloc = nalloc;
    dev.mbc = mbc;
    dev.mbg = mbg;
    dev.mbf = mbf;
    dev.mtiles[0] = mtiles[0];
    dev.mtiles[1] = mtiles[1];
    dev.max_nxo = 0;
    dev.max_face = 0;
    for (const BoxDev& B : hbox) {
        const int64_t n0 = B.hi[0] - B.lo[0] + 1, n1 = B.hi[1] - B.lo[1] + 1, n2 = B.hi[2] - B.lo[2] + 1;
        dev.max_face = std::max(dev.max_face, std::max(n0 * n1, std::max(n1 * n2, n0 * n2)));
    }
    for (const BoxDev& B : hbox)
        for (int f = 0; f < 6; ++f) dev.max_nxo = std::max(dev.max_nxo, B.nxo[f]);
}

// poly_interp_coeff (AMReX_LOUtil_K.H:24-37)
static void lagrange(double xInt, const double* x, int N, double* c) {
    for (int j = 0; j < N; ++j) {
        double num = 1.0, den = 1.0;
        for (int i = 0; i < N; ++i)
            if (i != j) { num *= xInt - x[i]; den *= x[j] - x[i]; }
        c[j] = num / den;
    }
}

// Per box and face: BC type where the face lies on a non-periodic domain
// boundary (MLMGBndryT::setBoxBC, AMReX_MLMGBndry.H:105-150, bcl = 0), the
// Dirichlet closure weights for NX = min(length+1, maxorder)
// (AMReX_MLLinOp_K.H:96-100) and the GSRB weight of the first interior cell
// (mllinop_comp_interp_coef0_*, AMReX_MLLinOp_K.H:329-406).
void LevelPlan::set_bc(const int* lobc, const int* hibc, int maxorder, const CfFaces* cf) {
    const Box dom = geom.Domain();
    for (BoxDev& B : hbox) {
        B.cfmask = 0;
        for (int f = 0; f < 6; ++f) {
            const int d = f % 3, side = f / 3;
            const bool at_dom = side == 0 ? B.lo[d] == dom.smallEnd(d) : B.hi[d] == dom.bigEnd(d);
            B.bct[f] = BC_INTERIOR;
            B.nxo[f] = 0;
            B.cf[f] = 0.0;
            for (double& c : B.bcc[f]) c = 0.0;
            int bct = BC_INTERIOR;
            double x0 = -0.0;  // physical faces: values on the face (bcl = 0)
            if (at_dom && !geom.isPeriodic(d)) {
                bct = side == 0 ? lobc[d] : hibc[d];
            } else if (cf) {
                // a face of the covered region that no periodic image covers
                const Box& R = cf->region;
                const bool on_region = side == 0 ? B.lo[d] == R.smallEnd(d) : B.hi[d] == R.bigEnd(d);
                const bool spans = geom.isPeriodic(d) && R.length(d) == dom.length(d);
                if (on_region && !spans) {
                    bct = cf->type;
                    x0 = cf->x0;
                    B.cfmask |= 1 << f;
                }
            }
            if (bct == BC_INTERIOR) continue;
            if (bct != BC_DIRICHLET && bct != BC_NEUMANN && bct != BC_REFLECT_ODD)
                throw std::runtime_error("mgx: unsupported LinOpBCType " + std::to_string(bct));
            B.bct[f] = bct;
            if (bct == BC_DIRICHLET) {
                const int blen = B.hi[d] - B.lo[d] + 1;
                const int NX = std::min(blen + 1, maxorder);
                const double x[4] = {x0, 0.5, 1.5, 2.5};
                lagrange(-0.5, x, NX, B.bcc[f]);
                B.nxo[f] = NX;
                B.cf[f] = B.bcc[f][1];
            } else {
                B.cf[f] = 1.0;  // Neumann and reflect-odd both store 1 (AMReX_MLLinOp_K.H:336-349)
            }
        }
    }
    upload_boxes();
}

// ---------------------------------------------------------------------------
// MG hierarchy: MLLinOpT::defineGrids for one AMR level (AMReX_MLLinOp.H:857-1110)
// ---------------------------------------------------------------------------
std::vector<std::pair<Box, BoxArray>> define_grids(const Geometry& geom, const BoxArray& grids, const amrex::LPInfo& info,
                                                   std::vector<int>* agg_out) {
    std::vector<std::pair<Box, BoxArray>> out;
    std::vector<int> agg_lev{0};
    const Box dom = geom.Domain();
    out.push_back({dom, grids});
    const int agg = info.agg_grid_size > 0 ? info.agg_grid_size : amrex::LPInfo::getDefaultAgglomerationGridSize();
    const bool covered = grids.numPts() == dom.numPts();
    Box aggbox;
    bool aggable = false;
    if (grids.size() > 1 && info.do_agglomeration) {
        if (covered) { aggbox = dom; aggable = true; }
        else { aggbox = grids.minimalBox(); aggable = aggbox.numPts() == grids.numPts(); }
    }
    auto dir_coarsenable = [](const Box& b, int d) {
        // coarsenable by 2 along d only, minimum width 2 (refrat*min_width = 4)
        if (b.length(d) < 4) return false;
        return (b.smallEnd(d) % 2 == 0) && ((b.bigEnd(d) + 1) % 2 == 0);
    };
    if (aggable) {
        Box dbx = dom, bbx = aggbox;
        const double nbxs = (double)grids.size();
        const double thr = (double)agg * agg * agg;
        std::vector<Box> domainboxes{dbx}, boundboxes{bbx};
        std::vector<int> agg_flag{0}, acc{1};
        for (int lev = 0; lev < info.max_coarsening_level; ++lev) {
            bool ok = true;
            for (int d = 0; d < 3; ++d) ok = ok && dir_coarsenable(dbx, d) && dir_coarsenable(bbx, d);
            if (!ok) break;
            acc.push_back(acc.back() * 2);
            dbx.coarsen(2);
            bbx.coarsen(2);
            domainboxes.push_back(dbx);
            boundboxes.push_back(bbx);
            agg_flag.push_back((double)bbx.numPts() / nbxs < 0.999 * thr);
        }
        bool agged = false;
        for (size_t lev = 1; lev < domainboxes.size(); ++lev) {
            if (!agged && !agg_flag[lev] && grids.coarsenable(acc[lev], 2)) {
                BoxArray c = grids;
                c.coarsen(acc[lev]);
                out.push_back({domainboxes[lev], c});
                agg_lev.push_back(0);
            } else {
                if (!out.back().second.coarsenable(2)) break;
                BoxArray c(boundboxes[lev]);
                c.maxSize(agg);
                out.push_back({domainboxes[lev], c});
                agg_lev.push_back(1);
                agged = true;
            }
        }
    } else {
        BoxArray cur = grids;
        Box dbx = dom;
        for (int lev = 0; lev < info.max_coarsening_level; ++lev) {
            bool ok = true;
            for (int d = 0; d < 3; ++d) ok = ok && dir_coarsenable(dbx, d);
            ok = ok && cur.coarsenable(2, 2);
            if (!ok) break;
            dbx.coarsen(2);
            cur.coarsen(2);
            out.push_back({dbx, cur});
            agg_lev.push_back(0);
        }
    }
    if (agg_out) *agg_out = agg_lev;
    return out;
}

// ---------------------------------------------------------------------------
// Linear operator
// ---------------------------------------------------------------------------
template <class T>
static T* dptr(const std::unique_ptr<amrex::DeviceBuffer>& b) {
    return b ? static_cast<T*>(b->ptr) : nullptr;
}

static bool g_dense_default = true;
void set_dense_default(bool d) { g_dense_default = d; }

void LinOpImpl::define(const Geometry& g, const BoxArray& ba, const amrex::LPInfo& inf) {
    info = inf;
    dense_requested = g_dense_default;
    auto levels = define_grids(g, ba, info, &agglomerated);
    lev.resize(levels.size());
    // Levels that cover the domain are stored as one ghost-grown FAB (dense);
    // the reference BoxArray of each level is kept for the call surface.
    // Level-0 boxes that leave part of the domain uncovered (a patch solved
    // with coarse/fine boundary data) must form one box; they keep box storage.
    const bool covers = ba.numPts() == g.Domain().numPts();
    needs_cf = !covers;
    if (needs_cf) {
        cf_region = ba.minimalBox();
        if (cf_region.numPts() != ba.numPts())
            throw std::runtime_error("MLLinOp: level-0 boxes that do not cover the domain must form one box");
    }
    dense = covers && dense_requested;
    for (size_t l = 0; l < levels.size(); ++l) {
        Geometry gl(levels[l].first, &g.ProbDomain(), g.Coord(), g.isPeriodic().data());
        lev[l].plan.build(gl, dense ? BoxArray(levels[l].first) : levels[l].second);
        lev[l].plan.ref_ba = levels[l].second;
    }
    same_layout.assign(lev.size(), 0);
    for (size_t l = 1; l < lev.size(); ++l) {
        BoxArray c = lev[l - 1].plan.ba;
        c.coarsen(2);
        same_layout[l] = (c == lev[l].plan.ba) ? 1 : 0;
    }
    for (auto& L : lev) {
        L.a64.reset(new amrex::DeviceBuffer(sizeof(double) * (size_t)L.plan.nalloc));
        for (int d = 0; d < 3; ++d)
            L.b64[d].reset(new amrex::DeviceBuffer(sizeof(double) * (size_t)std::max<int64_t>(L.plan.nface[d], 1)));
        amrex::Gpu::gpuStream();
        set_val<double>(dptr<double>(L.a64), L.plan.nalloc, 1.0, amrex::Gpu::gpuStream());
        for (int d = 0; d < 3; ++d) set_val<double>(dptr<double>(L.b64[d]), L.plan.nface[d], 1.0, amrex::Gpu::gpuStream());
    }
    prepared = false;
}

// averageDownCoeffs, update_singular_flags, computeVolInv, BC tables.
// Fine AMR level of a composite operator: refinement ratio 2, boxes
// coarsenable by 2, anywhere in the domain.  A face ghost beyond a periodic
// domain face is the periodic image (another fine box or a coarse/fine face);
// beyond a physical face it takes that face's closure with the fine level's
// boundary data (setLevelBC(1, data)).
void LinOpImpl::define_fine(const Geometry& g, const BoxArray& ba) {
    const Geometry& g0 = lev[0].plan.geom;
    for (int d = 0; d < 3; ++d)
        if (g.Domain().length(d) != 2 * g0.Domain().length(d))
            throw std::runtime_error("MLLinOp: the AMR refinement ratio must be 2");
    if (!lev[0].plan.covered || lev[0].plan.dev.nbox != 1)
        throw std::runtime_error("MLLinOp: two AMR levels need the dense coarse level");
    for (const Box& b : ba.boxes())
        for (int d = 0; d < 3; ++d) {
            if (b.smallEnd(d) % 2 != 0 || b.length(d) % 2 != 0)
                throw std::runtime_error("MLLinOp: fine boxes must be coarsenable by 2");
            if (b.smallEnd(d) < 0 || b.bigEnd(d) > g.Domain().bigEnd(d))
                throw std::runtime_error("MLLinOp: fine boxes must lie inside the domain");
        }
    namr = 2;
    fine.reset(new AmrFine);
    AmrFine& F = *fine;
    F.plan.build(g, ba);
    F.plan.ref_ba = ba;
    for (int m = 0; m < 5; ++m) for (int c = 0; c < 4; ++c) F.cc.c[m][c] = 0.0;
    for (int m = 0; m < 5; ++m) for (int c = 0; c < 4; ++c) F.cc.p[m][c] = 0.0;
    for (int NX = 1; NX <= 4; ++NX) {
        const double x[4] = {-1.0, 0.5, 1.5, 2.5};  // bcl = 0.5 * ratio * dx_f
        lagrange(-0.5, x, NX, F.cc.c[NX]);
        const double xp[4] = {-0.0, 0.5, 1.5, 2.5};  // physical face: bcl = 0
        lagrange(-0.5, xp, NX, F.cc.p[NX]);
    }
    for (int f = 0; f < 6; ++f) F.cc.bct[f] = BC_INTERIOR;  // physical types: apply_bc
    // ABecLaplacian coefficients of the fine boxes (unit until set)
    hipStream_t s0 = amrex::Gpu::gpuStream();
    F.a64.reset(new amrex::DeviceBuffer(sizeof(double) * (size_t)F.plan.nalloc));
    set_val<double>(dptr<double>(F.a64), F.plan.nalloc, 1.0, s0);
    for (int d = 0; d < 3; ++d) {
        F.b64[d].reset(new amrex::DeviceBuffer(sizeof(double) * (size_t)std::max<int64_t>(F.plan.nface[d], 1)));
        set_val<double>(dptr<double>(F.b64[d]), F.plan.nface[d], 1.0, s0);
    }
    // coarse cells: covered by the fine level, and the ones next to it (reflux)
    const int* n = lev[0].plan.dev.n;
    const int64_t nc = (int64_t)n[0] * n[1] * n[2];
    std::vector<unsigned char> cov((size_t)nc + 1, 0), near((size_t)nc + 1, 0);  // entry nc: beyond a physical face
    std::vector<int> cbox((size_t)nc + 1, -1);
    // coarse cell of a position (periodic image), or nc beyond a physical face
    auto at = [&](int i, int j, int k) -> size_t {
        const int p[3] = {i, j, k};
        for (int d = 0; d < 3; ++d)
            if (!g0.isPeriodic(d) && (p[d] < 0 || p[d] >= n[d])) return (size_t)nc;
        i = ((i % n[0]) + n[0]) % n[0];
        j = ((j % n[1]) + n[1]) % n[1];
        k = ((k % n[2]) + n[2]) % n[2];
        return (size_t)(i + (int64_t)n[0] * (j + (int64_t)n[1] * k));
    };
    for (int q = 0; q < (int)ba.size(); ++q) {
        const Box b = ba[q];
        for (int K = b.smallEnd(2) / 2 - 1; K <= b.bigEnd(2) / 2 + 1; ++K)
            for (int J = b.smallEnd(1) / 2 - 1; J <= b.bigEnd(1) / 2 + 1; ++J)
                for (int I = b.smallEnd(0) / 2 - 1; I <= b.bigEnd(0) / 2 + 1; ++I) near[at(I, J, K)] = 1;
        for (int K = b.smallEnd(2) / 2; K <= b.bigEnd(2) / 2; ++K)
            for (int J = b.smallEnd(1) / 2; J <= b.bigEnd(1) / 2; ++J)
                for (int I = b.smallEnd(0) / 2; I <= b.bigEnd(0) / 2; ++I) {
                    cov[at(I, J, K)] = 1;
                    cbox[at(I, J, K)] = q;
                }
    }
    F.hreflux.clear();
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const size_t c = at(i, j, k);
                if (!near[c] || cov[c]) continue;
                RefluxCell R{i, j, k, {-1, -1, -1, -1, -1, -1}};
                bool any = false;
                const int p[3] = {i, j, k};
                for (int f = 0; f < 6; ++f) {
                    int q[3] = {p[0], p[1], p[2]};
                    q[f % 3] += f < 3 ? -1 : 1;
                    const size_t cq = at(q[0], q[1], q[2]);
                    if (cov[cq]) { R.fbox[f] = cbox[cq]; any = true; }
                }
                if (any) F.hreflux.push_back(R);
            }
    F.d_reflux.reset(new amrex::DeviceBuffer(sizeof(RefluxCell) * std::max<size_t>(F.hreflux.size(), 1)));
    if (!F.hreflux.empty())
        MGX_HIP_CHECK(hipMemcpy(F.d_reflux->ptr, F.hreflux.data(), sizeof(RefluxCell) * F.hreflux.size(), hipMemcpyHostToDevice));
    F.d_covered.reset(new amrex::DeviceBuffer((size_t)nc));
    MGX_HIP_CHECK(hipMemcpy(F.d_covered->ptr, cov.data(), (size_t)nc, hipMemcpyHostToDevice));
}

// Boundary types, closure weights and coarse/fine face flags of every MG level.
void LinOpImpl::apply_bc() {
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) { lo[d] = (int)lobc[d]; hi[d] = (int)hibc[d]; }
    if (fine) {
        // the fine level's closures beyond physical faces
        const Geometry& g0 = lev[0].plan.geom;
        for (int f = 0; f < 6; ++f) {
            const int d = f % 3;
            fine->cc.bct[f] = g0.isPeriodic(d) ? BC_INTERIOR : (f < 3 ? lo[d] : hi[d]);
        }
    }
    for (size_t l = 0; l < lev.size(); ++l) {
        auto& P = lev[l].plan;
        if (!needs_cf) {
            P.set_bc(lo, hi, maxorder);
            continue;
        }
        // c/f location 0.5*ratio*dx of level 0 in this level's cells
        // (MLMGBndry.H:138-145 with the level-0 dx, MLCellLinOp.H:586-592)
        const double scale = (double)P.geom.Domain().length(0) / (double)lev[0].plan.geom.Domain().length(0);
        CfFaces cf{P.ref_ba.minimalBox(), cf_type, cf_ratio > 0 ? -0.5 * cf_ratio * scale : -0.0};
        P.set_bc(lo, hi, maxorder, &cf);
    }
}

void LinOpImpl::prepare() {
    hipStream_t s = amrex::Gpu::gpuStream();
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) { lo[d] = (int)lobc[d]; hi[d] = (int)hibc[d]; }
    if (dense && maxorder > 3) {
        // NX = min(box length + 1, maxorder) uses the reference boxes; dense
        // storage sees the domain length instead, which agrees unless a box
        // shorter than maxorder-1 touches a Dirichlet face.
        for (auto& L : lev)
            for (const Box& b : L.plan.ref_ba.boxes())
                for (int d = 0; d < 3; ++d)
                    if (b.length(d) + 1 < maxorder && L.plan.geom.Domain().length(d) + 1 >= maxorder &&
                        !L.plan.geom.isPeriodic(d) && (lo[d] == BC_DIRICHLET || hi[d] == BC_DIRICHLET))
                        throw std::runtime_error("maxorder 4 with boxes shorter than 3 cells needs box storage "
                                                 "(mgx_new_abeclaplacian/mgx_new_poisson with dense = 0)");
    }
    apply_bc();
    // averageDownCoeffsToCoarseAmrLevel (AMReX_MLABecLaplacian.H:694-712): the
    // coarse AMR level's a (when alpha != 0) and b under the fine boxes, b on
    // their boundary faces too, from the fine level's coefficients
    if (namr == 2 && kind == 0) {
        const auto& PF = fine->plan;
        const auto& P0 = lev[0].plan;
        if (a_scalar != 0.0) fine_avgdown<double>(PF.dev, P0.dev, dptr<double>(lev[0].a64), dptr<double>(fine->a64), s);
        for (int d = 0; d < 3; ++d)
            fine_avgdown_faces(PF.dev, P0.dev, PF.dev.nbox, d, dptr<double>(lev[0].b64[d]), dptr<double>(fine->b64[d]), s);
    }
    for (size_t l = 1; l < lev.size(); ++l) {
        auto& C = lev[l];
        auto& F = lev[l - 1];
        if (kind != 0) break;
        if (a_scalar == 0.0) set_val<double>(dptr<double>(C.a64), C.plan.nalloc, 0.0, s);
        else restrict_avg<double>(C.plan.dev, (int)C.plan.mbc, F.plan.dev, dptr<double>(C.a64), dptr<double>(F.a64), same_layout[l], s);
        for (int d = 0; d < 3; ++d)
            avgdown_faces<double>(C.plan.dev, F.plan.dev, dptr<double>(C.b64[d]), dptr<double>(F.b64[d]), d, same_layout[l], s);
    }
    // FP32 copies of the coefficients for the mixed-precision V-cycle
    for (auto& L : lev) {
        if (!L.a32) {
            L.a32.reset(new amrex::DeviceBuffer(sizeof(float) * (size_t)L.plan.nalloc));
            for (int d = 0; d < 3; ++d)
                L.b32[d].reset(new amrex::DeviceBuffer(sizeof(float) * (size_t)std::max<int64_t>(L.plan.nface[d], 1)));
        }
        convert<float, double>(dptr<float>(L.a32), dptr<double>(L.a64), L.plan.nalloc, s);
        for (int d = 0; d < 3; ++d) convert<float, double>(dptr<float>(L.b32[d]), dptr<double>(L.b64[d]), L.plan.nface[d], s);
    }
    if (namr == 2 && kind == 0) {
        AmrFine& F = *fine;
        if (!F.a32) {
            F.a32.reset(new amrex::DeviceBuffer(sizeof(float) * (size_t)F.plan.nalloc));
            for (int d = 0; d < 3; ++d)
                F.b32[d].reset(new amrex::DeviceBuffer(sizeof(float) * (size_t)std::max<int64_t>(F.plan.nface[d], 1)));
        }
        convert<float, double>(dptr<float>(F.a32), dptr<double>(F.a64), F.plan.nalloc, s);
        for (int d = 0; d < 3; ++d) convert<float, double>(dptr<float>(F.b32[d]), dptr<double>(F.b64[d]), F.plan.nface[d], s);
    }
    // singular flags (MLABecLaplacian.H:716-773, MLPoisson.H:140-178): without
    // Dirichlet domain faces a level covering the domain is singular when
    // a = 0; with coarse/fine data of Neumann type, when the level-0 region
    // grown at the Dirichlet domain faces stays inside the domain
    bool dirichlet = false;
    for (int d = 0; d < 3; ++d)
        if (!lev[0].plan.geom.isPeriodic(d) && (lobc[d] == amrex::LinOpBCType::Dirichlet || hibc[d] == amrex::LinOpBCType::Dirichlet))
            dirichlet = true;
    auto a_is_zero = [&]() -> bool {
        if (kind == 1 || a_scalar == 0.0) return true;
        // sum and max of the level-0 a coefficient
        DeviceScratch& sc = scratch(lev[0].plan);
        reduce<double>(lev[0].plan.dev, (int)lev[0].plan.mbc, 1, dptr<double>(lev[0].a64), nullptr, sc.partials(), sc.scalar(0), s);
        reduce<double>(lev[0].plan.dev, (int)lev[0].plan.mbc, 0, dptr<double>(lev[0].a64), nullptr, sc.partials(), sc.scalar(1), s);
        double h[2];
        MGX_HIP_CHECK(hipMemcpyAsync(h, sc.scalar(0), 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        MGX_HIP_CHECK(hipStreamSynchronize(s));
        return std::abs(h[0]) <= h[1] * 1.e-12;
    };
    singular = false;
    if (!dirichlet && !needs_cf) singular = a_is_zero();
    if (!singular && needs_cf && cf_type == BC_NEUMANN) {
        Box bb = cf_region;
        for (int d = 0; d < 3; ++d) {
            if (lobc[d] == amrex::LinOpBCType::Dirichlet) bb.growLo(d, 1);
            if (hibc[d] == amrex::LinOpBCType::Dirichlet) bb.growHi(d, 1);
        }
        if (lev[0].plan.geom.Domain().contains(bb)) singular = a_is_zero();
    }
    volinv0 = 1.0 / (double)lev.front().plan.geom.Domain().numPts();
    volinv_bot = 1.0 / (double)lev.back().plan.geom.Domain().numPts();
    ++epoch;
    prepared = true;
}

// InterpBndryData::setBndryValues from the coarse MultiFab (MLCellLinOp.H:
// 535-568): the coarse cells next to the level-0 region (coarsened by the
// ratio, grown by one for the tangential stencil) are gathered into a window,
// then every coarse/fine face ghost of the level-0 boxes is interpolated.
void LinOpImpl::fill_cf_values() {
    const auto& P = lev[0].plan;
    const int r = cf_ratio;
    if (r < 1) throw std::runtime_error("MLLinOp::setCoarseFineBC: the coarsening ratio must be positive");
    const Box dom = P.geom.Domain();
    for (int d = 0; d < 3; ++d)
        if (cf_region.smallEnd(d) % r != 0 || (cf_region.bigEnd(d) + 1) % r != 0 || dom.length(d) % r != 0)
            throw std::runtime_error("MLLinOp::setCoarseFineBC: the level-0 boxes and the domain must be coarsenable by the ratio");
    Box w = amrex::coarsen(cf_region, r);
    w.grow(1);
    int wlo[3], wn[3], nc[3], per[3], rlo[3], rhi[3];
    for (int d = 0; d < 3; ++d) {
        wlo[d] = w.smallEnd(d);
        wn[d] = w.length(d);
        nc[d] = dom.length(d) / r;
        per[d] = P.geom.isPeriodic(d) ? 1 : 0;
        rlo[d] = cf_region.smallEnd(d);
        rhi[d] = cf_region.bigEnd(d);
    }
    // the coarse/fine faces of the level-0 boxes (set_bc) must be known now:
    // the values are interpolated here, before the operator is prepared
    if (!bc_set) throw std::runtime_error("MLLinOp::setLevelBC: setDomainBC must be called first");
    apply_bc();
    hipStream_t s = amrex::Gpu::gpuStream();
    amrex::DeviceBuffer win(sizeof(double) * (size_t)w.numPts());
    const amrex::MultiFab& C = *cf_crse;
    cf_window(C.dataPtr(), static_cast<const UserBox*>(C.devBoxes()), (int)C.size(), C.nGrow(), wlo, wn, nc, per,
              static_cast<double*>(win.ptr), s);
    cf_bndry(P.dev, static_cast<const double*>(win.ptr), wlo, wn, r, rlo, rhi, P.dev.max_face,
             static_cast<double*>(bcv->ptr), s);
    MGX_HIP_CHECK(hipStreamSynchronize(s));  // the window is freed on return
}

DeviceScratch& LinOpImpl::scratch(const LevelPlan& P) {
    if (!m_scratch) m_scratch.reset(new DeviceScratch);
    m_scratch->ensure(reduce_partials(P.dev, (int)P.mbc));
    return *m_scratch;
}

void DeviceScratch::ensure(int64_t npart) {
    if (!buf || npart + 64 > cap) {
        cap = std::max<int64_t>(npart + 64, 4096);
        buf.reset(new amrex::DeviceBuffer(sizeof(double) * (size_t)cap));
    }
}

namespace {
// Operator coefficients of one level in precision T: the scalars are stored in
// the solve precision first (RT m_a_scalar, m_b_scalar).
template <class T>
OpCoef<T> level_coef(int kind, double a_scalar, double b_scalar, const Geometry& geom, const amrex::DeviceBuffer* a64,
                     const std::unique_ptr<amrex::DeviceBuffer>* b64, const amrex::DeviceBuffer* a32,
                     const std::unique_ptr<amrex::DeviceBuffer>* b32) {
    OpCoef<T> op{};
    op.kind = kind;
    const T alpha = static_cast<T>(kind == 0 ? a_scalar : 0.0);
    const T beta = static_cast<T>(kind == 0 ? b_scalar : 1.0);
    op.alpha = alpha;
    for (int d = 0; d < 3; ++d) {
        const double h = geom.CellSize(d), hi = geom.InvCellSize(d);
        op.dhs[d] = beta / static_cast<T>(h * h);
        const T dxi = static_cast<T>(hi);
        op.dha[d] = beta * dxi * dxi;
        op.dpo[d] = static_cast<T>(hi * hi);
    }
    if constexpr (std::is_same<T, double>::value) {
        op.a = a64 ? static_cast<const double*>(a64->ptr) : nullptr;
        for (int d = 0; d < 3; ++d) op.b[d] = b64[d] ? static_cast<const double*>(b64[d]->ptr) : nullptr;
    } else {
        op.a = a32 ? static_cast<const float*>(a32->ptr) : nullptr;
        for (int d = 0; d < 3; ++d) op.b[d] = b32[d] ? static_cast<const float*>(b32[d]->ptr) : nullptr;
    }
    return op;
}
}  // namespace

template <class T>
OpCoef<T> LinOpImpl::coef(int l) const {
    const auto& L = lev[(size_t)l];
    return level_coef<T>(kind, a_scalar, b_scalar, L.plan.geom, L.a64.get(), L.b64, L.a32.get(), L.b32);
}
template OpCoef<double> LinOpImpl::coef<double>(int) const;
template OpCoef<float> LinOpImpl::coef<float>(int) const;

template <class T>
OpCoef<T> LinOpImpl::fine_coef() const {
    const AmrFine& F = *fine;
    return level_coef<T>(kind, a_scalar, b_scalar, F.plan.geom, F.a64.get(), F.b64, F.a32.get(), F.b32);
}
template OpCoef<double> LinOpImpl::fine_coef<double>() const;
template OpCoef<float> LinOpImpl::fine_coef<float>() const;

RefluxCoef LinOpImpl::reflux_coef() const {
    RefluxCoef rc{1.0, {nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    if (kind == 0) {
        rc.scale = -b_scalar;
        for (int d = 0; d < 3; ++d) {
            rc.bc[d] = dptr<double>(lev[0].b64[d]);
            rc.bf[d] = dptr<double>(fine->b64[d]);
        }
    }
    return rc;
}

}  // namespace mgx

namespace amrex {

MLLinOp::MLLinOp(const Vector<Geometry>& a_geom, const Vector<BoxArray>& a_grids, const Vector<DistributionMapping>&,
                 const LPInfo& a_info, int kind)
    : m_impl(new mgx::LinOpImpl) {
    if (a_geom.empty() || a_geom.size() > 2 || a_grids.size() != a_geom.size())
        throw std::runtime_error("MLLinOp: one or two AMR levels are supported");
    m_impl->kind = kind;
    m_impl->define(a_geom[0], a_grids[0], a_info);
    if (a_geom.size() == 2) m_impl->define_fine(a_geom[1], a_grids[1]);
}
MLLinOp::~MLLinOp() = default;

void MLLinOp::setDomainBC(const std::array<LinOpBCType, 3>& lobc, const std::array<LinOpBCType, 3>& hibc) {
    const Geometry& g = m_impl->lev[0].plan.geom;
    for (int d = 0; d < 3; ++d) {
        // periodic directions must say Periodic, non-periodic ones must not (MLLinOp.H:1191-1239)
        const bool per = (lobc[d] == LinOpBCType::Periodic);
        if (per != g.isPeriodic(d) || (hibc[d] == LinOpBCType::Periodic) != g.isPeriodic(d))
            throw std::runtime_error("MLLinOp::setDomainBC: periodicity mismatch with Geometry");
    }
    m_impl->lobc = lobc;
    m_impl->hibc = hibc;
    m_impl->bc_set = true;
    m_impl->prepared = false;
}
void MLLinOp::setCoarseFineBC(const MultiFab* crse, int crse_ratio, LinOpBCType bc_type) {
    if (bc_type != LinOpBCType::Dirichlet && bc_type != LinOpBCType::Neumann)
        throw std::runtime_error("MLLinOp::setCoarseFineBC: the coarse/fine BC type must be Dirichlet or Neumann");
    if (crse && crse->nComp() < 1) throw std::runtime_error("MLLinOp::setCoarseFineBC: empty coarse MultiFab");
    m_impl->cf_crse = crse;
    m_impl->cf_ratio = crse_ratio;
    m_impl->cf_type = (int)bc_type;
    m_impl->prepared = false;
}
void MLLinOp::setLevelBC(int amrlev, const MultiFab* mf) {
    if (amrlev == 1 && m_impl->namr == 2) {
        // the fine level's Dirichlet values beyond physical faces, read from the
        // ghost cells of mf there (InterpBndryData::setBndryValues takes the
        // physical faces from the level's own data, AMReX_InterpBndryData.H:161-260)
        auto& F = *m_impl->fine;
        if (mf && !one_comp_on(*mf, F.plan.ref_ba))
            throw std::runtime_error("MLLinOp::setLevelBC: BoxArray mismatch (one component on the fine level's BoxArray expected)");
        m_impl->prepared = false;
        bool on_face = false;  // a fine box touches a physical face
        const Box dom = F.plan.geom.Domain();
        for (const Box& b : F.plan.ref_ba.boxes())
            for (int d = 0; d < 3; ++d)
                if (!F.plan.geom.isPeriodic(d) && (b.smallEnd(d) == dom.smallEnd(d) || b.bigEnd(d) == dom.bigEnd(d)))
                    on_face = true;
        if (!mf || !on_face) { F.d_pbc.reset(); return; }
        if (mf->nGrow() < 1) throw std::runtime_error("MLLinOp::setLevelBC: the fine level's MultiFab needs one ghost cell");
        if (!F.d_pbc) F.d_pbc.reset(new DeviceBuffer(sizeof(double) * (size_t)F.plan.nalloc));
        mgx::set_val<double>(static_cast<double*>(F.d_pbc->ptr), F.plan.nalloc, 0.0, Gpu::gpuStream());
        mgx::copy_user_grown_to_level<double, double>(F.plan.dev, static_cast<double*>(F.d_pbc->ptr), mf->dataPtr(),
                                                      static_cast<const mgx::UserBox*>(mf->devBoxes()), (int)mf->size(),
                                                      mf->maxGrownCells(), mf->nGrow(), Gpu::gpuStream());
        return;
    }
    if (amrlev != 0) throw std::runtime_error("MLLinOp::setLevelBC: bad AMR level");
    auto& I = *m_impl;
    auto& P = I.lev[0].plan;
    I.prepared = false;  // solvers drop captured work that refers to the old boundary data
    // coarse/fine values are read from the coarse data now (MLCellLinOp.H:535-568)
    const bool cf_values = I.needs_cf && I.cf_crse && I.cf_type == (int)LinOpBCType::Dirichlet;
    if (!mf && !cf_values) { I.bcv.reset(); return; }
    if (mf && mf->nGrow() < 1) throw std::runtime_error("MLLinOp::setLevelBC: the MultiFab needs one ghost cell");
    if (mf && !one_comp_on(*mf, P.ref_ba)) throw std::runtime_error("MLLinOp::setLevelBC: BoxArray mismatch (one component on the operator's BoxArray expected)");
    if (!I.bcv) I.bcv.reset(new DeviceBuffer(sizeof(double) * (size_t)P.nalloc));
    mgx::set_val<double>(static_cast<double*>(I.bcv->ptr), P.nalloc, 0.0, Gpu::gpuStream());
    if (mf)
        mgx::copy_user_grown_to_level<double, double>(P.dev, static_cast<double*>(I.bcv->ptr), mf->dataPtr(),
                                                      static_cast<const mgx::UserBox*>(mf->devBoxes()), (int)mf->size(),
                                                      mf->maxGrownCells(), mf->nGrow(), Gpu::gpuStream());
    if (cf_values) I.fill_cf_values();
}
void MLLinOp::setMaxOrder(int o) {
    if (o < 2 || o > 4) throw std::runtime_error("MLLinOp::setMaxOrder: order must be 2, 3 or 4");
    m_impl->maxorder = o;
    m_impl->prepared = false;
}
int MLLinOp::getMaxOrder() const { return m_impl->maxorder; }
void MLLinOp::setEnforceSingularSolvable(bool o) { m_impl->enforce_solvable = o; }
bool MLLinOp::getEnforceSingularSolvable() const { return m_impl->enforce_solvable; }
void MLLinOp::setGaussSeidel(bool flag) {
    if (!flag) throw std::runtime_error("MLLinOp::setGaussSeidel(false): only red-black Gauss-Seidel is implemented");
}
int MLLinOp::NAMRLevels() const { return m_impl->namr; }
int MLLinOp::NMGLevels(int amrlev) const { return amrlev == 0 ? (int)m_impl->lev.size() : 1; }
const Geometry& MLLinOp::Geom(int, int mglev) const { return m_impl->lev.at((size_t)mglev).plan.geom; }
const BoxArray& MLLinOp::boxArray(int, int mglev) const { return m_impl->lev.at((size_t)mglev).plan.ref_ba; }
bool MLLinOp::isSingular(int) const { return m_impl->singular; }

MLABecLaplacian::MLABecLaplacian(const Vector<Geometry>& g, const Vector<BoxArray>& b,
                                 const Vector<DistributionMapping>& dm, const LPInfo& info, int ncomp)
    : MLLinOp(g, b, dm, info, 0) {
    if (ncomp != 1) throw std::runtime_error("MLABecLaplacian: ncomp must be 1");
}
void MLABecLaplacian::setScalars(Real a, Real b) {
    m_impl->a_scalar = a;
    m_impl->b_scalar = b;
    m_impl->prepared = false;
}
namespace {
// The coefficient arrays of AMR level amrlev (0: the MG hierarchy's first
// level, 1: the fine level of a two-level operator).
struct CoefTarget {
    const mgx::LevelPlan* plan;
    double* a;
    double* b[3];
};
CoefTarget coef_target(mgx::LinOpImpl& I, int amrlev, const char* who) {
    if (amrlev == 0) {
        auto& L = I.lev[0];
        return {&L.plan, static_cast<double*>(L.a64->ptr),
                {static_cast<double*>(L.b64[0]->ptr), static_cast<double*>(L.b64[1]->ptr), static_cast<double*>(L.b64[2]->ptr)}};
    }
    if (amrlev == 1 && I.namr == 2) {
        auto& F = *I.fine;
        return {&F.plan, static_cast<double*>(F.a64->ptr),
                {static_cast<double*>(F.b64[0]->ptr), static_cast<double*>(F.b64[1]->ptr), static_cast<double*>(F.b64[2]->ptr)}};
    }
    throw std::runtime_error(std::string(who) + ": bad AMR level");
}
}  // namespace

void MLABecLaplacian::setACoeffs(int amrlev, const MultiFab& alpha) {
    const CoefTarget t = coef_target(*m_impl, amrlev, "setACoeffs");
    if (!one_comp_on(alpha, t.plan->ref_ba)) throw std::runtime_error("setACoeffs: BoxArray mismatch (one component on the operator's BoxArray expected)");
    mgx::copy_user_to_level<double, double>(t.plan->dev, t.a, alpha.dataPtr(),
                                            static_cast<const mgx::UserBox*>(alpha.devBoxes()), (int)alpha.size(),
                                            alpha.maxBoxCells(), alpha.nGrow(), Gpu::gpuStream());
    m_impl->prepared = false;
}
void MLABecLaplacian::setACoeffs(int amrlev, Real alpha) {
    const CoefTarget t = coef_target(*m_impl, amrlev, "setACoeffs");
    mgx::set_val<double>(t.a, t.plan->nalloc, alpha, Gpu::gpuStream());
    m_impl->prepared = false;
}
void MLABecLaplacian::setBCoeffs(int amrlev, const std::array<MultiFab const*, 3>& beta) {
    const CoefTarget t = coef_target(*m_impl, amrlev, "setBCoeffs");
    for (int d = 0; d < 3; ++d) {
        if (!one_comp_on(*beta[d], convert(t.plan->ref_ba, IntVect::TheDimensionVector(d))))
            throw std::runtime_error("setBCoeffs: face BoxArray mismatch (one component on the operator's BoxArray expected)");
        mgx::copy_user_faces_to_level<double, double>(t.plan->dev, t.b[d], beta[d]->dataPtr(),
                                                      static_cast<const mgx::UserBox*>(beta[d]->devBoxes()),
                                                      (int)beta[d]->size(), beta[d]->maxBoxCells(), beta[d]->nGrow(), d,
                                                      Gpu::gpuStream());
    }
    m_impl->prepared = false;
}
void MLABecLaplacian::setBCoeffs(int amrlev, Real beta) {
    const CoefTarget t = coef_target(*m_impl, amrlev, "setBCoeffs");
    for (int d = 0; d < 3; ++d) mgx::set_val<double>(t.b[d], t.plan->nface[d], beta, Gpu::gpuStream());
    m_impl->prepared = false;
}

void MLABecLaplacian::compCoeffGradient(int amrlev, const MultiFab& lam, const MultiFab& phi, MultiFab* grad_a,
                                        const std::array<MultiFab*, 3>& grad_b, bool accumulate) {
    if (amrlev != 0) throw std::runtime_error("compCoeffGradient: single AMR level");
    auto& I = *m_impl;
    if (std::isnan(I.a_scalar) || std::isnan(I.b_scalar))
        throw std::runtime_error("compCoeffGradient: setScalars must be called first");
    if (!I.prepared) I.prepare();  // boundary closure tables of level 0
    auto& P = I.lev[0].plan;
    if (I.needs_cf) throw std::runtime_error("compCoeffGradient: not available with setCoarseFineBC");
    if (!one_comp_on(lam, P.ref_ba) || !one_comp_on(phi, P.ref_ba))
        throw std::runtime_error("compCoeffGradient: BoxArray mismatch (one component on the operator's BoxArray expected)");
    if (grad_a && !one_comp_on(*grad_a, P.ref_ba)) throw std::runtime_error("compCoeffGradient: grad_a BoxArray mismatch (one component on the operator's BoxArray expected)");
    for (int d = 0; d < 3; ++d)
        if (grad_b[d] && !one_comp_on(*grad_b[d], convert(P.ref_ba, IntVect::TheDimensionVector(d))))
            throw std::runtime_error("compCoeffGradient: face BoxArray mismatch (one component on the operator's BoxArray expected)");
    hipStream_t s = Gpu::gpuStream();
    if (!I.grad_lam) {
        I.grad_lam.reset(new DeviceBuffer(sizeof(double) * (size_t)P.nalloc));
        I.grad_phi.reset(new DeviceBuffer(sizeof(double) * (size_t)P.nalloc));
        mgx::set_val<double>(static_cast<double*>(I.grad_lam->ptr), P.nalloc, 0.0, s);
        mgx::set_val<double>(static_cast<double*>(I.grad_phi->ptr), P.nalloc, 0.0, s);
    }
    double* l = static_cast<double*>(I.grad_lam->ptr);
    double* f = static_cast<double*>(I.grad_phi->ptr);
    mgx::copy_user_to_level<double, double>(P.dev, l, lam.dataPtr(), static_cast<const mgx::UserBox*>(lam.devBoxes()),
                                            (int)lam.size(), lam.maxBoxCells(), lam.nGrow(), s);
    mgx::copy_user_to_level<double, double>(P.dev, f, phi.dataPtr(), static_cast<const mgx::UserBox*>(phi.devBoxes()),
                                            (int)phi.size(), phi.maxBoxCells(), phi.nGrow(), s);
    const double* bcv = I.bcv ? static_cast<const double*>(I.bcv->ptr) : nullptr;
    const mgx::OpCoef<double> op = I.coef<double>(0);
    if (grad_a)
        mgx::grad_a(P.dev, op.alpha, l, f, static_cast<const mgx::UserBox*>(grad_a->devBoxes()), (int)grad_a->size(),
                    grad_a->maxBoxCells(), grad_a->nGrow(), grad_a->dataPtr(), accumulate ? 1 : 0, s);
    for (int d = 0; d < 3; ++d)
        if (grad_b[d])
            mgx::grad_b(P.dev, op, l, f, bcv, bcv ? 1 : 0, static_cast<const mgx::UserBox*>(grad_b[d]->devBoxes()),
                        (int)grad_b[d]->size(), grad_b[d]->maxBoxCells(), grad_b[d]->nGrow(), d, grad_b[d]->dataPtr(),
                        accumulate ? 1 : 0, s);
    MGX_HIP_CHECK(hipGetLastError());
    MGX_HIP_CHECK(hipStreamSynchronize(s));
}

MLPoisson::MLPoisson(const Vector<Geometry>& g, const Vector<BoxArray>& b, const Vector<DistributionMapping>& dm,
                     const LPInfo& info)
    : MLLinOp(g, b, dm, info, 1) {}

}  // namespace amrex
