// Minimal MI355X-native equivalents of the AMReX base types used by the MLMG
// call surface: IntVect, Box, BoxArray, DistributionMapping, RealBox,
// Geometry and FabArray (MultiFab / fMultiFab) with device storage.
//
// Names and argument meaning follow the reference (Src/Base/AMReX_Box.H,
// AMReX_BoxArray.H, AMReX_Geometry.H, AMReX_MultiFab.H) so caller code ports
// unchanged; the implementation is this library's own.  One process drives
// one GPU, so DistributionMapping is trivial and every box is local.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#define MGX_HIP_CHECK(expr)                                                                         \
    do {                                                                                            \
        hipError_t mgx_err_ = (expr);                                                               \
        if (mgx_err_ != hipSuccess)                                                                 \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(mgx_err_) +       \
                                     " at " __FILE__ ":" + std::to_string(__LINE__));               \
    } while (0)

namespace amrex {

using Real = double;
using Long = long long;
template <class T>
using Vector = std::vector<T>;

struct IntVect {
    int v[3] = {0, 0, 0};
    IntVect() = default;
    explicit IntVect(int a) : v{a, a, a} {}
    IntVect(int a, int b, int c) : v{a, b, c} {}
    explicit IntVect(const int* p) : v{p[0], p[1], p[2]} {}
    int& operator[](int d) { return v[d]; }
    int operator[](int d) const { return v[d]; }
    bool operator==(const IntVect& o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2]; }
    bool operator!=(const IntVect& o) const { return !(*this == o); }
    static IntVect TheDimensionVector(int d) { IntVect r; r.v[d] = 1; return r; }
    static IntVect TheZeroVector() { return IntVect(0); }
};

// Box with an index type: nodal[d] = 1 for face/node centring in d.
class Box {
  public:
    Box() = default;
    Box(const IntVect& lo, const IntVect& hi, const IntVect& nodal = IntVect(0)) : m_lo(lo), m_hi(hi), m_nodal(nodal) {}
    const IntVect& smallEnd() const { return m_lo; }
    const IntVect& bigEnd() const { return m_hi; }
    int smallEnd(int d) const { return m_lo[d]; }
    int bigEnd(int d) const { return m_hi[d]; }
    const IntVect& ixType() const { return m_nodal; }
    int length(int d) const { return m_hi[d] - m_lo[d] + 1; }
    IntVect length() const { return IntVect(length(0), length(1), length(2)); }
    Long numPts() const { return (Long)length(0) * length(1) * length(2); }
    bool ok() const { return length(0) > 0 && length(1) > 0 && length(2) > 0; }
    bool contains(int i, int j, int k) const {
        return i >= m_lo[0] && i <= m_hi[0] && j >= m_lo[1] && j <= m_hi[1] && k >= m_lo[2] && k <= m_hi[2];
    }
    bool contains(const Box& b) const { return contains(b.m_lo[0], b.m_lo[1], b.m_lo[2]) && contains(b.m_hi[0], b.m_hi[1], b.m_hi[2]); }
    bool operator==(const Box& o) const { return m_lo == o.m_lo && m_hi == o.m_hi && m_nodal == o.m_nodal; }
    bool operator!=(const Box& o) const { return !(*this == o); }
    Box& coarsen(int r);
    Box& refine(int r);
    Box& grow(int n) { for (int d = 0; d < 3; ++d) { m_lo[d] -= n; m_hi[d] += n; } return *this; }
    Box& grow(int d, int n) { m_lo[d] -= n; m_hi[d] += n; return *this; }
    Box& growLo(int d, int n) { m_lo[d] -= n; return *this; }
    Box& growHi(int d, int n) { m_hi[d] += n; return *this; }
    Box& surroundingNodes(int d) { if (!m_nodal[d]) { m_nodal[d] = 1; m_hi[d] += 1; } return *this; }
    Box& enclosedCells() { for (int d = 0; d < 3; ++d) if (m_nodal[d]) { m_nodal[d] = 0; m_hi[d] -= 1; } return *this; }
    bool coarsenable(int r, int min_width = 1) const;
    Box operator&(const Box& o) const;

  private:
    IntVect m_lo, m_hi, m_nodal;
};

Box coarsen(const Box& b, int r);
Box refine(const Box& b, int r);

class BoxArray {
  public:
    BoxArray() = default;
    explicit BoxArray(const Box& b) : m_boxes{b} {}
    explicit BoxArray(std::vector<Box> bl) : m_boxes(std::move(bl)) {}
    BoxArray& maxSize(int chunk) { return maxSize(IntVect(chunk)); }
    BoxArray& maxSize(const IntVect& chunk);
    Long size() const { return (Long)m_boxes.size(); }
    bool empty() const { return m_boxes.empty(); }
    Box operator[](Long i) const { return m_boxes[(size_t)i]; }
    Long numPts() const;
    Box minimalBox() const;
    BoxArray& coarsen(int r);
    bool coarsenable(int r, int min_width = 1) const;
    bool operator==(const BoxArray& o) const { return m_boxes == o.m_boxes; }
    bool operator!=(const BoxArray& o) const { return !(*this == o); }
    const std::vector<Box>& boxes() const { return m_boxes; }

  private:
    std::vector<Box> m_boxes;
};

BoxArray convert(const BoxArray& ba, const IntVect& nodal);

class DistributionMapping {
  public:
    DistributionMapping() = default;
    explicit DistributionMapping(const BoxArray& ba) : m_n((int)ba.size()) {}
    int size() const { return m_n; }
    int operator[](int) const { return 0; }  // every box lives on this GPU

  private:
    int m_n = 0;
};

struct RealBox {
    double m_lo[3] = {0.0, 0.0, 0.0}, m_hi[3] = {1.0, 1.0, 1.0};
    RealBox() = default;
    RealBox(const std::array<double, 3>& lo, const std::array<double, 3>& hi) {
        for (int d = 0; d < 3; ++d) { m_lo[d] = lo[d]; m_hi[d] = hi[d]; }
    }
    double lo(int d) const { return m_lo[d]; }
    double hi(int d) const { return m_hi[d]; }
    double length(int d) const { return m_hi[d] - m_lo[d]; }
};

class Geometry {
  public:
    Geometry() = default;
    explicit Geometry(const Box& dom, const RealBox* rb = nullptr, int coord = 0, const int* is_per = nullptr);
    Geometry(const Box& dom, const RealBox& rb, int coord, const std::array<int, 3>& is_per)
        : Geometry(dom, &rb, coord, is_per.data()) {}
    const Box& Domain() const { return m_domain; }
    const RealBox& ProbDomain() const { return m_rb; }
    double ProbLo(int d) const { return m_rb.lo(d); }
    double ProbHi(int d) const { return m_rb.hi(d); }
    int Coord() const { return m_coord; }
    // dx = prob length / cells and dxinv = 1/dx (AMReX_Geometry.cpp:520-521)
    double CellSize(int d) const { return m_dx[d]; }
    double InvCellSize(int d) const { return m_dxinv[d]; }
    const double* CellSize() const { return m_dx; }
    const double* InvCellSize() const { return m_dxinv; }
    bool isPeriodic(int d) const { return m_per[d] != 0; }
    std::array<int, 3> isPeriodic() const { return {m_per[0], m_per[1], m_per[2]}; }
    bool isAllPeriodic() const { return m_per[0] && m_per[1] && m_per[2]; }

  private:
    Box m_domain;
    RealBox m_rb;
    int m_coord = 0;
    int m_per[3] = {0, 0, 0};
    double m_dx[3] = {1, 1, 1}, m_dxinv[3] = {1, 1, 1};
};

namespace Gpu {
// Stream all MLMG work is issued on (default: the null stream of the device).
hipStream_t gpuStream();
void setStream(hipStream_t s);
void streamSynchronize();
}  // namespace Gpu

// Device allocation owned by a shared handle.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t nbytes);
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

// FabArray: device storage, boxes packed one after another, each stored
// ghost-grown with x fastest and the components one after another
// (AMReX_Array4.H:88-90).  The solvers take single-component FabArrays; more
// components serve I/O (plotfiles) and cell-centred flux output.
template <class T>
class FabArray {
  public:
    using value_type = T;
    FabArray() = default;
    FabArray(const BoxArray& ba, const DistributionMapping& dm, int ncomp, int ngrow) { define(ba, dm, ncomp, ngrow); }
    FabArray(const BoxArray& ba, const DistributionMapping& dm, int ncomp, const IntVect& ngrow) {
        define(ba, dm, ncomp, ngrow[0]);
    }
    void define(const BoxArray& ba, const DistributionMapping& dm, int ncomp, int ngrow);
    bool ok() const { return m_buf != nullptr; }
    const BoxArray& boxArray() const { return m_ba; }
    const DistributionMapping& DistributionMap() const { return m_dm; }
    int nComp() const { return m_ncomp; }
    int nGrow() const { return m_ngrow; }
    IntVect nGrowVect() const { return IntVect(m_ngrow); }
    Long size() const { return m_ba.size(); }
    // storage box (ghost-grown) of box i and its first element
    Box fabbox(Long i) const { Box b = m_ba[i]; b.grow(m_ngrow); return b; }
    T* dataPtr() { return static_cast<T*>(m_buf->ptr); }
    const T* dataPtr() const { return static_cast<const T*>(m_buf->ptr); }
    T* fabPtr(Long i) { return dataPtr() + m_off[(size_t)i]; }
    // elements of one component of box i's FAB, and where component c of it starts
    int64_t fabSize(Long i) const { return fabbox(i).numPts(); }
    int64_t compOffset(Long i, int c) const { return m_off[(size_t)i] + (int64_t)c * fabSize(i); }
    int64_t numElements() const { return m_total; }
    const std::vector<int64_t>& offsets() const { return m_off; }
    const int64_t* devOffsets() const { return static_cast<const int64_t*>(m_doff->ptr); }
    // device table of the boxes (valid box in this FabArray's index type, first element)
    const void* devBoxes() const { return m_dub->ptr; }
    int64_t maxBoxCells() const { return m_maxcells; }     // largest valid box
    int64_t maxGrownCells() const { return m_maxgrown; }   // largest box grown by one cell
    void setVal(T v);  // valid and ghost cells
    void setVal(T v, int comp, int ncomp);
    void copyToHost(std::vector<T>& out) const;
    void copyFromHost(const std::vector<T>& in);
    T norminf(int comp = 0) const;  // valid cells, reduced on the device
    T sum(int comp = 0) const;
    // out_dev[2c], out_dev[2c+1] = sum and max |x| of component comp + c over
    // the valid cells, asynchronous on the library stream; partials_dev holds
    // mgx::kMfReducePartials doubles
    void reduce(int comp, int ncomp, double* partials_dev, double* out_dev) const;

  private:
    BoxArray m_ba;
    DistributionMapping m_dm;
    int m_ncomp = 1, m_ngrow = 0;
    std::vector<int64_t> m_off;
    int64_t m_total = 0;
    int64_t m_maxcells = 0, m_maxgrown = 0;
    void sum_max(int comp, double out[2]) const;
    std::shared_ptr<DeviceBuffer> m_buf, m_doff, m_dub;
};

using MultiFab = FabArray<double>;
using fMultiFab = FabArray<float>;

// m is a single-component FabArray on ba (what the solvers accept)
template <class T>
bool one_comp_on(const FabArray<T>& m, const BoxArray& ba) {
    return m.nComp() == 1 && m.boxArray() == ba;
}

template <class T>
Vector<T*> GetVecOfPtrs(Vector<T>& v) {
    Vector<T*> r;
    for (auto& x : v) r.push_back(&x);
    return r;
}
template <class T>
Vector<const T*> GetVecOfConstPtrs(const Vector<T>& v) {
    Vector<const T*> r;
    for (auto& x : v) r.push_back(&x);
    return r;
}

}  // namespace amrex
