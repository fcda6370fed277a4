// VisMF, plotfile and MultiFab::Copy for device-resident MultiFabs.
//
// Text layouts restate AMReX's writers: VisMF::Header operator<<
// (AMReX_VisMF.cpp:126-330), BoxArray::writeOn (AMReX_BoxArray.cpp:475-505),
// the FAB header of FABio_binary (AMReX_FArrayBox.cpp:95-105, 905-912), the
// RealDescriptor of native doubles (AMReX_FabConv.cpp:864-897, AMReX_FPC.cpp:61-110)
// and WriteGenericPlotfileHeader (AMReX_PlotFileUtil.cpp:73-156).  Readers
// follow VisMF::Read and PlotFileDataImpl (AMReX_PlotFileDataImpl.cpp:16-103).
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <cerrno>
#include <cfloat>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <random>
#include <sstream>
#include <stdexcept>

#include "amrex/AMReX_MultiFabUtil.H"
#include "amrex/AMReX_PlotFileUtil.H"
#include "mgx_io.H"

namespace amrex {

namespace {

using mgx::IoFab;
using mgx::IoRegion;

std::string box_str(const Box& b) {
    std::ostringstream os;
    const IntVect& lo = b.smallEnd();
    const IntVect& hi = b.bigEnd();
    const IntVect& t = b.ixType();
    os << "((" << lo[0] << ',' << lo[1] << ',' << lo[2] << ") (" << hi[0] << ',' << hi[1] << ',' << hi[2] << ") ("
       << t[0] << ',' << t[1] << ',' << t[2] << "))";
    return os.str();
}

bool little_endian() {
    const uint16_t one = 1;
    unsigned char c;
    std::memcpy(&c, &one, 1);
    return c == 1;
}

// FPC::NativeRealDescriptor for IEEE doubles (format ieee_double, byte order of this machine)
std::string native_rd() {
    return little_endian() ? "((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"
                           : "((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))";
}

std::string fab_header(const Box& fab, int ncomp) {
    return "FAB " + native_rd() + box_str(fab) + " " + std::to_string(ncomp) + "\n";
}

std::string base_name(const std::string& path) {
    const size_t p = path.rfind('/');
    return p == std::string::npos ? path : path.substr(p + 1);
}
std::string dir_name(const std::string& path) {
    const size_t p = path.rfind('/');
    return p == std::string::npos ? std::string() : path.substr(0, p + 1);
}

// Device <-> file through two pinned buffers: the copy of one chunk overlaps
// the file access of the other.
class Stager {
  public:
    static constexpr size_t kChunk = size_t(64) << 20;
    Stager() {
        for (int i = 0; i < 2; ++i) {
            MGX_HIP_CHECK(hipHostMalloc((void**)&buf_[i], kChunk, hipHostMallocDefault));
            MGX_HIP_CHECK(hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming));
        }
        s_ = Gpu::gpuStream();
    }
    ~Stager() {
        for (int i = 0; i < 2; ++i) {
            (void)hipHostFree(buf_[i]);
            (void)hipEventDestroy(ev_[i]);
        }
    }
    Stager(const Stager&) = delete;
    Stager& operator=(const Stager&) = delete;

    // One piece of a file: optional host bytes (a FAB header) followed by at
    // most kChunk device bytes.
    struct Piece {
        const std::string* head;
        const char* dev;
        size_t bytes;
    };
    // Append the device ranges of a FAB to pieces, the first one after head.
    static void add_fab(std::vector<Piece>& pieces, const std::string* head, const void* dev, size_t bytes) {
        const char* d = static_cast<const char*>(dev);
        size_t off = 0;
        do {
            const size_t len = std::min(kChunk, bytes - off);
            pieces.push_back({off == 0 ? head : nullptr, d + off, len});
            off += len;
        } while (off < bytes);
    }
    // Write the pieces to os in order.  Each device range is copied to the start
    // of a pinned buffer (aligned, so the copy engine runs at full rate) while
    // the previous piece is written out.
    void write_pieces(std::ostream& os, const std::vector<Piece>& pieces) {
        const size_t n = pieces.size();
        for (size_t k = 0; k <= n; ++k) {
            if (k < n && pieces[k].bytes > 0) {
                MGX_HIP_CHECK(hipMemcpyAsync(buf_[k & 1], pieces[k].dev, pieces[k].bytes, hipMemcpyDeviceToHost, s_));
                MGX_HIP_CHECK(hipEventRecord(ev_[k & 1], s_));
            }
            if (k > 0) {
                const Piece& p = pieces[k - 1];
                if (p.head) os.write(p.head->data(), (std::streamsize)p.head->size());
                if (p.bytes > 0) flush_(os, (int)((k - 1) & 1), p.bytes);
            }
        }
    }
    // Read bytes from is into device memory in chunks: the read of a chunk
    // overlaps the copy to the device of the previous ones, across calls too.
    // The device memory is complete after finish().
    void from_file(std::istream& is, void* dev, size_t bytes) {
        char* d = static_cast<char*>(dev);
        for (size_t off = 0; off < bytes; off += kChunk) {
            const size_t len = std::min(kChunk, bytes - off);
            const int i = (int)(nread_ & 1);
            if (nread_ >= 2) MGX_HIP_CHECK(hipEventSynchronize(ev_[i]));  // the copy that last used this buffer
            is.read(buf_[i], (std::streamsize)len);
            if (!is) throw std::runtime_error("VisMF: data file ended early");
            MGX_HIP_CHECK(hipMemcpyAsync(d + off, buf_[i], len, hipMemcpyHostToDevice, s_));
            MGX_HIP_CHECK(hipEventRecord(ev_[i], s_));
            ++nread_;
        }
    }
    void finish() { MGX_HIP_CHECK(hipStreamSynchronize(s_)); }

  private:
    void flush_(std::ostream& os, int i, size_t len) {
        MGX_HIP_CHECK(hipEventSynchronize(ev_[i]));
        os.write(buf_[i], (std::streamsize)len);
        if (!os) throw std::runtime_error("VisMF: write failed");
    }
    char* buf_[2] = {nullptr, nullptr};
    hipEvent_t ev_[2] = {nullptr, nullptr};
    hipStream_t s_ = nullptr;
    uint64_t nread_ = 0;
};

template <class T>
T* dev_copy(const std::vector<T>& v, std::unique_ptr<DeviceBuffer>& keep) {
    keep.reset(new DeviceBuffer(sizeof(T) * std::max<size_t>(v.size(), 1)));
    MGX_HIP_CHECK(hipMemcpy(keep->ptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return static_cast<T*>(keep->ptr);
}

IoFab io_fab(const MultiFab& m, Long b, int comp) {
    const Box fb = m.fabbox(b);
    IoFab f{};
    for (int d = 0; d < 3; ++d) f.lo[d] = fb.smallEnd(d);
    f.nx = fb.length(0);
    f.ny = fb.length(1);
    f.off = m.compOffset(b, comp);
    return f;
}
IoRegion io_region(const Box& b) {
    IoRegion r{};
    for (int d = 0; d < 3; ++d) {
        r.lo[d] = b.smallEnd(d);
        r.hi[d] = b.bigEnd(d);
    }
    return r;
}

// mn/mx[b][c] over the valid cells of box b, component c
void box_minmax_all(const MultiFab& m, std::vector<std::vector<double>>& mn, std::vector<std::vector<double>>& mx) {
    const Long n = m.size();
    mn.assign((size_t)n, std::vector<double>((size_t)m.nComp(), 0.0));
    mx = mn;
    if (n == 0) return;
    std::vector<IoRegion> reg((size_t)n);
    int64_t maxr = 0;
    for (Long b = 0; b < n; ++b) {
        reg[(size_t)b] = io_region(m.boxArray()[b]);
        maxr = std::max<int64_t>(maxr, m.boxArray()[b].numPts());
    }
    std::unique_ptr<DeviceBuffer> kr, kf, kp;
    const IoRegion* dreg = dev_copy(reg, kr);
    const int nblk = mgx::io_minmax_blocks(maxr);
    kp.reset(new DeviceBuffer(sizeof(double) * 2 * (size_t)nblk * (size_t)n));
    std::vector<double> part((size_t)2 * nblk * n);
    hipStream_t s = Gpu::gpuStream();
    for (int c = 0; c < m.nComp(); ++c) {
        std::vector<IoFab> fabs((size_t)n);
        for (Long b = 0; b < n; ++b) fabs[(size_t)b] = io_fab(m, b, c);
        const IoFab* dfab = dev_copy(fabs, kf);
        mgx::box_minmax(m.dataPtr(), dfab, dreg, (int)n, maxr, static_cast<double*>(kp->ptr), s);
        MGX_HIP_CHECK(hipMemcpyAsync(part.data(), kp->ptr, sizeof(double) * part.size(), hipMemcpyDeviceToHost, s));
        MGX_HIP_CHECK(hipStreamSynchronize(s));
        for (Long b = 0; b < n; ++b) {
            double lo = DBL_MAX, hi = -DBL_MAX;
            for (int k = 0; k < nblk; ++k) {
                lo = std::min(lo, part[((size_t)b * nblk + k) * 2]);
                hi = std::max(hi, part[((size_t)b * nblk + k) * 2 + 1]);
            }
            mn[(size_t)b][(size_t)c] = lo;
            mx[(size_t)b][(size_t)c] = hi;
        }
    }
}

// ---- header parsing (the reference's operator>> of each piece)
void expect(std::istream& is, char want) {
    char c;
    is >> c;
    if (!is || c != want) throw std::runtime_error(std::string("VisMF: malformed header, expected '") + want + "'");
}
IntVect read_iv(std::istream& is) {
    IntVect v;
    expect(is, '(');
    is >> v[0];
    expect(is, ',');
    is >> v[1];
    expect(is, ',');
    is >> v[2];
    expect(is, ')');
    return v;
}
Box read_box(std::istream& is) {
    expect(is, '(');
    const IntVect lo = read_iv(is), hi = read_iv(is), t = read_iv(is);
    expect(is, ')');
    return Box(lo, hi, t);
}
BoxArray read_ba(std::istream& is) {
    expect(is, '(');
    long n, hash;
    is >> n >> hash;
    std::vector<Box> v;
    for (long i = 0; i < n; ++i) v.push_back(read_box(is));
    expect(is, ')');
    return BoxArray(std::move(v));
}

struct VisMFHeader {
    int vers = 0, how = 1, ncomp = 0, ngrow = 0;
    BoxArray ba;
    std::vector<std::string> fname;
    std::vector<long long> offset;
    std::string rd;  // descriptor of headerless data (versions 2-4)
};

VisMFHeader read_header(const std::string& name) {
    std::ifstream is(name + "_H");
    if (!is) throw std::runtime_error("VisMF::Read: cannot open " + name + "_H");
    VisMFHeader h;
    is >> h.vers >> h.how >> h.ncomp;
    if (h.vers < 1 || h.vers > 4) throw std::runtime_error("VisMF::Read: unknown header version");
    is >> std::ws;
    if (is.peek() == '(') {
        const IntVect g = read_iv(is);
        if (g[0] != g[1] || g[0] != g[2]) throw std::runtime_error("VisMF::Read: unequal ghost widths are not supported");
        h.ngrow = g[0];
    } else {
        is >> h.ngrow;
    }
    h.ba = read_ba(is);
    long n;
    is >> n;
    if (n != h.ba.size()) throw std::runtime_error("VisMF::Read: FabOnDisk count differs from the BoxArray");
    for (long i = 0; i < n; ++i) {
        std::string tag, f;
        long long off;
        is >> tag >> f >> off;
        if (tag != "FabOnDisk:") throw std::runtime_error("VisMF::Read: malformed FabOnDisk entry");
        h.fname.push_back(f);
        h.offset.push_back(off);
    }
    if (h.vers == 1 || h.vers == 3) {  // per-box minima and maxima
        for (int t = 0; t < 2; ++t) {
            long nn, mm;
            char ch;
            is >> nn >> ch >> mm;
            for (long i = 0; i < nn * mm; ++i) {
                double v;
                is >> v >> ch;
            }
        }
    }
    if (h.vers == 4) {
        for (int t = 0; t < 2; ++t)
            for (int c = 0; c < h.ncomp; ++c) {
                double v;
                char ch;
                is >> v >> ch;
            }
    }
    if (h.vers >= 2) {
        is >> std::ws;
        std::getline(is, h.rd);
        while (!h.rd.empty() && std::isspace((unsigned char)h.rd.back())) h.rd.pop_back();
    }
    if (!is) throw std::runtime_error("VisMF::Read: header read failed");
    return h;
}

std::string strip(std::string s) {
    while (!s.empty() && std::isspace((unsigned char)s.back())) s.pop_back();
    size_t i = 0;
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    return s.substr(i);
}

// Read FAB b of a VisMF into dst (components [c0, c0+nc) of the file into
// components [d0, d0+nc) of dst, cells the two FABs share).
void read_fab(const std::string& name, const VisMFHeader& h, Long b, MultiFab& dst, int c0, int d0, int nc,
              Stager& st, std::map<std::string, std::unique_ptr<std::ifstream>>& files,
              std::unique_ptr<DeviceBuffer>& tmp) {
    const std::string path = dir_name(name) + h.fname[(size_t)b];
    auto& f = files[path];
    if (!f) {
        f.reset(new std::ifstream(path, std::ios::binary));
        if (!*f) throw std::runtime_error("VisMF::Read: cannot open " + path);
    }
    std::ifstream& is = *f;
    is.clear();
    is.seekg(h.offset[(size_t)b]);
    Box fab = h.ba[b];
    fab.grow(h.ngrow);
    if (h.vers == 1) {
        std::string line;
        std::getline(is, line);
        const std::string want = strip(fab_header(fab, h.ncomp));
        if (strip(line) != want) {
            // accept any ncomp / box spelling only if the descriptor is native
            if (line.compare(0, 4 + native_rd().size(), "FAB " + native_rd()) != 0)
                throw std::runtime_error("VisMF::Read: only native IEEE double data is supported");
        }
    } else if (h.rd != native_rd()) {
        throw std::runtime_error("VisMF::Read: only native IEEE double data is supported");
    }
    const int64_t npts = fab.numPts();
    is.seekg((std::streamoff)(sizeof(double) * (size_t)npts * (size_t)c0), std::ios::cur);
    const size_t bytes = sizeof(double) * (size_t)npts * (size_t)nc;
    const bool same = dst.nGrow() == h.ngrow;
    if (same) {  // straight into dst's FAB
        st.from_file(is, dst.dataPtr() + dst.compOffset(b, d0), bytes);
        return;
    }
    if (!tmp || tmp->bytes < bytes) tmp.reset(new DeviceBuffer(bytes));
    st.from_file(is, tmp->ptr, bytes);
    const int ng = std::min(dst.nGrow(), h.ngrow);
    Box r = h.ba[b];
    r.grow(ng);
    std::vector<IoFab> df(1), sf(1);
    std::vector<IoRegion> rr(1, io_region(r));
    std::unique_ptr<DeviceBuffer> k1, k2, k3;
    for (int c = 0; c < nc; ++c) {
        df[0] = io_fab(dst, b, d0 + c);
        IoFab s{};
        for (int d = 0; d < 3; ++d) s.lo[d] = fab.smallEnd(d);
        s.nx = fab.length(0);
        s.ny = fab.length(1);
        s.off = (int64_t)c * npts;
        sf[0] = s;
        mgx::copy_region(dst.dataPtr(), dev_copy(df, k1), static_cast<const double*>(tmp->ptr), dev_copy(sf, k2),
                         dev_copy(rr, k3), 1, r.numPts(), Gpu::gpuStream());
        MGX_HIP_CHECK(hipStreamSynchronize(Gpu::gpuStream()));
    }
}

bool file_exists(const std::string& p) {
    struct stat sb;
    return ::stat(p.c_str(), &sb) == 0;
}

std::string unique_string() {
    static std::mt19937_64 rng((uint64_t)std::chrono::steady_clock::now().time_since_epoch().count());
    static const char al[] = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789";
    std::string s;
    for (int i = 0; i < 8; ++i) s += al[rng() % (sizeof(al) - 1)];
    return s;
}

}  // namespace

void UtilCreateCleanDirectory(const std::string& path, bool) {
    if (file_exists(path)) {
        const std::string old = path + ".old." + unique_string();
        if (std::rename(path.c_str(), old.c_str()) != 0)
            throw std::runtime_error("UtilCreateCleanDirectory: rename of " + path + " failed");
    }
    // create each missing component (UtilCreateDirectory)
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part;
        if (!file_exists(cur) && ::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
            throw std::runtime_error("UtilCreateCleanDirectory: cannot create " + cur);
        cur += "/";
    }
}

// VisMF::Write (AMReX_VisMF.cpp:934-1126) for one process: Version_v1 header,
// data file <name>_D_00000 holding each FAB (header line + components).
Long VisMF::Write(const MultiFab& mf, const std::string& name, How how, bool set_ghost) {
    const Long n = mf.size();
    const int ncomp = mf.nComp(), ng = mf.nGrow();
    std::vector<std::vector<double>> mn, mx;
    box_minmax_all(mf, mn, mx);
    if (set_ghost && ng > 0 && n > 0) {
        // every ghost cell of a FAB = (min + max)/2 of its valid cells, per
        // component, in the MultiFab itself (AMReX_VisMF.cpp:945-970)
        std::vector<IoRegion> valid((size_t)n), whole((size_t)n);
        int64_t maxr = 0;
        for (Long b = 0; b < n; ++b) {
            valid[(size_t)b] = io_region(mf.boxArray()[b]);
            whole[(size_t)b] = io_region(mf.fabbox(b));
            maxr = std::max<int64_t>(maxr, mf.fabbox(b).numPts());
        }
        std::unique_ptr<DeviceBuffer> kv, kw, kf, kd;
        const IoRegion* dv = dev_copy(valid, kv);
        const IoRegion* dw = dev_copy(whole, kw);
        hipStream_t s = Gpu::gpuStream();
        for (int c = 0; c < ncomp; ++c) {
            std::vector<IoFab> fabs((size_t)n);
            std::vector<double> vals((size_t)n);
            for (Long b = 0; b < n; ++b) {
                fabs[(size_t)b] = io_fab(mf, b, c);
                vals[(size_t)b] = (mn[(size_t)b][(size_t)c] + mx[(size_t)b][(size_t)c]) / 2.0;
            }
            mgx::set_complement(const_cast<double*>(mf.dataPtr()), dev_copy(fabs, kf), dv, dw, dev_copy(vals, kd),
                                (int)n, maxr, s);
            MGX_HIP_CHECK(hipStreamSynchronize(s));  // the tables above are freed at the next component
        }
    }
    const std::string dfile = name + "_D_00000";
    std::ofstream os(dfile, std::ios::binary | std::ios::trunc);
    if (!os) throw std::runtime_error("VisMF::Write: cannot open " + dfile);
    std::vector<long long> offset((size_t)n);
    std::vector<std::string> hdrs((size_t)n);
    std::vector<Stager::Piece> pieces;
    Long bytes = 0;
    for (Long b = 0; b < n; ++b) {
        offset[(size_t)b] = (long long)bytes;
        hdrs[(size_t)b] = fab_header(mf.fabbox(b), ncomp);
        const size_t nbytes = sizeof(double) * (size_t)mf.fabSize(b) * (size_t)ncomp;
        Stager::add_fab(pieces, &hdrs[(size_t)b], mf.dataPtr() + mf.compOffset(b, 0), nbytes);
        bytes += (Long)(hdrs[(size_t)b].size() + nbytes);
    }
    MGX_HIP_CHECK(hipStreamSynchronize(Gpu::gpuStream()));
    {
        Stager st;
        st.write_pieces(os, pieces);
    }
    os.close();
    if (!os) throw std::runtime_error("VisMF::Write: write of " + dfile + " failed");
    // the header (VisMF::Header operator<<, Version_v1)
    std::ostringstream h;
    h.setf(std::ios::floatfield, std::ios::scientific);
    h.precision(17);
    h << 1 << '\n' << int(how) << '\n' << ncomp << '\n' << ng << '\n';
    h << '(' << n << ' ' << 0 << '\n';
    for (Long b = 0; b < n; ++b) h << box_str(mf.boxArray()[b]) << '\n';
    h << ')' << '\n';
    h << n << '\n';
    for (Long b = 0; b < n; ++b) h << "FabOnDisk: " << base_name(dfile) << ' ' << offset[(size_t)b] << '\n';
    h << '\n';
    for (const auto* t : {&mn, &mx}) {
        h << n << ',' << ncomp << '\n';
        for (Long b = 0; b < n; ++b) {
            for (int c = 0; c < ncomp; ++c) h << (*t)[(size_t)b][(size_t)c] << ',';
            h << '\n';
        }
        h << '\n';
    }
    const std::string htext = h.str();
    std::ofstream hs(name + "_H", std::ios::binary | std::ios::trunc);
    if (!hs) throw std::runtime_error("VisMF::Write: cannot open " + name + "_H");
    hs.write(htext.data(), (std::streamsize)htext.size());
    hs.close();
    if (!hs) throw std::runtime_error("VisMF::Write: write of " + name + "_H failed");
    return bytes + (Long)htext.size();
}

void VisMF::Read(MultiFab& mf, const std::string& name) {
    const VisMFHeader h = read_header(name);
    if (!mf.ok()) {
        mf.define(h.ba, DistributionMapping(h.ba), h.ncomp, h.ngrow);
    } else if (mf.boxArray() != h.ba || mf.nComp() != h.ncomp) {
        throw std::runtime_error("VisMF::Read: the MultiFab's BoxArray or component count differs from " + name);
    }
    Stager st;
    std::map<std::string, std::unique_ptr<std::ifstream>> files;
    std::unique_ptr<DeviceBuffer> tmp;
    for (Long b = 0; b < h.ba.size(); ++b) read_fab(name, h, b, mf, 0, 0, h.ncomp, st, files, tmp);
    st.finish();
}

void average_face_to_cellcenter(MultiFab& cc, int dcomp, const std::array<const MultiFab*, 3>& fc, int ngrow) {
    if (dcomp < 0 || dcomp + 3 > cc.nComp())
        throw std::runtime_error("average_face_to_cellcenter: three components from dcomp expected");
    if (ngrow < 0 || ngrow > cc.nGrow()) throw std::runtime_error("average_face_to_cellcenter: ngrow too large");
    const Long n = cc.size();
    for (int d = 0; d < 3; ++d) {
        if (!fc[(size_t)d] || fc[(size_t)d]->nComp() != 1 || fc[(size_t)d]->nGrow() < ngrow ||
            fc[(size_t)d]->boxArray() != amrex::convert(cc.boxArray(), IntVect::TheDimensionVector(d)))
            throw std::runtime_error(
                "average_face_to_cellcenter: one face component on the cell BoxArray converted to faces expected");
    }
    if (n == 0) return;
    std::vector<IoFab> dfab((size_t)(3 * n)), sfab((size_t)(3 * n));
    std::vector<IoRegion> reg((size_t)n);
    int64_t maxr = 0;
    for (Long b = 0; b < n; ++b) {
        Box r = cc.boxArray()[b];
        r.grow(ngrow);
        reg[(size_t)b] = io_region(r);
        maxr = std::max<int64_t>(maxr, r.numPts());
        for (int d = 0; d < 3; ++d) {
            dfab[(size_t)(3 * b + d)] = io_fab(cc, b, dcomp + d);
            sfab[(size_t)(3 * b + d)] = io_fab(*fc[(size_t)d], b, 0);
        }
    }
    std::unique_ptr<DeviceBuffer> kd, ks, kr;
    mgx::FaceSrc src{{fc[0]->dataPtr(), fc[1]->dataPtr(), fc[2]->dataPtr()}};
    hipStream_t s = Gpu::gpuStream();
    mgx::face_to_cell(cc.dataPtr(), dev_copy(dfab, kd), src, dev_copy(sfab, ks), dev_copy(reg, kr), (int)n, maxr, s);
    MGX_HIP_CHECK(hipStreamSynchronize(s));  // the tables are freed on return
}

void Copy(MultiFab& dst, const MultiFab& src, int srccomp, int dstcomp, int numcomp, int nghost) {
    if (dst.boxArray() != src.boxArray()) throw std::runtime_error("MultiFab::Copy: BoxArrays differ");
    if (srccomp < 0 || dstcomp < 0 || numcomp < 1 || srccomp + numcomp > src.nComp() || dstcomp + numcomp > dst.nComp())
        throw std::runtime_error("MultiFab::Copy: bad component range");
    if (nghost > dst.nGrow() || nghost > src.nGrow()) throw std::runtime_error("MultiFab::Copy: nghost too large");
    const Long n = dst.size();
    if (n == 0) return;
    std::vector<IoRegion> reg((size_t)n);
    int64_t maxr = 0;
    for (Long b = 0; b < n; ++b) {
        Box r = dst.boxArray()[b];
        r.grow(nghost);
        reg[(size_t)b] = io_region(r);
        maxr = std::max<int64_t>(maxr, r.numPts());
    }
    std::unique_ptr<DeviceBuffer> kr, kd, ks;
    const IoRegion* dreg = dev_copy(reg, kr);
    hipStream_t s = Gpu::gpuStream();
    for (int c = 0; c < numcomp; ++c) {
        std::vector<IoFab> df((size_t)n), sf((size_t)n);
        for (Long b = 0; b < n; ++b) {
            df[(size_t)b] = io_fab(dst, b, dstcomp + c);
            sf[(size_t)b] = io_fab(src, b, srccomp + c);
        }
        mgx::copy_region(dst.dataPtr(), dev_copy(df, kd), src.dataPtr(), dev_copy(sf, ks), dreg, (int)n, maxr, s);
        MGX_HIP_CHECK(hipStreamSynchronize(s));  // the tables above are freed at the next iteration
    }
}

// ---- plotfiles

void WriteMultiLevelPlotfile(const std::string& plotfilename, int nlevels, const Vector<const MultiFab*>& mf,
                             const Vector<std::string>& varnames, const Vector<Geometry>& geom, Real time,
                             const Vector<int>& level_steps, const Vector<IntVect>& ref_ratio,
                             const std::string& versionName, const std::string& levelPrefix,
                             const std::string& mfPrefix, const Vector<std::string>& extra_dirs) {
    if (nlevels < 1 || (int)mf.size() < nlevels || (int)geom.size() < nlevels || (int)level_steps.size() < nlevels ||
        (int)ref_ratio.size() + 1 < nlevels)
        throw std::runtime_error("WriteMultiLevelPlotfile: fewer MultiFabs, geometries, steps or ratios than levels");
    for (int l = 0; l < nlevels; ++l)
        if (mf[(size_t)l]->nComp() != (int)varnames.size())
            throw std::runtime_error("WriteMultiLevelPlotfile: component count differs from the variable names");
    auto level_path = [&](const std::string& root, int l) { return root + "/" + levelPrefix + std::to_string(l); };
    // PreBuildDirectorHierarchy (AMReX_PlotFileUtil.cpp:56-70)
    UtilCreateCleanDirectory(plotfilename, false);
    for (int l = 0; l < nlevels; ++l) UtilCreateCleanDirectory(level_path(plotfilename, l), false);
    for (const auto& d : extra_dirs) {
        UtilCreateCleanDirectory(plotfilename + "/" + d, false);
        for (int l = 0; l < nlevels; ++l) UtilCreateCleanDirectory(level_path(plotfilename + "/" + d, l), false);
    }
    // WriteGenericPlotfileHeader
    std::ostringstream h;
    h.precision(17);
    const int finest = nlevels - 1;
    h << versionName << '\n' << varnames.size() << '\n';
    for (const auto& v : varnames) h << v << "\n";
    h << 3 << '\n' << time << '\n' << finest << '\n';
    for (int i = 0; i < 3; ++i) h << geom[0].ProbLo(i) << ' ';
    h << '\n';
    for (int i = 0; i < 3; ++i) h << geom[0].ProbHi(i) << ' ';
    h << '\n';
    for (int i = 0; i < finest; ++i) h << ref_ratio[(size_t)i][0] << ' ';
    h << '\n';
    for (int i = 0; i <= finest; ++i) h << box_str(geom[(size_t)i].Domain()) << ' ';
    h << '\n';
    for (int i = 0; i <= finest; ++i) h << level_steps[(size_t)i] << ' ';
    h << '\n';
    for (int i = 0; i <= finest; ++i) {
        for (int k = 0; k < 3; ++k) h << geom[(size_t)i].CellSize(k) << ' ';
        h << '\n';
    }
    h << geom[0].Coord() << '\n' << "0\n";
    for (int l = 0; l <= finest; ++l) {
        const BoxArray& ba = mf[(size_t)l]->boxArray();
        const Geometry& g = geom[(size_t)l];
        h << l << ' ' << ba.size() << ' ' << time << '\n' << level_steps[(size_t)l] << '\n';
        const IntVect& dlo = g.Domain().smallEnd();
        for (Long i = 0; i < ba.size(); ++i) {
            const Box b = ba[i];
            for (int d = 0; d < 3; ++d) {
                // RealBox(shift(b, -domain_lo), dx, prob_lo) (AMReX_RealBox.cpp:10-22)
                const int shft = b.ixType()[d] ? 0 : 1;
                const double lo = g.ProbLo(d) + g.CellSize(d) * static_cast<double>(b.smallEnd(d) - dlo[d]);
                const double hi = g.ProbLo(d) + g.CellSize(d) * static_cast<double>(b.bigEnd(d) - dlo[d] + shft);
                h << lo << ' ' << hi << '\n';
            }
        }
        h << levelPrefix << l << '/' << mfPrefix << '\n';
    }
    const std::string htext = h.str();
    {
        std::ofstream hs(plotfilename + "/Header", std::ios::binary | std::ios::trunc);
        if (!hs) throw std::runtime_error("WriteMultiLevelPlotfile: cannot open the Header");
        hs.write(htext.data(), (std::streamsize)htext.size());
        if (!hs) throw std::runtime_error("WriteMultiLevelPlotfile: Header write failed");
    }
    for (int l = 0; l <= finest; ++l) {
        const MultiFab& m = *mf[(size_t)l];
        const std::string prefix = level_path(plotfilename, l) + "/" + mfPrefix;
        if (m.nGrow() != 0) {  // a ghost-free copy, as the reference makes
            MultiFab tmp(m.boxArray(), m.DistributionMap(), m.nComp(), 0);
            Copy(tmp, m, 0, 0, m.nComp(), 0);
            VisMF::Write(tmp, prefix);
        } else {
            VisMF::Write(m, prefix);
        }
    }
}

void WriteSingleLevelPlotfile(const std::string& plotfilename, const MultiFab& mf, const Vector<std::string>& varnames,
                              const Geometry& geom, Real time, int level_step, const std::string& versionName,
                              const std::string& levelPrefix, const std::string& mfPrefix,
                              const Vector<std::string>& extra_dirs) {
    WriteMultiLevelPlotfile(plotfilename, 1, Vector<const MultiFab*>{&mf}, varnames, Vector<Geometry>{geom}, time,
                            Vector<int>{level_step}, Vector<IntVect>(), versionName, levelPrefix, mfPrefix, extra_dirs);
}

struct PlotFileData::Impl {
    std::string name;
    std::string version;
    int ncomp = 0, spacedim = 3, finest = 0, coord = 0;
    double time = 0.0;
    Vector<std::string> vars;
    std::array<Real, 3> plo{{0, 0, 0}}, phi{{1, 1, 1}};
    Vector<int> ratio, steps;
    Vector<Box> domain;
    Vector<std::array<Real, 3>> dx;
    Vector<std::string> mfname;
    Vector<BoxArray> ba;
    Vector<DistributionMapping> dm;
    Vector<int> ngrow;
};

PlotFileData::PlotFileData(std::string const& plotfile_name) : m_impl(new Impl) {
    Impl& P = *m_impl;
    P.name = plotfile_name;
    std::ifstream is(plotfile_name + "/Header");
    if (!is) throw std::runtime_error("PlotFileData: cannot open " + plotfile_name + "/Header");
    is >> P.version >> P.ncomp;
    is.ignore(100000, '\n');
    for (int i = 0; i < P.ncomp; ++i) {
        std::string t;
        std::getline(is, t);
        P.vars.push_back(strip(t));
    }
    is >> P.spacedim >> P.time >> P.finest;
    if (P.spacedim != 3) throw std::runtime_error("PlotFileData: only 3-D plotfiles are supported");
    const int nl = P.finest + 1;
    for (int i = 0; i < 3; ++i) is >> P.plo[(size_t)i];
    for (int i = 0; i < 3; ++i) is >> P.phi[(size_t)i];
    P.ratio.assign((size_t)nl, 0);
    for (int i = 0; i < P.finest; ++i) is >> P.ratio[(size_t)i];
    is.ignore(100000, '\n');
    for (int i = 0; i < nl; ++i) P.domain.push_back(read_box(is));
    P.steps.assign((size_t)nl, 0);
    for (int i = 0; i < nl; ++i) is >> P.steps[(size_t)i];
    P.dx.assign((size_t)nl, std::array<Real, 3>{{1, 1, 1}});
    for (int l = 0; l < nl; ++l)
        for (int d = 0; d < 3; ++d) is >> P.dx[(size_t)l][(size_t)d];
    int bwidth;
    is >> P.coord >> bwidth;
    for (int l = 0; l < nl; ++l) {
        int lev, ngrids, lstep;
        double gt;
        is >> lev >> ngrids >> gt >> lstep;
        for (int g = 0; g < ngrids; ++g)
            for (int d = 0; d < 3; ++d) {
                double a, b;
                is >> a >> b;
            }
        std::string rel;
        is >> rel;
        P.mfname.push_back(plotfile_name + "/" + rel);
    }
    if (!is) throw std::runtime_error("PlotFileData: malformed Header in " + plotfile_name);
    for (int l = 0; l < nl; ++l) {
        const VisMFHeader h = read_header(P.mfname[(size_t)l]);
        P.ba.push_back(h.ba);
        P.dm.push_back(DistributionMapping(h.ba));
        P.ngrow.push_back(h.ngrow);
    }
}
PlotFileData::~PlotFileData() = default;
int PlotFileData::spaceDim() const noexcept { return m_impl->spacedim; }
Real PlotFileData::time() const noexcept { return m_impl->time; }
int PlotFileData::finestLevel() const noexcept { return m_impl->finest; }
int PlotFileData::refRatio(int level) const noexcept { return m_impl->ratio[(size_t)level]; }
int PlotFileData::levelStep(int level) const noexcept { return m_impl->steps[(size_t)level]; }
const BoxArray& PlotFileData::boxArray(int level) const noexcept { return m_impl->ba[(size_t)level]; }
const DistributionMapping& PlotFileData::DistributionMap(int level) const noexcept { return m_impl->dm[(size_t)level]; }
int PlotFileData::coordSys() const noexcept { return m_impl->coord; }
Box PlotFileData::probDomain(int level) const noexcept { return m_impl->domain[(size_t)level]; }
std::array<Real, 3> PlotFileData::probSize() const noexcept {
    std::array<Real, 3> r;
    for (size_t d = 0; d < 3; ++d) r[d] = m_impl->phi[d] - m_impl->plo[d];
    return r;
}
std::array<Real, 3> PlotFileData::probLo() const noexcept { return m_impl->plo; }
std::array<Real, 3> PlotFileData::probHi() const noexcept { return m_impl->phi; }
std::array<Real, 3> PlotFileData::cellSize(int level) const noexcept { return m_impl->dx[(size_t)level]; }
const Vector<std::string>& PlotFileData::varNames() const noexcept { return m_impl->vars; }
int PlotFileData::nComp() const noexcept { return m_impl->ncomp; }
IntVect PlotFileData::nGrowVect(int level) const noexcept { return IntVect(m_impl->ngrow[(size_t)level]); }

MultiFab PlotFileData::get(int level) {
    MultiFab m;
    VisMF::Read(m, m_impl->mfname[(size_t)level]);
    return m;
}

MultiFab PlotFileData::get(int level, std::string const& varname) {
    const auto& v = m_impl->vars;
    const auto it = std::find(v.begin(), v.end(), varname);
    if (it == v.end()) throw std::runtime_error("PlotFileData::get: no variable " + varname);
    const int comp = (int)(it - v.begin());
    const std::string& name = m_impl->mfname[(size_t)level];
    const VisMFHeader h = read_header(name);
    MultiFab m(h.ba, DistributionMapping(h.ba), 1, h.ngrow);
    Stager st;
    std::map<std::string, std::unique_ptr<std::ifstream>> files;
    std::unique_ptr<DeviceBuffer> tmp;
    for (Long b = 0; b < h.ba.size(); ++b) read_fab(name, h, b, m, comp, 0, 1, st, files, tmp);
    st.finish();
    return m;
}

}  // namespace amrex
