// On-disk interchange in AMReX's formats: VisMF (AMReX_VisMF.H/.cpp), plotfiles
// (AMReX_PlotFileUtil.H/.cpp, AMReX_PlotFileDataImpl.cpp) and MultiFab::Copy.
//
// Files written here read back in AMReX and the reverse: the VisMF header of
// version 1 (one FAB header per box in the data file, per-box minima and
// maxima in the header), one data file per MultiFab (the one-process case of
// NFiles), IEEE doubles in the machine's byte order, plotfile headers
// "HyperCLaw-V1.1".  Data moves from HBM through pinned staging buffers in
// chunks, the copy of one chunk overlapping the file write of the previous.
#pragma once

#include <array>
#include <memory>
#include <string>

#include "AMReX_Base.H"

namespace amrex {

class VisMF {
  public:
    enum How { OneFilePerCPU, NFiles };  // AMReX_VisMF.H:41
    // Write mf as <name>_H and <name>_D_00000; returns the bytes written.  With
    // set_ghost, every ghost cell of mf itself is first set to (min + max)/2 of
    // its box's valid cells, per component, as the reference does through a
    // const_cast (AMReX_VisMF.cpp:945-970).
    static Long Write(const MultiFab& mf, const std::string& name, How how = NFiles, bool set_ghost = false);
    // Read <name>: into mf when it is defined (same BoxArray and components;
    // the ghost cells both have are read too), else define mf from the header.
    // Reads versions 1-4 of the header with native doubles.
    static void Read(MultiFab& mf, const std::string& name);
};

// mf components [srccomp, srccomp+numcomp) -> dst [dstcomp, ...), valid cells
// grown by nghost; the two must share the BoxArray (MultiFab::Copy).
void Copy(MultiFab& dst, const MultiFab& src, int srccomp, int dstcomp, int numcomp, int nghost);

// AMReX_PlotFileUtil.cpp:159-248.  mf[lev] holds varnames.size() components;
// ghost cells are not written.
void WriteMultiLevelPlotfile(const std::string& plotfilename, int nlevels, const Vector<const MultiFab*>& mf,
                             const Vector<std::string>& varnames, const Vector<Geometry>& geom, Real time,
                             const Vector<int>& level_steps, const Vector<IntVect>& ref_ratio,
                             const std::string& versionName = "HyperCLaw-V1.1",
                             const std::string& levelPrefix = "Level_", const std::string& mfPrefix = "Cell",
                             const Vector<std::string>& extra_dirs = Vector<std::string>());
void WriteSingleLevelPlotfile(const std::string& plotfilename, const MultiFab& mf, const Vector<std::string>& varnames,
                              const Geometry& geom, Real time, int level_step,
                              const std::string& versionName = "HyperCLaw-V1.1",
                              const std::string& levelPrefix = "Level_", const std::string& mfPrefix = "Cell",
                              const Vector<std::string>& extra_dirs = Vector<std::string>());

// Reader of a plotfile (AMReX_PlotFileUtil.H:154-197, AMReX_PlotFileDataImpl.cpp).
class PlotFileData {
  public:
    explicit PlotFileData(std::string const& plotfile_name);
    ~PlotFileData();
    int spaceDim() const noexcept;
    Real time() const noexcept;
    int finestLevel() const noexcept;
    int refRatio(int level) const noexcept;
    int levelStep(int level) const noexcept;
    const BoxArray& boxArray(int level) const noexcept;
    const DistributionMapping& DistributionMap(int level) const noexcept;
    int coordSys() const noexcept;
    Box probDomain(int level) const noexcept;
    std::array<Real, 3> probSize() const noexcept;
    std::array<Real, 3> probLo() const noexcept;
    std::array<Real, 3> probHi() const noexcept;
    std::array<Real, 3> cellSize(int level) const noexcept;
    const Vector<std::string>& varNames() const noexcept;
    int nComp() const noexcept;
    IntVect nGrowVect(int level) const noexcept;
    MultiFab get(int level);                              // every component
    MultiFab get(int level, std::string const& varname);  // one component

    struct Impl;

  private:
    std::unique_ptr<Impl> m_impl;
};

// Create path as a fresh directory; an existing one is renamed to
// path.old.<unique string> (UtilCreateCleanDirectory, AMReX_Utility.cpp:161-182).
void UtilCreateCleanDirectory(const std::string& path, bool callbarrier = true);

}  // namespace amrex
