// Internal state of the linear operator and the MLMG solver (not part of the
// public call surface).
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <limits>
#include <memory>
#include <vector>

#include "amrex/AMReX_MLMG_Core.H"
#include "mgx_amr.H"
#include "mgx_device.H"

namespace mgx {

// Host copy and device upload of one MG level's box layout.
struct LevelPlan {
    amrex::Geometry geom;
    amrex::BoxArray ba;      // storage boxes (one domain box when dense)
    amrex::BoxArray ref_ba;  // the reference hierarchy's BoxArray of this level
    std::vector<BoxDev> hbox;
    std::vector<int> slab_of[3], slab_lo[3], box_of;
    int nslab[3] = {0, 0, 0};
    std::vector<int64_t> cell_start;
    int64_t nalloc = 0, nface[3] = {0, 0, 0}, mbc = 0, mbg = 0, mbf = 0;
    int mtiles[2] = {0, 0};
    bool covered = false;
    std::vector<std::unique_ptr<amrex::DeviceBuffer>> d_slab;
    std::unique_ptr<amrex::DeviceBuffer> d_boxof, d_cellstart, d_box;
    LevelDev dev{};

    void build(const amrex::Geometry& g, const amrex::BoxArray& b);
    // cf (may be null): box faces on the boundary of cf->region that are not
    // physical faces and not covered by a periodic image are coarse/fine faces
    void set_bc(const int* lobc, const int* hibc, int maxorder, const struct CfFaces* cf = nullptr);
    void upload_boxes();
};

// Coarse/fine faces of a level whose boxes do not cover the domain
// (MLLinOp::setCoarseFineBC): their union `region` is one box, and each face
// of it that is neither physical nor periodic-covered carries BC type `type`
// (Dirichlet or Neumann) located x0 cells of this level outside the face.
struct CfFaces {
    amrex::Box region;
    int type;
    double x0;
};

// Reduction partials (after 64 scalar slots).
struct DeviceScratch {
    std::unique_ptr<amrex::DeviceBuffer> buf;
    int64_t cap = 0;
    void ensure(int64_t npart);
    double* scalar(int i) { return static_cast<double*>(buf->ptr) + i; }
    double* partials() { return static_cast<double*>(buf->ptr) + 64; }
};

struct LinOpLevel {
    LevelPlan plan;
    std::unique_ptr<amrex::DeviceBuffer> a64, b64[3], a32, b32[3];
};

// Fine AMR level of a two-level composite operator (ratio 2, boxes inside the
// domain, box storage).
struct AmrFine {
    LevelPlan plan;
    CfCoef cc{};
    std::vector<RefluxCell> hreflux;
    std::unique_ptr<amrex::DeviceBuffer> d_reflux, d_covered;  // reflux cells; coarse cells under fine boxes
    // setLevelBC(1, data): the fine level's Dirichlet values beyond physical
    // faces (fine layout, in the ghost cells); null: homogeneous
    std::unique_ptr<amrex::DeviceBuffer> d_pbc;
    // MLABecLaplacian coefficients of the fine level (box layout), FP32 copies
    // for the mixed form's fine smoothing
    std::unique_ptr<amrex::DeviceBuffer> a64, b64[3], a32, b32[3];
};

struct LinOpImpl {
    int kind = 0;  // 0 ABecLaplacian, 1 Poisson
    int namr = 1;  // AMR levels
    std::unique_ptr<AmrFine> fine;
    amrex::LPInfo info;
    std::array<amrex::LinOpBCType, 3> lobc{{amrex::LinOpBCType::Dirichlet, amrex::LinOpBCType::Dirichlet,
                                            amrex::LinOpBCType::Dirichlet}};
    std::array<amrex::LinOpBCType, 3> hibc = lobc;
    bool bc_set = false;
    int maxorder = 3;  // AMReX_MLLinOp.H:579
    bool enforce_solvable = true;
    bool dense_requested = true;  // store domain-covering levels as one FAB
    bool dense = false;
    double a_scalar = std::numeric_limits<double>::quiet_NaN();
    double b_scalar = std::numeric_limits<double>::quiet_NaN();
    std::vector<LinOpLevel> lev;
    std::vector<int> same_layout;  // lev[l] BoxArray == coarsened lev[l-1] BoxArray
    std::vector<int> agglomerated;  // lev[l]'s reference BoxArray came from agglomeration
    bool prepared = false, singular = false;
    // bumped by every prepare(): work captured against an older epoch (graphs
    // holding by-value coefficients, boundary pointers, branch choices) is stale
    uint64_t epoch = 0;
    double volinv0 = 0, volinv_bot = 0;
    std::unique_ptr<amrex::DeviceBuffer> bcv;  // inhomogeneous Dirichlet values (level-0 layout)
    // setCoarseFineBC (AMReX_MLLinOp.H:200-212, :1427-1441): the level-0 boxes
    // leave part of the domain uncovered (their union is one box); its faces
    // inside the domain take coarse data at 0.5*ratio*dx (ratio -1: never set,
    // location at the face), Dirichlet or Neumann
    bool needs_cf = false;
    amrex::Box cf_region;  // union of the level-0 boxes when needs_cf
    const amrex::MultiFab* cf_crse = nullptr;
    int cf_ratio = -1;
    int cf_type = 101;
    std::unique_ptr<DeviceScratch> m_scratch;
    std::unique_ptr<amrex::DeviceBuffer> grad_lam, grad_phi;  // level-0 copies for the coefficient gradient

    void define(const amrex::Geometry& g, const amrex::BoxArray& ba, const amrex::LPInfo& info);
    void define_fine(const amrex::Geometry& g, const amrex::BoxArray& ba);
    void prepare();
    // boundary values of the level-0 coarse/fine faces from cf_crse into bcv
    void fill_cf_values();
    void apply_bc();  // set_bc on every MG level (coarse/fine faces of a patch included)
    DeviceScratch& scratch(const LevelPlan& P);
    template <class T>
    OpCoef<T> coef(int l) const;
    // the fine AMR level's operator, and the reflux flux scaling
    template <class T>
    OpCoef<T> fine_coef() const;
    RefluxCoef reflux_coef() const;
};

// Whether consecutive fused sweeps may run as one two-sweep launch (default on).
void set_pair_sweeps(bool on);
bool pair_sweeps_enabled();

// Storage of operators constructed from now on: dense (one FAB per
// domain-covering level) or the reference BoxArrays.
void set_dense_default(bool dense);

// MG levels of one AMR level: (domain, reference BoxArray) per level; agg[l]
// (when given) = 1 where level l's BoxArray was made by agglomeration rather
// than by coarsening the level-0 BoxArray.
std::vector<std::pair<amrex::Box, amrex::BoxArray>> define_grids(const amrex::Geometry& geom,
                                                                const amrex::BoxArray& grids,
                                                                const amrex::LPInfo& info,
                                                                std::vector<int>* agg = nullptr);

// Per-level arrays of a V-cycle in precision V.
template <class V>
struct Cycle {
    std::vector<std::unique_ptr<amrex::DeviceBuffer>> own;
    std::vector<V*> cor, res, rescor;
    std::vector<V*> tmp;  // second correction buffer of levels smoothed out of place
    std::vector<V*> cur;  // buffer holding the level's correction after the cycle
    std::vector<V*> hold;  // interpolated correction of the F-cycle (allocated on first use)
    std::vector<OpCoef<V>> op;
};

struct SolverImpl {
    amrex::MLMG& mg;
    LinOpImpl& lp;
    bool allocated = false;
    amrex::MLMG::Precision alloc_prec = amrex::MLMG::Precision::Double;
    std::unique_ptr<amrex::DeviceBuffer> sol, rhs, res;  // level-0 FP64 state
    std::unique_ptr<amrex::DeviceBuffer> sol2;           // second solution buffer (dense: sol ping-pongs)
    Cycle<double> c64;
    Cycle<float> c32;
    std::unique_ptr<amrex::DeviceBuffer> bot_work, bot_x, bot_b, status;
    DeviceScratch scr;
    double* hpin = nullptr;  // pinned host scalars
    int* hpin_status = nullptr;
    // captured iterations: [4*(V-cycles - 1) + 2*transpose + parity]
    static constexpr int kGraphs = 8;
    hipGraphExec_t exec[kGraphs] = {};
    hipGraph_t graph[kGraphs] = {};
    bool graph_ok[kGraphs] = {};
    std::array<double, 8> graph_key{};
    uint64_t graph_epoch = ~uint64_t(0);  // operator epoch the captured graphs were made against
    bool transpose = false;  // the solve in progress is the adjoint (L^T x = b)
    // two-level composite solve: fine-level FP64 state and the FP64 coarse correction
    std::unique_ptr<amrex::DeviceBuffer> f_sol, f_rhs, f_res, f_cor, f_bsol, f_bcor, c_cor64;
    std::unique_ptr<amrex::DeviceBuffer> f_cor32, f_res32;  // FP32 fine smoothing (mixed form)
    hipGraphExec_t amr_exec = nullptr;
    hipGraph_t amr_graph = nullptr;
    // statistics of the last solve
    int last_bottom_status = 0;
    // where the last forward solve left its solution: 0 sol, 1 sol2, -1 none
    // (the two-level solve keeps the coarse level in sol, the fine in f_sol)
    int sol_at = -1;

    SolverImpl(amrex::MLMG& m, LinOpImpl& l) : mg(m), lp(l) {}
    ~SolverImpl();
    void release_graph();
    // prepare the operator if a setter changed it, and drop captured graphs
    // made against an earlier state of the operator
    void sync_linop();
    void allocate();
    void refresh_ops();
    // from MG level top; plain_top: that level smoothed by the two-colour in-place sweeps
    template <class V>
    void vcycle(Cycle<V>& C, hipStream_t s, int top = 0, bool plain_top = false, bool warm = false);
    template <class V>
    void fcycle(Cycle<V>& C, hipStream_t s);
    template <class V>
    void smooth(Cycle<V>& C, int l, bool skip_first_fill, hipStream_t s);
    bool fused(int l) const;  // level l smoothed by the fused out-of-place sweep
    bool tcycle() const;      // the cycles of the adjoint solve in progress work on L^T
    template <class V>
    bool pair_sweeps(int l) const;  // two consecutive sweeps of level l in one launch
    template <class V>
    void bottom(Cycle<V>& C, hipStream_t s);
    void iteration(hipStream_t s, int parity, bool fmg = false, int ncycles = 1);
    void allocate_amr();
    void amr_fine_residual(hipStream_t s);   // f_res = rhs1 - L1 sol1, boundary from sol0
    void amr_coarse_residual(hipStream_t s); // res = rhs0 - L0 sol0, refluxed
    void amr_iteration(hipStream_t s, bool fmg = false);  // oneIter + computeResidual(1) + fine norm
    amrex::Real solve_amr(const amrex::Vector<amrex::MultiFab*>& a_sol,
                          const amrex::Vector<amrex::MultiFab const*>& a_rhs, amrex::Real tol_rel, amrex::Real tol_abs);
    // prepare the operator and the cycle arrays for V-cycles outside MLMG::solve
    // (the bottom-solver check, prepareLinOp and the allocations of prepareMGcycle)
    void prepare_cycles();
    // the bottom solver of mg resolved and checked (Default: bicgstab), and its
    // code written to the device before a solve's cycles; after them, a
    // bicgcg/cgbicg switch made on the device is taken over by mg
    void check_bottom();
    void start_bottom(hipStream_t s);
    void end_bottom();
    double* sol_buf(int parity);                 // solution read by iteration of this parity
    double* spare_sol();                         // level-0 buffer not holding the last solution
    // face gradients (flux false) or fluxes of AMR level amrlev into out[d],
    // from usol or, when null, the last solution
    void face_values(int amrlev, const std::array<amrex::MultiFab*, 3>& out, const amrex::MultiFab* usol, bool flux,
                     hipStream_t s);
    template <class V>
    void initial_residual(Cycle<V>& C, hipStream_t s);
    void residual(double* r, double* x, const double* b, hipStream_t s);
    double norminf(const double* x, int l, hipStream_t s);  // synchronous
};


}  // namespace mgx
