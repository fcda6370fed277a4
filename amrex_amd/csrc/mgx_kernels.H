// Host-callable launchers of the MLMG HIP kernels (gfx950).  Every launcher
// is asynchronous on the given stream and takes device pointers in the level
// layout of mgx_device.H.
#pragma once
#include <hip/hip_runtime.h>

#include "mgx_device.H"

namespace mgx {

// Partials buffer needed by the reductions for a level (in doubles).
int64_t reduce_partials(const LevelDev& L, int max_box_cells);

template <class T> void fill_ghosts(const LevelDev& L, T* phi, const T* bcv, int inhomog, hipStream_t s);
template <class T> void gsrb(const LevelDev& L, int max_box_cells, const OpCoef<T>& op, T* phi, const T* rhs, int redblack, hipStream_t s);
// out = L(x) (b == nullptr) or out = b - L(x); ghosts of x must be current.
template <class T> void apply(const LevelDev& L, int max_box_cells, const OpCoef<T>& op, T* out, const T* x, const T* b, hipStream_t s);
template <class T> void restrict_avg(const LevelDev& Lc, int max_box_cells_c, const LevelDev& Lf, T* crse, const T* fine, int same_layout, hipStream_t s);
template <class T> void prolong_add(const LevelDev& Lf, int max_box_cells_f, const LevelDev& Lc, T* fine, const T* crse, int same_layout, hipStream_t s);
// F-cycle interpolation fine = trilinear(crse) on the valid fine cells; ghosts
// outside one physical face are read from crse's ghost layer (phys != 0) or 0.
template <class T> void fmg_interp(const LevelDev& Lf, const LevelDev& Lc, T* fine, const T* crse, int phys, hipStream_t s);
// y += x on the valid cells only.
template <class T> void add_valid(const LevelDev& L, T* y, const T* x, hipStream_t s);
template <class T> void avgdown_faces(const LevelDev& Lc, const LevelDev& Lf, T* bc, const T* bf, int dir, int same_layout, hipStream_t s);
template <class T> void normalize(const LevelDev& L, int max_box_cells, const OpCoef<T>& op, T* x, hipStream_t s);

// Whole-allocation vector operations (ghost cells included).
template <class T> void set_val(T* x, int64_t n, T v, hipStream_t s);
template <class T> void axpy(T* y, const T* x, T a, int64_t n, hipStream_t s);          // y += a*x
template <class TD, class TS> void convert(TD* y, const TS* x, int64_t n, hipStream_t s);  // y = x
template <class TD, class TS> void add_convert(TD* y, const TS* x, int64_t n, hipStream_t s);  // y += x
// x += -(*scalar_dev) * factor on valid cells (solvability fix without a host round trip)
template <class T> void add_scaled_dev_scalar(const LevelDev& L, int max_box_cells, T* x, const double* scalar_dev, double factor, hipStream_t s);

// Reductions over valid cells; result written to out_dev[0] as double.
// op: 0 max|x|, 1 sum x, 2 sum x*y.
template <class T> void reduce(const LevelDev& L, int max_box_cells, int op, const T* x, const T* y, double* partials, double* out_dev, hipStream_t s);

// Global-array interchange: global layouts of the workload generator.
template <class TL, class TG> void scatter_cells(const LevelDev& L, int max_box_cells, TL* lvl, const TG* glob, int grown, hipStream_t s);
template <class TL, class TG> void gather_cells(const LevelDev& L, int max_box_cells, const TL* lvl, TG* glob, int grown, hipStream_t s);
template <class TL, class TG> void scatter_faces(const LevelDev& L, TL* lvl, const TG* glob, int dir, hipStream_t s);
template <class TL, class TG> void gather_faces(const LevelDev& L, const TL* lvl, TG* glob, int dir, hipStream_t s);
// User FAB set (any box layout, any ghost width) <-> level storage, valid cells.
// ub: device table of the user's boxes, max_cells: largest user box.
template <class TL, class TU> void copy_user_to_level(const LevelDev& L, TL* lvl, const TU* user, const UserBox* ub, int nbox, int64_t max_cells, int user_ng, hipStream_t s);
template <class TL, class TU> void copy_level_to_user(const LevelDev& L, const TL* lvl, TU* user, const UserBox* ub, int nbox, int64_t max_cells, int user_ng, hipStream_t s);
// Face ghosts (one layer) of the user boxes from a level with filled ghosts
// (MLMG::setFinalFillBC); max_face: the largest face of a user box.
void face_ghosts_to_user(const LevelDev& L, const double* lvl, double* user, const UserBox* ub, int nbox,
                         int64_t max_face, int ung, hipStream_t s);
// Zero the edge and corner cells of the first ghost layer of the user boxes;
// max_items: the largest 4(nx+ny+nz)+8 over the boxes.
void zero_user_edge_ghosts(double* user, const UserBox* ub, int nbox, int64_t max_items, int ung, hipStream_t s);
// Dirichlet values in the ghost cells at physical faces (user_ng >= 1).
template <class TL, class TU> void copy_user_grown_to_level(const LevelDev& L, TL* lvl, const TU* user, const UserBox* ub, int nbox, int64_t max_grown_cells, int user_ng, hipStream_t s);
// Face values of direction dir from user face FABs (boxes nodal in dir).
template <class TL, class TU> void copy_user_faces_to_level(const LevelDev& L, TL* lvl, const TU* user, const UserBox* ub, int nbox, int64_t max_cells, int user_ng, int dir, hipStream_t s);
// Face values of direction dir into user face FABs (boxes nodal in dir whose
// cells lie on the level): c * diff (b == nullptr) or ((-c) * b) * diff,
// times post when scale != 0; diff = x(p) - x(p - e_dir), ghosts of x current.
template <class T> void face_values(const LevelDev& L, const T* x, const T* b, const UserBox* ub, int nbox, int64_t max_cells, int user_ng, int dir, T c, int scale, T post, T* user, hipStream_t s);

// Copy between a user FAB set and a global array of extents G (plus gg ghost
// layers): the region of each box grown by rg, clipped to the global array.
template <class TU, class TG>
void user_global(const UserBox* ub_dev, int nbox, int64_t max_region, int ung, TU* user, TG* glob, const int* G, int gg,
                 int rg, int to_user, hipStream_t s);

// Krylov bottom solve in one work-group, iteration loops on the device
// (MLMGT::actualBottomSolve with MLCGSolverT::solve_bicgstab / solve_cg).
// work holds 5 level arrays (p, r, rh, v, t) plus one for the solvable copy
// of b.  status_dev[2] = the bottom solver in force (amrex::BottomSolver
// codes 2 bicgstab, 3 cg, 4 bicgcg, 5 cgbicg), read on entry and rewritten
// when bicgcg/cgbicg switch for good; status_dev[0] = return code,
// status_dev[1] = iterations of every attempt.  On failure the correction is
// zeroed and nsmooth_fail smooths are applied, as actualBottomSolve does.
template <class T> void bottom_solve(const LevelDev& L, const OpCoef<T>& op, T* x, const T* b, T* work,
                                     T eps_rel, T eps_abs, int maxiter, int singular, double volinv,
                                     int nsmooth_ok, int nsmooth_fail, int* status_dev, int transpose,
                                     hipStream_t s);
// The adjoint cycle on dense levels (maxorder <= 3): one colour of the
// transposed sweep x_p += omega / g_m_d (b_p - (L^T x)_p), g_m_d the forward
// sweep's divisor; out = b - L^T x; both with homogeneous boundary data.
template <class T> void gsrb_T(const LevelDev& L, const OpCoef<T>& op, T* phi, const T* rhs, int rb, hipStream_t s);
template <class T> void resid_T(const LevelDev& L, const OpCoef<T>& op, T* out, const T* x, const T* b, hipStream_t s);

}  // namespace mgx
