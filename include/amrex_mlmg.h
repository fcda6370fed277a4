/*
 * C ABI of the MI355X MLMG library (libmgx.so).
 *
 * The amrex_fi_* entry points are the ones the reference exposes for this
 * path (Src/F_Interfaces, bound by its Fortran modules with bind(c)); they
 * keep the reference names, argument order and meaning, so existing bindings
 * link against this library unchanged.  Handles are opaque pointers to this
 * library's own objects.  MultiFab data live in GPU memory.
 *
 * Error behaviour: the reference aborts the process on errors.  Here a failing
 * call records the error (mgx_last_error) and returns; functions returning a
 * value return NaN / 0 on failure.  mgx_set_abort_on_error(1) restores the
 * reference's abort.
 *
 * The mgx_* entry points are extensions with no reference counterpart
 * (device/stream control, precision policy, statistics, interchange with
 * global arrays, per-kernel hooks used by the parity tests).
 */
#ifndef AMREX_MLMG_H_
#define AMREX_MLMG_H_

#ifdef __cplusplus
extern "C" {
#endif

/* ---- BoxArray: Src/F_Interfaces/Base/AMReX_boxarray_fi.cpp -------------- */
void amrex_fi_new_boxarray(void** ba, int lo[3], int hi[3]);                        /* :10 */
void amrex_fi_new_boxarray_from_bxfarr(void** ba, const int* bxs, int nsides,       /* :16 */
                                       int ndims, int nbxs);
void amrex_fi_delete_boxarray(void* ba);                                             /* :33 */
void amrex_fi_clone_boxarray(void** bao, const void* bai);                           /* :38 */
void amrex_fi_boxarray_maxsize(void* ba, int sz[3]);                                 /* :44 */
long long amrex_fi_boxarray_nboxes(const void* ba);                                  /* :50 */
void amrex_fi_boxarray_get_box(const void* ba, int i, int lo[3], int hi[3]);         /* :55 */
long long amrex_fi_boxarray_numpts(const void* ba);                                  /* :79 */

/* ---- Geometry: AMReX_geometry_fi.cpp ------------------------------------- */
void amrex_fi_new_geometry(void** geom, int lo[3], int hi[3]);                       /* :7  */
void amrex_fi_delete_geometry(void* geom);                                           /* :14 */
void amrex_fi_geometry_get_intdomain(const void* geom, int lo[3], int hi[3]);       /* :36 */

/* ---- DistributionMapping: AMReX_distromap_fi.cpp ------------------------- */
void amrex_fi_new_distromap(void** dm, const void* ba);                              /* :9  */
void amrex_fi_delete_distromap(void* dm);                                            /* :20 */

/* ---- MultiFab: AMReX_multifab_fi.cpp ------------------------------------- */
void amrex_fi_new_multifab(void** mf, const void** ba, const void** dm, int nc,     /* :10 */
                           const int* ng, const int* nodal);
void amrex_fi_delete_multifab(void* mf);                                             /* :24 */
int amrex_fi_multifab_ncomp(const void* mf);                                         /* :29 */
void amrex_fi_multifab_ngrow(const void* mf, int* ngv);                              /* :34 */
void amrex_fi_multifab_dataptr_int(void* mf, int igrd, double** dp, int lo[3],      /* :63 */
                                   int hi[3]);
double amrex_fi_multifab_sum(const void* mf, int comp);                              /* :86 */
double amrex_fi_multifab_norm0(const void* mf, int comp);                            /* :91 */
void amrex_fi_multifab_setval(void* mf, double val, int ic, int nc, const int* ng); /* :106 */
/* No counterpart in the reference: amrex_fi_multifab_sum and amrex_fi_multifab_norm0
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight loops, harnesses that forbid hipMalloc).  For components comp..comp+ncomp-1
 * (ncomp at most 16) writes out_dev[2c] = sum and out_dev[2c+1] = max |x| over
 * the valid cells into the caller's device buffer of 2*ncomp doubles,
 * asynchronously on the library stream (no host synchronisation).  The partials
 * live in a fixed 256 KiB device array per GPU, held until the process ends;
 * calls that overlap on one GPU race. */
int mgx_multifab_sum_norm0_noalloc(const void* mf, int comp, int ncomp, double* out_dev);
/* ---- On-disk interchange: Src/F_Interfaces/Base/AMReX_vismf_fi.cpp, AMReX_plotfile_fi.cpp
 * VisMF version-1 files (<name>_H, <name>_D_00000) and "HyperCLaw-V1.1" plotfiles. */
void amrex_fi_write_multifab(const void* mf, const char* name);                      /* vismf_fi.cpp:9 */
void amrex_fi_read_multifab(void* mf, const char* name);                             /* vismf_fi.cpp:14 */
void amrex_fi_write_plotfile(const char* name, int nlevs, const void* mf[],          /* plotfile_fi.cpp:8 */
                             const char* varname[], const void* geom[], double time,
                             const int level_steps[], const int ref_ratio[]);
/* Extensions: a MultiFab defined from a VisMF file; its layout; MultiFab::Copy;
 * PlotFileData (AMReX_PlotFileUtil.H:154-197); MLMG::checkPoint (MLMG.H:1979-2032,
 * which then reports the operator's checkPoint as not implemented, status 2). */
int mgx_read_multifab_new(void** mf, const char* name);
/* VisMF::Write(mf, name, NFiles, set_ghost) (AMReX_VisMF.H:191-194); with set_ghost the
 * MultiFab's ghost cells are overwritten first, as in the reference. */
int mgx_write_multifab_set_ghost(void* mf, const char* name, int set_ghost);
int mgx_multifab_info(const void* mf, int* nbox, int* ncomp, int* ngrow, int nodal[3]);
int mgx_multifab_box(const void* mf, int i, int lo[3], int hi[3]);
int mgx_multifab_copy(void* dst, const void* src, int srccomp, int dstcomp, int numcomp, int nghost);
int mgx_new_plotfile_data(void** pf, const char* name);
void mgx_delete_plotfile_data(void* pf);
int mgx_plotfile_data_info(void* pf, int* ncomp, int* finest_level, double* time, int* coord);
int mgx_plotfile_data_varname(void* pf, int i, char* buf, int len);
int mgx_plotfile_data_level(void* pf, int level, int dom_lo[3], int dom_hi[3], double prob_lo[3], double prob_hi[3],
                            double dx[3], int* ref_ratio, int* level_step);
/* varname NULL: every component */
int mgx_plotfile_data_get(void* pf, int level, const char* varname, void** mf);
int mgx_multigrid_checkpoint(void* mlmg, void* a_sol[], void* a_rhs[], double tol_rel, double tol_abs,
                             const char* file_name);

/* ---- Linear operators: Src/F_Interfaces/LinearSolvers -------------------- */
/* nlevels 1 or 2 (a dense coarse level and fine boxes of ratio 2); amrlev of the
   coefficient setters 0 or, on a two-level operator, 1 (the fine level) */
void amrex_fi_new_abeclaplacian(void** linop, int nlevels, const void* geom[],      /* abeclaplacian_fi.cpp:8 */
                                const void* ba[], const void* dm[], int metric_term,
                                int agglomeration, int consolidation, int max_coarsening_level);
void amrex_fi_abeclap_set_scalars(void* linop, double a, double b);                  /* abeclaplacian_fi.cpp:31 */
void amrex_fi_abeclap_set_acoeffs(void* linop, int amrlev, const void* alpha);       /* abeclaplacian_fi.cpp:37 */
void amrex_fi_abeclap_set_bcoeffs(void* linop, int amrlev, const void* beta[]);      /* abeclaplacian_fi.cpp:43 */
void amrex_fi_new_poisson(void** linop, int nlevels, const void* geom[],            /* poisson_fi.cpp:8 */
                          const void* ba[], const void* dm[], int metric_term,
                          int agglomeration, int consolidation, int max_coarsening_level);
void amrex_fi_delete_linop(void* linop);                                             /* linop_fi.cpp:8 */
void amrex_fi_linop_set_maxorder(void* linop, int ord);                              /* linop_fi.cpp:13 */
void amrex_fi_linop_set_domain_bc(void* linop, const int* ilobc, const int* ihibc); /* linop_fi.cpp:18 */
/* Level-0 boxes that leave part of the domain uncovered (their union must be
 * one box, coarsenable by ratio) get Dirichlet values at 0.5*ratio*dx outside
 * their coarse/fine faces, interpolated from crse (one-component cell data on
 * the domain coarsened by ratio; NULL: zero) at the next set_level_bc call.
 * MLLinOp::setCoarseFineBC, AMReX_MLLinOp.H:200-212. */
void amrex_fi_linop_set_coarse_fine_bc(void* linop, const void* crse, int ratio);   /* linop_fi.cpp:31 */
/* The same with the coarse/fine BC type (101 Dirichlet, 102 homogeneous Neumann,
 * crse then unused): the bc_type argument of MLLinOp::setCoarseFineBC, which the
 * reference's Fortran interface does not expose. */
int mgx_linop_set_coarse_fine_bc_type(void* linop, const void* crse, int ratio, int bc_type);
void amrex_fi_linop_set_level_bc(void* linop, int amrlev, const void* levelbcdata); /* linop_fi.cpp:36 */

/* ---- MLMG: Src/F_Interfaces/LinearSolvers/AMReX_multigrid_fi.cpp --------- */
void amrex_fi_new_multigrid(void** mlmg, void* linop);                               /* :8  */
void amrex_fi_delete_multigrid(void* mlmg);                                          /* :13 */
double amrex_fi_multigrid_solve(void* mlmg, void* a_sol[], void* a_rhs[],           /* :18 */
                                double tol_rel, double tol_abs);
/* a_grad_sol / a_fluxes: 3 face MultiFabs (x, y, z) per AMR level, level-major;
 * face values of the last solve's solution (MLMG::getGradSolution, getFluxes). */
void amrex_fi_multigrid_get_grad_solution(void* mlmg, void* a_grad_sol[]);           /* :27 */
void amrex_fi_multigrid_get_fluxes(void* mlmg, void* a_fluxes[]);                    /* :40 */
void amrex_fi_multigrid_comp_residual(void* mlmg, void* a_res[], void* a_sol[],     /* :53 */
                                      void* a_rhs[]);
void amrex_fi_multigrid_set_verbose(void* mlmg, int v);                              /* :62 */
void amrex_fi_multigrid_set_max_iter(void* mlmg, int n);                             /* :67 */
void amrex_fi_multigrid_set_max_fmg_iter(void* mlmg, int n);                         /* :72 */
void amrex_fi_multigrid_set_fixed_iter(void* mlmg, int n);                           /* :77 */
void amrex_fi_multigrid_set_bottom_solver(void* mlmg, int s);                        /* :82 */
void amrex_fi_multigrid_set_bottom_verbose(void* mlmg, int n);                       /* :99 */
void amrex_fi_multigrid_set_always_use_bnorm(void* mlmg, int f);                     /* :104 */
void amrex_fi_multigrid_set_final_fill_bc(void* mlmg, int f);                        /* :109 */

/* ---- Extensions ---------------------------------------------------------- */
typedef struct mgx_solve_stats {
    int niters;            /* outer MLMG iterations */
    int nlevels;           /* MG levels */
    int last_bottom_status;
    int last_bottom_iters;
    double init_resnorm;   /* ||b - L x0||_inf */
    double rhsnorm;        /* ||b||_inf (after the solvability fix) */
    double final_resnorm;
    double solve_seconds;
    double iter_seconds;
    double res_history[256];
    int nvcycles;          /* V- and F-cycles run (more than niters with refinement V-cycles 0 or 2) */
} mgx_solve_stats;

int mgx_last_error(char* buf, int len);          /* status of the last failing call, message copied */
void mgx_clear_error(void);
void mgx_set_abort_on_error(int flag);
int mgx_set_stream(void* hip_stream);            /* NULL: the library's own stream */
void* mgx_get_stream(void);
int mgx_synchronize(void);
int mgx_abi_version(void);
/* storage of operators constructed afterwards: 1 (default) one FAB per MG level
 * covering the domain, smoothed by the fused single-pass sweep; 0 the
 * reference BoxArrays with box-batched two-colour kernels */
int mgx_set_dense_storage(int dense);

void mgx_new_geometry(void** geom, const int lo[3], const int hi[3], const double prob_lo[3],
                      const double prob_hi[3], const int is_periodic[3]);
void mgx_new_abeclaplacian(void** linop, const void* geom, const void* ba, int agg_grid_size,
                           int agglomeration, int max_coarsening_level);
void mgx_new_poisson(void** linop, const void* geom, const void* ba, int agg_grid_size,
                     int agglomeration, int max_coarsening_level);
/* packed device storage of a MultiFab: all boxes, ghost-grown, x fastest */
void* mgx_multifab_data(void* mf, long long* nelem);
/* global array (device pointer, extents of the index type, gg ghost layers)
 * <-> MultiFab, over each box grown by rg; to_mf selects the direction */
int mgx_multifab_copy_global(void* mf, double* glob, const int dom_n[3], int gg, int rg, int to_mf);
/* whole packed storage of a MultiFab <-> host buffer of n doubles */
int mgx_multifab_host(void* mf, double* host, long long n, int to_mf);
int mgx_multigrid_set_precision(void* mlmg, int mixed);  /* 0 FP64, 1 FP32 V-cycle in FP64 refinement */
/* FP32 V-cycles per FP64 refinement step of the mixed form: 1 (default), 2, or 0 (two, or one when
 * predicted to suffice from the measured contraction); status 2 for other values */
int mgx_multigrid_set_refinement_vcycles(void* mlmg, int n);
int mgx_multigrid_set_use_graph(void* mlmg, int use);
/* Bottom solver by its amrex::BottomSolver code (AMReX_MLLinOp.H:30-32: 0 Default,
 * 1 smoother, 2 bicgstab, 3 cg, 4 bicgcg, 5 cgbicg; 6 hypre and 7 petsc are
 * refused at solve), the C++ MLMG::setBottomSolver, which reaches bicgcg and
 * cgbicg (amrex_fi_multigrid_set_bottom_solver has no code for them).  The
 * getter returns the solver in force: after a solve Default is resolved and a
 * bicgcg/cgbicg switch (AMReX_MLMG.H:1577-1592) is taken over. */
int mgx_multigrid_set_bottom_solver_code(void* mlmg, int code);
int mgx_multigrid_get_bottom_solver_code(void* mlmg);
int mgx_multigrid_set_bottom_tolerance(void* mlmg, double reltol, double abstol);
int mgx_multigrid_set_bottom_max_iter(void* mlmg, int n);
int mgx_multigrid_set_pre_post_smooth(void* mlmg, int nu1, int nu2);
int mgx_multigrid_get_stats(void* mlmg, mgx_solve_stats* st);
int mgx_linop_num_mg_levels(void* linop);
/* level mglev: extents n[3], number of boxes; boxes (6 ints each) if non-NULL */
int mgx_linop_level_info(void* linop, int mglev, int n[3], int* nboxes, int* boxes, int max_boxes);
/* Single operations on MG level mglev (parity hooks).  Device pointers to
 * global arrays: cell (n)^3, grown cell (n+2)^3, faces (n+e_d).
 *   which 0: io (grown) smoothed `count` times against in1 (homogeneous BC)
 *   which 1: out = L(io) (io grown; inhomogeneous level BC on level 0)
 *   which 2: out = in1 - L(io)
 *   which 3: one GSRB colour (count = red/black) after refreshing ghosts
 *   which 4: out = a | bx | by | bz of the level
 *   which 5: io = BiCGStab solution of L x = in1 on the level; out[0..1] = status, iterations
 *   which 6: out (level mglev+1) = restriction of in1 (level mglev)
 *   which 7: io (level mglev) += prolongation of in1 (level mglev+1)
 *   which 8: io (grown) after `count` fused single-pass sweeps (dense storage)
 *   which 9: out = in1 - L(io) by the fused level-0 residual (FP64, dense), then
 *            out[ncell], out[ncell+1] = max|res|, sum res; count = 1: the
 *            transposed residual in1 - L^T(io), homogeneous BC
 *   which 10: out (level mglev+1) = restriction of in1 - L(io), fused (dense)
 *   which 13: io = the Krylov bottom solve of L x = in1; count = solver code
 *             (2 bicgstab, 3 cg, 4 bicgcg, 5 cgbicg) + 16 * iteration limit;
 *             out[0..2] = status, iterations, solver code in force after
 *   which 14: io (grown) after `count` two-sweep launches (dense levels of
 *             whole 64x16 tiles)
 *   which 15: one colour (count = red/black) of the adjoint cycle's transposed
 *             sweep (dense levels)
 *   which 16: out = in1 - L^T(io), homogeneous BC (dense levels)
 *   which 17: 13 on the transposed operator (dense levels)
 * precision: 0 FP64, 1 FP32 kernels (inputs and outputs stay double). */
int mgx_linop_level_op(void* linop, int mglev, int which, int precision, const double* in1, double* io,
                       double* out, int count);

/* Adjoint solve: L^T sol = rhs with homogeneous boundary data; same
 * convergence test, precision policy and return value as
 * amrex_fi_multigrid_solve.  Needs the dense storage of the finest level (the
 * default).  With maxorder <= 3 the V-cycles work on L^T (transposed red-black
 * sweeps, residuals of corrections and bottom solve), so they take the forward
 * solve's iterations; with a higher closure order they are the V-cycles of L
 * with the outer residual rhs - L^T sol. */
double mgx_multigrid_solve_adjoint(void* mlmg, void* a_sol[], void* a_rhs[], double tol_rel, double tol_abs);
/* GMRES with the MLMG V-cycle as right preconditioner, amrex::GMRESMLMG
 * (Src/LinearSolvers/AMReX_GMRES_MLMG.H:18-231; the reference has no Fortran
 * interface for it).  One AMR level.  mgx_gmres_mlmg_solve: sol (initial
 * guess in, solution out) and rhs are cell MultiFabs on the operator's
 * BoxArray; returns the final 2-norm residual estimate (NaN on error).  The
 * preconditioner runs in the MLMG object's precision. */
int mgx_new_gmres_mlmg(void** gmres, void* mlmg);                                    /* GMRES_MLMG.H:111 */
void mgx_delete_gmres_mlmg(void* gmres);
double mgx_gmres_mlmg_solve(void* gmres, void* sol, const void* rhs, double tol_rel, double tol_abs); /* :213 */
void mgx_gmres_mlmg_set_verbose(void* gmres, int v);                                  /* :37 */
void mgx_gmres_mlmg_set_max_iters(void* gmres, int n);                                /* :40 */
void mgx_gmres_mlmg_set_restart_length(void* gmres, int n);                           /* GMRES.H:67 */
void mgx_gmres_mlmg_set_precond_num_iters(void* gmres, int n);                        /* :97 */
int mgx_gmres_mlmg_use_precond(void* gmres, int flag);  /* :94; returns the previous flag */
void mgx_gmres_mlmg_set_property_of_zero(void* gmres, int flag);                      /* :62 */
/* niters, status (0 converged, 1 iteration limit, -1 not run), residual norm and
 * up to max_history per-iteration residual estimates; returns the history length. */
int mgx_gmres_mlmg_stats(void* gmres, int* niters, int* status, double* resnorm, double* history, int max_history);
/* Face fluxes of a given solution, MLMG::getFluxes(a_flux, a_sol)
 * (AMReX_MLMG.H:73-76, no Fortran-interface counterpart): a_fluxes as for
 * amrex_fi_multigrid_get_fluxes, a_sol one cell MultiFab per AMR level (not
 * written).  Coarse/fine ghosts of a fine level come from the last solve. */
int mgx_multigrid_get_fluxes_sol(void* mlmg, void* a_fluxes[], void* a_sol[]);
/* MLMG::getFluxes(Vector<MultiFab*>, [a_sol,] Location::CellCenter) (AMReX_MLMG.H:667-745): one
 * cell MultiFab of >= 3 components per AMR level on the level's BoxArray; a_sol NULL: the last
 * solution.  The reference FFI has no cell-centred form. */
int mgx_multigrid_get_cell_fluxes(void* mlmg, void* a_fluxes[], void* a_sol[]);
/* Gradient of J(phi) through the solve L(a, b) phi = rhs of an ABecLaplacian,
 * given lam = L^{-T} dJ/dphi: grad_a (cell MultiFab or NULL) = -alpha*lam*phi,
 * grad_b[d] (face MultiFabs or NULL; grad_b itself may be NULL) =
 * -lam . d(L phi)/db_d including the boundary closure and inhomogeneous level
 * BC of phi.  Outputs are overwritten, or added to when accumulate != 0. */
int mgx_abeclap_coeff_gradient(void* linop, const void* lam, const void* phi, void* grad_a, void* grad_b[3],
                               int accumulate);

/* Times nsweeps smoothing sweeps of level mglev with HIP events on the library
 * stream: out_ms[0] per red+black sweep (the fused launch on dense storage,
 * else ghost fills + both colours), out_ms[1] per GSRB launch, out_ms[2] per
 * ghost-fill launch. */
int mgx_linop_time_smoother(void* linop, int mglev, int precision, int nsweeps, double* out_ms);
/* Times npairs launches of the two-sweep GSRB march on level mglev (a dense
 * level of whole 64x16 tiles); out_ms = mean time of one launch (two sweeps). */
int mgx_linop_time_sweep_pairs(void* linop, int mglev, int precision, int npairs, double* out_ms);
/* Whether V-cycles run two consecutive smoothing sweeps of a level as one
 * launch where the level allows (default 1); returns the previous setting.
 * Results are the same either way; the switch exists for timing. */
int mgx_set_pair_sweeps(int on);
/* HBM streaming probe, the practical bandwidth ceiling of the stencil kernels:
 * mode 0 dst = src, 1 read src, 2 write dst (16-byte aligned device buffers,
 * nbytes a multiple of 16), non-temporal 16-byte accesses; out_ms = mean time
 * of reps launches after one warm-up, HIP events on the library stream. */
int mgx_hbm_probe(int mode, const void* src, void* dst, long long nbytes, int reps, double* out_ms);

#ifdef __cplusplus
}
#endif
#endif /* AMREX_MLMG_H_ */
