// Coarse/fine kernels of the two-level AMR composite solve (mgx_amr.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mgx_device.H"

namespace mgx {

// Ghost closures of the fine level by closure order NX: c, the coarse/fine
// Dirichlet closure (nodes -1, 1/2, 3/2, 5/2 in fine cells from the face,
// evaluated at -1/2); at a physical domain face (dir + 3*side) the boundary
// type bct and, for Dirichlet, the weights p with the value at the face
// (nodes 0, 1/2, 3/2, 5/2).
struct CfCoef {
    double c[5][4];
    double p[5][4];
    int bct[6];
};

// A coarse cell next to the fine region: fbox[f] = fine box across face f
// (dir + 3*side) when that face is a coarse/fine face, else -1.
struct RefluxCell {
    int i, j, k;
    int fbox[6];
};

// Flux scaling of the reflux: MLPoisson scale 1, bc/bf null (unit b); an
// ABecLaplacian scale -b_scalar with the coarse level's face coefficients bc
// (dense layout) and the fine level's bf (box layout).
struct RefluxCoef {
    double scale;
    const double* bc[3];
    const double* bf[3];
};

int64_t masked_max_partials();

// Face ghosts of the fine level Lf: covered ghosts from the covering box,
// others the closure through bcv (null: homogeneous).
template <class T>
void fine_fill(const LevelDev& Lf, T* phi, const T* bcv, const CfCoef& cc, int maxorder, hipStream_t s);
// One GSRB colour of the fine-level operator (op: the fine level's Poisson or
// ABecLaplacian coefficients).
template <class T>
void fine_gsrb(const LevelDev& Lf, const OpCoef<T>& op, T* phi, const T* rhs, int rb, const CfCoef& cc, int maxorder,
               hipStream_t s);
// One GSRB colour of the fine level with homogeneous boundary data: the face
// ghosts formed inside the colour's pass (closure order <= 3), else fine_fill
// then fine_gsrb.  The face ghosts of phi are left as they were.
template <class T>
void fine_smooth_colour(const LevelDev& Lf, const OpCoef<T>& op, T* phi, const T* rhs, int rb, const CfCoef& cc,
                        int maxorder, hipStream_t s);
// Boundary values of the fine level's face ghosts: coarse/fine ones from the
// dense coarse level Lc, ghosts beyond a physical face from phys (the fine
// level's boundary data in its own layout; null: 0).
void fine_bndry(const LevelDev& Lf, const LevelDev& Lc, const double* crse, const double* phys, double* bcv,
                hipStream_t s);
// crse (cells under the fine boxes) = mean of the 8 fine children.
template <class T>
void fine_avgdown(const LevelDev& Lf, const LevelDev& Lc, T* crse, const T* fine, hipStream_t s);
// Coarse data of a window of coarse cells (lo wlo, extents wn; x fastest) from
// the valid cells of a user FAB set on the coarse domain (extents nc);
// positions beyond a periodic face take the periodic image, others outside
// the domain or not covered by a box are 0.
void cf_window(const double* user, const UserBox* ub, int nbox, int ung, const int* wlo, const int* wn, const int* nc,
               const int* periodic, double* win, hipStream_t s);
// Coarse/fine boundary values (interpbndrydata_*_o3, AMReX_InterpBndryData_3D_K.H:
// 23-118, refinement ratio r) into the face ghosts of the level-0 boxes' faces
// flagged in cfmask; region = union of the level-0 boxes (mask: a position in
// it or its periodic image is covered, beyond a non-periodic face outside).
void cf_bndry(const LevelDev& L, const double* win, const int* wlo, const int* wn, int r, const int* rlo,
              const int* rhi, int64_t max_face, double* bcv, hipStream_t s);
// bc (dense coarse faces along d) = mean of the 4 fine faces of bf for every
// coarse face of the coarsened fine boxes, boxes in order (nbox of them).
void fine_avgdown_faces(const LevelDev& Lf, const LevelDev& Lc, int nbox, int d, double* bc, const double* bf,
                        hipStream_t s);
// fine = trilinear interpolation of crse.
void fine_interp(const LevelDev& Lf, const LevelDev& Lc, double* fine, const double* crse, hipStream_t s);
// res (coarse) += reflux of the operator's flux through the coarse/fine faces;
// sol0 and sol1 need current ghosts.
void reflux(const LevelDev& Lc, const LevelDev& Lf, const RefluxCell* cells, int ncell, double* res,
            const double* sol0, const double* sol1, const RefluxCoef& rc, hipStream_t s);
// out[0] = max |x| over coarse cells with covered == 0 (covered: n^3 bytes, x fastest).
void masked_max(const LevelDev& Lc, const unsigned char* covered, const double* x, double* partials, double* out,
                hipStream_t s);

}  // namespace mgx
