// Device-side descriptions of an MG level, shared by the host layer and the
// HIP kernels.
//
// Layout contract (same as AMReX's Array4 layout, AMReX_Array4.H:88-90): every
// cell array of a level stores each box as one ghost-grown FAB, x fastest,
// (i - lo.x + 1) + sy*(j - lo.y + 1) + sz*(k - lo.z + 1), boxes packed one
// after another.  All internal cell arrays of a level (solution, correction,
// residual, rhs, a-coefficient) share these offsets.  Face arrays of
// direction d store box extents + 1 along d without ghosts.
#pragma once
#include <cstdint>

namespace mgx {

enum BcType : int { BC_INTERIOR = 0, BC_DIRICHLET = 101, BC_NEUMANN = 102, BC_REFLECT_ODD = 103 };

struct BoxDev {
    int lo[3], hi[3];
    int64_t off;       // first element of this box's ghost-grown cell FAB
    int64_t sy, sz;    // cell strides: nx+2, (nx+2)(ny+2)
    int64_t foff[3];   // first element of the face FABs
    int bct[6];        // per face (dir + 3*side): BC_INTERIOR when the ghost layer
                       // is covered by other boxes / periodic images (mask 0),
                       // else the physical BC type of that domain face (mask 2)
    int nxo[6];        // Dirichlet closure order NX = min(box length + 1, maxorder)
    double bcc[6][4];  // Lagrange weights of the closure: ghost = sum_{m>=1} c_m phi_m + c_0 g
    double cf[6];      // weight of the first interior cell (undrrelxr) used by GSRB
    int cfmask;        // bit f: face f is a coarse/fine face (boundary data from a coarser level)
};

struct LevelDev {
    const BoxDev* box;     // nbox entries
    int nbox;
    int n[3];              // domain extents, domain lo = 0
    int periodic[3];
    const int* slab_of[3]; // per dim: global cell index -> slab index (length n[d])
    const int* box_of;     // [s2][s1][s0] -> owning box id
    int nslab[3];
    const int64_t* cell_start;  // prefix of valid-cell counts, nbox+1 entries
    int64_t ncell;         // valid cells on the level
    int64_t nalloc;        // length of a cell array (ghost-grown FABs)
    int64_t nface[3];      // lengths of the face arrays
    int64_t mbc, mbg, mbf; // largest box: valid cells, ghost-grown cells, face values of one direction
    int mtiles[2];         // largest per-box tile count of the cell kernels (kTileZ, kTileZRed planes)
    int max_nxo;           // largest Dirichlet closure order NX over the boxes' faces (0 without Dirichlet faces)
    int64_t max_face;      // largest face (cells) of any box
};

// Tiles of the level cell kernels: 64 (x) x 4 (y) threads, kTileZ planes in z
// (kTileZRed for reductions).
constexpr int kTileX = 64, kTileY = 4, kTileZ = 8, kTileZRed = 32;
inline int tiles_of(int nx, int ny, int nz, int tz) {
    return ((nx + kTileX - 1) / kTileX) * ((ny + kTileY - 1) / kTileY) * ((nz + tz - 1) / tz);
}

// One box of a user FAB set: valid box in its own index type, first element.
struct UserBox {
    int lo[3], hi[3];
    int64_t off;
};

// Tile of a cell kernel: box id and tile origin.
struct Tile {
    int box, x0, y0, z0;
};

// Operator scalars of one level, already cast the way the reference casts them.
// Compute units of the current device (queried once; 256 on MI355X).  Large
// z-marching launches size their z chunks so that the grid holds one round of
// work-groups.
int device_compute_units();

template <class T>
struct OpCoef {
    int kind;          // 0 ABecLaplacian, 1 Poisson
    T alpha;           // a-scalar
    T dhs[3];          // smoother factor b_scalar / (h*h)  (MLABecLaplacian Fsmooth)
    T dha[3];          // apply factor  b_scalar * dxinv * dxinv (mlabeclap_adotx)
    T dpo[3];          // Poisson factor dxinv*dxinv
    const T* a;        // cell array (level layout)
    const T* b[3];     // face arrays
};

}  // namespace mgx
