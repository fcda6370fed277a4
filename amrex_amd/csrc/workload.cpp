// Seeded synthetic workloads for the MLMG benchmark configurations.
//
// Every random draw is a pure function of (seed, array id, global index), so
// the same field comes out whatever the box decomposition, thread count or
// consumer (CPU oracle, GPU solver, benchmark).  Generator: SplitMix64
// (Steele, Lea & Flood, "Fast splittable pseudorandom number generators",
// OOPSLA 2014) applied to a counter key; normals by Box-Muller on counters
// 2n and 2n+1.
//
// Global layouts (domain lo = 0, x fastest):
//   cell  (nx,ny,nz):        i + nx*(j + ny*k)
//   x-face (nx+1,ny,nz):     i + (nx+1)*(j + ny*k)
//   y-face (nx,ny+1,nz):     i + nx*(j + (ny+1)*k)
//   z-face (nx,ny,nz+1):     i + nx*(j + ny*k)
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <vector>

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;  // 2^64 / golden ratio

inline uint64_t splitmix64(uint64_t z) {
    z += kGolden;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t key(uint64_t seed, uint32_t array_id, uint64_t idx) {
    return (seed * kGolden) ^ (static_cast<uint64_t>(array_id) << 48) ^ idx;
}

inline double uniform01(uint64_t seed, uint32_t id, uint64_t idx) {
    return static_cast<double>(splitmix64(key(seed, id, idx)) >> 11) * 0x1.0p-53;
}

inline double normal01(uint64_t seed, uint32_t id, uint64_t n) {
    const double u1 = uniform01(seed, id, 2 * n);
    const double u2 = uniform01(seed, id, 2 * n + 1);
    // 1-u1 lies in (0,1], so the log is finite.
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * M_PI * u2);
}

// One pass of the 7-point box filter (mean of a point and its in-range face
// neighbours) on an (mx,my,mz) grid, x fastest.  Neighbours outside the grid
// are dropped from both the sum and the count.
void box_filter_pass(const std::vector<double>& in, std::vector<double>& out,
                     int mx, int my, int mz) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < mz; ++k) {
        for (int j = 0; j < my; ++j) {
            for (int i = 0; i < mx; ++i) {
                const int64_t p = i + (int64_t)mx * (j + (int64_t)my * k);
                double s = in[p];
                int c = 1;
                if (i > 0)      { s += in[p - 1]; ++c; }
                if (i < mx - 1) { s += in[p + 1]; ++c; }
                if (j > 0)      { s += in[p - mx]; ++c; }
                if (j < my - 1) { s += in[p + mx]; ++c; }
                if (k > 0)      { s += in[p - (int64_t)mx * my]; ++c; }
                if (k < mz - 1) { s += in[p + (int64_t)mx * my]; ++c; }
                out[p] = s / c;
            }
        }
    }
}

}  // namespace

extern "C" {

uint64_t mgx_wl_splitmix64(uint64_t x) { return splitmix64(x); }

void mgx_wl_uniform(uint64_t seed, int array_id, int64_t n, double lo, double hi,
                    double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        out[i] = lo + (hi - lo) * uniform01(seed, (uint32_t)array_id, (uint64_t)i);
    }
}

void mgx_wl_normal(uint64_t seed, int array_id, int64_t n, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        out[i] = normal01(seed, (uint32_t)array_id, (uint64_t)i);
    }
}

// Face coefficient field for direction dir on an (nx,ny,nz) cell domain:
// b = exp(0.5*g), g = N(0,1) per face smoothed by 3 box-filter passes on the
// face grid of that direction.  Array ids 10,11,12 for x,y,z.
void mgx_wl_bface(uint64_t seed, const int* n, int dir, double* out) {
    int m[3] = {n[0], n[1], n[2]};
    m[dir] += 1;
    const int64_t cnt = (int64_t)m[0] * m[1] * m[2];
    std::vector<double> g(cnt), h(cnt);
    mgx_wl_normal(seed, 10 + dir, cnt, g.data());
    for (int pass = 0; pass < 3; ++pass) {
        box_filter_pass(g, h, m[0], m[1], m[2]);
        g.swap(h);
    }
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < cnt; ++p) out[p] = std::exp(0.5 * g[p]);
}

// Variable-coefficient ABecLaplacian recipe: a ~ U[0.5,1.5] per cell (id 1),
// rhs ~ N(0,1) per cell (id 2), face b as in mgx_wl_bface.
void mgx_wl_abec(uint64_t seed, const int* n, double* a, double* bx, double* by,
                 double* bz, double* rhs) {
    const int64_t nc = (int64_t)n[0] * n[1] * n[2];
    if (a) mgx_wl_uniform(seed, 1, nc, 0.5, 1.5, a);
    if (rhs) mgx_wl_normal(seed, 2, nc, rhs);
    if (bx) mgx_wl_bface(seed, n, 0, bx);
    if (by) mgx_wl_bface(seed, n, 1, by);
    if (bz) mgx_wl_bface(seed, n, 2, bz);
}

// Periodic Poisson right-hand side: sum of nmodes Fourier modes
// A_m * {sin|cos}(2*pi*k_m.x), k_m uniform integers in [1,8]^3, A_m ~ N(0,1),
// sin or cos by a fair coin, evaluated at cell centres of the unit cube, then
// the discrete mean subtracted.  Mode m uses counters 8m..8m+5 of array id 20.
// modes_out (may be null) receives 5 doubles per mode: kx, ky, kz, is_sin, A.
void mgx_wl_fourier_rhs(uint64_t seed, const int* n, int nmodes, double* rhs,
                        double* modes_out) {
    std::vector<double> md(5 * nmodes);
    for (int m = 0; m < nmodes; ++m) {
        for (int d = 0; d < 3; ++d) {
            md[5 * m + d] = 1.0 + std::floor(8.0 * uniform01(seed, 20, 8 * m + d));
        }
        md[5 * m + 3] = uniform01(seed, 20, 8 * m + 3) < 0.5 ? 1.0 : 0.0;
        md[5 * m + 4] = normal01(seed, 20, 4 * m + 2);  // counters 8m+4, 8m+5
    }
    const int64_t nc = (int64_t)n[0] * n[1] * n[2];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k) {
        for (int j = 0; j < n[1]; ++j) {
            for (int i = 0; i < n[0]; ++i) {
                const double x = (i + 0.5) / n[0], y = (j + 0.5) / n[1], z = (k + 0.5) / n[2];
                double f = 0.0;
                for (int m = 0; m < nmodes; ++m) {
                    const double arg = 2.0 * M_PI *
                        (md[5 * m] * x + md[5 * m + 1] * y + md[5 * m + 2] * z);
                    f += md[5 * m + 4] * (md[5 * m + 3] != 0.0 ? std::sin(arg) : std::cos(arg));
                }
                rhs[i + (int64_t)n[0] * (j + (int64_t)n[1] * k)] = f;
            }
        }
    }
    double s = 0.0;
    for (int64_t p = 0; p < nc; ++p) s += rhs[p];
    const double mean = s / (double)nc;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < nc; ++p) rhs[p] -= mean;
    if (modes_out) std::memcpy(modes_out, md.data(), sizeof(double) * 5 * nmodes);
}

// Fine boxes of the two-level AMR workload: nbox non-overlapping boxes in the
// fine index space (refinement ratio 2 of an (nc^3) coarse domain), edge
// lengths drawn from {16, 24, 32, 48, 64}, lower corners at even indices, each
// box inside the central half [nf/4, 3nf/4) of the fine domain.  Candidate m
// draws counters 6m..6m+5 of array id 40 (three edges, three corners); a
// candidate overlapping an accepted box is rejected.  boxes receives 6 ints
// per box (lo xyz, hi xyz).  Returns the number of boxes placed.
int mgx_wl_amr_boxes(uint64_t seed, int ncoarse, int nbox, int* boxes) {
    static const int edges[5] = {16, 24, 32, 48, 64};
    const int nf = 2 * ncoarse;
    const int lo_min = nf / 4, hi_max = 3 * nf / 4;  // boxes inside [lo_min, hi_max)
    int placed = 0;
    for (uint64_t m = 0; placed < nbox && m < 1000000; ++m) {
        int lo[3], hi[3];
        bool ok = true;
        for (int d = 0; d < 3; ++d) {
            const int e = edges[std::min(4, (int)(5.0 * uniform01(seed, 40, 6 * m + d)))];
            const int slots = (hi_max - lo_min - e) / 2 + 1;  // even lower corners that fit
            if (slots <= 0) { ok = false; break; }
            const int c = std::min(slots - 1, (int)(slots * uniform01(seed, 40, 6 * m + 3 + d)));
            lo[d] = lo_min + 2 * c;
            hi[d] = lo[d] + e - 1;
        }
        if (!ok) continue;
        for (int q = 0; q < placed && ok; ++q) {
            const int* b = boxes + 6 * q;
            bool sep = false;
            for (int d = 0; d < 3; ++d)
                if (hi[d] < b[d] || lo[d] > b[3 + d]) sep = true;
            if (!sep) ok = false;
        }
        if (!ok) continue;
        for (int d = 0; d < 3; ++d) { boxes[6 * placed + d] = lo[d]; boxes[6 * placed + 3 + d] = hi[d]; }
        ++placed;
    }
    return placed;
}

// Gaussian blobs of the AMR workload: nblob blobs of width sigma, centres
// uniform in [0.25, 0.75]^3 (counters 8m..8m+2 of array id 41), amplitudes
// N(0,1) (counters 8m+4, 8m+5).  blobs receives 4 doubles per blob (cx, cy,
// cz, A).
void mgx_wl_blobs(uint64_t seed, int nblob, double* blobs) {
    for (int m = 0; m < nblob; ++m) {
        for (int d = 0; d < 3; ++d) blobs[4 * m + d] = 0.25 + 0.5 * uniform01(seed, 41, 8 * m + d);
        blobs[4 * m + 3] = normal01(seed, 41, 4 * m + 2);
    }
}

// Sum of the blobs at the cell centres of a box [lo, hi] of a level with cell
// size h (unit cube), minus `shift`; out is the box's cells, x fastest.
void mgx_wl_blob_field(int nblob, const double* blobs, double sigma, const int* lo, const int* hi, double h,
                       double shift, double* out) {
    const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    const double inv2s2 = 1.0 / (2.0 * sigma * sigma);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k) {
        for (int j = 0; j < ny; ++j) {
            for (int i = 0; i < nx; ++i) {
                const double x = (lo[0] + i + 0.5) * h, y = (lo[1] + j + 0.5) * h, z = (lo[2] + k + 0.5) * h;
                double f = 0.0;
                for (int m = 0; m < nblob; ++m) {
                    const double dx = x - blobs[4 * m], dy = y - blobs[4 * m + 1], dz = z - blobs[4 * m + 2];
                    f += blobs[4 * m + 3] * std::exp(-(dx * dx + dy * dy + dz * dz) * inv2s2);
                }
                out[i + (int64_t)nx * (j + (int64_t)ny * k)] = f - shift;
            }
        }
    }
}

}  // extern "C"
