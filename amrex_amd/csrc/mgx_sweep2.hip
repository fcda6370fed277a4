// Launch of the fused two-sweep GSRB march (mgx_sweep2.H) on dense levels.
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "mgx_buffer.H"
#include "mgx_device.H"
#include "mgx_smoother.H"
#include "mgx_sweep2.H"

namespace mgx {
namespace {

// Device memory access of the march: raw buffer loads and stores through
// whole-array scalar resources, the plane as a scalar offset.
struct DevMem {
    using Rsrc = mgx::Rsrc;
    __device__ __forceinline__ static Rsrc rsrc(const char* b, uint32_t n) { return make_rsrc_n(b, n); }
    template <class T>
    __device__ __forceinline__ static T ld(Rsrc r, uint32_t off, uint32_t soff) {
        return bld<T>(r, off, soff);
    }
    template <class T>
    __device__ __forceinline__ static Two<T> ld2(Rsrc r, uint32_t off, uint32_t soff) {
        return bld2<T>(r, off, soff);
    }
    template <class T>
    __device__ __forceinline__ static void st2(Rsrc r, uint32_t off, uint32_t soff, T a, T b) {
        bst2<T>(r, off, soff, a, b);
    }
    __device__ __forceinline__ static bool any(bool b) { return __builtin_amdgcn_ballot_w64(b) != 0; }
};

constexpr int kTX = 64, kTY = 16;
constexpr int kCz = 512;  // planes per z chunk (each chunk re-runs six planes below it)
constexpr int kNT = sw2::Geo<kTX, kTY>::NT;  // 512 tile threads + 256 ring threads
static_assert(kNT == 768, "the launch bounds below");

// The prologue, then steps in groups of four (the plane buffers' period); up
// to three steps past z1 write no output.  The parity of the first step is
// the same on every lane of a wave.
template <class TH, class T>
__device__ __forceinline__ void march(TH& th, T* S) {
    th.prologue(S);
    __syncthreads();
    if (__builtin_amdgcn_readfirstlane(th.p0)) {
        for (int k = th.z0 - 6; k < th.z1; k += 4) {
            th.template step<0, 1>(S, k);
            __syncthreads();
            th.template step<1, 0>(S, k + 1);
            __syncthreads();
            th.template step<2, 1>(S, k + 2);
            __syncthreads();
            th.template step<3, 0>(S, k + 3);
            __syncthreads();
        }
    } else {
        for (int k = th.z0 - 6; k < th.z1; k += 4) {
            th.template step<0, 0>(S, k);
            __syncthreads();
            th.template step<1, 1>(S, k + 1);
            __syncthreads();
            th.template step<2, 0>(S, k + 2);
            __syncthreads();
            th.template step<3, 1>(S, k + 3);
            __syncthreads();
        }
    }
}

template <class T, int KIND, bool ZERO, bool CRSE>
__global__ __launch_bounds__(kNT) void k_gsrb_sweep2(sw2::Args<T> args) {
    using G = sw2::Geo<kTX, kTY>;
    __shared__ T S[G::NSLOT * G::PS];
    __shared__ sw2::Closure<T> CL;
    if (threadIdx.x < 6) {
        CL.k1[threadIdx.x] = args.cl.k1[threadIdx.x];
        CL.k2[threadIdx.x] = args.cl.k2[threadIdx.x];
        CL.cf[threadIdx.x] = args.cl.cf[threadIdx.x];
    }
    // waves 0 .. 7 own the tile's columns, waves 8 .. 11 the ring's cells;
    // the role is the same on every lane of a wave, and every wave passes the
    // same barriers: the prologue's, then one per step of the same march
    const int t = (int)threadIdx.x;
    if (__builtin_amdgcn_readfirstlane(t) < G::NTT) {
        sw2::TileThread<T, KIND, kTX, kTY, ZERO, CRSE, DevMem> th;
        th.init(args, &CL, (int)blockIdx.x, (int)gridDim.x, t);
        march(th, S);
    } else {
        sw2::RingThread<T, KIND, kTX, kTY, ZERO, CRSE, DevMem> th;
        th.init(args, &CL, (int)blockIdx.x, (int)gridDim.x, t - G::NTT);
        march(th, S);
    }
}

template <class T, int KIND>
void launch(const sw2::Args<T>& a, bool zero, bool crse, hipStream_t s) {
    const int nz = (a.n2 + a.cz - 1) / a.cz;
    const dim3 grid((unsigned)((a.n0 / kTX) * (a.n1 / kTY) * nz)), block(kNT);
    if (zero && crse)
        hipLaunchKernelGGL((k_gsrb_sweep2<T, KIND, true, true>), grid, block, 0, s, a);
    else if (crse)
        hipLaunchKernelGGL((k_gsrb_sweep2<T, KIND, false, true>), grid, block, 0, s, a);
    else if (zero)
        hipLaunchKernelGGL((k_gsrb_sweep2<T, KIND, true, false>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((k_gsrb_sweep2<T, KIND, false, false>), grid, block, 0, s, a);
}

}  // namespace

// z chunk of a launch: at most 512 planes, halved (down to 16) while the grid
// has fewer work-groups than the GPU has CUs; each chunk re-runs six planes
// below it, so the longest chunks that still fill the CUs re-read least
// (512^3: one chunk per tile, 256 work-groups).
// One work-group of 768 threads at 168 VGPRs fills a CU (three waves per
// SIMD), so a launch of one work-group per CU runs in a single round.
int device_compute_units() {
    static int n = 0;
    if (n == 0) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            n = cus;
        else
            n = 256;  // MI355X
    }
    return n;
}
static int chunk_planes(const LevelDev& L) {
    const int64_t tiles = (int64_t)(L.n[0] / kTX) * (L.n[1] / kTY);
    int cz = L.n[2] < kCz ? L.n[2] : kCz;
    while (cz > 16 && tiles * ((L.n[2] + cz - 1) / cz) < device_compute_units()) cz = (cz + 1) / 2;
    return cz;
}

bool gsrb_sweep2_fits(const LevelDev& L) {
    if (L.nbox != 1) return false;
    const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
    // whole tiles in x and y (even extents, so colours agree across a periodic wrap)
    if (n0 % kTX != 0 || n1 % kTY != 0 || n2 < 8) return false;
    if (L.periodic[2] && (n2 % 2) != 0) return false;
    // 32-bit byte offsets into every array (FP64 bound, ghost-grown FAB and faces)
    if ((int64_t)(n0 + 2) * (n1 + 2) * (n2 + 2) * 8 >= 0xffffffffLL) return false;
    return true;
}

bool gsrb_sweep2_pays(const LevelDev& L) {
    if (!gsrb_sweep2_fits(L)) return false;
    const int cz = chunk_planes(L);
    const int64_t wgs = (int64_t)(L.n[0] / kTX) * (L.n[1] / kTY) * ((L.n[2] + cz - 1) / cz);
    return wgs >= 256 && cz >= 16;
}

template <class T>
void gsrb_sweep2(const LevelDev& L, const BoxDev& B, const OpCoef<T>& op, const T* pin, T* pout, const T* rhs,
                 int zero_in, const T* crse, const BoxDev* Bc, hipStream_t s) {
    if (!gsrb_sweep2_fits(L)) throw std::runtime_error("gsrb_sweep2: level shape not supported");
    if (crse && !Bc) throw std::runtime_error("gsrb_sweep2: coarse correction without its level");
    sw2::Args<T> a{};
    a.n0 = L.n[0];
    a.n1 = L.n[1];
    a.n2 = L.n[2];
    a.per0 = L.periodic[0];
    a.per1 = L.periodic[1];
    a.per2 = L.periodic[2];
    for (int f = 0; f < 6; ++f) {
        a.face_on[f] = B.bct[f] != BC_INTERIOR;
        T k1 = T(0), k2 = T(0);
        if (B.bct[f] == BC_NEUMANN) {
            k1 = T(1);
        } else if (B.bct[f] == BC_REFLECT_ODD) {
            k1 = T(-1);
        } else {
            if (B.nxo[f] > 1) k1 = T(B.bcc[f][1]);
            if (B.nxo[f] > 2) k2 = T(B.bcc[f][2]);
        }
        a.cl.k1[f] = k1;
        a.cl.k2[f] = k2;
        a.cl.cf[f] = T(B.cf[f]);
    }
    a.alpha = op.alpha;
    for (int d = 0; d < 3; ++d) a.dh[d] = op.kind == 0 ? op.dhs[d] : op.dpo[d];
    a.pin = pin + B.off;
    a.pout = pout + B.off;
    a.rhs = rhs + B.off;
    a.a = op.kind == 0 ? op.a + B.off : nullptr;
    for (int d = 0; d < 3; ++d) a.b[d] = op.kind == 0 ? op.b[d] + B.foff[d] : nullptr;
    a.sy = (int)B.sy;
    a.sz = B.sz;
    if (crse) {
        const BoxDev& C = *Bc;
        a.crse = crse + C.off;
        a.csy = (int)C.sy;
        a.csz = C.sz;
    } else {
        a.crse = pin + B.off;  // never read
        a.csy = a.sy;
        a.csz = a.sz;
    }
    a.cz = chunk_planes(L);
    if (op.kind == 0)
        launch<T, 0>(a, zero_in != 0, crse != nullptr, s);
    else
        launch<T, 1>(a, zero_in != 0, crse != nullptr, s);
}

template void gsrb_sweep2<double>(const LevelDev&, const BoxDev&, const OpCoef<double>&, const double*, double*,
                                  const double*, int, const double*, const BoxDev*, hipStream_t);
template void gsrb_sweep2<float>(const LevelDev&, const BoxDev&, const OpCoef<float>&, const float*, float*,
                                 const float*, int, const float*, const BoxDev*, hipStream_t);

}  // namespace mgx
