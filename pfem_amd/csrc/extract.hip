// Extractors (srcs/simulation/extractors): the problem's global "mass" and the smallest or
// largest coordinate of the fluid nodes, reduced on the device.  Each workgroup writes one
// partial; the host adds them in workgroup order, so a result does not vary between calls.
#include <cfloat>
#include <mutex>

#include "common.hpp"
#include "hostutil.hpp"
#include "p1geom.hpp"
#include "launch.hpp"

namespace pfem {

// Per-workgroup partial sums of the allocation-free mass entry point
// (pfem_case_mass_noalloc): a fixed device array, one per GPU for the life of the process.
// Host code never passes `scratch` itself to a kernel (its host-side name is not a device
// address); it passes the per-device address that hipGetSymbolAddress returned
// (scratch_partials).  Every call on one GPU shares it.
static __device__ __align__(256) double scratch[kScratchPartials];

namespace {

std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Sum over elements of |e| (rho null) or of |e| times the mean density of the element's
// nodes (IncompNewton/Problem.cpp:173-190, WCompNewton/Problem.cpp:216-240).
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_mass(int N, int E, const int* __restrict__ conn,
                                                 const double* __restrict__ pos, const S* __restrict__ rho,
                                                 double* __restrict__ partials) {
    constexpr int NE = D + 1;
    double acc[1] = {0.0};
    for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < E; e += (long long)gridDim.x * kBlock) {
        int nd[NE];
#pragma unroll
        for (int a = 0; a < NE; ++a) nd[a] = conn[(long long)a * E + e];
        double x[D][NE];
        WGeom<D> G;
        wc_geometry<D>(pos, N, nd, x, G);
        if (rho) {
            double middle = 0;
#pragma unroll
            for (int a = 0; a < NE; ++a) middle += (double)rho[nd[a]];
            middle /= (double)NE;
            acc[0] += middle * G.vol;
        } else {
            acc[0] += G.vol;
        }
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

// Smallest (want_max = 0, from DBL_MAX) or largest (from 0, as MinMaxExtractor.cpp:45-53)
// of x over the nodes that are neither bound nor free.
__global__ __launch_bounds__(kBlock) void k_minmax(int N, const uint8_t* __restrict__ flags,
                                                   const double* __restrict__ x, int want_max,
                                                   double* __restrict__ partials) {
    __shared__ double red[kWaves];
    double v = want_max ? 0.0 : DBL_MAX;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < N; i += (long long)gridDim.x * kBlock) {
        if (flags[i] & (NODE_BOUND | NODE_FREE)) continue;
        v = want_max ? fmax(v, x[i]) : fmin(v, x[i]);
    }
    v = want_max ? wave_max(v) : wave_min(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) r = want_max ? fmax(r, red[w]) : fmin(r, red[w]);
        partials[blockIdx.x] = r;
    }
}

// The partials of one k_mass launch added in workgroup order by one thread: the same
// sequence of additions as pfem_case_mass makes on the host, so the same bits.
__global__ void k_sum_partials(const double* __restrict__ partials, int n, double* __restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += partials[i];
    *out = s;
}

}  // namespace

void init_scratch(int device) {
    PFEM_REQUIRE(device >= 0 && device < kMaxDevices, E_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    int cur = -1;
    PFEM_HIP(hipGetDevice(&cur));
    PFEM_REQUIRE(cur == device, E_STATE, "init_scratch needs its device to be current");
    void* p = nullptr;
    PFEM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PFEM_REQUIRE(p != nullptr, E_HIP, "no device address for the scratch array");
    g_scratch_ptr[device] = static_cast<double*>(p);
}

double* scratch_partials(int device) {
    PFEM_REQUIRE(device >= 0 && device < kMaxDevices, E_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    PFEM_REQUIRE(g_scratch_ptr[device] != nullptr, E_STATE, "scratch of this device not set up (pfem_ctx_create)");
    return g_scratch_ptr[device];
}

void launch_sum_partials(const double* partials, int n, double* out, hipStream_t s) {
    k_sum_partials<<<1, 64, 0, s>>>(partials, n, out);
}

void launch_mass(int dim, int N, int E, const int* conn, const double* pos, const void* rho, int rho_fp32,
                 double* partials, int nblocks, hipStream_t s) {
    if (dim == 2) {
        if (rho_fp32) k_mass<2, float><<<nblocks, kBlock, 0, s>>>(N, E, conn, pos, (const float*)rho, partials);
        else k_mass<2, double><<<nblocks, kBlock, 0, s>>>(N, E, conn, pos, (const double*)rho, partials);
    } else {
        if (rho_fp32) k_mass<3, float><<<nblocks, kBlock, 0, s>>>(N, E, conn, pos, (const float*)rho, partials);
        else k_mass<3, double><<<nblocks, kBlock, 0, s>>>(N, E, conn, pos, (const double*)rho, partials);
    }
}

void launch_minmax(int N, const uint8_t* flags, const double* x, int want_max, double* partials, int nblocks,
                   hipStream_t s) {
    k_minmax<<<nblocks, kBlock, 0, s>>>(N, flags, x, want_max, partials);
}

}  // namespace pfem
