// Instrumentation kernels: the streaming-read and copy ceilings the SpMV is compared
// with, and a probe of which XCD each workgroup of a launch runs on.
#include "common.hpp"
#include "views.hpp"

namespace pfem {

// Sum of n doubles (one partial per workgroup): the HBM read ceiling for the matrix values,
// read with the non-temporal loads the SpMVs use for them (its known byte count also
// calibrates the FETCH_SIZE counter for such loads).
__global__ __launch_bounds__(kBlock) void k_stream_read(const double* __restrict__ a, long long n,
                                                        double* __restrict__ out) {
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        acc += ld_stream(a + i);
    double v[1] = {acc};
    block_sum<1>(v);
    if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

__global__ __launch_bounds__(kBlock) void k_copy(const double* __restrict__ a, double* __restrict__ b, long long n) {
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        b[i] = a[i];
}

// XCC (XCD) id of every workgroup: decoded from the hardware id register as in
// hip's __smid() for gfx94x/gfx95x (xcc id above the shader-engine and CU fields).
__global__ void k_xcd_probe(int* __restrict__ out) {
    if (threadIdx.x == 0) out[blockIdx.x] = (int)(__smid() >> (HW_ID_SE_ID_SIZE + HW_ID_CU_ID_SIZE));
}

void launch_stream_read(const double* a, long long n, double* out, int nblocks, hipStream_t s) {
    k_stream_read<<<nblocks, kBlock, 0, s>>>(a, n, out);
}

void launch_copy(const double* a, double* b, long long n, int nblocks, hipStream_t s) {
    k_copy<<<nblocks, kBlock, 0, s>>>(a, b, n);
}

void launch_xcd_probe(int* out, int nblocks, hipStream_t s) { k_xcd_probe<<<nblocks, 64, 0, s>>>(out); }

}  // namespace pfem
