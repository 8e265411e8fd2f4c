// Helpers for stacks of independent instances: the instance of every row
// (row ids of instance i start at bases[i]) and one FP64 mass per instance.
#include "stx_common.h"
#include "stx_host.h"

using namespace stx;

namespace {

constexpr int kT = 256;

__global__ void row_instance_kernel(const int32_t* __restrict__ rows, int64_t R, const int32_t* __restrict__ bases,
                                    int32_t ninst, int32_t* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= R) return;
    const int32_t v = rows[k];
    int32_t lo = 0, hi = ninst;  // last i with bases[i] <= v
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (bases[mid] <= v) lo = mid; else hi = mid;
    }
    out[k] = lo;
}

// Block per instance: cells [cstart[i], cstart[i+1]) summed with a fixed
// order that depends only on the instance's cell count n (so an instance's
// mass is the same in any stack): thread t adds the 16-byte vectors t,
// t + kMassThreads, ... of the instance's cells in turn, thread t < n mod V
// then the tail cell n - n mod V + t, and a fixed tree combines the threads.
// 1024 threads and dword-aligned vector loads keep enough bytes in flight
// for a few blocks to stream large instances.
constexpr int kMassThreads = 1024;

template <typename V, typename T>
__device__ __forceinline__ V load_dwords(const T* p) {
    V v;
    __builtin_memcpy(&v, __builtin_assume_aligned(p, 4), sizeof(V));
    return v;
}

template <typename T>
__global__ __launch_bounds__(kMassThreads) void instance_mass_kernel(const T* __restrict__ u,
                                                                    const int64_t* __restrict__ cstart,
                                                                    const double* __restrict__ scale,
                                                                    double* __restrict__ out) {
    constexpr int V = 16 / (int)sizeof(T);
    typedef T VT __attribute__((ext_vector_type(V)));
    __shared__ double s[kMassThreads];
    const int32_t i = blockIdx.x;
    const int64_t c0 = cstart[i], n = cstart[i + 1] - c0;
    const int64_t nv = n / V;
    const T* base = u + c0;
    double acc = 0.0;
#pragma unroll 4
    for (int64_t k = threadIdx.x; k < nv; k += kMassThreads) {
        const VT v = load_dwords<VT>(base + k * V);
#pragma unroll
        for (int j = 0; j < V; ++j) acc += (double)v[j];
    }
    if ((int64_t)threadIdx.x < n - nv * V) acc += (double)base[nv * V + threadIdx.x];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kMassThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[i] = s[0] * scale[i];
}

}  // namespace

extern "C" {

int32_t stx_row_instance(const int32_t* rows, int64_t R, const int32_t* bases, int32_t ninst, int32_t* out,
                         hipStream_t stream) {
    if (R <= 0) return STX_OK;
    if (ninst <= 0) return stx_fail(STX_EINVAL, "row_instance: no instances");
    hipLaunchKernelGGL(row_instance_kernel, dim3((unsigned)((R + kT - 1) / kT)), dim3(kT), 0, stream, rows, R, bases,
                       ninst, out);
    return stx_launch_status("row_instance");
}

int32_t stx_instance_mass(int32_t dtype, const void* u, const int64_t* cstart, int32_t ninst, const double* scale,
                          double* out, hipStream_t stream) {
    if (ninst <= 0) return STX_OK;
    if (dtype == 4)
        hipLaunchKernelGGL(instance_mass_kernel<float>, dim3(ninst), dim3(kMassThreads), 0, stream, (const float*)u,
                           cstart, scale, out);
    else if (dtype == 8)
        hipLaunchKernelGGL(instance_mass_kernel<double>, dim3(ninst), dim3(kMassThreads), 0, stream, (const double*)u,
                           cstart, scale, out);
    else
        return stx_fail(STX_EINVAL, "instance_mass: dtype must be 4 or 8");
    return stx_launch_status("instance_mass");
}

}  // extern "C"
