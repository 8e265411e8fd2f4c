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
// order (strided per thread, then a fixed tree), times scale[i].
template <typename T>
__global__ __launch_bounds__(kT) void instance_mass_kernel(const T* __restrict__ u, const int64_t* __restrict__ cstart,
                                                           const double* __restrict__ scale, double* __restrict__ out) {
    __shared__ double s[kT];
    const int32_t i = blockIdx.x;
    const int64_t c0 = cstart[i], c1 = cstart[i + 1];
    double acc = 0.0;
    for (int64_t c = c0 + threadIdx.x; c < c1; c += kT) acc += (double)u[c];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
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
        hipLaunchKernelGGL(instance_mass_kernel<float>, dim3(ninst), dim3(kT), 0, stream, (const float*)u, cstart, scale,
                           out);
    else if (dtype == 8)
        hipLaunchKernelGGL(instance_mass_kernel<double>, dim3(ninst), dim3(kT), 0, stream, (const double*)u, cstart,
                           scale, out);
    else
        return stx_fail(STX_EINVAL, "instance_mass: dtype must be 4 or 8");
    return stx_launch_status("instance_mass");
}

}  // extern "C"
