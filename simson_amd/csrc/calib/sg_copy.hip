// Plain device copy of a known number of bytes, the reference against which
// rocprofv3 FETCH_SIZE / WRITE_SIZE readings of the step kernels are scaled.
#include <hip/hip_runtime.h>

#include <cstddef>

namespace {

__global__ void k_fill(double *a, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        a[i] = 1.0;
}

__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

}  // namespace

// Copies nbytes (rounded down to whole doubles) reps times; 0 on success.
extern "C" int simson_calib_copy(size_t nbytes, int reps)
{
    const size_t n = nbytes / sizeof(double);
    double *a = nullptr, *b = nullptr;
    if (n == 0 || hipMalloc((void **)&a, n * sizeof(double)) != hipSuccess)
        return -1;
    if (hipMalloc((void **)&b, n * sizeof(double)) != hipSuccess) {
        (void)hipFree(a);
        return -1;
    }
    hipLaunchKernelGGL(k_fill, dim3(4096), dim3(256), 0, 0, a, n);
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
    const int st = hipDeviceSynchronize() == hipSuccess ? 0 : -2;
    (void)hipFree(a);
    (void)hipFree(b);
    return st;
}
