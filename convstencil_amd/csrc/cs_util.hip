// cs_util.hip -- small device kernels: seeded grid fill and zero fill.
#include "cs_internal.h"
#include "cs_rng.h"

namespace cs {

// One thread per interior cell: interior linear index r -> (z, y, x).
__global__ void fill_random_kernel(double *a, Geo g, uint64_t s0, int64_t total) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= total) return;
  const int64_t x = r % g.n[2];
  const int64_t t = r / g.n[2];
  const int64_t y = t % g.n[1];
  const int64_t z = t / g.n[1];
  a[((g.org[0] + z) * g.ext[1] + (g.org[1] + y)) * g.ext[2] + (g.org[2] + x)] = unit_data(splitmix_at(s0, (uint64_t)r));
}

__global__ void zero_kernel(double *a, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) a[i] = 0.0;
}

__global__ void copy_kernel(const double *__restrict__ src, double *__restrict__ dst, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = src[i];
}

hipError_t launch_copy(const double *src, double *dst, size_t count, hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, src, dst, count);
  return hipGetLastError();
}

hipError_t launch_zero(double *a, size_t count, hipStream_t s) {
  if (count == 0) return hipSuccess;
  size_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a, count);
  return hipGetLastError();
}

// Zeroes the whole padded array, then writes the interior.
hipError_t launch_fill_random(double *a, const Geo &g, uint64_t s0, hipStream_t s) {
  const size_t count = (size_t)(g.ext[0] * g.ext[1] * g.ext[2]);
  hipError_t e = launch_zero(a, count, s);
  if (e != hipSuccess) return e;
  const int64_t total = g.n[0] * g.n[1] * g.n[2];
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_random_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a, g, s0, total);
  return hipGetLastError();
}

}  // namespace cs
