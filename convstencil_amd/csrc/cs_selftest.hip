// cs_selftest.hip -- on-device check of the FP64 MFMA lane layouts the kernels assume.
//
// One wave per experiment: A is one-hot at lane p, B holds a distinct code per lane,
// so D shows which (row, k) lane p feeds and which B lane meets it in which output
// slot.  The expected layouts are documented in cs_ktable.h and checked by
// tests/test_gpu_mfma_layout.py.
#include "../../include/convstencil_debug.h"
#include "cs_internal.h"

namespace cs {

__device__ __forceinline__ double lane_code(int q) { return 1.0 + q / 64.0; }

__global__ void probe_f64_16x16x4(double *out) {
  const int p = blockIdx.x, l = threadIdx.x;
  d4 c = {0.0, 0.0, 0.0, 0.0};
  d4 d = __builtin_amdgcn_mfma_f64_16x16x4f64(l == p ? 1.0 : 0.0, lane_code(l), c, 0, 0, 0);
  for (int v = 0; v < 4; ++v) out[(p * 64 + l) * 4 + v] = d[v];
}

__global__ void probe_f64_4x4x4(double *out) {
  const int p = blockIdx.x, l = threadIdx.x;
  out[p * 64 + l] = __builtin_amdgcn_mfma_f64_4x4x4f64(l == p ? 1.0 : 0.0, lane_code(l), 0.0, 0, 0, 0);
}

}  // namespace cs

extern "C" int32_t cs_debug_mfma_layout(int32_t which, double *host_out) {
  if (!host_out || (which != 0 && which != 1)) return 1;
  const size_t n = which == 0 ? 64 * 64 * 4 : 64 * 64;
  double *d = nullptr;
  if (hipMalloc(&d, n * sizeof(double)) != hipSuccess) return 3;
  if (which == 0)
    hipLaunchKernelGGL(cs::probe_f64_16x16x4, dim3(64), dim3(64), 0, 0, d);
  else
    hipLaunchKernelGGL(cs::probe_f64_4x4x4, dim3(64), dim3(64), 0, 0, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(host_out, d, n * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? 0 : 4;
}
