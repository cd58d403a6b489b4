// k1d_ablation.hip -- the FMA kernels of the 1-D --breakdown ladder (FP64).
//
// The reference's 1-D ladder (src/1d/gpu_1r.cu:77-218, launchers :305-651, run by
// src/1d/main.cu:224-228) times the same pre-fused 7-point stencil with one optimisation
// added per rung.  On MI355X the rungs are:
//   1  every output summed with FP64 FMAs from neighbours read straight from HBM/L2
//      (k1d_abl_global, here);
//   2  the tile and its halo staged in LDS once, FP64 FMAs from LDS (k1d_abl_lds, here);
//   3  the v_mfma_f64_4x4x4 kernel on the reference's tiling, 1024 outputs per
//      256-thread workgroup (k1d_f64 with that tile, k1d_f64.hip);
//   4  the production single-step kernel (2048 outputs per 512-thread workgroup);
// and the production launcher last.  Like the 2-D ladder's kernels they run only when a
// plan is forced to them by id (Variant::mf 1 and 2 mark them).  Every kernel reads
// neighbours without bounds checks: plans require a halo of at least R cells.
#include "cs_internal.h"

namespace cs {

// Rung 1: one thread per output cell.
template <int R>
__global__ __launch_bounds__(256) void k1d_abl_global(const double *__restrict__ in, double *__restrict__ out,
                                                      const double *__restrict__ w, Geo g) {
  const int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (x >= g.n[2]) return;
  const int64_t c = g.org[2] + x;
  double acc = 0.0;
#pragma unroll
  for (int d = -R; d <= R; ++d) acc = fma(w[d + R], in[c + d], acc);
  out[c] = acc;
}

// Rung 2: L outputs and their halo loaded into LDS once (zero past the array), then every
// output summed with FP64 FMAs from LDS.
template <int R, int L>
__global__ __launch_bounds__(256) void k1d_abl_lds(const double *__restrict__ in, double *__restrict__ out,
                                                   const double *__restrict__ w, Geo g) {
  constexpr int E = L + 2 * R;
  __shared__ double tile[E];
  const int64_t X0 = (int64_t)blockIdx.x * L - R;  // interior coordinate of tile[0]
  for (int i = threadIdx.x; i < E; i += 256) {
    const int64_t ax = g.org[2] + X0 + i;  // >= 0: the halo is at least R cells
    tile[i] = ax < g.ext[2] ? in[ax] : 0.0;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < L; i += 256) {
    const int64_t x = X0 + R + i;
    if (x >= g.n[2]) break;  // later i are further right
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d <= 2 * R; ++d) acc = fma(w[d], tile[i + d], acc);
    out[g.org[2] + x] = acc;
  }
}

template <int R>
static hipError_t launch_abl1_global(const double *in, double *out, const double *w, const Geo &g, hipStream_t s) {
  const int64_t blocks = (g.n[2] + 255) / 256;
  if (blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((k1d_abl_global<R>), dim3((unsigned)blocks), dim3(256), 0, s, in, out, w, g);
  return hipGetLastError();
}

template <int R, int L>
static hipError_t launch_abl1_lds(const double *in, double *out, const double *w, const Geo &g, hipStream_t s) {
  const int64_t blocks = (g.n[2] + L - 1) / L;
  if (blocks > 0x7fffffff) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((k1d_abl_lds<R, L>), dim3((unsigned)blocks), dim3(256), 0, s, in, out, w, g);
  return hipGetLastError();
}

static hipError_t prepare_none1() { return hipSuccess; }

// ids 910000 + rung * 10000 + R * 1000
#define CS_VABL1G(ID, R)                                                                                     \
  Variant {                                                                                                  \
    ID, 1, R, SHAPE_BOX, 1, 1, 256, 4, 0, 0, 0, &launch_abl1_global<R>, &prepare_none1, "k1d_abl_global<r" #R ">", \
        0                                                                                                    \
  }
#define CS_VABL1L(ID, R, L)                                                                                  \
  Variant {                                                                                                  \
    ID, 2, R, SHAPE_BOX, 1, 1, L, 4, (size_t)(L + 2 * R) * 8, L + 2 * R, 0, &launch_abl1_lds<R, L>,          \
        &prepare_none1, "k1d_abl_lds<r" #R "," #L ">", 0                                                     \
  }

static const Variant kAblation1D[] = {
    CS_VABL1G(911000, 1), CS_VABL1G(913000, 3), CS_VABL1L(921000, 1, 1024), CS_VABL1L(923000, 3, 1024),
};

const Variant *variants_1d_ablation(int *count) {
  *count = (int)(sizeof(kAblation1D) / sizeof(kAblation1D[0]));
  return kAblation1D;
}

}  // namespace cs
