// k2d_ablation.hip -- the simple kernels of the --breakdown ladder (2-D, FP64).
//
// The reference's ladder (src/2d/gpu.cu:72-236, launchers :434-801, run by
// src/2d/main.cu:349-357) times the same pre-fused stencil with one optimisation added per
// rung.  On MI355X the rungs are:
//   1  every output summed with FP64 FMAs from neighbours read straight from HBM/L2
//      (k2d_abl_global, here);
//   2  the tile and its halo staged in LDS once, FP64 FMAs from LDS (k2d_abl_lds, here);
//   3  the dilated v_mfma_f64_4x4x4 tables on the tile in LDS with rows aligned to 128
//      bytes, the layout a straightforward port uses (k2d_f64 with that pitch, k2d_f64.hip);
//   4  the production single-step kernel: the pitch the LDS bank model picks, and the
//      output staged in LDS and written row by row (k2d_f64 with ST 2);
// and the production launcher last.  These kernels compute the same stencil as every
// other variant (halo never written); they run only when a plan is forced to them by id
// (Variant::mf 1 and 2 mark them), never by the plan's choice or the autotuner.
#include "cs_internal.h"

namespace cs {

// Rung 1: one thread per output cell, 4 rows x 64 columns per workgroup.
template <int R>
__global__ __launch_bounds__(256) void k2d_abl_global(const double *__restrict__ in, double *__restrict__ out,
                                                      const double *__restrict__ w, Geo g) {
  constexpr int D = 2 * R + 1;
  const int64_t x = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t y = (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= g.n[2] || y >= g.n[1]) return;
  const int64_t pitch = g.ext[2], c = (g.org[1] + y) * pitch + g.org[2] + x;
  double acc = 0.0;
#pragma unroll
  for (int dy = -R; dy <= R; ++dy)
#pragma unroll
    for (int dx = -R; dx <= R; ++dx) acc = fma(w[(dy + R) * D + dx + R], in[c + dy * pitch + dx], acc);
  out[c] = acc;
}

// Rung 2: a BY x BX tile plus its halo loaded into LDS once (zero outside the array), then
// every output summed with FP64 FMAs from LDS.
template <int R, int BY, int BX>
__global__ __launch_bounds__(256) void k2d_abl_lds(const double *__restrict__ in, double *__restrict__ out,
                                                   const double *__restrict__ w, Geo g) {
  constexpr int D = 2 * R + 1, EY = BY + 2 * R, EX = BX + 2 * R;
  __shared__ double tile[EY * EX];
  const int64_t pitch = g.ext[2];
  const int64_t Y0 = (int64_t)blockIdx.y * BY - R, X0 = (int64_t)blockIdx.x * BX - R;  // interior coordinates
  for (int i = threadIdx.x; i < EY * EX; i += 256) {
    const int r = i / EX, cc = i - r * EX;
    const int64_t ay = g.org[1] + Y0 + r, ax = g.org[2] + X0 + cc;
    tile[i] = (ay >= 0 && ay < g.ext[1] && ax >= 0 && ax < pitch) ? in[ay * pitch + ax] : 0.0;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < BY * BX; i += 256) {
    const int ty = i / BX, tx = i - ty * BX;
    const int64_t y = Y0 + R + ty, x = X0 + R + tx;
    if (y >= g.n[1] || x >= g.n[2]) continue;
    double acc = 0.0;
#pragma unroll
    for (int dy = 0; dy < D; ++dy)
#pragma unroll
      for (int dx = 0; dx < D; ++dx) acc = fma(w[dy * D + dx], tile[(ty + dy) * EX + tx + dx], acc);
    out[(g.org[1] + y) * pitch + g.org[2] + x] = acc;
  }
}

template <int R>
static hipError_t launch_abl_global(const double *in, double *out, const double *w, const Geo &g, hipStream_t s) {
  if ((g.n[1] + 3) / 4 > 65535) return hipErrorInvalidConfiguration;  // grid y limit
  const dim3 grid((unsigned)((g.n[2] + 63) / 64), (unsigned)((g.n[1] + 3) / 4));
  hipLaunchKernelGGL((k2d_abl_global<R>), grid, dim3(256), 0, s, in, out, w, g);
  return hipGetLastError();
}

template <int R, int BY, int BX>
static hipError_t launch_abl_lds(const double *in, double *out, const double *w, const Geo &g, hipStream_t s) {
  if ((g.n[1] + BY - 1) / BY > 65535) return hipErrorInvalidConfiguration;  // grid y limit
  const dim3 grid((unsigned)((g.n[2] + BX - 1) / BX), (unsigned)((g.n[1] + BY - 1) / BY));
  hipLaunchKernelGGL((k2d_abl_lds<R, BY, BX>), grid, dim3(256), 0, s, in, out, w, g);
  return hipGetLastError();
}

static hipError_t prepare_none() { return hipSuccess; }

// ids 910000 + rung * 10000 + R * 1000 (+ 10 for the star shape class)
#define CS_VABLG(ID, R, SH)                                                                              \
  Variant {                                                                                              \
    ID, 1, R, SH, 1, 4, 64, 4, 0, 0, 0, &launch_abl_global<R>, &prepare_none,                            \
        "k2d_abl_global<r" #R "," #SH ">", 0                                                             \
  }
#define CS_VABLL(ID, R, SH, BY, BX)                                                                      \
  Variant {                                                                                              \
    ID, 2, R, SH, 1, BY, BX, 4, (size_t)(BY + 2 * R) * (BX + 2 * R) * 8, BX + 2 * R, 0,                  \
        &launch_abl_lds<R, BY, BX>, &prepare_none, "k2d_abl_lds<r" #R "," #SH "," #BY "x" #BX ">", 0      \
  }

static const Variant kAblation[] = {
    CS_VABLG(911000, 1, SHAPE_BOX), CS_VABLG(913000, 3, SHAPE_BOX),
    CS_VABLG(911010, 1, SHAPE_STAR), CS_VABLG(913010, 3, SHAPE_STAR),
    CS_VABLL(921000, 1, SHAPE_BOX, 32, 64), CS_VABLL(923000, 3, SHAPE_BOX, 32, 64),
    CS_VABLL(921010, 1, SHAPE_STAR, 32, 64), CS_VABLL(923010, 3, SHAPE_STAR, 32, 64),
};

const Variant *variants_2d_ablation(int *count) {
  *count = (int)(sizeof(kAblation) / sizeof(kAblation[0]));
  return kAblation;
}

}  // namespace cs
