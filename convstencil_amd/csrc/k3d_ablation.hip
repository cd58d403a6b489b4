// k3d_ablation.hip -- the FMA kernels of the 3-D --breakdown ladder (FP64, box).
//
// The reference's 3-D ladder (src/3d/gpu_box.cu:240-442, launchers :538-904, run by
// src/3d/main.cu:623-628 for box3d1r) times the same pre-fused 343-point box stencil
// with one optimisation added per rung.  On MI355X the rungs are:
//   1  every output summed with FP64 FMAs from neighbours read straight from HBM/L2
//      (k3d_abl_global, here);
//   2  2.5-D blocking: a BY x BX column tile streams its planes through a ring of 2R+1
//      LDS planes, each input plane loaded once, FP64 FMAs from LDS (k3d_abl_lds, here);
//   3  the v_mfma_f64_4x4x4 strip kernel on the same ring with 128-byte aligned LDS rows,
//      the layout a straightforward port uses (k3d_f64 with that pitch, k3d_f64.hip);
//   4  the production kernel: register-blocked FP64 FMAs, each thread keeping partial sums
//      for the 2R+1 output planes an input plane feeds (k3d_fma.hip; at 686 flops per
//      point the MFMA strips lose to it, as FP64 FMA and FP64 MFMA peak alike on MI355X);
// and the production launcher last.  They run only when a plan is forced to them by id
// (Variant::mf 1 and 2 mark them).  Neighbours are read without bounds checks in z and
// on the low side in y and x: plans require a halo of at least R cells.
#include "cs_internal.h"

namespace cs {

// Rung 1: one thread per output cell, 4 rows x 64 columns of one plane per workgroup.
template <int R>
__global__ __launch_bounds__(256) void k3d_abl_global(const double *__restrict__ in, double *__restrict__ out,
                                                      const double *__restrict__ w, Geo g) {
  constexpr int D = 2 * R + 1;
  const int64_t x = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int64_t y = (int64_t)blockIdx.y * 4 + (threadIdx.x >> 6);
  const int64_t z = blockIdx.z;
  if (x >= g.n[2] || y >= g.n[1]) return;
  const int64_t p1 = g.ext[2], p0 = g.ext[1] * g.ext[2];
  const int64_t c = (g.org[0] + z) * p0 + (g.org[1] + y) * p1 + g.org[2] + x;
  double acc = 0.0;
  for (int dz = -R; dz <= R; ++dz)
#pragma unroll
    for (int dy = -R; dy <= R; ++dy)
#pragma unroll
      for (int dx = -R; dx <= R; ++dx)
        acc = fma(w[((dz + R) * D + dy + R) * D + dx + R], in[c + dz * p0 + dy * p1 + dx], acc);
  out[c] = acc;
}

// Rung 2: the workgroup owns a BY x BX column tile and ZC output planes; plane p of the
// tile (with its halo, zero past the array) sits in ring slot (p - Z0 + R) mod (2R+1).
template <int R, int BY, int BX, int ZC>
__global__ __launch_bounds__(256) void k3d_abl_lds(const double *__restrict__ in, double *__restrict__ out,
                                                   const double *__restrict__ w, Geo g) {
  constexpr int D = 2 * R + 1, EY = BY + 2 * R, EX = BX + 2 * R, PL = EY * EX;
  extern __shared__ double ring[];
  const int tid = threadIdx.x;
  const int64_t p1 = g.ext[2], p0 = g.ext[1] * g.ext[2];
  const int64_t Y0 = (int64_t)blockIdx.y * BY - R, X0 = (int64_t)blockIdx.x * BX - R;  // interior coordinates
  const int64_t Z0 = (int64_t)blockIdx.z * ZC;
  const int64_t Z1 = Z0 + ZC < g.n[0] ? Z0 + ZC : g.n[0];
  auto slot = [&](int64_t p) { return (int)((p - Z0 + R) % D); };
  auto load = [&](int64_t p) {  // interior plane p, -R <= p < n0 + R
    double *dst = ring + slot(p) * PL;
    const double *src = in + (g.org[0] + p) * p0;
    for (int i = tid; i < PL; i += 256) {
      const int r = i / EX, cc = i - r * EX;
      const int64_t ay = g.org[1] + Y0 + r, ax = g.org[2] + X0 + cc;  // both >= 0
      dst[i] = (ay < g.ext[1] && ax < p1) ? src[ay * p1 + ax] : 0.0;
    }
  };
  for (int64_t p = Z0 - R; p < Z0 + R; ++p) load(p);
  for (int64_t z = Z0; z < Z1; ++z) {
    load(z + R);
    __syncthreads();  // planes z - R .. z + R are in the ring
    for (int i = tid; i < BY * BX; i += 256) {
      const int ty = i / BX, tx = i - ty * BX;
      const int64_t y = Y0 + R + ty, x = X0 + R + tx;
      if (y >= g.n[1] || x >= g.n[2]) continue;
      double acc = 0.0;
      for (int dz = 0; dz < D; ++dz) {
        const double *pl = ring + slot(z - R + dz) * PL + ty * EX + tx;
#pragma unroll
        for (int dy = 0; dy < D; ++dy)
#pragma unroll
          for (int dx = 0; dx < D; ++dx) acc = fma(w[(dz * D + dy) * D + dx], pl[dy * EX + dx], acc);
      }
      out[(g.org[0] + z) * p0 + (g.org[1] + y) * p1 + g.org[2] + x] = acc;
    }
    __syncthreads();  // every thread is done with plane z - R before its slot is refilled
  }
}

template <int R>
static hipError_t launch_abl3_global(const double *in, double *out, const double *w, const Geo &g, hipStream_t s) {
  const int64_t gy = (g.n[1] + 3) / 4;
  if (gy > 65535 || g.n[0] > 65535 || g.n[0] < 1) return hipErrorInvalidConfiguration;  // grid y / z limits
  const dim3 grid((unsigned)((g.n[2] + 63) / 64), (unsigned)gy, (unsigned)g.n[0]);
  hipLaunchKernelGGL((k3d_abl_global<R>), grid, dim3(256), 0, s, in, out, w, g);
  return hipGetLastError();
}

template <int R, int BY, int BX, int ZC>
static constexpr size_t abl3_lds() {
  return (size_t)(2 * R + 1) * (BY + 2 * R) * (BX + 2 * R) * sizeof(double);
}

template <int R, int BY, int BX, int ZC>
static hipError_t prepare_abl3_lds() {
  static_assert(abl3_lds<R, BY, BX, ZC>() <= 160 * 1024, "ring does not fit gfx950 LDS");
  return hipFuncSetAttribute((const void *)k3d_abl_lds<R, BY, BX, ZC>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)abl3_lds<R, BY, BX, ZC>());
}

template <int R, int BY, int BX, int ZC>
static hipError_t launch_abl3_lds(const double *in, double *out, const double *w, const Geo &g, hipStream_t s) {
  const int64_t gy = (g.n[1] + BY - 1) / BY, gz = (g.n[0] + ZC - 1) / ZC;
  if (gy > 65535 || gz > 65535 || gz < 1) return hipErrorInvalidConfiguration;  // grid y / z limits
  const dim3 grid((unsigned)((g.n[2] + BX - 1) / BX), (unsigned)gy, (unsigned)gz);
  constexpr size_t lds = abl3_lds<R, BY, BX, ZC>();
  hipLaunchKernelGGL((k3d_abl_lds<R, BY, BX, ZC>), grid, dim3(256), lds, s, in, out, w, g);
  return hipGetLastError();
}

static hipError_t prepare_none3() { return hipSuccess; }

// ids 910000 + rung * 10000 + R * 1000 (+ 10 for the star shape class)
#define CS_VABL3G(ID, R, SH)                                                                                 \
  Variant {                                                                                                  \
    ID, 1, R, SH, 1, 4, 64, 4, 0, 0, 0, &launch_abl3_global<R>, &prepare_none3, "k3d_abl_global<r" #R "," #SH ">", \
        0                                                                                                    \
  }
#define CS_VABL3L(ID, R, SH, BY, BX, ZC)                                                                     \
  Variant {                                                                                                  \
    ID, 2, R, SH, 1, BY, BX, 4, abl3_lds<R, BY, BX, ZC>(), BX + 2 * R, ZC, &launch_abl3_lds<R, BY, BX, ZC>,  \
        &prepare_abl3_lds<R, BY, BX, ZC>, "k3d_abl_lds<r" #R "," #SH "," #BY "x" #BX ",zc" #ZC ">", 0        \
  }

static const Variant kAblation3D[] = {
    CS_VABL3G(911000, 1, SHAPE3_BOX),          CS_VABL3G(913000, 3, SHAPE3_BOX),
    CS_VABL3G(911010, 1, SHAPE3_STAR),         CS_VABL3G(913010, 3, SHAPE3_STAR),
    CS_VABL3L(921000, 1, SHAPE3_BOX, 8, 64, 16), CS_VABL3L(923000, 3, SHAPE3_BOX, 8, 64, 16),
    CS_VABL3L(921010, 1, SHAPE3_STAR, 8, 64, 16), CS_VABL3L(923010, 3, SHAPE3_STAR, 8, 64, 16),
};

const Variant *variants_3d_ablation(int *count) {
  *count = (int)(sizeof(kAblation3D) / sizeof(kAblation3D[0]));
  return kAblation3D;
}

}  // namespace cs
