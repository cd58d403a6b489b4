// k3d_fma.hip -- dense FP64 3-D stencils of radius 2-3 with register-blocked FMAs.
//
// Replaces gpu_box_3d1r_kernel and gpu_star_3d1r_step3_kernel (reference
// src/3d/gpu_box.cu:186-238, src/3d/gpu_star.cu:178-236) for the reference's pre-fused
// weights: a 343-point box and a 63-point octahedron of radius 3.  At 686 flops per point
// these are compute-bound, and the FP64 rate of v_fma_f64 equals that of the FP64 MFMA on
// MI355X, so the kernel is built for operand reuse instead of matrix shape:
//   * a workgroup owns a BY x BX column tile and ZC output planes and reads each input
//     plane of its column once, into one of two LDS plane buffers (the next plane is in
//     registers while the current one is used);
//   * each thread owns a TY x TX block of outputs and keeps partial sums for the 2R+1
//     output planes an input plane feeds (output plane p - R + j takes weight plane
//     2R - j of input plane p); when plane p is done, output plane p - R is complete,
//     is written, and the partial sums shift by one plane;
//   * a row of TX + 2R input values is read from LDS once and feeds up to
//     TY (2R+1)^2 TX FMAs (about 34 per value for the box); the weights have wave-uniform
//     addresses and come through scalar loads (SGPRs), which measured 10-20% faster than
//     staging them in LDS and reading them as broadcasts;
//   * the stencil's shape class (box, axes, octahedron) drops zero taps at compile time.
// Plans use it where the 8 x 8 block family (k3d2_f64.hip) has no kernel for the radius
// (at radius 1 that family is faster: 27 taps are memory-bound, not compute-bound).
#include "cs_internal.h"

namespace cs {

template <int R, int SH>
__host__ __device__ constexpr bool tap_live3(int dz, int dy, int dx) {
  const int az = dz < 0 ? -dz : dz, ay = dy < 0 ? -dy : dy, ax = dx < 0 ? -dx : dx;
  return SH == SHAPE3_BOX ? true : SH == SHAPE3_STAR ? (az != 0) + (ay != 0) + (ax != 0) <= 1 : az + ay + ax <= R;
}

template <int R, int SH, int BY, int BX, int ZC, int TY, int TX>
struct CfgF3 {
  static constexpr int D = 2 * R + 1;
  static constexpr int EY = BY + 2 * R, EX = BX + 2 * R;
  static constexpr int PX = (EX + 1) / 2 * 2;  // even: row starts stay 16-byte aligned
  static constexpr int PLANE = EY * PX;
  static constexpr int NT = (BY / TY) * (BX / TX);  // threads
  static constexpr int PF = (EY * EX + NT - 1) / NT;  // plane loads per thread
  static constexpr size_t LDS = (size_t)2 * PLANE * sizeof(double);
  static_assert(NT == 256, "256 threads per workgroup");
  static_assert(BX / TX == 16 && BY % TY == 0, "16 threads across a row of the tile");
  static_assert(LDS <= 160 * 1024, "planes do not fit gfx950 LDS");
};

template <int R, int SH, int BY, int BX, int ZC, int TY, int TX>
__global__ __launch_bounds__(256) void k3d_fma(const double *__restrict__ in, double *__restrict__ out,
                                               const double *__restrict__ w, Geo g) {
  using C = CfgF3<R, SH, BY, BX, ZC, TY, TX>;
  constexpr int D = C::D, EX = C::EX, PX = C::PX, PLANE = C::PLANE, PF = C::PF;
  extern __shared__ __align__(16) double lds[];
  const int tid = threadIdx.x;
  const int txi = tid & 15, tyi = tid >> 4;
  const int64_t ext1 = g.ext[1], ext2 = g.ext[2], pz = ext1 * ext2;
  const int64_t n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
  const int64_t Y0 = (int64_t)blockIdx.y * BY - R, X0 = (int64_t)blockIdx.x * BX - R;  // LDS cell (0, 0)
  const int64_t Z0 = (int64_t)blockIdx.z * ZC;
  const int64_t Z1 = Z0 + ZC < n0 ? Z0 + ZC : n0;
  const int64_t ay0 = g.org[1] + Y0, ax0 = g.org[2] + X0;  // >= 0: the halo is at least R cells
  const bool all_in = ay0 + C::EY <= ext1 && ax0 + EX <= ext2;

  // plane p (interior z, -R <= p < n0 + R) of the tile: zero past the array
  double pf[PF];
  auto issue = [&](int64_t p) {
    const double *base = in + (g.org[0] + p) * pz + ay0 * ext2 + ax0;
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int idx = tid + i * 256;
      const int r = idx / EX, c = idx - r * EX;
      const bool ok = idx < C::EY * EX && (all_in || (ay0 + r < ext1 && ax0 + c < ext2));
      const double v = *(ok ? base + (int64_t)r * ext2 + c : in);
      pf[i] = ok ? v : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int idx = tid + i * 256;
      const int r = idx / EX, c = idx - r * EX;
      if (idx < C::EY * EX) lds[buf * PLANE + r * PX + c] = pf[i];
    }
  };

  double acc[D][TY * TX];
#pragma unroll
  for (int j = 0; j < D; ++j)
#pragma unroll
    for (int e = 0; e < TY * TX; ++e) acc[j][e] = 0.0;

  const int64_t P0 = Z0 - R, P1 = Z1 + R;  // input planes [P0, P1)
  issue(P0);
  store(0);
  __syncthreads();
  int buf = 0;
  for (int64_t p = P0; p < P1; ++p) {
    const bool more = p + 1 < P1;
    if (more) issue(p + 1);  // in flight while plane p is used
    // output plane p - R + j exists for j in [jlo, jhi)
    const int jlo = (int)(Z0 - (p - R) > 0 ? Z0 - (p - R) : 0);
    const int jhi = (int)(Z1 - (p - R) < D ? Z1 - (p - R) : D);
    const double *B = lds + buf * PLANE + (TY * tyi) * PX + TX * txi;
#pragma unroll
    for (int iy = 0; iy < TY + 2 * R; ++iy) {
      double v[TX + 2 * R];
#pragma unroll
      for (int c = 0; c < TX + 2 * R; ++c) v[c] = B[iy * PX + c];
#pragma unroll
      for (int oy = 0; oy < TY; ++oy) {
        const int dy = iy - oy;
        if (dy < 0 || dy > 2 * R) continue;
#pragma unroll
        for (int j = 0; j < D; ++j) {
          if (j < jlo || j >= jhi) continue;  // uniform
          const int kz = 2 * R - j;
#pragma unroll
          for (int dx = 0; dx < D; ++dx) {
            if (!tap_live3<R, SH>(kz - R, dy - R, dx - R)) continue;
            const double wv = w[(kz * D + dy) * D + dx];  // uniform: a scalar load
#pragma unroll
            for (int ox = 0; ox < TX; ++ox) acc[j][oy * TX + ox] = fma(wv, v[ox + dx], acc[j][oy * TX + ox]);
          }
        }
      }
    }
    const int64_t q = p - R;  // complete now
    if (q >= Z0) {
#pragma unroll
      for (int oy = 0; oy < TY; ++oy) {
        const int64_t y = Y0 + R + TY * tyi + oy;
        double *row = out + (g.org[0] + q) * pz + (g.org[1] + y) * ext2 + g.org[2];
#pragma unroll
        for (int ox = 0; ox < TX; ++ox) {
          const int64_t x = X0 + R + TX * txi + ox;
          if (y < n1 && x < n2) row[x] = acc[0][oy * TX + ox];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < D - 1; ++j)
#pragma unroll
      for (int e = 0; e < TY * TX; ++e) acc[j][e] = acc[j + 1][e];
#pragma unroll
    for (int e = 0; e < TY * TX; ++e) acc[D - 1][e] = 0.0;
    if (more) store(buf ^ 1);  // its last readers finished before the previous barrier
    __syncthreads();
    buf ^= 1;
  }
}

template <int R, int SH, int BY, int BX, int ZC, int TY, int TX>
static hipError_t prepare_k3fma() {
  using C = CfgF3<R, SH, BY, BX, ZC, TY, TX>;
  return hipFuncSetAttribute((const void *)k3d_fma<R, SH, BY, BX, ZC, TY, TX>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS);
}

template <int R, int SH, int BY, int BX, int ZC, int TY, int TX>
static hipError_t launch_k3fma(const double *in, double *out, const double *w, const Geo &g, hipStream_t s) {
  using C = CfgF3<R, SH, BY, BX, ZC, TY, TX>;
  const int64_t gy = (g.n[1] + BY - 1) / BY, gz = (g.n[0] + ZC - 1) / ZC;
  if (gy > 65535 || gz > 65535 || gz < 1) return hipErrorInvalidConfiguration;  // grid y / z limits
  const dim3 grid((unsigned)((g.n[2] + BX - 1) / BX), (unsigned)gy, (unsigned)gz);
  hipLaunchKernelGGL((k3d_fma<R, SH, BY, BX, ZC, TY, TX>), grid, dim3(256), C::LDS, s, in, out, w, g);
  return hipGetLastError();
}

// Variant::mf = 3 marks this family (dense weights, like the ladder's FMA kernels);
// ids 35000 + R * 100 + shape * 10 + tile alternative.
#define CS_V3FMA(ID, R, SH, BY, BX, ZC, TY, TX)                                                               \
  Variant {                                                                                                   \
    ID, 3, R, SH, 1, BY, BX, 4, CfgF3<R, SH, BY, BX, ZC, TY, TX>::LDS, CfgF3<R, SH, BY, BX, ZC, TY, TX>::PX, ZC, \
        &launch_k3fma<R, SH, BY, BX, ZC, TY, TX>, &prepare_k3fma<R, SH, BY, BX, ZC, TY, TX>,                   \
        "k3d_fma<r" #R "," #SH "," #BY "x" #BX ",zc" #ZC ",t" #TY "x" #TX ">", 0                              \
  }

static const Variant kVariants3F[] = {
    // 1 x 4 outputs per thread (4 waves per SIMD) first: 59 against 50 GStencil/s for
    // 2 x 4 on the 343-point box at 256^3
    CS_V3FMA(35302, 3, SHAPE3_BOX, 16, 64, 16, 1, 4),     CS_V3FMA(35322, 3, SHAPE3_DIAMOND, 16, 64, 16, 1, 4),
    CS_V3FMA(35312, 3, SHAPE3_STAR, 16, 64, 16, 1, 4),    CS_V3FMA(35303, 3, SHAPE3_BOX, 16, 64, 32, 1, 4),
    CS_V3FMA(35323, 3, SHAPE3_DIAMOND, 16, 64, 32, 1, 4), CS_V3FMA(35305, 3, SHAPE3_BOX, 16, 32, 16, 1, 2),
    CS_V3FMA(35325, 3, SHAPE3_DIAMOND, 16, 32, 16, 1, 2), CS_V3FMA(35300, 3, SHAPE3_BOX, 32, 64, 16, 2, 4),
    CS_V3FMA(35310, 3, SHAPE3_STAR, 32, 64, 16, 2, 4),    CS_V3FMA(35320, 3, SHAPE3_DIAMOND, 32, 64, 16, 2, 4),
    CS_V3FMA(35301, 3, SHAPE3_BOX, 32, 64, 32, 2, 4),     CS_V3FMA(35321, 3, SHAPE3_DIAMOND, 32, 64, 32, 2, 4),
    CS_V3FMA(35202, 2, SHAPE3_BOX, 16, 64, 16, 1, 4),     CS_V3FMA(35212, 2, SHAPE3_STAR, 16, 64, 16, 1, 4),
    CS_V3FMA(35222, 2, SHAPE3_DIAMOND, 16, 64, 16, 1, 4), CS_V3FMA(35200, 2, SHAPE3_BOX, 32, 64, 16, 2, 4),
    CS_V3FMA(35210, 2, SHAPE3_STAR, 32, 64, 16, 2, 4),    CS_V3FMA(35220, 2, SHAPE3_DIAMOND, 32, 64, 16, 2, 4),
    CS_V3FMA(35306, 3, SHAPE3_BOX, 16, 128, 16, 1, 8),
};

const Variant *variants_3d_fma(int *count) {
  *count = (int)(sizeof(kVariants3F) / sizeof(kVariants3F[0]));
  return kVariants3F;
}

}  // namespace cs
