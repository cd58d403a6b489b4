// k3d_f64.hip -- plain FP64 3-D ConvStencil kernels for gfx950 (v_mfma_f64_4x4x4).
//
// Replaces gpu_box_3d1r_kernel (reference src/3d/gpu_box.cu:186-238) and
// gpu_star_3d1r_step3_kernel (src/3d/gpu_star.cu:178-236).  The reference streams
// z-planes through a 19-slot ring of per-plane partial sums that assumes
// z-symmetric weights (4 distinct planes for 7, src/3d/gpu_box.cu:221-237).  Here:
//   * a workgroup owns a BY x BX column tile and a chunk of ZC output planes and
//     streams its input planes through a ring of 2R+1 LDS planes (2.5-D blocking):
//     each input plane is loaded from HBM once per column tile, the next plane is
//     prefetched into registers while the current output plane is computed;
//   * an output plane is a sum over dz of 2-D stencil products, done as one MFMA
//     chain per 4 x 16 output strip whose k-steps are grouped by dz (cs_ktable.h),
//     so the LDS plane a k-step reads is uniform over the wave -- any weights, no
//     z-symmetry assumption;
//   * B (banded Toeplitz of the weights) is held in registers when the stencil
//     needs at most BREG_MAX k-steps; the 343-point radius-3 box reloads each dz
//     group's fragments once per plane and shares them across the wave's blocks.
#include <vector>

#include "cs_internal.h"

namespace cs {

// AL = 1: rows aligned to 128 bytes (the --breakdown ladder's rung 3) instead of the odd
// pitch, which puts the 16 rows a strip's lanes read in distinct LDS banks.
template <int R, int SH, int BY, int BX, int ZC, int NW, int AL = 0>
struct Cfg3D {
  static constexpr int EY = BY + 2 * R;
  static constexpr int EX = BX + 2 * R;
  static constexpr int PX = AL ? (EX + 15) / 16 * 16 : (EX % 2 == 0) ? EX + 1 : EX;
  static constexpr int PLANE = EY * PX;
  static constexpr int NP = 2 * R + 1;
  static constexpr size_t LDS = (size_t)NP * PLANE * sizeof(double);
  static constexpr KTable3D KT = make_ktable3d(R, SH);
  static constexpr int NKS = KT.n;
  static constexpr bool BREG = NKS <= 48;
  static constexpr int NBLK = (BY / 4) * (BX / 16);
  static constexpr int PER = (NBLK + NW - 1) / NW;
  static constexpr int PF = (PLANE + NW * 64 - 1) / (NW * 64);
  static constexpr int first_ks(int dz) {
    int i = 0;
    while (i < KT.n && KT.dz[i] < dz) ++i;
    return i;
  }
  static constexpr int count_ks(int dz) {
    int n = 0;
    for (int i = 0; i < KT.n; ++i) n += KT.dz[i] == dz;
    return n;
  }
  static constexpr int max_ks_dz() {
    int m = 0;
    for (int dz = -R; dz <= R; ++dz) m = count_ks(dz) > m ? count_ks(dz) : m;
    return m;
  }
  static_assert(BY % 4 == 0 && BX % 16 == 0, "tile must be a multiple of the 4 x 16 strip");
  static_assert(LDS <= 160 * 1024, "tile does not fit gfx950 LDS");
};

template <int R, int SH, int BY, int BX, int ZC, int NW, int AL>
__global__ __launch_bounds__(NW * 64) void k3d_f64(const double *__restrict__ in, double *__restrict__ out,
                                                   const double *__restrict__ bfrag, Geo g) {
  using C = Cfg3D<R, SH, BY, BX, ZC, NW, AL>;
  constexpr KTable3D kt = C::KT;
  constexpr int PX = C::PX;
  constexpr int NBR = C::BREG ? C::NKS : C::max_ks_dz();
  extern __shared__ double lds[];

  const int64_t tiles_yx = g.tiles[1] * g.tiles[2];
  const int64_t ntiles = g.tiles[0] * tiles_yx;
  const int64_t per = (ntiles + 7) / 8;
  const int64_t tile = (blockIdx.x % 8) * per + blockIdx.x / 8;  // XCD-contiguous tile bands
  if (tile >= ntiles) return;
  const int64_t zc = tile / tiles_yx, rem = tile % tiles_yx;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t ext1 = g.ext[1], ext2 = g.ext[2];
  const int64_t nz = g.n[0], ny = g.n[1], nx = g.n[2];
  const int64_t Y0 = (rem / g.tiles[2]) * BY - R, X0 = (rem % g.tiles[2]) * BX - R;
  const int64_t ay0 = g.org[1] + Y0, ax0 = g.org[2] + X0;
  const int64_t Z0 = zc * ZC;
  const int64_t Z1 = Z0 + ZC < nz ? Z0 + ZC : nz;
  const bool full = Y0 + R + BY <= ny && X0 + R + BX <= nx;
  const bool rows_in = ay0 + C::EY <= ext1 && ax0 + C::EX <= ext2;

  double b[NBR];
  if constexpr (C::BREG) {
#pragma unroll
    for (int ks = 0; ks < C::NKS; ++ks) b[ks] = bfrag[ks * 64 + lane];
  }
  const int k = lane >> 4, b4 = (lane >> 2) & 3, i4 = lane & 3;

  auto plane_cell = [&](int64_t p, int idx) -> double {
    const int r = idx / PX, c = idx - r * PX;
    const int64_t ay = ay0 + r, ax = ax0 + c;
    if (c >= C::EX || (!rows_in && (ay >= ext1 || ax >= ext2))) return 0.0;
    return in[((g.org[0] + p) * ext1 + ay) * ext2 + ax];
  };
  auto slot = [&](int64_t p) -> int { return (int)((p - (Z0 - R)) % C::NP); };

  // prologue: planes Z0 - R .. Z0 + R
  for (int64_t p = Z0 - R; p <= Z0 + R; ++p) {
    double *dst = lds + slot(p) * C::PLANE;
    for (int idx = tid; idx < C::PLANE; idx += NW * 64) dst[idx] = plane_cell(p, idx);
  }
  __syncthreads();

  for (int64_t z = Z0; z < Z1; ++z) {
    const bool more = z + 1 < Z1;
    double pf[C::PF];
    if (more) {
#pragma unroll
      for (int q = 0; q < C::PF; ++q) {
        const int idx = tid + q * NW * 64;
        pf[q] = idx < C::PLANE ? plane_cell(z + R + 1, idx) : 0.0;
      }
    }
    double acc[C::PER];
#pragma unroll
    for (int j = 0; j < C::PER; ++j) acc[j] = 0.0;
#pragma unroll
    for (int dz = -R; dz <= R; ++dz) {
      constexpr int dummy = 0;
      (void)dummy;
      const int k0 = C::first_ks(dz), kn = C::count_ks(dz);
      const double *pl = lds + slot(z + dz) * C::PLANE;
      if constexpr (!C::BREG) {
#pragma unroll
        for (int q = 0; q < NBR; ++q)
          if (q < kn) b[q] = bfrag[(k0 + q) * 64 + lane];
      }
#pragma unroll
      for (int j = 0; j < C::PER; ++j) {
        const int blk = j * NW + wave;
        if (blk < C::NBLK) {
          const int y = R + 4 * (blk / (BX / 16)), x = R + 16 * (blk % (BX / 16));
          const double *p = pl + (y + i4) * PX + x + 4 * b4;
          double a = acc[j];
#pragma unroll
          for (int q = 0; q < NBR; ++q) {
            if (q < kn) {
              const int ks = k0 + q;
              const int o = kt.dy[ks][k] * PX + kt.c[ks][k];
              a = __builtin_amdgcn_mfma_f64_4x4x4f64(p[o], C::BREG ? b[ks < NBR ? ks : 0] : b[q], a, 0, 0, 0);
            }
          }
          acc[j] = a;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < C::PER; ++j) {
      const int blk = j * NW + wave;
      if (blk < C::NBLK) {
        const int ty = R + 4 * (blk / (BX / 16)) + (lane >> 4), tx = R + 16 * (blk % (BX / 16)) + (lane & 15);
        if (full || (Y0 + ty < ny && X0 + tx < nx))
          out[((g.org[0] + z) * ext1 + ay0 + ty) * ext2 + ax0 + tx] = acc[j];
      }
    }
    __syncthreads();  // every wave is done with plane z - R
    if (more) {
      double *dst = lds + slot(z + R + 1) * C::PLANE;
#pragma unroll
      for (int q = 0; q < C::PF; ++q) {
        const int idx = tid + q * NW * 64;
        if (idx < C::PLANE) dst[idx] = pf[q];
      }
    }
    __syncthreads();
  }
}

template <int R, int SH, int BY, int BX, int ZC, int NW, int AL>
static hipError_t prepare_k3d() {
  using C = Cfg3D<R, SH, BY, BX, ZC, NW, AL>;
  return hipFuncSetAttribute((const void *)k3d_f64<R, SH, BY, BX, ZC, NW, AL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)C::LDS);
}

template <int R, int SH, int BY, int BX, int ZC, int NW, int AL>
static hipError_t launch_k3d(const double *in, double *out, const double *bfrag, const Geo &g0, hipStream_t s) {
  using C = Cfg3D<R, SH, BY, BX, ZC, NW, AL>;
  Geo g = g0;
  g.tiles[0] = (g.n[0] + ZC - 1) / ZC;
  const int64_t ntiles = g.tiles[0] * g.tiles[1] * g.tiles[2];
  const int64_t grid = ((ntiles + 7) / 8) * 8;
  hipLaunchKernelGGL((k3d_f64<R, SH, BY, BX, ZC, NW, AL>), dim3((unsigned)grid), dim3(NW * 64), C::LDS, s, in, out,
                     bfrag, g);
  return hipGetLastError();
}

#define CS_V3DA(ID, R, SH, BY, BX, ZC, NW, AL, SUF)                                                            \
  Variant {                                                                                                       \
    ID, 4, R, SH, 1, BY, BX, NW, Cfg3D<R, SH, BY, BX, ZC, NW, AL>::LDS, Cfg3D<R, SH, BY, BX, ZC, NW, AL>::PX, ZC, \
        &launch_k3d<R, SH, BY, BX, ZC, NW, AL>, &prepare_k3d<R, SH, BY, BX, ZC, NW, AL>,                          \
        "k3d_f64<r" #R "," #SH "," #BY "x" #BX ",zc" #ZC ",w" #NW SUF ">"                                        \
  }
#define CS_V3D(ID, R, SH, BY, BX, ZC, NW) CS_V3DA(ID, R, SH, BY, BX, ZC, NW, 0, "")

// For 3-D variants the Variant::st field carries the z-chunk length.
static const Variant kVariants3D[] = {
    CS_V3D(31001, 1, SHAPE3_BOX, 32, 64, 32, 8),
    CS_V3D(31011, 1, SHAPE3_STAR, 32, 64, 32, 8),
    CS_V3D(31002, 1, SHAPE3_BOX, 16, 64, 32, 4),
    CS_V3D(31012, 1, SHAPE3_STAR, 16, 64, 32, 4),
    CS_V3D(33001, 3, SHAPE3_BOX, 16, 64, 32, 8),
    CS_V3D(33011, 3, SHAPE3_STAR, 16, 64, 32, 8),
    CS_V3D(33021, 3, SHAPE3_DIAMOND, 16, 64, 32, 8),
    // the --breakdown ladder's rung 3 (listed last: never the first match)
    CS_V3DA(933001, 3, SHAPE3_BOX, 16, 64, 32, 8, 1, ",aligned128"),
    CS_V3DA(933011, 3, SHAPE3_STAR, 16, 64, 32, 8, 1, ",aligned128"),
};

const Variant *variants_3d_f64(int *count) {
  // the 8 x 8 block family (k3d2_f64.hip) first: the host takes the first match
  static const std::vector<Variant> all = [] {
    int n2 = 0;
    const Variant *v2 = variants_3d2_f64(&n2);
    std::vector<Variant> v(v2, v2 + n2);
    v.insert(v.end(), kVariants3D, kVariants3D + sizeof(kVariants3D) / sizeof(kVariants3D[0]));
    return v;
  }();
  *count = (int)all.size();
  return all.data();
}

int bfrag_ksteps_3d_f64(int R, int shape) { return make_ktable3d(R, shape).n; }

void build_bfrag_3d_f64(int R, int shape, const double *w, double *bfrag) {
  const KTable3D t = make_ktable3d(R, shape);
  const int D = 2 * R + 1;
  for (int ks = 0; ks < t.n; ++ks)
    for (int lane = 0; lane < 64; ++lane) {
      const int k = lane >> 4, j = lane & 3;
      const int dx = t.c[ks][k] - j;
      double v = 0.0;
      if (t.live[ks][k] && dx >= -R && dx <= R) v = w[((t.dz[ks] + R) * D + (t.dy[ks][k] + R)) * D + (dx + R)];
      bfrag[ks * 64 + lane] = v;
    }
}

}  // namespace cs
