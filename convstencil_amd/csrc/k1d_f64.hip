// k1d_f64.hip -- plain FP64 1-D ConvStencil kernels for gfx950 (v_mfma_f64_4x4x4).
//
// Replaces gpu_1d1r_kernel (reference src/1d/gpu_1r.cu:26-75) and
// gpu_star_1d2r_step2_kernel (src/1d/gpu_2r.cu:23-77).  A wave computes 64
// consecutive outputs per MFMA chain: the four 4x4 blocks of v_mfma_f64_4x4x4 are
// laid out so that lane l produces output x + l (block b = (l>>2)&3, row i = l>>4,
// column j = l&3 stand for position 16 i + 4 b + j).  K is the flattened list of
// input offsets c in [-R, 3 + R] that a 4-wide output block needs, padded to a
// multiple of 4 with zero weights:
//   A_b[i][k] = tile[x + 16 i + 4 b + c_f],   B[k][j] = w[c_f - j]   (f = 4 ks + k).
// The reference's stencil2row scatter (lookup tables + a shared "dump" slot that
// several threads write concurrently, src/1d/gpu_1r.cu:243-264) is not needed.
//
// Temporal fusion: a workgroup loads L + 2 KF R cells and advances them KF steps in
// LDS, keeping halo cells of the domain fixed, then writes the central L outputs.
// With L = 2048 and KF = 32 the 2^20-point grid needs two launches for 64 steps.
#include "cs_internal.h"

namespace cs {

template <int R>
struct Table1D {
  static constexpr int NE = ((4 + 2 * R) + 3) / 4 * 4;  // padded flattened entries
  static constexpr int NKS = NE / 4;
  static constexpr int c(int f) { return f < 4 + 2 * R ? f - R : 0; }
  static constexpr bool live(int f) { return f < 4 + 2 * R; }
};

template <int R, int KF, int L, int NW>
struct Cfg1D {
  static constexpr int E = L + 2 * KF * R;
  static constexpr int P = E + 1;  // padded buffer length
  static constexpr int NBUF = KF > 1 ? 2 : 1;
  static constexpr int SPARE = NBUF * P;  // one scratch cell after the buffers
  static constexpr size_t LDS = (size_t)(NBUF * P + 2) * sizeof(double);
  static_assert(L % 64 == 0, "tile must be a multiple of 64");
  static_assert(LDS <= 160 * 1024, "tile does not fit gfx950 LDS");
};

template <int R, int KF, int L, int NW>
__global__ __launch_bounds__(NW * 64) void k1d_f64(const double *__restrict__ in, double *__restrict__ out,
                                                   const double *__restrict__ bfrag, Geo g) {
  using C = Cfg1D<R, KF, L, NW>;
  using T = Table1D<R>;
  extern __shared__ double lds[];
  const int64_t ntiles = g.tiles[2];
  const int64_t tile = blockIdx.x;
  if (tile >= ntiles) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t nx = g.n[2], ext = g.ext[2];
  const int64_t X0 = tile * L - KF * R;  // interior coordinate of buffer cell 0
  const int64_t ax0 = g.org[2] + X0;

  double b[T::NKS];
  int off[T::NKS];
  const int k = lane >> 4, b4 = (lane >> 2) & 3, i = lane & 3;
#pragma unroll
  for (int ks = 0; ks < T::NKS; ++ks) {
    b[ks] = bfrag[ks * 64 + lane];
    const int f = 4 * ks + k;
    off[ks] = 16 * i + 4 * b4 + T::c(f);
  }

  // all loads in flight before the first LDS store; a cell that is not read loads in[0]
  // (always valid) and is replaced by 0, threads past the buffer store to the spare cell
  {
    constexpr int NT = NW * 64, NIT = (C::P + NT - 1) / NT;
    double v[NIT];
#pragma unroll
    for (int i = 0; i < NIT; ++i) {
      const int idx = tid + i * NT;
      const int64_t ax = ax0 + idx;
      const bool ok = idx < C::E && ax >= 0 && ax < ext;
      const double x = *(ok ? in + ax : in);
      v[i] = ok ? x : 0.0;
    }
#pragma unroll
    for (int i = 0; i < NIT; ++i) {
      const int idx = tid + i * NT;
      const bool in_buf = i < NIT - 1 || idx < C::P;
      lds[in_buf ? idx : C::SPARE] = v[i];
      if constexpr (KF > 1) lds[in_buf ? C::P + idx : C::SPARE] = v[i];
    }
  }
  __syncthreads();

  const bool full = X0 >= 0 && X0 + C::E <= nx;
  for (int s = 0; s < KF; ++s) {
    const double *src = lds + (s & 1) * C::P;
    double *dst = lds + ((s + 1) & 1) * C::P;
    const int W = L + 2 * R * (KF - 1 - s);
    const int base = R * (s + 1);
    const int nb = (W + 63) / 64;
    for (int blk = wave; blk < nb; blk += NW) {
      const int x = base + min(64 * blk, W - 64);
      const double *p = src + x;
      double acc = 0.0;
#pragma unroll
      for (int ks = 0; ks < T::NKS; ++ks) acc = __builtin_amdgcn_mfma_f64_4x4x4f64(p[off[ks]], b[ks], acc, 0, 0, 0);
      const int tx = x + lane;
      const bool inside = full || (X0 + tx >= 0 && X0 + tx < nx);
      if (inside) {
        if (s < KF - 1)
          dst[tx] = acc;
        else
          out[ax0 + tx] = acc;
      }
    }
    if (s < KF - 1) __syncthreads();
  }
}

template <int R, int KF, int L, int NW>
static hipError_t prepare_k1d() {
  using C = Cfg1D<R, KF, L, NW>;
  return hipFuncSetAttribute((const void *)k1d_f64<R, KF, L, NW>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)C::LDS);
}

template <int R, int KF, int L, int NW>
static hipError_t launch_k1d(const double *in, double *out, const double *bfrag, const Geo &g, hipStream_t s) {
  using C = Cfg1D<R, KF, L, NW>;
  hipLaunchKernelGGL((k1d_f64<R, KF, L, NW>), dim3((unsigned)g.tiles[2]), dim3(NW * 64), C::LDS, s, in, out, bfrag,
                     g);
  return hipGetLastError();
}

#define CS_V1D(ID, R, KF, L, NW)                                                                                 \
  Variant {                                                                                                      \
    ID, 4, R, SHAPE_BOX, KF, 1, L, NW, Cfg1D<R, KF, L, NW>::LDS, Cfg1D<R, KF, L, NW>::P, 0, &launch_k1d<R, KF, L, NW>, \
        &prepare_k1d<R, KF, L, NW>, "k1d_f64<r" #R ",kf" #KF "," #L ",w" #NW ">"                                 \
  }

static const Variant kVariants1D[] = {
    // 1024 outputs per 512 threads first at depth 16: 969 against 865 GStencil/s at 2^20
    // points (twice the workgroups), 1486 against 1528 at 10.24e6
    CS_V1D(521016, 1, 16, 1024, 8),
    CS_V1D(11001, 1, 1, 2048, 8),  CS_V1D(11002, 1, 2, 2048, 8),  CS_V1D(11004, 1, 4, 2048, 8),
    CS_V1D(11008, 1, 8, 2048, 8),  CS_V1D(11016, 1, 16, 2048, 8), CS_V1D(11032, 1, 32, 2048, 8),
    CS_V1D(12001, 2, 1, 2048, 8),  CS_V1D(12002, 2, 2, 2048, 8),  CS_V1D(12004, 2, 4, 2048, 8),
    CS_V1D(12008, 2, 8, 2048, 8),  CS_V1D(12016, 2, 16, 2048, 8), CS_V1D(13001, 3, 1, 2048, 8),
    CS_V1D(13002, 3, 2, 2048, 8),  CS_V1D(13004, 3, 4, 2048, 8),  CS_V1D(13008, 3, 8, 2048, 8),
    CS_V1D(14001, 4, 1, 2048, 8),  CS_V1D(14002, 4, 2, 2048, 8),  CS_V1D(14004, 4, 4, 2048, 8),
    CS_V1D(14008, 4, 8, 2048, 8),
    // smaller tiles (more workgroups for grids of a few million points): autotuner candidates
    CS_V1D(511016, 1, 16, 1024, 4), CS_V1D(531016, 1, 16, 512, 4),
    CS_V1D(511008, 1, 8, 1024, 4),  CS_V1D(511032, 1, 32, 1024, 4), CS_V1D(531008, 1, 8, 512, 4),
    // the --breakdown ladder's rung 3: the reference's tiling, 1024 outputs per 256 threads
    // (listed last: never the first match)
    CS_V1D(933001, 3, 1, 1024, 4),
};

const Variant *variants_1d_f64(int *count) {
  *count = (int)(sizeof(kVariants1D) / sizeof(kVariants1D[0]));
  return kVariants1D;
}

int bfrag_ksteps_1d_f64(int R) { return (4 + 2 * R + 3) / 4; }

void build_bfrag_1d_f64(int R, const double *w, double *bfrag) {
  const int ne = 4 + 2 * R, nks = (ne + 3) / 4;
  for (int ks = 0; ks < nks; ++ks)
    for (int lane = 0; lane < 64; ++lane) {
      const int f = 4 * ks + (lane >> 4), j = lane & 3;
      double v = 0.0;
      if (f < ne) {
        const int dx = (f - R) - j;
        if (dx >= -R && dx <= R) v = w[dx + R];
      }
      bfrag[ks * 64 + lane] = v;
    }
}

}  // namespace cs
