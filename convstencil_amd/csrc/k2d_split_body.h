// k2d_split_body.h -- split-precision 2-D ConvStencil tile body for gfx950.
//
// The state stays FP64 in HBM.  Inside a fused launch each tile value u and each
// weight w is split into slices of a narrow type,
//   BF16x3 : u = u0 + u1 + u2 (bf16),                  v_mfma_f32_16x16x32_bf16
//   FP16x2 : u = u0 + u1 * 2^-11 (fp16, rescaled),     v_mfma_f32_16x16x32_f16
//   FP32x2 : u = u0 + u1 (fp32 hi / lo),               v_mfma_f32_16x16x4f32
// and the stencil is a sum of slice products accumulated in FP32, one
// accumulator per precision level s + t (levels above the kept ones are dropped:
// BF16x3 keeps s + t <= 2, i.e. 6 of 9 products; the 2-slice forms keep 3 of 4).
// Products of two 8-bit (bf16) or 11-bit (fp16) significands are exact in FP32,
// so the only FP32 rounding is in the accumulation.  After every fused step the
// levels are recombined in FP64 (the rescaling is a power of two), halo cells are
// left untouched, and the FP64 value is re-split for the next step; the last step
// writes FP64 to HBM.
//
// GEMM form (output block of 16 rows x 16 columns): A rows are output columns and
// B columns are output rows, with the Toeplitz along y:
//   A[i][(dx, rr)] = u(y - R + rr, x + i + dx),   B[(dx, rr)][n] = w(rr - R - n, dx).
// Slice planes are stored column-major in LDS, so the eight consecutive rows a
// lane feeds to one MFMA are one aligned 16-byte read.  Row storage is offset so
// that every block's first input row y - R starts an 8-row chunk; 3 chunks (24
// rows) then cover the 16 + 2R rows a block needs (R <= 4).
#pragma once
#include "cs_internal.h"

namespace cs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

struct TrBF16 {
  using T = __bf16;
  using V = bf16x8;
  static constexpr int CK = 8, KI = 32, NSU = 3, NSW = 3, NL = 3;
  static constexpr double SCALE = 1.0;  // slice s carries weight SCALE^-s
  __device__ static f4 mma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};
struct TrF16 {
  using T = _Float16;
  using V = f16x8;
  static constexpr int CK = 8, KI = 32, NSU = 2, NSW = 2, NL = 2;
  static constexpr double SCALE = 2048.0;
  __device__ static f4 mma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};
// FP16 with three weight slices and two state slices: the weights carry 33 bits, the
// state 22, and four products are kept (w0 u0; w0 u1, w1 u0; w2 u0) -- the state's
// rounding error stays random from step to step, while a two-slice weight would bias
// every step the same way.  The third weight slice is stored at the second's scale
// (the host's split), so its product shares the level-1 accumulator.
struct TrF16W3 {
  using T = _Float16;
  using V = f16x8;
  static constexpr int CK = 8, KI = 32, NSU = 2, NSW = 3, NL = 3;
  static constexpr double SCALE = 2048.0;
  __device__ static f4 mma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};
struct TrF32 {
  using T = float;
  using V = float;
  static constexpr int CK = 1, KI = 4, NSU = 2, NSW = 2, NL = 2;
  static constexpr double SCALE = 1.0;
  __device__ static f4 mma(V a, V b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

// Slice products the dilated form keeps below level 0 (state slice s times weight
// slice q, level s + q): every level below NL, except w1 u1 for TrF16W3.
template <class Tr>
constexpr bool keep_product(int s, int q) {
  if (s + q < 1 || s + q >= Tr::NL) return false;
  if constexpr (Tr::NSU == 2 && Tr::NSW == 3) return s == 0 || q == 0;
  return true;
}

// Splits v into NS slices of Tr::T; slice s is scaled by SCALE^s before rounding.
template <class Tr>
__host__ __device__ inline void split_value(double v, typename Tr::T *s) {
  double rem = v, sc = 1.0;
#pragma unroll
  for (int k = 0; k < Tr::NSU; ++k) {
    const typename Tr::T t = (typename Tr::T)(float)(rem * sc);
    s[k] = t;
    rem -= (double)(float)t / sc;
    sc *= Tr::SCALE;
  }
}

// The weights' slices of the dilated forms (host side): as the state (split_value) with
// NSW slices, except for TrF16W3, whose third slice is the remainder after two slices
// stored at the second slice's scale (its products join the level-1 accumulator).
template <class Tr>
inline void split_weight(double v, typename Tr::T *sl) {
  using T = typename Tr::T;
  if constexpr (Tr::NSU == 2 && Tr::NSW == 3) {
    const T h = (T)(float)v;
    const double r = (v - (double)(float)h) * Tr::SCALE;
    const T m = (T)(float)r;
    sl[0] = h;
    sl[1] = m;
    sl[2] = (T)(float)(r - (double)(float)m);
  } else {
    T s[Tr::NSU];
    split_value<Tr>(v, s);
    for (int k = 0; k < Tr::NSW; ++k) sl[k] = s[k];
  }
}

template <class Tr, int R, int KF, int BY, int BX, int NW>
struct CfgS {
  using T = typename Tr::T;
  static constexpr int EY = BY + 2 * KF * R;
  static constexpr int EX = BX + 2 * KF * R;
  // rows entries per dx, and the storage offset that aligns every block's first input row
  static constexpr int RPD = Tr::CK == 8 ? 24 : ((16 + 2 * R + 3) / 4) * 4;
  static constexpr int ALIGN = Tr::CK == 8 ? 8 : 1;
  static constexpr int TOP0 = ((R - KF * R) % 8 + 8) % 8;  // TOP == R - KF*R (mod 8)
  static constexpr int TOP = Tr::CK == 8 ? TOP0 + 8 : 8;
  static constexpr int PR0 = TOP + EY + RPD - 2 * R - 1;   // rows incl. the last block's overrun
  static constexpr int PR = (PR0 + 7) / 8 * 8;
  static constexpr int PLANE = EX * PR;                     // elements per slice plane
  static constexpr int NSU = Tr::NSU;
  static constexpr int BUF = NSU * PLANE;                   // one buffer: all slices
  static constexpr int SPARE = 2 * BUF;                     // scratch cells after the buffers
  static constexpr size_t LDS = (2 * (size_t)BUF + 8) * sizeof(T);
  static constexpr int NENT = (2 * R + 1) * RPD;            // K entries (dx, rr)
  static constexpr int NINS = (NENT + Tr::KI - 1) / Tr::KI; // MFMA instructions per product
  static_assert(BY % 16 == 0 && BX % 16 == 0, "tile must be a multiple of 16");
  static_assert(R <= 4, "3 chunks of 8 rows cover 16 + 2R rows only for R <= 4");
  static_assert(LDS <= 160 * 1024, "tile does not fit gfx950 LDS");
  // K entry e -> (dx, rr); padding entries map to (0, 0) with weight 0
  static constexpr int ent_dx(int e) { return e < NENT ? e / RPD - R : 0; }
  static constexpr int ent_rr(int e) { return e < NENT ? e % RPD : 0; }
};

// One workgroup advances tile `tile` of grid g by KF steps (smem: the kernel's dynamic LDS).
template <class Tr, int R, int KF, int BY, int BX, int NW, int LG = 64>
__device__ __forceinline__ void tile_split(const double *__restrict__ in, double *__restrict__ out,
                                           const typename Tr::T *__restrict__ bfrag, const Geo &g, int64_t tile,
                                           unsigned char *__restrict__ smem, int tid) {
  using C = CfgS<Tr, R, KF, BY, BX, NW>;
  using T = typename Tr::T;
  using V = typename Tr::V;
  constexpr int NINS = C::NINS, NSU = Tr::NSU, NSW = Tr::NSW, NL = Tr::NL, CK = Tr::CK;
  T *lds = reinterpret_cast<T *>(smem);
  const int lane = tid & 63, wave = tid >> 6;
  const int64_t pitch = g.ext[2], ny = g.n[1], nx = g.n[2];
  const int64_t Y0 = (tile / g.tiles[2]) * BY - KF * R;
  const int64_t X0 = (tile % g.tiles[2]) * BX - KF * R;
  const int64_t ay0 = g.org[1] + Y0, ax0 = g.org[2] + X0;
  const bool full = Y0 >= 0 && X0 >= 0 && Y0 + C::EY <= ny && X0 + C::EX <= nx;

  // B fragments: b[t][m] for weight slice t and instruction m
  V b[NSW][NINS];
#pragma unroll
  for (int t = 0; t < NSW; ++t)
#pragma unroll
    for (int m = 0; m < NINS; ++m) {
      const T *src = bfrag + ((size_t)(t * NINS + m) * 64 + lane) * CK;
      if constexpr (CK == 8)
        b[t][m] = *reinterpret_cast<const V *>(src);
      else
        b[t][m] = *src;
    }
  // per-lane A offsets (elements, within a slice plane) relative to the block's
  // (x, y - R + TOP) origin: lane group q = lane >> 4 takes K entries of chunk q
  int off[NINS];
  const int ai = lane & 15, q = lane >> 4;
#pragma unroll
  for (int m = 0; m < NINS; ++m) {
    const int e = m * Tr::KI + q * CK;
    off[m] = (ai + C::ent_dx(e)) * C::PR + C::ent_rr(e);
  }

  // ---- load: FP64 tile -> slices in both buffers (halo cells stay as loaded) ----
  // Consecutive threads walk a row (coalesced loads); each thread issues its loads in
  // groups of LG before the group's first LDS store.  A cell that is not read loads the
  // array's first element (always valid) and is replaced by 0; threads past the plane
  // store to the spare cells.
  const bool all_in = ay0 >= 0 && ay0 + C::EY <= g.ext[1] && ax0 >= 0 && ax0 + C::EX <= pitch;
  {
    constexpr int NT = NW * 64, NIT = (C::PLANE + NT - 1) / NT;
    constexpr int G = LG < NIT ? LG : NIT;
#pragma unroll
    for (int i0 = 0; i0 < NIT; i0 += G) {
      double v[G];
#pragma unroll
      for (int j = 0; j < G; ++j) {
        const int idx = tid + (i0 + j) * NT;
        const int sr = idx / C::EX, c = idx - sr * C::EX;  // storage row sr holds tile row sr - TOP
        const int r = sr - C::TOP;
        const int64_t ay = ay0 + r, ax = ax0 + c;
        const bool ok = i0 + j < NIT && idx < C::PLANE && r >= 0 && r < C::EY &&
                        (all_in || (ay >= 0 && ay < g.ext[1] && ax >= 0 && ax < pitch));
        const double x = *(ok ? in + (ay * pitch + ax) : in);
        v[j] = ok ? x : 0.0;
      }
#pragma unroll
      for (int j = 0; j < G; ++j) {
        const int idx = tid + (i0 + j) * NT;
        const int sr = idx / C::EX, c = idx - sr * C::EX;
        const bool in_plane = i0 + j < NIT - 1 || (i0 + j == NIT - 1 && idx < C::PLANE);
        const int li = c * C::PR + sr;  // column-major slice planes
        T s[NSU];
        split_value<Tr>(v[j], s);
#pragma unroll
        for (int k = 0; k < NSU; ++k) {
          lds[in_plane ? k * C::PLANE + li : C::SPARE] = s[k];
          lds[in_plane ? C::BUF + k * C::PLANE + li : C::SPARE] = s[k];
        }
      }
    }
  }
  __syncthreads();

  for (int st = 0; st < KF; ++st) {
    const T *src = lds + (st & 1) * C::BUF;
    T *dst = lds + ((st + 1) & 1) * C::BUF;
    const bool last = st == KF - 1;
    const int lo = R * (st + 1), hi = C::EY - R * (st + 1);  // valid rows (and columns) of this step
    const int W = C::EX - 2 * R * (st + 1);
    // first block row: the largest y <= lo with y == KF*R (mod 8) (alignment of y - R + TOP)
    const int y0 = lo - (((lo - KF * R) % 8) + 8) % 8;
    const int nby = (hi - y0 + 15) / 16, nbx = (W + 15) / 16;
    for (int blk = wave; blk < nby * nbx; blk += NW) {
      const int ba = blk / nbx, bb = blk - ba * nbx;
      const int y = y0 + 16 * ba;
      const int x = lo + min(16 * bb, W - 16);
      const int ubase = x * C::PR + (y - R + C::TOP);
      f4 acc[NL];
#pragma unroll
      for (int L = 0; L < NL; ++L) acc[L] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int m = 0; m < NINS; ++m) {
#pragma unroll
        for (int s = 0; s < NSU; ++s) {
          V a;
          if constexpr (CK == 8)
            a = *reinterpret_cast<const V *>(src + s * C::PLANE + ubase + off[m]);
          else
            a = src[s * C::PLANE + ubase + off[m]];
#pragma unroll
          for (int t = 0; t < NSW; ++t)
            if (s + t < NL) acc[s + t] = Tr::mma(a, b[t][m], acc[s + t]);
        }
      }
      // recombine in FP64: lane holds output column x + 4 (lane >> 4) + v, row y + (lane & 15)
      const int tyy = y + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int txx = x + 4 * (lane >> 4) + v;
        double val = 0.0, sc = 1.0;
#pragma unroll
        for (int L = 0; L < NL; ++L) {
          val += (double)acc[L][v] / sc;
          sc *= Tr::SCALE;
        }
        const bool in_region = tyy >= lo && tyy < hi;
        const bool inside = full || (Y0 + tyy >= 0 && Y0 + tyy < ny && X0 + txx >= 0 && X0 + txx < nx);
        if (!(in_region && inside)) continue;
        if (last) {
          out[(ay0 + tyy) * pitch + ax0 + txx] = val;
        } else {
          T sl[NSU];
          split_value<Tr>(val, sl);
#pragma unroll
          for (int k = 0; k < NSU; ++k) dst[k * C::PLANE + txx * C::PR + tyy + C::TOP] = sl[k];
        }
      }
    }
    if (!last) __syncthreads();
  }
}

}  // namespace cs
