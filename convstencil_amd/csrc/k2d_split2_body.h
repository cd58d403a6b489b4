// k2d_split2_body.h -- split-precision 2-D tile body, dilated form (v_mfma_f32_16x16x32).
//
// Same arithmetic contract as k2d_split_body.h -- FP64 state in HBM, slices of a
// narrow type inside a fused launch, FP32 MFMA accumulation -- with the GEMM laid
// out for the MFMA rather than for a Toeplitz along one axis (cs_ktable.h, CTable):
//   * A = the weights' slices (registers, whole kernel), B = the tile's slices read
//     from LDS with one 16-byte read per lane per MFMA, D = a 16 x 16 output block in
//     which every lane holds 4 adjacent columns of one row;
//   * box2d3r needs 4 MFMAs per slice product per block (the Toeplitz form needs 6),
//     star2d3r 3, box2d1r 2;
//   * two FP32 accumulators per block: level 0 (hi x hi) and every lower level
//     together (their terms are at most 2^-8 of the level-1 sum for BF16x3, so FP32
//     keeps them to about 2^-32 of the value);
//   * between fused steps the value is recombined in FP32 (one rounding, 2^-24) and
//     re-split exactly (three BF16 slices hold any FP32 value); the 4 adjacent results
//     of a lane go back to LDS as one 8-byte write per slice; the last step recombines
//     in FP64 and writes FP64;
//   * one LDS buffer: each wave computes all of its blocks of a step into registers,
//     the workgroup synchronises, then the results are written back.  Cells written
//     outside the step's region are never read again (the next step reads exactly
//     this step's region), so only cells outside the domain interior are masked.
#pragma once
#include "k2d_split_body.h"

namespace cs {

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

template <class Tr>
struct Quad;
template <>
struct Quad<TrBF16> {
  using V = bf16x4;
};
template <>
struct Quad<TrF16> {
  using V = f16x4;
};

// OS = 1: the last step's FP64 results are staged in LDS (over the slice planes, which
// are no longer read) and written out row by row, coalesced.
template <class Tr_, int R_, int SH_, int KF_, int BY_, int BX_, int NW_, int OS_ = 0>
struct CfgS2 {
  using Tr = Tr_;
  using T = typename Tr::T;
  static constexpr int R = R_, SH = SH_, KF = KF_, BY = BY_, BX = BX_, NW = NW_, OS = OS_;
  static constexpr int EY = BY + 2 * KF * R;
  static constexpr int EX = BX + 2 * KF * R;
  static constexpr CTable CT = make_ctable(R, SH);
  static constexpr int NINS = CT.n / 4;  // MFMAs per slice product per block
  static constexpr int NSU = Tr::NSU, NSW = Tr::NSW;
  // region of fused step s: rows [LO, LO + H), columns [LO, HI)
  static constexpr int H(int s) { return BY + 2 * R * (KF - 1 - s); }
  static constexpr int W(int s) { return BX + 2 * R * (KF - 1 - s); }
  static constexpr int LO(int s) { return R * (s + 1); }
  static constexpr int HI(int s) { return LO(s) + W(s); }
  // tile column c is stored at column c + CS; blocks start at columns x with
  // x + CS + c0 = 0 (mod 8), the first at most 7 columns left of the region
  static constexpr int lox(int s, int cs) { return LO(s) - (((LO(s) + cs + CT.c0) % 8) + 8) % 8; }
  static constexpr int nbx(int s, int cs) { return (HI(s) - lox(s, cs) + 15) / 16; }
  static constexpr int min_col(int cs) {
    int m = 1 << 20;
    for (int s = 0; s < KF; ++s) m = lox(s, cs) + CT.cmin < m ? lox(s, cs) + CT.cmin : m;
    return m;
  }
  static constexpr int find_cs() {
    for (int cs = 0; cs < 64; ++cs)
      if (min_col(cs) + cs >= 0) return cs;
    return 64;
  }
  static constexpr int CS = find_cs();
  static constexpr int LOX(int s) { return lox(s, CS); }
  static constexpr int NBX(int s) { return nbx(s, CS); }
  static constexpr int NBY(int s) { return (H(s) + 15) / 16; }
  static constexpr int NB(int s) { return NBY(s) * NBX(s); }
  static constexpr int PERS(int s) { return (NB(s) + NW - 1) / NW; }
  static constexpr int max_col() {  // largest tile column read (chunk + 7) or written (block + 15)
    int m = EX - 1;
    for (int s = 0; s < KF; ++s) {
      const int xl = LOX(s) + 16 * (NBX(s) - 1);
      m = xl + 8 + CT.cmax + 7 > m ? xl + 8 + CT.cmax + 7 : m;
      m = xl + 15 > m ? xl + 15 : m;
    }
    return m;
  }
  static constexpr int PX0 = CS + max_col() + 1;
  // a multiple of 8 elements (16-byte reads) with the fewest LDS bank conflicts for the
  // chunk reads and the result writes that keeps the most workgroups per CU
  // (cs_ktable.h, lds_cost)
  static constexpr int PX = best_pitch_ctable(CT, PX0, (int)sizeof(T), (long)NSU * EY * sizeof(T), 8 * sizeof(T));
  static constexpr int PLANE = EY * PX;
  static constexpr int BUF = NSU * PLANE;
  static constexpr int SPARE = BUF;  // scratch cells after the slice planes
  static constexpr size_t LDS = ((size_t)BUF + 8) * sizeof(T);
  static_assert(BY % 16 == 0 && BX % 16 == 0, "tile must be a multiple of 16");
  static_assert(CS < 64 && (CS + LOX(0) + CT.c0) % 8 == 0, "block columns must align the chunk reads");
  static_assert(PX % 8 == 0, "16-byte reads need a pitch that is a multiple of 8 elements");
  static_assert(LO(0) + CT.tmin >= 0 && LO(0) + H(0) - 16 + 14 + CT.tmax < EY, "block reads outside the tile");
  static_assert(LDS <= 160 * 1024, "tile does not fit gfx950 LDS");
  static_assert(OS == 0 || (size_t)BY * BX * sizeof(double) <= LDS, "staged output does not fit");
};

struct SplitCtx {
  int tid, lane, wave;
  int orow, ocol;  // output cells of this lane relative to a block: row orow, columns ocol .. ocol + 3
  int full;        // 1 if the whole tile (with halo) lies inside the domain interior
  int64_t Y0, X0, ny, nx, ay0, ax0, pitch;
};

// Loads the FP64 tile (EY x EX cells, zero outside the array and in pad columns) and
// stores its slices; loads in groups of LG per thread, all issued before the group's
// first LDS store.
template <class C, int LG>
__device__ __forceinline__ void load_tile_split(const double *__restrict__ in, const Geo &g, const SplitCtx &t,
                                                typename C::T *__restrict__ lds) {
  using T = typename C::T;
  constexpr int NT = C::NW * 64;
  constexpr int NIT = (C::PLANE + NT - 1) / NT;
  constexpr int G = LG < NIT ? LG : NIT;
  const bool all_in = t.ay0 >= 0 && t.ay0 + C::EY <= g.ext[1] && t.ax0 >= 0 && t.ax0 + C::EX <= t.pitch;
  const double *base = in + t.ay0 * t.pitch + t.ax0;
#pragma unroll
  for (int i0 = 0; i0 < NIT; i0 += G) {
    double v[G];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const int idx = t.tid + (i0 + j) * NT;
      const int r = idx / C::PX, c = idx - r * C::PX - C::CS;
      const int64_t ay = t.ay0 + r, ax = t.ax0 + c;
      const bool ok = i0 + j < NIT && idx < C::PLANE && c >= 0 && c < C::EX &&
                      (all_in || (ay >= 0 && ay < g.ext[1] && ax >= 0 && ax < t.pitch));
      const double x = *(ok ? base + ((int64_t)r * t.pitch + c) : in);
      v[j] = ok ? x : 0.0;
    }
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const int idx = t.tid + (i0 + j) * NT;
      const bool in_plane = i0 + j < NIT - 1 || (i0 + j == NIT - 1 && idx < C::PLANE);
      T s[C::NSU];
      split_value<typename C::Tr>(v[j], s);
#pragma unroll
      for (int k = 0; k < C::NSU; ++k) lds[in_plane ? k * C::PLANE + idx : C::SPARE] = s[k];
    }
  }
}

// The tile's FP64 loads, all issued into v[] (used by the persistent kernel to keep the
// next tile's loads in flight), and their split into LDS slices.
template <class C>
struct SplitRegs {
  static constexpr int NT = C::NW * 64;
  static constexpr int NIT = (C::PLANE + NT - 1) / NT;
  double v[NIT];
};

template <class C>
__device__ __forceinline__ void issue_tile_split(const double *__restrict__ in, const Geo &g, const SplitCtx &t,
                                                 SplitRegs<C> &rg) {
  constexpr int NT = SplitRegs<C>::NT, NIT = SplitRegs<C>::NIT;
  const bool all_in = t.ay0 >= 0 && t.ay0 + C::EY <= g.ext[1] && t.ax0 >= 0 && t.ax0 + C::EX <= t.pitch;
  const double *base = in + t.ay0 * t.pitch + t.ax0;
#pragma unroll
  for (int i = 0; i < NIT; ++i) {
    const int idx = t.tid + i * NT;
    const int r = idx / C::PX, c = idx - r * C::PX - C::CS;
    const int64_t ay = t.ay0 + r, ax = t.ax0 + c;
    const bool ok = idx < C::PLANE && c >= 0 && c < C::EX &&
                    (all_in || (ay >= 0 && ay < g.ext[1] && ax >= 0 && ax < t.pitch));
    const double x = *(ok ? base + ((int64_t)r * t.pitch + c) : in);
    rg.v[i] = ok ? x : 0.0;
  }
}

template <class C>
__device__ __forceinline__ void store_tile_split(const SplitCtx &t, const SplitRegs<C> &rg,
                                                 typename C::T *__restrict__ lds) {
  using T = typename C::T;
  constexpr int NT = SplitRegs<C>::NT, NIT = SplitRegs<C>::NIT;
#pragma unroll
  for (int i = 0; i < NIT; ++i) {
    const int idx = t.tid + i * NT;
    const bool in_plane = i < NIT - 1 || idx < C::PLANE;
    T sl[C::NSU];
    split_value<typename C::Tr>(rg.v[i], sl);
#pragma unroll
    for (int k = 0; k < C::NSU; ++k) lds[in_plane ? k * C::PLANE + idx : C::SPARE] = sl[k];
  }
}

template <class C>
__device__ __forceinline__ void split_tile_setup(const Geo &g, int64_t tile, SplitCtx &t) {
  t.pitch = g.ext[2];
  t.ny = g.n[1];
  t.nx = g.n[2];
  t.Y0 = (tile / g.tiles[2]) * C::BY - C::KF * C::R;
  t.X0 = (tile % g.tiles[2]) * C::BX - C::KF * C::R;
  t.ay0 = g.org[1] + t.Y0;
  t.ax0 = g.org[2] + t.X0;
  t.full = t.Y0 >= 0 && t.X0 >= 0 && t.Y0 + C::EY <= t.ny && t.X0 + C::EX <= t.nx;
}

template <class C>
__device__ __forceinline__ void split_lane_setup(const typename C::T *__restrict__ bfrag, int tid, SplitCtx &t,
                                                 typename C::Tr::V (*w)[C::NINS], int *off) {
  using V = typename C::Tr::V;
  t.tid = tid;
  t.lane = tid & 63;
  t.wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = t.lane, grp = lane >> 4, n = lane & 15;
  t.orow = 2 * (n & 7) + (grp >> 1);
  t.ocol = 8 * (n >> 3) + 4 * (grp & 1);
#pragma unroll
  for (int q = 0; q < C::NSW; ++q)
#pragma unroll
    for (int i = 0; i < C::NINS; ++i)
      w[q][i] = *reinterpret_cast<const V *>(bfrag + ((size_t)(q * C::NINS + i) * 64 + lane) * 8);
#pragma unroll
  for (int i = 0; i < C::NINS; ++i)
    off[i] = (2 * (n & 7) + C::CT.ty[4 * i + grp]) * C::PX + 8 * (n >> 3) + C::CT.tx[4 * i + grp] + C::CS;
}

// One block: D0 = level 0 (hi x hi), D1 = the kept lower levels, scaled by SCALE.
template <class C>
__device__ __forceinline__ void block_split(const typename C::T *p, const int *off,
                                            const typename C::Tr::V (*w)[C::NINS], f4 &d0, f4 &d1) {
  using Tr = typename C::Tr;
  using V = typename Tr::V;
  d0 = f4{0.f, 0.f, 0.f, 0.f};
  d1 = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < C::NINS; ++i) {
    V u[C::NSU];
#pragma unroll
    for (int s = 0; s < C::NSU; ++s) u[s] = *reinterpret_cast<const V *>(p + s * C::PLANE + off[i]);
    d0 = Tr::mma(w[0][i], u[0], d0);
#pragma unroll
    for (int s = 0; s < C::NSU; ++s)
#pragma unroll
      for (int q = 0; q < C::NSW; ++q)
        if (s + q >= 1 && s + q < Tr::NL) d1 = Tr::mma(w[q][i], u[s], d1);
  }
}

// Re-splits one FP32 value into the slices of Tr (exact for BF16x3).
template <class Tr>
__device__ __forceinline__ void split_f32(float v, typename Tr::T *s) {
  using T = typename Tr::T;
  if constexpr (Tr::NSU == 3) {
    const T h = (T)v;
    const float r1 = v - (float)h;
    const T m = (T)r1;
    s[0] = h;
    s[1] = m;
    s[2] = (T)(r1 - (float)m);
  } else {
    const T h = (T)v;
    s[0] = h;
    s[1] = (T)((v - (float)h) * (float)Tr::SCALE);
  }
}

template <class C, int S>
__device__ __forceinline__ void block_origin_s2(int blk, int &y, int &x) {
  constexpr int NBX = C::NBX(S);
  const int ba = blk / NBX, bb = blk - ba * NBX;
  y = C::LO(S) + min(16 * ba, C::H(S) - 16);
  x = C::LOX(S) + 16 * bb;
}

template <class C, int S>
__device__ __forceinline__ void split_step(const SplitCtx &t, const typename C::Tr::V (*w)[C::NINS], const int *off,
                                           typename C::T *lds, double *out) {
  using Tr = typename C::Tr;
  using T = typename C::T;
  using Q = typename Quad<Tr>::V;
  constexpr bool LAST = S == C::KF - 1;
  constexpr int NB = C::NB(S), PERS = C::PERS(S), NW = C::NW;
  constexpr float INV = 1.0f / (float)Tr::SCALE;
  f4 d0[PERS], d1[PERS];
#pragma unroll
  for (int j = 0; j < PERS; ++j) {
    int y, x;
    block_origin_s2<C, S>(min(j * NW + t.wave, NB - 1), y, x);
    block_split<C>(lds + y * C::PX + x, off, w, d0[j], d1[j]);
  }
  if constexpr (!LAST || C::OS == 1) __syncthreads();  // every wave has read this step's input
#pragma unroll
  for (int j = 0; j < PERS; ++j) {
    const int blk = j * NW + t.wave;
    if (!((NB % NW == 0 || j < PERS - 1) || blk < NB)) continue;
    int y, x;
    block_origin_s2<C, S>(blk, y, x);
    const int ty = y + t.orow, tx0 = x + t.ocol;
    const bool row_in = t.full || (t.Y0 + ty >= 0 && t.Y0 + ty < t.ny);
    if constexpr (LAST && C::OS == 1) {
      // staged: the output block cell (ty - LO, tx - LO) of a compact BY x BX FP64 array
      double *ob = reinterpret_cast<double *>(lds);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int tx = tx0 + v;
        if (tx >= C::LO(S) && tx < C::HI(S))
          ob[(ty - C::LO(S)) * C::BX + tx - C::LO(S)] = (double)d0[j][v] + (double)d1[j][v] * (double)INV;
      }
    } else if constexpr (LAST) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int tx = tx0 + v;
        const bool in = tx >= C::LO(S) && tx < C::HI(S) && row_in && (t.full || (t.X0 + tx >= 0 && t.X0 + tx < t.nx));
        if (in) out[(t.ay0 + ty) * t.pitch + t.ax0 + tx] = (double)d0[j][v] + (double)d1[j][v] * (double)INV;
      }
    } else {
      Q q[C::NSU];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        T s[C::NSU];
        split_f32<Tr>(d0[j][v] + d1[j][v] * INV, s);
#pragma unroll
        for (int k = 0; k < C::NSU; ++k) q[k][v] = s[k];
      }
      const int li = ty * C::PX + tx0 + C::CS;
      const bool all4 = t.full || (row_in && t.X0 + tx0 >= 0 && t.X0 + tx0 + 3 < t.nx);
      if (all4) {
#pragma unroll
        for (int k = 0; k < C::NSU; ++k) *reinterpret_cast<Q *>(lds + k * C::PLANE + li) = q[k];
      } else if (row_in) {  // a block at the domain edge: keep the halo cells
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (t.X0 + tx0 + v >= 0 && t.X0 + tx0 + v < t.nx)
#pragma unroll
            for (int k = 0; k < C::NSU; ++k) lds[k * C::PLANE + li + v] = q[k][v];
      }
    }
  }
  if constexpr (!LAST) __syncthreads();  // this step's output is complete before it is read
  if constexpr (LAST && C::OS == 1) {
    __syncthreads();  // the staged block is complete
    constexpr int NT = NW * 64, RP = NT / C::BX, OFF = C::KF * C::R;
    static_assert(NT % C::BX == 0, "a pass covers whole rows");
    const double *ob = reinterpret_cast<const double *>(lds);
    const int r0 = t.tid / C::BX, c = t.tid - r0 * C::BX;
    const bool col_in = t.full || t.X0 + OFF + c < t.nx;
    double *o = out + (t.ay0 + OFF + r0) * t.pitch + t.ax0 + OFF + c;
#pragma unroll
    for (int i = 0; i < C::BY / RP; ++i)
      if (col_in && (t.full || t.Y0 + OFF + r0 + RP * i < t.ny)) o[(int64_t)RP * i * t.pitch] = ob[(r0 + RP * i) * C::BX + c];
  }
}

template <class C, int S>
__device__ __forceinline__ void split_steps(const SplitCtx &t, const typename C::Tr::V (*w)[C::NINS], const int *off,
                                            typename C::T *lds, double *out) {
  if constexpr (S < C::KF) {
    split_step<C, S>(t, w, off, lds, out);
    split_steps<C, S + 1>(t, w, off, lds, out);
  }
}

// One workgroup advances tile `tile` of grid g by KF steps (smem: the kernel's dynamic
// LDS; bfrag: the weights' slices as A fragments; LG: tile loads in flight per thread;
// tid: threadIdx.x).
template <class Tr, int R, int SH, int KF, int BY, int BX, int NW, int LG = 64, int OS = 0>
__device__ __forceinline__ void tile_split2(const double *__restrict__ in, double *__restrict__ out,
                                            const typename Tr::T *__restrict__ bfrag, const Geo &g, int64_t tile,
                                            unsigned char *__restrict__ smem, int tid) {
  using C = CfgS2<Tr, R, SH, KF, BY, BX, NW, OS>;
  using T = typename Tr::T;
  using V = typename Tr::V;
  T *lds = reinterpret_cast<T *>(smem);
  SplitCtx t;
  t.tid = tid;
  t.lane = tid & 63;
  t.wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  t.pitch = g.ext[2];
  t.ny = g.n[1];
  t.nx = g.n[2];
  t.Y0 = (tile / g.tiles[2]) * BY - KF * R;
  t.X0 = (tile % g.tiles[2]) * BX - KF * R;
  t.ay0 = g.org[1] + t.Y0;
  t.ax0 = g.org[2] + t.X0;
  t.full = t.Y0 >= 0 && t.X0 >= 0 && t.Y0 + C::EY <= t.ny && t.X0 + C::EX <= t.nx;
  const int lane = t.lane, grp = lane >> 4, n = lane & 15;
  // D[m][n], m = 4 grp + v: cell (2 (n & 7) + (grp >> 1), 8 (n >> 3) + 4 (grp & 1) + v)
  t.orow = 2 * (n & 7) + (grp >> 1);
  t.ocol = 8 * (n >> 3) + 4 * (grp & 1);

  V w[C::NSW][C::NINS];
#pragma unroll
  for (int q = 0; q < C::NSW; ++q)
#pragma unroll
    for (int i = 0; i < C::NINS; ++i)
      w[q][i] = *reinterpret_cast<const V *>(bfrag + ((size_t)(q * C::NINS + i) * 64 + lane) * 8);
  // B offsets: chunk 4 i + grp of instruction i, anchor u_n = (2 (n & 7), 8 (n >> 3))
  int off[C::NINS];
#pragma unroll
  for (int i = 0; i < C::NINS; ++i)
    off[i] = (2 * (n & 7) + C::CT.ty[4 * i + grp]) * C::PX + 8 * (n >> 3) + C::CT.tx[4 * i + grp] + C::CS;

  load_tile_split<C, LG>(in, g, t, lds);
  __syncthreads();
  split_steps<C, 0>(t, w, off, lds, out);
}

}  // namespace cs
