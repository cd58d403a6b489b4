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
template <>
struct Quad<TrF16W3> {
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
  static constexpr CTable CT = CTab2<R, SH>::T;
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
  // the reads' shared vector addresses and immediate offsets at this pitch
  static constexpr RdPlan RPL = make_splan(CT, PX, (int)sizeof(T));
  // block `blk` of step s -> its first output cell (y, x); the last block row is moved
  // back to end at the region's edge, the last block column may extend past it
  static constexpr void origin(int s, int blk, int &y, int &x) {
    const int ba = blk / NBX(s), bb = blk - ba * NBX(s);
    y = LO(s) + (16 * ba < H(s) - 16 ? 16 * ba : H(s) - 16);
    x = LOX(s) + 16 * bb;
  }
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
  int vw;                  // LDS byte offset of this lane's first result cell from the block origin
  double *gout;            // output cell of tile cell (0, 0)
  int ry0, nry, cx0, ncx;  // tile rows [ry0, ry0 + nry) and columns [cx0, cx0 + ncx) are interior
};

// Splits one FP32 value into the slices of Tr (exact for BF16x3).  The FP64 state is
// rounded to FP32 once where a tile is loaded (the MFMAs accumulate in FP32 anyway) and
// split with this: a direct FP64 -> BF16 conversion would need round-to-odd fix-ups.
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

// Splits four FP32 values into slice quads (one 8-byte LDS store per slice).  BF16x3:
// each slice is the remainder rounded to nearest (v_cvt_pk_bf16_f32, two values per
// instruction): h = rn(v), m = rn(v - h), l = rn(v - h - m), exact (v = h + m + l) for
// every normal v.  Rounding to nearest keeps the lower slices centred on zero; slices
// cut by truncation carry the sign of v, so the products of level 3 that the kernel
// drops (w1 u2, w2 u1) would bias every step the same way.
template <class Tr>
__device__ __forceinline__ void split4(const float (&v)[4], typename Quad<Tr>::V (&q)[Tr::NSU]) {
  using T = typename Tr::T;
  if constexpr (Tr::NSU == 3) {
    typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
    typedef uint32_t u2v __attribute__((ext_vector_type(2)));
    float r[4] = {v[0], v[1], v[2], v[3]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const bf2 p0 = {(__bf16)r[0], (__bf16)r[1]};
      const bf2 p1 = {(__bf16)r[2], (__bf16)r[3]};
      const uint32_t w0 = __builtin_bit_cast(uint32_t, p0), w1 = __builtin_bit_cast(uint32_t, p1);
      const u2v w = {w0, w1};
      q[k] = __builtin_bit_cast(typename Quad<Tr>::V, w);
      if (k < 2) {  // remainders for the next slice (exact)
        r[0] -= __uint_as_float(w0 << 16);
        r[1] -= __uint_as_float(w0 & 0xffff0000u);
        r[2] -= __uint_as_float(w1 << 16);
        r[3] -= __uint_as_float(w1 & 0xffff0000u);
      }
    }
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      T s[Tr::NSU];
      split_f32<Tr>(v[e], s);
#pragma unroll
      for (int k = 0; k < Tr::NSU; ++k) q[k][e] = s[k];
    }
  }
}

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
      split_f32<typename C::Tr>((float)v[j], s);
#pragma unroll
      for (int k = 0; k < C::NSU; ++k) lds[in_plane ? k * C::PLANE + idx : C::SPARE] = s[k];
    }
  }
}

// A tile inside the domain interior: each thread takes four adjacent storage columns
// of a row (QPR = PX / 4 threads per row, RPP = NT / QPR rows per pass) and stores each
// slice of the four values with one 8-byte write; pad columns get 0.  The threads past
// RPP * QPR repeat the first row of the next pass (same cells, same values); in the last
// pass, threads past the tile's rows load valid cells and store to the spare cells.
template <class C>
__device__ __forceinline__ void load_tile_split_full(const double *__restrict__ in, const SplitCtx &t,
                                                     typename C::T *__restrict__ lds) {
  using Q = typename Quad<typename C::Tr>::V;
  constexpr int NT = C::NW * 64, QPR = C::PX / 4, RPP = NT / QPR;
  static_assert(C::PX % 4 == 0 && RPP >= 1, "quads tile the rows");
  constexpr int NP = (C::EY + RPP - 1) / RPP;
  const int r0 = t.tid / QPR, c4 = 4 * (t.tid - r0 * QPR);
  const bool lastok = r0 + RPP * (NP - 1) < C::EY;
  uint32_t voff[4];
  bool colok[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int cc = c4 + e - C::CS;
    colok[e] = cc >= 0 && cc < C::EX;
    voff[e] = (uint32_t)((r0 * t.pitch + (colok[e] ? cc : 0)) * 8);
  }
  const char *gb = reinterpret_cast<const char *>(in + t.ay0 * t.pitch + t.ax0);
  const int64_t pstep = (int64_t)RPP * t.pitch * 8;
  double v[NP][4];
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const bool ok = i < NP - 1 || lastok;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double x = *reinterpret_cast<const double *>(gb + i * pstep + (ok ? voff[e] : 0u));
      v[i][e] = colok[e] ? x : 0.0;
    }
  }
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const bool ok = i < NP - 1 || lastok;
    Q q[C::NSU];
    const float f[4] = {(float)v[i][0], (float)v[i][1], (float)v[i][2], (float)v[i][3]};
    split4<typename C::Tr>(f, q);
    const int idx = (r0 + RPP * i) * C::PX + c4;
#pragma unroll
    for (int k = 0; k < C::NSU; ++k) *reinterpret_cast<Q *>(lds + (ok ? k * C::PLANE + idx : C::SPARE)) = q[k];
  }
}

// The tile's FP64 loads, all issued into v[] (used by the persistent kernel to keep the
// next tile's loads in flight), and their split into LDS slices.
template <class C>
struct SplitRegs {
  static constexpr int NT = C::NW * 64;
  static constexpr int NIT = (C::PLANE + NT - 1) / NT;
  static_assert(NIT <= 64, "one validity bit per load");
  double v[NIT];
  uint64_t ok;  // bit i: v[i] is a cell of the array (else it is stored as 0)
};

template <class C>
__device__ __forceinline__ void issue_tile_split(const double *__restrict__ in, const Geo &g, const SplitCtx &t,
                                                 SplitRegs<C> &rg) {
  constexpr int NT = SplitRegs<C>::NT, NIT = SplitRegs<C>::NIT;
  const bool all_in = t.ay0 >= 0 && t.ay0 + C::EY <= g.ext[1] && t.ax0 >= 0 && t.ax0 + C::EX <= t.pitch;
  const double *base = in + t.ay0 * t.pitch + t.ax0;
  // the loaded values are not touched until store_tile_split (a select here would wait)
  rg.ok = 0;
#pragma unroll
  for (int i = 0; i < NIT; ++i) {
    const int idx = t.tid + i * NT;
    const int r = idx / C::PX, c = idx - r * C::PX - C::CS;
    const int64_t ay = t.ay0 + r, ax = t.ax0 + c;
    const bool ok = idx < C::PLANE && c >= 0 && c < C::EX &&
                    (all_in || (ay >= 0 && ay < g.ext[1] && ax >= 0 && ax < t.pitch));
    rg.v[i] = *(ok ? base + ((int64_t)r * t.pitch + c) : in);
    rg.ok |= (uint64_t)ok << i;
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
    split_f32<typename C::Tr>((rg.ok >> i) & 1 ? (float)rg.v[i] : 0.0f, sl);
#pragma unroll
    for (int k = 0; k < C::NSU; ++k) lds[in_plane ? k * C::PLANE + idx : C::SPARE] = sl[k];
  }
}

template <class C>
__device__ __forceinline__ void split_tile_setup(const Geo &g, int64_t tile, SplitCtx &t, double *out) {
  t.pitch = g.ext[2];
  t.ny = g.n[1];
  t.nx = g.n[2];
  t.Y0 = (tile / g.tiles[2]) * C::BY - C::KF * C::R;
  t.X0 = (tile % g.tiles[2]) * C::BX - C::KF * C::R;
  t.ay0 = g.org[1] + t.Y0;
  t.ax0 = g.org[2] + t.X0;
  t.full = t.Y0 >= 0 && t.X0 >= 0 && t.Y0 + C::EY <= t.ny && t.X0 + C::EX <= t.nx;
  t.gout = out + t.ay0 * t.pitch + t.ax0;
  const int64_t ry0 = t.Y0 < 0 ? -t.Y0 : 0, cx0 = t.X0 < 0 ? -t.X0 : 0;
  const int64_t ry1 = t.ny - t.Y0 < C::EY ? t.ny - t.Y0 : C::EY, cx1 = t.nx - t.X0 < C::EX ? t.nx - t.X0 : C::EX;
  t.ry0 = (int)ry0;
  t.cx0 = (int)cx0;
  t.nry = ry1 > ry0 ? (int)(ry1 - ry0) : 0;
  t.ncx = cx1 > cx0 ? (int)(cx1 - cx0) : 0;
}

// Per-lane constants: A fragments (the weights' slices), the byte offsets of the
// representative B reads, the lane's output cells.
template <class C>
__device__ __forceinline__ void split_lane_setup(const typename C::T *__restrict__ bfrag, int tid, SplitCtx &t,
                                                 typename C::Tr::V (*w)[C::NINS], int *off) {
  using V = typename C::Tr::V;
  using T = typename C::T;
  t.tid = tid;
  t.lane = tid & 63;
  t.wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = t.lane, grp = lane >> 4, n = lane & 15;
  // D[m][n], m = 4 grp + v: cell (2 arow + (grp >> 1), 8 aoct + 4 (grp & 1) + v) of the
  // block, where (arow, aoct) is the anchor of column n (cs_ktable.h, split_arow / split_aoct)
  const int arow = split_arow(C::CT.map, n), aoct = split_aoct(C::CT.map, n);
  t.orow = 2 * arow + (grp >> 1);
  t.ocol = 8 * aoct + 4 * (grp & 1);
  t.vw = (t.orow * C::PX + t.ocol + C::CS) * (int)sizeof(T);
#pragma unroll
  for (int q = 0; q < C::NSW; ++q)
#pragma unroll
    for (int i = 0; i < C::NINS; ++i)
      w[q][i] = *reinterpret_cast<const V *>(bfrag + ((size_t)(q * C::NINS + i) * 64 + lane) * 8);
  // B reads: chunk 4 i + grp of instruction i, anchor u_n = (2 arow, 8 aoct)
#pragma unroll
  for (int i = 0; i < C::NINS; ++i) {
    if (C::RPL.rep[i] != i) continue;
    off[i] = ((2 * arow + C::CT.ty[4 * i + grp]) * C::PX + 8 * aoct + C::CT.tx[4 * i + grp] + C::CS) *
             (int)sizeof(T);
  }
}

// Block origins of every fused step, per wave, as for the FP64 kernels (k2d_f64_body.h,
// Org4): o = LDS byte offset of the block's first cell, yx = (y << 16) | x.
template <class C>
struct OrgS {
  static constexpr int base(int s) {
    int n = 0;
    for (int k = 0; k < s; ++k) n += C::PERS(k);
    return n;
  }
  static constexpr int TOT = base(C::KF);
  int o[C::NW][TOT];
  int yx[C::NW][TOT];
};

template <class C>
constexpr OrgS<C> make_orgs() {
  OrgS<C> t{};
  for (int w = 0; w < C::NW; ++w)
    for (int s = 0; s < C::KF; ++s)
      for (int j = 0; j < C::PERS(s); ++j) {
        const int blk = j * C::NW + w < C::NB(s) ? j * C::NW + w : C::NB(s) - 1;
        int y = 0, x = 0;
        C::origin(s, blk, y, x);
        t.o[w][OrgS<C>::base(s) + j] = (y * C::PX + x) * (int)sizeof(typename C::T);
        t.yx[w][OrgS<C>::base(s) + j] = (y << 16) | x;
      }
  return t;
}

template <class C>
struct OrgSTab {
  static constexpr OrgS<C> t = make_orgs<C>();
};

// One block: D0 = level 0 (hi x hi), D1 = the kept lower levels, scaled by SCALE.
// p = LDS address of the block origin (slice 0).
template <class C>
__device__ __forceinline__ void block_split(const char *p, const int *off, const typename C::Tr::V (*w)[C::NINS],
                                            f4 &d0, f4 &d1) {
  using Tr = typename C::Tr;
  using V = typename Tr::V;
  constexpr RdPlan P = C::RPL;
  constexpr int SB = C::PLANE * (int)sizeof(typename C::T);  // bytes between slice planes
  d0 = f4{0.f, 0.f, 0.f, 0.f};
  d1 = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < C::NINS; ++i) {
    V u[C::NSU];
#pragma unroll
    for (int s = 0; s < C::NSU; ++s) u[s] = *reinterpret_cast<const V *>(p + off[P.rep[i]] + P.delta[i] + s * SB);
    d0 = Tr::mma(w[0][i], u[0], d0);
#pragma unroll
    for (int s = 0; s < C::NSU; ++s)
#pragma unroll
      for (int q = 0; q < C::NSW; ++q)
        if (keep_product<Tr>(s, q)) d1 = Tr::mma(w[q][i], u[s], d1);
  }
}

__device__ __forceinline__ bool split_in_interior(const SplitCtx &t, int ty, int tx) {
  return (unsigned)(ty - t.ry0) < (unsigned)t.nry && (unsigned)(tx - t.cx0) < (unsigned)t.ncx;
}

// Fused step S.  MASK = false (a tile inside the domain interior): intermediate results
// are stored without tests -- cells a block computes beyond the step's region are never
// read by a later step's region -- and the last step writes only the region's columns.
// MASK = true: cells outside the domain interior keep their value.
template <class C, int S, bool MASK>
__device__ __forceinline__ void split_step(const SplitCtx &t, const typename C::Tr::V (*w)[C::NINS], const int *off,
                                           char *lds) {
  using Tr = typename C::Tr;
  using T = typename C::T;
  using Q = typename Quad<Tr>::V;
  constexpr bool LAST = S == C::KF - 1;
  constexpr int PERS = C::PERS(S), B0 = OrgS<C>::base(S), NW = C::NW;
  constexpr int SB = C::PLANE * (int)sizeof(T);
  constexpr float INV = 1.0f / (float)Tr::SCALE;
  const int *ot = OrgSTab<C>::t.o[t.wave] + B0;
  const int *yt = OrgSTab<C>::t.yx[t.wave] + B0;
  f4 d0[PERS], d1[PERS];
#pragma unroll
  for (int j = 0; j < PERS; ++j) block_split<C>(lds + ot[j], off, w, d0[j], d1[j]);
  if constexpr (!LAST || C::OS == 1) lds_barrier();  // every wave has read this step's input
#pragma unroll
  for (int j = 0; j < PERS; ++j) {
    if constexpr (!LAST) {
      Q q[C::NSU];
      const float f[4] = {d0[j][0] + d1[j][0] * INV, d0[j][1] + d1[j][1] * INV, d0[j][2] + d1[j][2] * INV,
                          d0[j][3] + d1[j][3] * INV};
      split4<Tr>(f, q);
      char *dst = lds + ot[j] + t.vw;
      if constexpr (!MASK) {
#pragma unroll
        for (int k = 0; k < C::NSU; ++k) *reinterpret_cast<Q *>(dst + k * SB) = q[k];
      } else {
        const int yx = yt[j], ty = (yx >> 16) + t.orow, tx0 = (yx & 0xffff) + t.ocol;
        const bool row_in = (unsigned)(ty - t.ry0) < (unsigned)t.nry;
        if (row_in && (unsigned)(tx0 - t.cx0) < (unsigned)t.ncx && (unsigned)(tx0 + 3 - t.cx0) < (unsigned)t.ncx) {
#pragma unroll
          for (int k = 0; k < C::NSU; ++k) *reinterpret_cast<Q *>(dst + k * SB) = q[k];
        } else if (row_in) {  // a block at the domain edge: keep the halo cells
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if ((unsigned)(tx0 + v - t.cx0) < (unsigned)t.ncx)
#pragma unroll
              for (int k = 0; k < C::NSU; ++k) reinterpret_cast<T *>(dst + k * SB)[v] = q[k][v];
        }
      }
    } else {
      const int yx = yt[j], y = yx >> 16, x = yx & 0xffff, ty = y + t.orow, tx0 = x + t.ocol;
      if constexpr (C::OS == 1) {
        // staged: the output block cell (ty - LO, tx - LO) of a compact BY x BX FP64 array
        double *ob = reinterpret_cast<double *>(lds) + (ty - C::LO(S)) * C::BX + tx0 - C::LO(S);
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if ((unsigned)(tx0 + v - C::LO(S)) < (unsigned)(C::HI(S) - C::LO(S)))
            ob[v] = (double)d0[j][v] + (double)d1[j][v] * (double)INV;
      } else {
        double *o = t.gout + (int64_t)ty * t.pitch + tx0;
        const bool row_in = !MASK || (unsigned)(ty - t.ry0) < (unsigned)t.nry;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          bool in = row_in && (unsigned)(tx0 + v - C::LO(S)) < (unsigned)(C::HI(S) - C::LO(S));
          if constexpr (MASK) in = in && (unsigned)(tx0 + v - t.cx0) < (unsigned)t.ncx;
          if (in) o[v] = (double)d0[j][v] + (double)d1[j][v] * (double)INV;
        }
      }
    }
  }
  if constexpr (!LAST) lds_barrier();  // this step's output is complete before it is read
  if constexpr (LAST && C::OS == 1) {
    lds_barrier();  // the staged block is complete
    constexpr int NT = NW * 64, RP = NT / C::BX, OFF = C::KF * C::R;
    static_assert(NT % C::BX == 0, "a pass covers whole rows");
    const double *ob = reinterpret_cast<const double *>(lds);
    const int r0 = t.tid / C::BX, c = t.tid - r0 * C::BX;
    const bool col_in = !MASK || (unsigned)(OFF + c - t.cx0) < (unsigned)t.ncx;
    double *o = t.gout + (int64_t)(OFF + r0) * t.pitch + OFF + c;
#pragma unroll
    for (int i = 0; i < C::BY / RP; ++i)
      if (!MASK || (col_in && (unsigned)(OFF + r0 + RP * i - t.ry0) < (unsigned)t.nry))
        o[(int64_t)RP * i * t.pitch] = ob[(r0 + RP * i) * C::BX + c];
  }
}

template <class C, int S, bool MASK>
__device__ __forceinline__ void split_steps(const SplitCtx &t, const typename C::Tr::V (*w)[C::NINS], const int *off,
                                            char *lds) {
  if constexpr (S < C::KF) {
    split_step<C, S, MASK>(t, w, off, lds);
    split_steps<C, S + 1, MASK>(t, w, off, lds);
  }
}

// The fused steps of a tile already in LDS (the persistent kernels' tile loop).
template <class C>
__device__ __forceinline__ void run_split_steps(const SplitCtx &t, const typename C::Tr::V (*w)[C::NINS],
                                                const int *off, typename C::T *lds) {
  if (t.full)
    split_steps<C, 0, false>(t, w, off, reinterpret_cast<char *>(lds));
  else
    split_steps<C, 0, true>(t, w, off, reinterpret_cast<char *>(lds));
}

// One workgroup advances tile `tile` of grid g by KF steps (smem: the kernel's dynamic
// LDS; bfrag: the weights' slices as A fragments; LG: tile loads in flight per thread
// for tiles at the domain edge; tid: threadIdx.x).
template <class Tr, int R, int SH, int KF, int BY, int BX, int NW, int LG = 64, int OS = 0>
__device__ __forceinline__ void tile_split2(const double *__restrict__ in, double *__restrict__ out,
                                            const typename Tr::T *__restrict__ bfrag, const Geo &g, int64_t tile,
                                            unsigned char *__restrict__ smem, int tid) {
  using C = CfgS2<Tr, R, SH, KF, BY, BX, NW, OS>;
  using T = typename Tr::T;
  using V = typename Tr::V;
  T *lds = reinterpret_cast<T *>(smem);
  SplitCtx t;
  V w[C::NSW][C::NINS];
  int off[C::NINS];
  split_lane_setup<C>(bfrag, tid, t, w, off);
  split_tile_setup<C>(g, tile, t, out);
  if (t.full) {
    load_tile_split_full<C>(in, t, lds);
    lds_barrier();
    split_steps<C, 0, false>(t, w, off, reinterpret_cast<char *>(smem));
  } else {
    load_tile_split<C, LG>(in, g, t, lds);
    lds_barrier();
    split_steps<C, 0, true>(t, w, off, reinterpret_cast<char *>(smem));
  }
}

}  // namespace cs
