// k2d_f64_body.h -- plain FP64 2-D ConvStencil tile body for gfx950 (FP64 MFMA).
//
// Replaces kernel2d (reference src/2d/gpu.cu:30-70) and its launchers
// gpu_box_2d1r / gpu_box_2d3r (src/2d/gpu.cu:245-425).  What changes for CDNA4:
//   * the A operand is the input tile itself, read straight from LDS (see
//     cs_ktable.h), so the reference's lookup-table scatter into two stencil2row
//     copies is gone and each input element is stored in LDS once;
//   * B is a banded Toeplitz slice of the weights, held in registers for the
//     whole kernel, one f64 per lane per k-step;
//   * two MFMA shapes: v_mfma_f64_4x4x4 on 8 x 8 output blocks with the dilated
//     table (16-byte LDS reads, fewest wasted flops; the default) and
//     v_mfma_f64_16x16x4 on 16 x 16 blocks (for comparison);
//   * the time loop is fused inside the kernel: a workgroup loads a tile with a
//     halo of KF*R cells, advances it KF steps in LDS, keeping the domain's halo
//     cells fixed after every step, and writes only the final BY x BX block.  This
//     equals KF separate steps exactly in exact arithmetic -- unlike the
//     reference's pre-composed K^3 weights, which differ from stepping in the outer
//     (KF-1)*R ring of the domain;
//   * every fused step is its own compile-time instance (region, block grid and
//     blocks per wave are constants), and each wave computes all of its blocks of a
//     step into registers before any result is stored: the blocks' LDS reads and
//     MFMA chains are independent, so they overlap without branches between them;
//   * two staging modes: ST = 0 ping-pongs between two LDS buffers; ST = 1 keeps one
//     buffer and writes the registers back between two barriers (half the LDS, so
//     more workgroups share a CU and one's HBM loads overlap the others' MFMAs);
//   * workgroups are mapped to tiles so that each XCD (own L2) sweeps a contiguous
//     band of tile rows, keeping vertical halo re-reads inside one L2.
#pragma once
#include "cs_internal.h"

namespace cs {

constexpr int kAligned = 99;  // Cfg2D pitch override: rows aligned to 128 bytes

template <int MF_, int R_, int SH_, int KF_, int BY_, int BX_, int NW_, int ST_, int PXM_ = -1>
struct Cfg2D {
  static constexpr int MF = MF_, R = R_, SH = SH_, KF = KF_, BY = BY_, BX = BX_, NW = NW_, ST = ST_;
  static constexpr int EY = BY + 2 * KF * R;
  static constexpr int EX = BX + 2 * KF * R;
  static constexpr int BH = MF == 16 ? 16 : 8;  // output block rows
  static constexpr int BW = MF == 16 ? 16 : 8;  // output block columns
  static constexpr KTable KT = make_ktable(R, SH);  // 16x16x4
  static constexpr DTable DT = make_dtable(R, SH);  // 4x4x4
  static constexpr int NKS = MF == 16 ? KT.n : dtable_ksteps(DT);
  static constexpr int NOFF = MF == 16 ? KT.n : DT.np + DT.ns;  // LDS reads per block
  // 4x4x4: tile column c is stored at column c + CS; every block starts at a column x
  // with x + CS + pp even, so that all 16-byte pair reads are aligned.  The first block
  // of a step may start one column left of the region.
  static constexpr int CS = MF == 16 ? 0 : ((1 - R - DT.cmin) > 0 ? (1 - R - DT.cmin) : 0);
  static constexpr int PP = MF == 16 ? 0 : DT.pp;

  // region of fused step s: rows [LO, LO + H), columns [LO, HI)
  static constexpr int H(int s) { return BY + 2 * R * (KF - 1 - s); }
  static constexpr int W(int s) { return BX + 2 * R * (KF - 1 - s); }
  static constexpr int LO(int s) { return R * (s + 1); }
  static constexpr int HI(int s) { return LO(s) + W(s); }
  static constexpr int LOX(int s) { return MF == 16 ? LO(s) : LO(s) - ((LO(s) + CS + PP) & 1); }
  static constexpr int LASTX(int s) {
    return MF == 16 ? HI(s) - 16 : (HI(s) - 8) + (((HI(s) - 8) + CS + PP) & 1);
  }
  static constexpr int NBY(int s) { return (H(s) + BH - 1) / BH; }
  static constexpr int NBX(int s) { return (LASTX(s) - LOX(s) + BW - 1) / BW + 1; }
  static constexpr int NB(int s) { return NBY(s) * NBX(s); }
  static constexpr int PERS(int s) { return (NB(s) + NW - 1) / NW; }

  // columns and rows read relative to a block's (y, x)
  static constexpr int RD_C0 = MF == 16 ? 0 : DT.cmin;
  static constexpr int RD_C1 = MF == 16 ? 15 + ktable_overrun(KT) : 6 + DT.cmax;
  static constexpr int RD_R0 = MF == 16 ? -R : DT.tmin;
  static constexpr int RD_R1 = MF == 16 ? 15 + R : 6 + DT.tmax;
  static constexpr int max_col() {
    int m = 0;
    for (int s = 0; s < KF; ++s) m = LASTX(s) + RD_C1 > m ? LASTX(s) + RD_C1 : m;
    return m;
  }
  static constexpr int min_col() {
    int m = 1 << 20;
    for (int s = 0; s < KF; ++s) m = LOX(s) + RD_C0 < m ? LOX(s) + RD_C0 : m;
    return m;
  }
  static constexpr int PX0 = CS + (max_col() + 1 > EX ? max_col() + 1 : EX);
  // ST: 0 = two LDS buffers, 1 = one buffer + registers, 2 = as 1 with the last step's
  // results staged in LDS and written out row by row (coalesced)
  static constexpr int NBUF = (KF > 1 && ST == 0) ? 2 : 1;
  // 16x16x4: odd pitch; 4x4x4: the even pitch with the fewest LDS bank conflicts for
  // this table's reads and the result writes that keeps the most workgroups per CU
  // (cs_ktable.h, lds_cost); PXM_ >= 0 forces the smallest pitch >= PX0 that is PXM_ mod 16,
  // PXM_ = kAligned the smallest multiple of 16 (rows aligned to 128 bytes, no bank-aware
  // choice: the --breakdown ladder's rung 3)
  static constexpr int PX = MF == 16 ? (PX0 | 1)
                            : PXM_ == kAligned ? (PX0 + 15) / 16 * 16
                            : PXM_ >= 0 ? PX0 + ((PXM_ - PX0 % 16) + 16) % 16
                                        : best_pitch_dtables(&DT, 1, PX0, (long)NBUF * EY * 8, 16);
  static constexpr int BUF = EY * PX;
  static constexpr int SPARE = NBUF * BUF;  // one scratch cell after the buffers
  static constexpr size_t LDS = (size_t)(NBUF * BUF + 2) * sizeof(double);
  // 4x4x4: the reads' shared vector addresses and immediate offsets at this pitch
  static constexpr RdPlan RPL = make_rdplan(DT, PX);
  // block `blk` of step s -> its first output cell (y, x) in tile coordinates; the last
  // block row and column are moved back to end at the region's edge (they overlap)
  static constexpr void origin(int s, int blk, int &y, int &x) {
    const int ba = blk / NBX(s), bb = blk - ba * NBX(s);
    y = LO(s) + (BH * ba < H(s) - BH ? BH * ba : H(s) - BH);
    x = LOX(s) + BW * bb < LASTX(s) ? LOX(s) + BW * bb : LASTX(s);
  }
  static_assert(BY % 16 == 0 && BX % 16 == 0, "tile must be a multiple of 16");
  static_assert(min_col() + CS >= 0, "block reads left of the LDS row");
  static_assert(MF == 16 || (PX % 2 == 0 && (CS + PP + LOX(0)) % 2 == 0), "pair reads must be aligned");
  static_assert(LO(0) + RD_R0 >= 0 && LO(0) + H(0) - BH + RD_R1 < EY, "block reads outside the tile rows");
  static_assert(LDS <= 160 * 1024, "tile does not fit gfx950 LDS");
};

__device__ __forceinline__ int64_t xcd_tile(int64_t wg, int64_t ntiles) {
  // Workgroups are dispatched round-robin over the 8 XCDs; give XCD x the
  // contiguous tile range [x * per, (x + 1) * per).
  const int64_t per = (ntiles + 7) / 8;
  return (wg % 8) * per + wg / 8;
}

// MFMA result of one block: 4 f64 per lane (16x16x4) or 1 (4x4x4).
template <int MF>
struct Acc;
template <>
struct Acc<16> {
  d4 v;
};
template <>
struct Acc<4> {
  double v;
};

typedef double dbl2 __attribute__((ext_vector_type(2)));

// Per-workgroup context of one tile.
struct TileCtx {
  int tid, lane, wave;
  int orow, ocol;  // 4x4x4: the output cell of this lane relative to the block
  int full;        // 1 if the whole tile (with halo) lies inside the domain interior
  int64_t Y0, X0, ny, nx, ay0, ax0, pitch;
  // 4x4x4 only
  int vw;              // LDS byte offset of this lane's result cell from the block origin
  uint32_t vgo;        // global byte offset of this lane's result cell from the block origin
  double *gout;        // output cell of tile cell (0, 0)
  int ry0, nry, cx0, ncx;  // tile rows [ry0, ry0 + nry) and columns [cx0, cx0 + ncx) are interior
};

// One block's MFMA chain: A from LDS at p + off[], B from registers.
template <class C>
__device__ __forceinline__ Acc<C::MF> block_mfma(const double *p, const int *off, const double *b) {
  Acc<C::MF> a;
  if constexpr (C::MF == 16) {
    a.v = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < C::NKS; ++ks) a.v = __builtin_amdgcn_mfma_f64_16x16x4f64(p[off[ks]], b[ks], a.v, 0, 0, 0);
  } else {
    a.v = 0.0;
#pragma unroll
    for (int ps = 0; ps < C::DT.np; ++ps) {  // one 16-byte read feeds two k-steps
      const dbl2 pr = *reinterpret_cast<const dbl2 *>(p + off[ps]);
      a.v = __builtin_amdgcn_mfma_f64_4x4x4f64(pr.x, b[2 * ps], a.v, 0, 0, 0);
      a.v = __builtin_amdgcn_mfma_f64_4x4x4f64(pr.y, b[2 * ps + 1], a.v, 0, 0, 0);
    }
#pragma unroll
    for (int ss = 0; ss < C::DT.ns; ++ss)
      a.v = __builtin_amdgcn_mfma_f64_4x4x4f64(p[off[C::DT.np + ss]], b[2 * C::DT.np + ss], a.v, 0, 0, 0);
  }
  return a;
}

// Block `blk` of step S -> its first output cell (y, x) in tile coordinates.
template <class C, int S>
__device__ __forceinline__ void block_origin(int blk, int &y, int &x) {
  constexpr int NBX = C::NBX(S);
  const int ba = blk / NBX, bb = blk - ba * NBX;
  y = C::LO(S) + min(C::BH * ba, C::H(S) - C::BH);
  x = min(C::LOX(S) + C::BW * bb, C::LASTX(S));
}

// Stores a block's results to LDS (tile coordinates) or, in the last step, to the
// output grid.  Cells outside the domain interior are skipped (they keep their halo
// value), and so are columns outside the step's region [lo, hi).
template <class C, bool GLOBAL>
__device__ __forceinline__ void put_block(const Acc<C::MF> &a, int y, int x, int lo, int hi, const TileCtx &t,
                                          double *lds_dst, double *out) {
  if constexpr (C::MF == 16) {
    const int tx = x + (t.lane & 15);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int ty = y + 4 * v + (t.lane >> 4);
      if (t.full || (t.Y0 + ty >= 0 && t.Y0 + ty < t.ny && t.X0 + tx >= 0 && t.X0 + tx < t.nx)) {
        if constexpr (GLOBAL)
          out[(t.ay0 + ty) * t.pitch + t.ax0 + tx] = a.v[v];
        else
          lds_dst[ty * C::PX + tx] = a.v[v];
      }
    }
  } else {
    const int ty = y + t.orow, tx = x + t.ocol;
    if (tx >= lo && tx < hi && (t.full || (t.Y0 + ty >= 0 && t.Y0 + ty < t.ny && t.X0 + tx >= 0 && t.X0 + tx < t.nx))) {
      if constexpr (GLOBAL)
        out[(t.ay0 + ty) * t.pitch + t.ax0 + tx] = a.v;
      else
        lds_dst[ty * C::PX + tx + C::CS] = a.v;
    }
  }
}

// Fused step S: every wave computes its blocks (wave, wave + NW, ...) into registers,
// then stores them.  In the last round a wave without a block of its own recomputes
// the step's last block and stores nothing.
template <class C, int S>
__device__ __forceinline__ void fused_step(const TileCtx &t, const double *b, const int *off, double *lds,
                                           double *out) {
  constexpr bool LAST = S == C::KF - 1;
  constexpr int NB = C::NB(S), PERS = C::PERS(S), NW = C::NW;
  const double *src = lds + (C::NBUF == 2 ? (S & 1) * C::BUF : 0);
  double *dst = lds + (C::NBUF == 2 ? ((S + 1) & 1) * C::BUF : 0);
  Acc<C::MF> acc[PERS];
#pragma unroll
  for (int j = 0; j < PERS; ++j) {
    int y, x;
    block_origin<C, S>(min(j * NW + t.wave, NB - 1), y, x);
    acc[j] = block_mfma<C>(src + y * C::PX + x, off, b);
  }
  constexpr bool STAGE = LAST && C::ST == 2;  // last step through LDS, then coalesced rows
  if constexpr ((!LAST || STAGE) && C::NBUF == 1) lds_barrier();  // every wave has read this step's input
#pragma unroll
  for (int j = 0; j < PERS; ++j) {
    const int blk = j * NW + t.wave;
    if ((NB % NW == 0 || j < PERS - 1) || blk < NB) {
      int y, x;
      block_origin<C, S>(blk, y, x);
      put_block<C, LAST && !STAGE>(acc[j], y, x, C::LO(S), C::HI(S), t, dst, out);
    }
  }
  if constexpr (!LAST || STAGE) lds_barrier();  // this step's output is complete before it is read
  if constexpr (STAGE) {
    // the BY x BX output block, NT / BX rows per pass, one row segment per BX lanes;
    // cells outside the domain interior are skipped (they were not written)
    constexpr int NT = NW * 64, OFF = C::KF * C::R, RP = NT / C::BX;
    static_assert(NT % C::BX == 0, "a pass covers whole rows");
    const int r0 = t.tid / C::BX, c = t.tid - r0 * C::BX;
    const bool col_in = t.full || t.X0 + OFF + c < t.nx;
    double *o = out + (t.ay0 + OFF + r0) * t.pitch + t.ax0 + OFF + c;
    const double *l = dst + (OFF + r0) * C::PX + OFF + c + C::CS;
#pragma unroll
    for (int i = 0; i < C::BY / RP; ++i) {
      if (col_in && (t.full || t.Y0 + OFF + r0 + RP * i < t.ny)) o[(int64_t)RP * i * t.pitch] = l[RP * i * C::PX];
    }
  }
}

template <class C, int S>
__device__ __forceinline__ void fused_steps(const TileCtx &t, const double *b, const int *off, double *lds,
                                            double *out) {
  if constexpr (S < C::KF) {
    fused_step<C, S>(t, b, off, lds, out);
    fused_steps<C, S + 1>(t, b, off, lds, out);
  }
}

// ---- 4x4x4 path ----
//
// A wave's block origins in every fused step are compile-time functions of its wave
// index, so each kernel tabulates them: o = LDS byte offset of the block's first cell,
// yx = (y << 16) | x in tile coordinates.  The table sits in constant memory and is read
// with scalar loads (eight blocks per load); a block then costs one vector add per read
// class (cs_ktable.h, RdPlan) and one for its result, with no branches.  In the last
// round of a step a wave without a block of its own recomputes the step's last block
// and stores the same values again.
template <class C>
struct Org4 {
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
constexpr Org4<C> make_org4() {
  Org4<C> t{};
  for (int w = 0; w < C::NW; ++w)
    for (int s = 0; s < C::KF; ++s)
      for (int j = 0; j < C::PERS(s); ++j) {
        const int blk = j * C::NW + w < C::NB(s) ? j * C::NW + w : C::NB(s) - 1;
        int y = 0, x = 0;
        C::origin(s, blk, y, x);
        t.o[w][Org4<C>::base(s) + j] = (y * C::PX + x) * 8;
        t.yx[w][Org4<C>::base(s) + j] = (y << 16) | x;
      }
  return t;
}

template <class C>
struct Org4Tab {
  static constexpr Org4<C> t = make_org4<C>();
};

// One 8 x 8 block: p = LDS address of its origin; off = the lane's read offsets.
template <class C>
__device__ __forceinline__ double block4(const char *p, const int *off, const double *b) {
  constexpr RdPlan P = C::RPL;
  double a = 0.0;
#pragma unroll
  for (int ps = 0; ps < C::DT.np; ++ps) {  // one 16-byte read feeds two k-steps
    const dbl2 pr = *reinterpret_cast<const dbl2 *>(p + off[P.rep[ps]] + P.delta[ps]);
    a = __builtin_amdgcn_mfma_f64_4x4x4f64(pr.x, b[2 * ps], a, 0, 0, 0);
    a = __builtin_amdgcn_mfma_f64_4x4x4f64(pr.y, b[2 * ps + 1], a, 0, 0, 0);
  }
#pragma unroll
  for (int ss = 0; ss < C::DT.ns; ++ss) {
    const int i = C::DT.np + ss;
    const double v = *reinterpret_cast<const double *>(p + off[P.rep[i]] + P.delta[i]);
    a = __builtin_amdgcn_mfma_f64_4x4x4f64(v, b[2 * C::DT.np + ss], a, 0, 0, 0);
  }
  return a;
}

__device__ __forceinline__ bool in_interior(const TileCtx &t, int ty, int tx) {
  return (unsigned)(ty - t.ry0) < (unsigned)t.nry && (unsigned)(tx - t.cx0) < (unsigned)t.ncx;
}

// ST = 2: the BY x BX output block from LDS to the grid, NT / BX rows per pass, one row
// segment per BX lanes; with MASK, cells outside the domain interior are skipped.
template <class C, bool MASK>
__device__ __forceinline__ void copy_out4(const TileCtx &t, const char *buf) {
  constexpr int NT = C::NW * 64, OFF = C::KF * C::R, RP = NT / C::BX;
  static_assert(NT % C::BX == 0 && C::BY % RP == 0, "passes cover whole rows and the whole block");
  const int r0 = t.tid / C::BX, c = t.tid - r0 * C::BX;
  const bool col_in = !MASK || (unsigned)(OFF + c - t.cx0) < (unsigned)t.ncx;
  double *o = t.gout + (int64_t)(OFF + r0) * t.pitch + OFF + c;
  const double *l = reinterpret_cast<const double *>(buf) + (OFF + r0) * C::PX + OFF + c + C::CS;
#pragma unroll
  for (int i = 0; i < C::BY / RP; ++i) {
    if (!MASK || (col_in && (unsigned)(OFF + r0 + RP * i - t.ry0) < (unsigned)t.nry))
      o[(int64_t)RP * i * t.pitch] = l[RP * i * C::PX];
  }
}

// Fused step S.  MASK = false (a tile inside the domain interior): results are stored
// without tests -- cells a block computes beyond the step's region are never read by
// a later step's region -- except in the last step, which writes only the region's
// columns to the grid.  MASK = true: cells outside the domain interior keep their value.
template <class C, int S, bool MASK>
__device__ __forceinline__ void step4(const TileCtx &t, const double *b, const int *off, char *lds) {
  constexpr bool LAST = S == C::KF - 1, STAGE = LAST && C::ST == 2, GLOBAL = LAST && !STAGE;
  constexpr int PERS = C::PERS(S), B0 = Org4<C>::base(S);
  const char *src = lds + (C::NBUF == 2 ? (S & 1) * C::BUF * 8 : 0);
  char *dst = lds + (C::NBUF == 2 ? ((S + 1) & 1) * C::BUF * 8 : 0);
  const int *ot = Org4Tab<C>::t.o[t.wave] + B0;
  const int *yt = Org4Tab<C>::t.yx[t.wave] + B0;
  double acc[PERS];
#pragma unroll
  for (int j = 0; j < PERS; ++j) acc[j] = block4<C>(src + ot[j], off, b);
  if constexpr ((!LAST || STAGE) && C::NBUF == 1) lds_barrier();  // every wave has read this step's input
#pragma unroll
  for (int j = 0; j < PERS; ++j) {
    if constexpr (!GLOBAL) {
      bool ok = true;
      if constexpr (MASK) {
        const int yx = yt[j];
        ok = in_interior(t, (yx >> 16) + t.orow, (yx & 0xffff) + t.ocol);
      }
      if (ok) *reinterpret_cast<double *>(dst + ot[j] + t.vw) = acc[j];
    } else {
      const int yx = yt[j], y = yx >> 16, x = yx & 0xffff;
      bool ok = (unsigned)(t.ocol + x - C::LO(S)) < (unsigned)(C::HI(S) - C::LO(S));
      if constexpr (MASK) ok = ok && in_interior(t, y + t.orow, x + t.ocol);
      if (ok) *reinterpret_cast<double *>(reinterpret_cast<char *>(t.gout + (int64_t)y * t.pitch + x) + t.vgo) = acc[j];
    }
  }
  if constexpr (!LAST || STAGE) lds_barrier();  // this step's output is complete before it is read
  if constexpr (STAGE) copy_out4<C, MASK>(t, dst);
}

template <class C, int S, bool MASK>
__device__ __forceinline__ void steps4(const TileCtx &t, const double *b, const int *off, char *lds) {
  if constexpr (S < C::KF) {
    step4<C, S, MASK>(t, b, off, lds);
    steps4<C, S + 1, MASK>(t, b, off, lds);
  }
}

// A tile inside the domain interior into LDS, row by row: RPP = NT / PX rows per pass,
// one thread per storage column (pad columns get 0).  The threads past RPP * PX repeat
// the first row of the next pass (same cells, same values); in the last pass, threads
// past the tile's rows load a valid cell and store it to the spare cell.
template <class C>
__device__ __forceinline__ void load_tile_full(const double *__restrict__ in, const TileCtx &t,
                                               double *__restrict__ lds) {
  constexpr int NT = C::NW * 64, PX = C::PX, RPP = NT / PX;
  static_assert(RPP >= 1, "a pass covers at least one row");
  constexpr int NP = (C::EY + RPP - 1) / RPP;
  const int r0 = t.tid / PX, c = t.tid - r0 * PX, cc = c - C::CS;
  const bool colok = cc >= 0 && cc < C::EX;
  const bool lastok = r0 + RPP * (NP - 1) < C::EY;
  const uint32_t voff = (uint32_t)((r0 * t.pitch + (colok ? cc : 0)) * 8);
  const char *gb = reinterpret_cast<const char *>(in + t.ay0 * t.pitch + t.ax0);
  const int64_t pstep = (int64_t)RPP * t.pitch * 8;
  double v[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const bool ok = i < NP - 1 || lastok;
    const double x = *reinterpret_cast<const double *>(gb + i * pstep + (ok ? voff : 0u));
    v[i] = colok ? x : 0.0;
  }
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const bool ok = i < NP - 1 || lastok;
    const int idx = (r0 + RPP * i) * PX + c;
    lds[ok ? idx : C::SPARE] = v[i];
    if constexpr (C::NBUF == 2) lds[ok ? C::BUF + idx : C::SPARE] = v[i];
  }
}

// Copies the tile with its halo (EY x EX cells from (ay0, ax0)) into LDS, zero for cells
// outside the array and for pad columns.  Each thread issues its loads in groups of LG
// (all of them by default) before the group's first LDS store, so a whole group of the
// tile is in flight at once; a cell that is not read loads the array's first element
// (always valid) and is replaced by 0.
template <class C, int LG>
__device__ __forceinline__ void load_tile(const double *__restrict__ in, const Geo &g, const TileCtx &t,
                                          double *__restrict__ lds) {
  constexpr int NT = C::NW * 64;
  constexpr int NIT = (C::BUF + NT - 1) / NT;
  constexpr int G = LG < NIT ? LG : NIT;
  const int tid = t.tid;
  const bool all_in = t.ay0 >= 0 && t.ay0 + C::EY <= g.ext[1] && t.ax0 >= 0 && t.ax0 + C::EX <= t.pitch;
  const double *base = in + t.ay0 * t.pitch + t.ax0;
#pragma unroll
  for (int i0 = 0; i0 < NIT; i0 += G) {
    double v[G];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const int idx = tid + (i0 + j) * NT;
      const int r = idx / C::PX, c = idx - r * C::PX - C::CS;  // storage column idx - r*PX holds tile column c
      const int64_t ay = t.ay0 + r, ax = t.ax0 + c;
      const bool ok = i0 + j < NIT && idx < C::BUF && c >= 0 && c < C::EX &&
                      (all_in || (ay >= 0 && ay < g.ext[1] && ax >= 0 && ax < t.pitch));
      const double x = *(ok ? base + ((int64_t)r * t.pitch + c) : in);
      v[j] = ok ? x : 0.0;
    }
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const int idx = tid + (i0 + j) * NT;
      // threads past the buffer store to the spare cell after the buffers (unconditional
      // stores keep every load of the group ahead of the first wait)
      const bool in_buf = i0 + j < NIT - 1 || (i0 + j == NIT - 1 && idx < C::BUF);
      lds[in_buf ? idx : C::SPARE] = v[j];
      if constexpr (C::NBUF == 2) lds[in_buf ? C::BUF + idx : C::SPARE] = v[j];
    }
  }
}

// The tile's loads, all issued into v[] (one HBM round trip), and their LDS stores:
// used by the persistent kernel to keep the next tile's loads in flight while the
// current tile is computed.
template <class C>
struct TileRegs {
  static constexpr int NT = C::NW * 64;
  static constexpr int NIT = (C::BUF + NT - 1) / NT;
  static_assert(NIT <= 64, "one validity bit per load");
  double v[NIT];
  uint64_t ok;  // bit i: v[i] is a cell of the array (else it is stored as 0)
};

template <class C>
__device__ __forceinline__ void issue_tile(const double *__restrict__ in, const Geo &g, const TileCtx &t,
                                           TileRegs<C> &rg) {
  constexpr int NT = TileRegs<C>::NT, NIT = TileRegs<C>::NIT;
  const bool all_in = t.ay0 >= 0 && t.ay0 + C::EY <= g.ext[1] && t.ax0 >= 0 && t.ax0 + C::EX <= t.pitch;
  const double *base = in + t.ay0 * t.pitch + t.ax0;
  // the loaded values are not touched until store_tile (a select here would wait for them)
  rg.ok = 0;
#pragma unroll
  for (int i = 0; i < NIT; ++i) {
    const int idx = t.tid + i * NT;
    const int r = idx / C::PX, c = idx - r * C::PX - C::CS;
    const int64_t ay = t.ay0 + r, ax = t.ax0 + c;
    const bool ok = idx < C::BUF && c >= 0 && c < C::EX && (all_in || (ay >= 0 && ay < g.ext[1] && ax >= 0 && ax < t.pitch));
    rg.v[i] = *(ok ? base + ((int64_t)r * t.pitch + c) : in);
    rg.ok |= (uint64_t)ok << i;
  }
}

template <class C>
__device__ __forceinline__ void store_tile(const TileCtx &t, const TileRegs<C> &rg, double *__restrict__ lds) {
  constexpr int NT = TileRegs<C>::NT, NIT = TileRegs<C>::NIT;
#pragma unroll
  for (int i = 0; i < NIT; ++i) {
    const int idx = t.tid + i * NT;
    const bool in_buf = i < NIT - 1 || idx < C::BUF;
    const double v = (rg.ok >> i) & 1 ? rg.v[i] : 0.0;
    lds[in_buf ? idx : C::SPARE] = v;
    if constexpr (C::NBUF == 2) lds[in_buf ? C::BUF + idx : C::SPARE] = v;
  }
}

// Per-lane constants of a kernel: B fragments, A offsets, output cell of the lane.
template <class C>
__device__ __forceinline__ void lane_setup(const double *__restrict__ bfrag, int tid, TileCtx &t, double *b,
                                           int *off) {
  constexpr int PX = C::PX;
  t.tid = tid;
  t.lane = tid & 63;
  t.wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = t.lane;
#pragma unroll
  for (int ks = 0; ks < C::NKS; ++ks) b[ks] = bfrag[ks * 64 + lane];
  if constexpr (C::MF == 16) {
#pragma unroll
    for (int ks = 0; ks < C::NKS; ++ks) off[ks] = ((lane & 15) + C::KT.dy[ks]) * PX + C::KT.c0[ks] + (lane >> 4);
    t.orow = 0;
    t.ocol = 0;
  } else {
    // off[i]: byte offset of read i's cell for this lane from the block origin, set for
    // the representative reads of the read plan only (the others are immediates).  Lane l
    // feeds A_b[l & 3][l >> 4] of block b = (l >> 2) & 3: tap group q = l >> 4 of the 2 x 2
    // sub-block at (dil_row, dil_col)(b, l & 3) of the 8 x 8 block (cs_ktable.h)
    const int q = lane >> 4, b = (lane >> 2) & 3, ai = lane & 3;
    const int arow = 2 * dil_row(C::DT.map, b, ai), acol = 2 * dil_col(C::DT.map, b, ai);
#pragma unroll
    for (int i = 0; i < C::RPL.n; ++i) {
      if (C::RPL.rep[i] != i) continue;
      const int ty = i < C::DT.np ? C::DT.pty[i][q] : C::DT.sty[i - C::DT.np][q];
      const int tx = i < C::DT.np ? C::DT.ptx[i][q] : C::DT.stx[i - C::DT.np][q];
      off[i] = ((arow + ty) * PX + acol + tx + C::CS) * 8;
    }
    // the lane's result D_b[l >> 4][n], n = l & 3: cell (2 row + (n >> 1), 2 col + (n & 1))
    // of the sub-block of (b, l >> 4)
    const int di = lane >> 4, n = lane & 3;
    t.orow = 2 * dil_row(C::DT.map, b, di) + (n >> 1);
    t.ocol = 2 * dil_col(C::DT.map, b, di) + (n & 1);
    t.vw = (t.orow * PX + t.ocol + C::CS) * 8;
  }
}

// Per-tile fields of the context.
template <class C>
__device__ __forceinline__ void tile_setup(const Geo &g, int64_t tile, TileCtx &t, double *out) {
  t.pitch = g.ext[2];
  t.ny = g.n[1];
  t.nx = g.n[2];
  t.Y0 = (tile / g.tiles[2]) * C::BY - C::KF * C::R;
  t.X0 = (tile % g.tiles[2]) * C::BX - C::KF * C::R;
  t.ay0 = g.org[1] + t.Y0;
  t.ax0 = g.org[2] + t.X0;
  t.full = t.Y0 >= 0 && t.X0 >= 0 && t.Y0 + C::EY <= t.ny && t.X0 + C::EX <= t.nx;
  if constexpr (C::MF == 4) {
    t.gout = out + t.ay0 * t.pitch + t.ax0;
    t.vgo = (uint32_t)((t.orow * t.pitch + t.ocol) * 8);
    const int64_t ry0 = t.Y0 < 0 ? -t.Y0 : 0, cx0 = t.X0 < 0 ? -t.X0 : 0;
    const int64_t ry1 = t.ny - t.Y0 < C::EY ? t.ny - t.Y0 : C::EY, cx1 = t.nx - t.X0 < C::EX ? t.nx - t.X0 : C::EX;
    t.ry0 = (int)ry0;
    t.cx0 = (int)cx0;
    t.nry = ry1 > ry0 ? (int)(ry1 - ry0) : 0;
    t.ncx = cx1 > cx0 ? (int)(cx1 - cx0) : 0;
  }
}

// The fused steps of a tile already in LDS (the persistent kernels' tile loop).
template <class C>
__device__ __forceinline__ void run_steps(const TileCtx &t, const double *b, const int *off, double *lds,
                                          double *out) {
  if constexpr (C::MF == 4) {
    if (t.full)
      steps4<C, 0, false>(t, b, off, reinterpret_cast<char *>(lds));
    else
      steps4<C, 0, true>(t, b, off, reinterpret_cast<char *>(lds));
  } else {
    fused_steps<C, 0>(t, b, off, lds, out);
  }
}

// One workgroup advances tile `tile` of grid g by KF steps (lds: the kernel's dynamic LDS;
// LG: tile loads in flight per thread; tid: threadIdx.x).
template <int MF, int R, int SH, int KF, int BY, int BX, int NW, int ST, int LG = 64, int PXM = -1>
__device__ __forceinline__ void tile_f64(const double *__restrict__ in, double *__restrict__ out,
                                         const double *__restrict__ bfrag, const Geo &g, int64_t tile,
                                         double *__restrict__ lds, int tid) {
  using C = Cfg2D<MF, R, SH, KF, BY, BX, NW, ST, PXM>;
  TileCtx t;
  double b[C::NKS];
  int off[C::NOFF];
  lane_setup<C>(bfrag, tid, t, b, off);
  tile_setup<C>(g, tile, t, out);

  // ---- load the tile (and its halo); cells outside the array and pad columns are 0 ----
  if constexpr (C::MF == 4) {
    if (t.full) {
      load_tile_full<C>(in, t, lds);
      lds_barrier();
      steps4<C, 0, false>(t, b, off, reinterpret_cast<char *>(lds));
    } else {
      load_tile<C, LG>(in, g, t, lds);
      lds_barrier();
      steps4<C, 0, true>(t, b, off, reinterpret_cast<char *>(lds));
    }
  } else {
    load_tile<C, LG>(in, g, t, lds);
    lds_barrier();
    fused_steps<C, 0>(t, b, off, lds, out);
  }
}

}  // namespace cs
