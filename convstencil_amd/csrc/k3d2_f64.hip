// k3d2_f64.hip -- FP64 3-D ConvStencil kernels on 8 x 8 blocks (dilated v_mfma_f64_4x4x4).
//
// Replaces gpu_box_3d1r_kernel (reference src/3d/gpu_box.cu:186-238) and
// gpu_star_3d1r_step3_kernel (src/3d/gpu_star.cu:178-236) for radius 1 (the 2.5-D
// kernel in k3d_f64.hip remains for larger radii):
//   * a workgroup owns a BY x BX column tile and a chunk of ZC output planes and
//     streams its input planes through a ring of 2R+1 LDS planes; the next plane's
//     loads are all issued (into registers) before the current plane is computed, so
//     HBM traffic overlaps the MFMAs;
//   * an output plane is a sum over dz of 2-D products, each with the dilated table of
//     its plane (cs_ktable.h, DTable3): one v_mfma_f64_4x4x4 chain per 8 x 8 output
//     block, 16-byte LDS reads, 12 k-steps per 64 outputs for box3d1r, 5 for star3d1r;
//     all planes share one pair parity so a block's column alignment holds for each dz;
//   * the plane loop over dz is unrolled at compile time (every table entry, k-step
//     count and LDS read kind is a constant), B stays in registers for the whole kernel.
#include "cs_internal.h"

namespace cs {

typedef double dbl2v __attribute__((ext_vector_type(2)));

// Tables of a 3-D stencil on 8 x 8 blocks: plane dz's k-steps and LDS reads start at
// kbase(dz) / rbase(dz) in the per-lane arrays b[] and off[].
template <int R_, int SH_>
struct Tab3D {
  static constexpr int R = R_, SH = SH_;
  static constexpr DTable3 DT = make_dtable3(R, SH);
  static constexpr int PP = DT.pp;
  static constexpr int NKS = dtable3_ksteps(DT);
  static constexpr int kbase(int dz) {
    int n = 0;
    for (int d = -R; d < dz; ++d) n += dtable_ksteps(DT.t[d + R]);
    return n;
  }
  static constexpr int rbase(int dz) {
    int n = 0;
    for (int d = -R; d < dz; ++d) n += DT.t[d + R].np + DT.t[d + R].ns;
    return n;
  }
  static constexpr int NRD = rbase(R + 1);
  static constexpr int cmin() {
    int m = 1 << 20;
    for (int d = -R; d <= R; ++d) m = DT.t[d + R].cmin < m ? DT.t[d + R].cmin : m;
    return m;
  }
  static constexpr int cmax() {
    int m = -(1 << 20);
    for (int d = -R; d <= R; ++d) m = DT.t[d + R].cmax > m ? DT.t[d + R].cmax : m;
    return m;
  }
  static constexpr int tmin() {
    int m = 1 << 20;
    for (int d = -R; d <= R; ++d) m = DT.t[d + R].tmin < m ? DT.t[d + R].tmin : m;
    return m;
  }
  static constexpr int tmax() {
    int m = -(1 << 20);
    for (int d = -R; d <= R; ++d) m = DT.t[d + R].tmax > m ? DT.t[d + R].tmax : m;
    return m;
  }
};

template <int R_, int SH_, int BY_, int BX_, int ZC_, int NW_>
struct Cfg3D2 : Tab3D<R_, SH_> {
  using T3 = Tab3D<R_, SH_>;
  using T3::cmax;
  using T3::cmin;
  using T3::DT;
  using T3::PP;
  using T3::tmax;
  using T3::tmin;
  static constexpr int R = R_, SH = SH_, BY = BY_, BX = BX_, ZC = ZC_, NW = NW_;
  static constexpr int EY = BY + 2 * R;
  static constexpr int EX = BX + 2 * R;
  // region: rows [R, R + BY), columns [R, R + BX); tile column c is stored at c + CS and
  // blocks start at columns x with x + CS + PP even
  static constexpr int lox(int cs) { return R - ((R + cs + PP) & 1); }
  static constexpr int find_cs() {
    for (int cs = 0; cs < 16; ++cs)
      if (lox(cs) + cmin() + cs >= 0) return cs;
    return 16;
  }
  static constexpr int CS = find_cs();
  static constexpr int LOX = lox(CS);
  static constexpr int LASTX = (R + BX - 8) + (((R + BX - 8) + CS + PP) & 1);
  static constexpr int NBX = (LASTX - LOX + 7) / 8 + 1;
  static constexpr int NBY = BY / 8;
  static constexpr int NB = NBY * NBX;
  static constexpr int PER = (NB + NW - 1) / NW;
  static constexpr int PX0 = CS + (LASTX + 6 + cmax() + 1 > EX ? LASTX + 6 + cmax() + 1 : EX);
  // fewest LDS bank conflicts among the pitches that keep the most workgroups per CU
  static constexpr int PX = best_pitch_dtables(DT.t, 2 * R + 1, PX0, (long)(2 * R + 1) * EY * 8, 16);
  static constexpr int PLANE = EY * PX;
  static constexpr int NP = 2 * R + 1;
  static constexpr int SPARE = NP * PLANE;
  static constexpr size_t LDS = ((size_t)NP * PLANE + 2) * sizeof(double);
  static constexpr int PF = (PLANE + NW * 64 - 1) / (NW * 64);  // plane loads per thread
  static_assert(BY % 8 == 0 && BX % 16 == 0, "tile must be a multiple of the 8 x 8 block");
  static_assert(CS < 16 && PX % 2 == 0 && (LOX + CS + PP) % 2 == 0, "pair reads must be aligned");
  static_assert(R + tmin() >= 0 && R + BY - 8 + 6 + tmax() < EY, "block reads outside the plane");
  static_assert(LDS <= 160 * 1024, "ring does not fit gfx950 LDS");
};

template <class C>
__device__ __forceinline__ void block_origin3(int blk, int &y, int &x) {
  const int ba = blk / C::NBX, bb = blk - ba * C::NBX;
  y = C::R + 8 * ba;
  x = min(C::LOX + 8 * bb, C::LASTX);
}

// Adds plane dz's products to the block's accumulator (plane: the ring slot of z + dz).
template <class C, int DZ>
__device__ __forceinline__ double plane_chain(const double *plane, int boff, const int *off, const double *b,
                                              double a) {
  constexpr DTable T = C::DT.t[DZ + C::R];
  constexpr int K0 = C::kbase(DZ), R0 = C::rbase(DZ);
  const double *p = plane + boff;
#pragma unroll
  for (int ps = 0; ps < T.np; ++ps) {
    const dbl2v pr = *reinterpret_cast<const dbl2v *>(p + off[R0 + ps]);
    a = __builtin_amdgcn_mfma_f64_4x4x4f64(pr.x, b[K0 + 2 * ps], a, 0, 0, 0);
    a = __builtin_amdgcn_mfma_f64_4x4x4f64(pr.y, b[K0 + 2 * ps + 1], a, 0, 0, 0);
  }
#pragma unroll
  for (int ss = 0; ss < T.ns; ++ss)
    a = __builtin_amdgcn_mfma_f64_4x4x4f64(p[off[R0 + T.np + ss]], b[K0 + 2 * T.np + ss], a, 0, 0, 0);
  return a;
}

template <class C, int DZ>
__device__ __forceinline__ double all_planes(const double *lds, const int *slot, int boff, const int *off,
                                             const double *b, double a) {
  if constexpr (DZ > C::R) {
    return a;
  } else {
    a = plane_chain<C, DZ>(lds + slot[DZ + C::R] * C::PLANE, boff, off, b, a);
    return all_planes<C, DZ + 1>(lds, slot, boff, off, b, a);
  }
}

// OS = 1: each output plane is staged in LDS (in the ring slot of plane z - R, which the
// next plane replaces) and written row by row, coalesced, instead of from the 8 x 8 blocks.
template <int R, int SH, int BY, int BX, int ZC, int NW, int OS = 0>
__global__ __launch_bounds__(NW * 64) void k3d2_f64(const double *__restrict__ in, double *__restrict__ out,
                                                    const double *__restrict__ bfrag, Geo g) {
  using C = Cfg3D2<R, SH, BY, BX, ZC, NW>;
  constexpr int SP = BX + 2;  // staging pitch: the 8 rows of a block's writes hit distinct banks
  static_assert(OS == 0 || BY * SP <= C::PLANE, "the staged plane fits a ring slot");
  constexpr int PX = C::PX, NT = NW * 64;
  extern __shared__ __align__(16) double lds[];

  const int64_t tiles_yx = g.tiles[1] * g.tiles[2];
  const int64_t ntiles = g.tiles[0] * tiles_yx;
  const int64_t per = (ntiles + 7) / 8;
  const int64_t tile = (blockIdx.x % 8) * per + blockIdx.x / 8;  // XCD-contiguous tile bands
  if (tile >= ntiles) return;
  const int64_t zc = tile / tiles_yx, rem = tile % tiles_yx;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t ext1 = g.ext[1], ext2 = g.ext[2];
  const int64_t nz = g.n[0], ny = g.n[1], nx = g.n[2];
  const int64_t Y0 = (rem / g.tiles[2]) * BY - R, X0 = (rem % g.tiles[2]) * BX - R;
  const int64_t ay0 = g.org[1] + Y0, ax0 = g.org[2] + X0;
  const int64_t Z0 = zc * ZC;
  const int64_t Z1 = Z0 + ZC < nz ? Z0 + ZC : nz;
  const bool full = Y0 + C::EY <= ny && X0 + C::EX <= nx && Y0 >= 0 && X0 >= 0;
  const bool all_in = ay0 >= 0 && ay0 + C::EY <= ext1 && ax0 >= 0 && ax0 + C::EX <= ext2;

  double b[C::NKS];
#pragma unroll
  for (int ks = 0; ks < C::NKS; ++ks) b[ks] = bfrag[ks * 64 + lane];
  int off[C::NRD];
  {
    const int q = lane >> 4, b4 = (lane >> 2) & 3, m = lane & 3;
#pragma unroll
    for (int dz = -R; dz <= R; ++dz) {
      const DTable &T = C::DT.t[dz + R];
      const int r0 = C::rbase(dz);
#pragma unroll
      for (int ps = 0; ps < T.np; ++ps) off[r0 + ps] = (2 * b4 + T.pty[ps][q]) * PX + 2 * m + T.ptx[ps][q] + C::CS;
#pragma unroll
      for (int ss = 0; ss < T.ns; ++ss)
        off[r0 + T.np + ss] = (2 * b4 + T.sty[ss][q]) * PX + 2 * m + T.stx[ss][q] + C::CS;
    }
  }
  const int orow = 2 * ((lane >> 2) & 3) + ((lane & 3) >> 1), ocol = 2 * (lane >> 4) + (lane & 1);

  // plane p (interior z) of the tile's column, cell idx of the LDS plane layout
  // (the loaded values are not touched until store(): a select here would wait for them)
  static_assert(C::PF <= 64, "one validity bit per load");
  uint64_t okm = 0;
  auto issue = [&](int64_t p, double *v) {
    const double *base = in + ((g.org[0] + p) * ext1 + ay0) * ext2 + ax0;
    okm = 0;
#pragma unroll
    for (int i = 0; i < C::PF; ++i) {
      const int idx = tid + i * NT;
      const int r = idx / PX, c = idx - r * PX - C::CS;
      const bool ok = idx < C::PLANE && c >= 0 && c < C::EX &&
                      (all_in || (ay0 + r >= 0 && ay0 + r < ext1 && ax0 + c >= 0 && ax0 + c < ext2));
      v[i] = *(ok ? base + ((int64_t)r * ext2 + c) : in);
      okm |= (uint64_t)ok << i;
    }
  };
  auto store = [&](int s, const double *v) {
    double *dst = lds + s * C::PLANE;
#pragma unroll
    for (int i = 0; i < C::PF; ++i) {
      const int idx = tid + i * NT;
      dst[(i < C::PF - 1 || idx < C::PLANE) ? idx : C::SPARE - s * C::PLANE] = (okm >> i) & 1 ? v[i] : 0.0;
    }
  };
  auto slot_of = [&](int64_t p) -> int { return (int)((p - (Z0 - R)) % C::NP); };

  // prologue: planes Z0 - R .. Z0 + R - 1 into the ring, plane Z0 + R in flight
  double pf[C::PF];
  for (int64_t p = Z0 - R; p < Z0 + R; ++p) {
    issue(p, pf);
    store(slot_of(p), pf);
  }
  issue(Z0 + R, pf);

  for (int64_t z = Z0; z < Z1; ++z) {
    store(slot_of(z + R), pf);
    lds_barrier();  // planes z - R .. z + R are in the ring
    if (z + 1 < Z1) issue(z + R + 1, pf);
    int slot[C::NP];
#pragma unroll
    for (int d = 0; d < C::NP; ++d) slot[d] = __builtin_amdgcn_readfirstlane(slot_of(z - R + d));
    double acc[C::PER];
#pragma unroll
    for (int j = 0; j < C::PER; ++j) {
      int y, x;
      block_origin3<C>(min(j * NW + wave, C::NB - 1), y, x);
      acc[j] = all_planes<C, -R>(lds, slot, y * PX + x, off, b, 0.0);
    }
    double *oplane = out + (g.org[0] + z) * ext1 * ext2;
    if constexpr (OS == 0) {
#pragma unroll
      for (int j = 0; j < C::PER; ++j) {
        const int blk = j * NW + wave;
        if ((C::NB % NW == 0 || j < C::PER - 1) || blk < C::NB) {
          int y, x;
          block_origin3<C>(blk, y, x);
          const int ty = y + orow, tx = x + ocol;
          if (tx >= R && tx < R + BX && (full || (Y0 + ty < ny && X0 + tx < nx)))
            oplane[(ay0 + ty) * ext2 + ax0 + tx] = acc[j];
        }
      }
    } else {
      lds_barrier();  // every wave has read planes z - R .. z + R
      double *stg = lds + slot[0] * C::PLANE;  // plane z - R's slot: BY rows of pitch SP
#pragma unroll
      for (int j = 0; j < C::PER; ++j) {
        const int blk = j * NW + wave;
        if ((C::NB % NW == 0 || j < C::PER - 1) || blk < C::NB) {
          int y, x;
          block_origin3<C>(blk, y, x);
          const int ty = y + orow, tx = x + ocol;
          if (tx >= R && tx < R + BX) stg[(ty - R) * SP + tx - R] = acc[j];
        }
      }
      lds_barrier();  // the staged plane is complete
      constexpr int NT2 = NW * 64;
      for (int idx = tid; idx < BY * BX; idx += NT2) {
        const int r = idx / BX, c = idx - r * BX;
        if (full || (Y0 + R + r < ny && X0 + R + c < nx)) oplane[(ay0 + R + r) * ext2 + ax0 + R + c] = stg[r * SP + c];
      }
    }
    lds_barrier();  // every wave is done with plane z - R before its slot is refilled
  }
}

// ---- temporal fusion: KF time levels streamed through z together ----
//
// Level 0 is the input; level l (1..KF) holds the state after l steps on a region
// that shrinks by R per level (rows / columns [l R, E - l R) of the level-0 tile).
// Levels 0 .. KF-1 each keep a ring of 2R+1 LDS planes.  Iteration p stores input
// plane p (prefetched during the previous iteration) and then computes, level by
// level, plane p - l R of level l from planes p - (l+1) R .. p - (l-1) R of level
// l - 1; level KF writes output plane p - KF R to HBM.  A fused launch is exact
// only for a zero halo (the plan fuses only then): halo planes and cells of every
// level are 0, and the rings of levels >= 1 start zeroed, so the cells a level never
// writes stay 0 (halo) or finite (outside its region, read only with zero weight).
template <int R_, int SH_, int BY_, int BX_, int KF_, int ZC_, int NW_>
struct Cfg3F : Tab3D<R_, SH_> {
  using T3 = Tab3D<R_, SH_>;
  using T3::cmax;
  using T3::cmin;
  using T3::DT;
  using T3::PP;
  using T3::tmax;
  using T3::tmin;
  static constexpr int R = R_, SH = SH_, BY = BY_, BX = BX_, KF = KF_, ZC = ZC_, NW = NW_;
  static constexpr int EY = BY + 2 * KF * R;
  static constexpr int EX = BX + 2 * KF * R;
  static constexpr int LO(int l) { return R * l; }
  static constexpr int H(int l) { return BY + 2 * R * (KF - l); }
  static constexpr int HI(int l) { return LO(l) + BX + 2 * R * (KF - l); }
  static constexpr int lox(int l, int cs) { return LO(l) - ((LO(l) + cs + PP) & 1); }
  static constexpr int lastx(int l, int cs) { return (HI(l) - 8) + (((HI(l) - 8) + cs + PP) & 1); }
  static constexpr int find_cs() {
    for (int cs = 0; cs < 16; ++cs) {
      bool ok = true;
      for (int l = 1; l <= KF; ++l) ok = ok && lox(l, cs) + cmin() + cs >= 0;
      if (ok) return cs;
    }
    return 16;
  }
  static constexpr int CS = find_cs();
  static constexpr int LOX(int l) { return lox(l, CS); }
  static constexpr int LASTX(int l) { return lastx(l, CS); }
  static constexpr int NBY(int l) { return (H(l) + 7) / 8; }
  static constexpr int NBX(int l) { return (LASTX(l) - LOX(l) + 7) / 8 + 1; }
  static constexpr int NB(int l) { return NBY(l) * NBX(l); }
  static constexpr int PERS(int l) { return (NB(l) + NW - 1) / NW; }
  static constexpr int max_col() {
    int m = EX - 1;
    for (int l = 1; l <= KF; ++l) m = LASTX(l) + 6 + cmax() > m ? LASTX(l) + 6 + cmax() : m;
    return m;
  }
  static constexpr int PX0 = CS + max_col() + 1;
  // fewest LDS bank conflicts among the pitches that keep the most workgroups per CU
  static constexpr int PX = best_pitch_dtables(DT.t, 2 * R + 1, PX0, (long)KF * (2 * R + 1) * EY * 8, 16);
  static constexpr int PLANE = EY * PX;
  static constexpr int NP = 2 * R + 1;
  static constexpr int RING = NP * PLANE;
  static constexpr int SPARE = KF * RING;
  static constexpr size_t LDS = ((size_t)KF * RING + 2) * sizeof(double);
  static constexpr int PF = (PLANE + NW * 64 - 1) / (NW * 64);
  static_assert(KF >= 2, "the fused kernel runs two or more levels");
  static_assert(BY % 8 == 0 && BX % 16 == 0, "tile must be a multiple of the 8 x 8 block");
  static_assert(CS < 16 && PX % 2 == 0 && (LOX(1) + CS + PP) % 2 == 0, "pair reads must be aligned");
  static_assert(LO(1) + tmin() >= 0 && LO(1) + H(1) - 8 + 6 + tmax() < EY, "block reads outside the plane");
  static_assert(LDS <= 160 * 1024, "rings do not fit gfx950 LDS");
};

struct FusedCtx {
  int tid, wave, orow, ocol;
  bool full;
  int64_t Y0, X0, ny, nx, nz, ay0, ax0, ext1, ext2, org0, Z0, Z1, base;
};

// Level L of iteration p (compile-time L), then the levels above it.
template <class C, int L>
__device__ __forceinline__ void fused_level(const FusedCtx &t, int64_t p, double *lds, double *out, const int *off,
                                            const double *b) {
  if constexpr (L <= C::KF) {
    constexpr int R = C::R, NW = C::NW, PX = C::PX;
    const int64_t q = p - (int64_t)L * R;
    if (q >= t.Z0 - (int64_t)(C::KF - L) * R && q <= t.Z1 - 1 + (int64_t)(C::KF - L) * R) {
      if (L < C::KF && (q < 0 || q >= t.nz)) {  // a halo plane of level L: zero
        double *dst = lds + L * C::RING + (int)((q - t.base) % C::NP) * C::PLANE;
        for (int idx = t.tid; idx < C::PLANE; idx += NW * 64) dst[idx] = 0.0;
      } else {
        int sl[C::NP];
#pragma unroll
        for (int d = 0; d < C::NP; ++d) sl[d] = __builtin_amdgcn_readfirstlane((int)((q - R + d - t.base) % C::NP));
        const double *src = lds + (L - 1) * C::RING;
        constexpr int NB = C::NB(L), PERS = C::PERS(L);
        double acc[PERS];
#pragma unroll
        for (int j = 0; j < PERS; ++j) {
          const int blk = min(j * NW + t.wave, NB - 1);
          const int ba = blk / C::NBX(L), bb = blk - ba * C::NBX(L);
          const int y = C::LO(L) + min(8 * ba, C::H(L) - 8), x = min(C::LOX(L) + 8 * bb, C::LASTX(L));
          acc[j] = all_planes<C, -R>(src, sl, y * PX + x, off, b, 0.0);
        }
        double *dst = L < C::KF ? lds + L * C::RING + (int)((q - t.base) % C::NP) * C::PLANE : nullptr;
#pragma unroll
        for (int j = 0; j < PERS; ++j) {
          const int blk = j * NW + t.wave;
          if ((NB % NW == 0 || j < PERS - 1) || blk < NB) {
            const int ba = blk / C::NBX(L), bb = blk - ba * C::NBX(L);
            const int y = C::LO(L) + min(8 * ba, C::H(L) - 8), x = min(C::LOX(L) + 8 * bb, C::LASTX(L));
            const int ty = y + t.orow, tx = x + t.ocol;
            const bool in = tx >= C::LO(L) && tx < C::HI(L) &&
                            (t.full || (t.Y0 + ty >= 0 && t.Y0 + ty < t.ny && t.X0 + tx >= 0 && t.X0 + tx < t.nx));
            if (in) {
              if constexpr (L < C::KF)
                dst[ty * PX + tx + C::CS] = acc[j];
              else
                out[((t.org0 + q) * t.ext1 + t.ay0 + ty) * t.ext2 + t.ax0 + tx] = acc[j];
            }
          }
        }
      }
    }
    lds_barrier();  // level L of plane q is complete before level L + 1 reads it
    fused_level<C, L + 1>(t, p, lds, out, off, b);
  }
}

template <int R, int SH, int BY, int BX, int KF, int ZC, int NW>
__global__ __launch_bounds__(NW * 64) void k3f_f64(const double *__restrict__ in, double *__restrict__ out,
                                                   const double *__restrict__ bfrag, Geo g) {
  using C = Cfg3F<R, SH, BY, BX, KF, ZC, NW>;
  constexpr int PX = C::PX, NT = NW * 64;
  extern __shared__ __align__(16) double lds[];
  const int64_t tiles_yx = g.tiles[1] * g.tiles[2];
  const int64_t ntiles = g.tiles[0] * tiles_yx;
  const int64_t per = (ntiles + 7) / 8;
  const int64_t tile = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (tile >= ntiles) return;
  const int64_t zc = tile / tiles_yx, rem = tile % tiles_yx;
  FusedCtx t;
  t.tid = threadIdx.x;
  const int lane = t.tid & 63;
  t.wave = __builtin_amdgcn_readfirstlane(t.tid >> 6);
  t.ext1 = g.ext[1];
  t.ext2 = g.ext[2];
  t.org0 = g.org[0];
  t.nz = g.n[0];
  t.ny = g.n[1];
  t.nx = g.n[2];
  t.Y0 = (rem / g.tiles[2]) * BY - KF * R;
  t.X0 = (rem % g.tiles[2]) * BX - KF * R;
  t.ay0 = g.org[1] + t.Y0;
  t.ax0 = g.org[2] + t.X0;
  t.Z0 = zc * ZC;
  t.Z1 = t.Z0 + ZC < t.nz ? t.Z0 + ZC : t.nz;
  t.base = t.Z0 - (int64_t)KF * R;
  t.full = t.Y0 >= 0 && t.X0 >= 0 && t.Y0 + C::EY <= t.ny && t.X0 + C::EX <= t.nx;
  const bool all_in = t.ay0 >= 0 && t.ay0 + C::EY <= t.ext1 && t.ax0 >= 0 && t.ax0 + C::EX <= t.ext2;
  t.orow = 2 * ((lane >> 2) & 3) + ((lane & 3) >> 1);
  t.ocol = 2 * (lane >> 4) + (lane & 1);

  double b[C::NKS];
#pragma unroll
  for (int ks = 0; ks < C::NKS; ++ks) b[ks] = bfrag[ks * 64 + lane];
  int off[C::NRD];
  {
    const int q = lane >> 4, b4 = (lane >> 2) & 3, m = lane & 3;
#pragma unroll
    for (int dz = -R; dz <= R; ++dz) {
      const DTable &T = C::DT.t[dz + R];
      const int r0 = C::rbase(dz);
#pragma unroll
      for (int ps = 0; ps < T.np; ++ps) off[r0 + ps] = (2 * b4 + T.pty[ps][q]) * PX + 2 * m + T.ptx[ps][q] + C::CS;
#pragma unroll
      for (int ss = 0; ss < T.ns; ++ss)
        off[r0 + T.np + ss] = (2 * b4 + T.sty[ss][q]) * PX + 2 * m + T.stx[ss][q] + C::CS;
    }
  }
  // the rings of levels >= 1 start at zero
  for (int idx = t.tid; idx < (KF - 1) * C::RING; idx += NT) lds[C::RING + idx] = 0.0;

  // input plane p (interior z), zero outside the array
  // (the loaded values are not touched until store(): a select here would wait for them)
  static_assert(C::PF <= 64, "one validity bit per load");
  uint64_t okm = 0;
  auto issue = [&](int64_t p, double *v) {
    const bool plane_ok = t.org0 + p >= 0 && t.org0 + p < g.ext[0];
    okm = 0;
#pragma unroll
    for (int i = 0; i < C::PF; ++i) {
      const int idx = t.tid + i * NT;
      const int r = idx / PX, c = idx - r * PX - C::CS;
      const bool ok = plane_ok && idx < C::PLANE && c >= 0 && c < C::EX &&
                      (all_in || (t.ay0 + r >= 0 && t.ay0 + r < t.ext1 && t.ax0 + c >= 0 && t.ax0 + c < t.ext2));
      const int64_t o = ((t.org0 + p) * t.ext1 + t.ay0 + r) * t.ext2 + t.ax0 + c;
      v[i] = *(ok ? in + o : in);
      okm |= (uint64_t)ok << i;
    }
  };
  auto store = [&](int64_t p, const double *v) {
    double *dst = lds + (int)((p - t.base) % C::NP) * C::PLANE;
#pragma unroll
    for (int i = 0; i < C::PF; ++i) {
      const int idx = t.tid + i * NT;
      dst[(i < C::PF - 1 || idx < C::PLANE) ? idx : C::SPARE - (int)(dst - lds)] = (okm >> i) & 1 ? v[i] : 0.0;
    }
  };

  double pf[C::PF];
  const int64_t p0 = t.base, p1 = t.Z1 - 1 + (int64_t)KF * R;
  issue(p0, pf);
  for (int64_t p = p0; p <= p1; ++p) {
    store(p, pf);
    lds_barrier();  // input plane p is in ring 0 (and, at p0, the zeroed rings are visible)
    if (p < p1) issue(p + 1, pf);
    fused_level<C, 1>(t, p, lds, out, off, b);
  }
}

template <int R, int SH, int BY, int BX, int KF, int ZC, int NW>
static hipError_t prepare_k3f() {
  using C = Cfg3F<R, SH, BY, BX, KF, ZC, NW>;
  return hipFuncSetAttribute((const void *)k3f_f64<R, SH, BY, BX, KF, ZC, NW>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS);
}

template <int R, int SH, int BY, int BX, int KF, int ZC, int NW>
static hipError_t launch_k3f(const double *in, double *out, const double *bfrag, const Geo &g0, hipStream_t s) {
  using C = Cfg3F<R, SH, BY, BX, KF, ZC, NW>;
  Geo g = g0;
  g.tiles[0] = (g.n[0] + ZC - 1) / ZC;
  const int64_t ntiles = g.tiles[0] * g.tiles[1] * g.tiles[2];
  const int64_t grid = ((ntiles + 7) / 8) * 8;
  hipLaunchKernelGGL((k3f_f64<R, SH, BY, BX, KF, ZC, NW>), dim3((unsigned)grid), dim3(NW * 64), C::LDS, s, in, out,
                     bfrag, g);
  return hipGetLastError();
}

template <int R, int SH, int BY, int BX, int ZC, int NW, int OS>
static hipError_t prepare_k3d2() {
  using C = Cfg3D2<R, SH, BY, BX, ZC, NW>;
  return hipFuncSetAttribute((const void *)k3d2_f64<R, SH, BY, BX, ZC, NW, OS>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS);
}

template <int R, int SH, int BY, int BX, int ZC, int NW, int OS>
static hipError_t launch_k3d2(const double *in, double *out, const double *bfrag, const Geo &g0, hipStream_t s) {
  using C = Cfg3D2<R, SH, BY, BX, ZC, NW>;
  Geo g = g0;
  g.tiles[0] = (g.n[0] + ZC - 1) / ZC;
  const int64_t ntiles = g.tiles[0] * g.tiles[1] * g.tiles[2];
  const int64_t grid = ((ntiles + 7) / 8) * 8;
  hipLaunchKernelGGL((k3d2_f64<R, SH, BY, BX, ZC, NW, OS>), dim3((unsigned)grid), dim3(NW * 64), C::LDS, s, in,
                     out, bfrag, g);
  return hipGetLastError();
}

// Variant::mf = 8 marks this family (8 x 8 blocks, dilated tables); st = z-chunk length.
#define CS_V3D2O(ID, R, SH, BY, BX, ZC, NW, OS, OSN)                                                          \
  Variant {                                                                                                    \
    ID, 8, R, SH, 1, BY, BX, NW, Cfg3D2<R, SH, BY, BX, ZC, NW>::LDS, Cfg3D2<R, SH, BY, BX, ZC, NW>::PX, ZC,    \
        &launch_k3d2<R, SH, BY, BX, ZC, NW, OS>, &prepare_k3d2<R, SH, BY, BX, ZC, NW, OS>,                     \
        "k3d2_f64<r" #R "," #SH "," #BY "x" #BX ",zc" #ZC ",w" #NW OSN ">", Cfg3D2<R, SH, BY, BX, ZC, NW>::CS \
  }
#define CS_V3D2(ID, R, SH, BY, BX, ZC, NW) CS_V3D2O(ID, R, SH, BY, BX, ZC, NW, 0, "")

// Fused family: same tables and B fragments (mf 8), fusion depth KF; st = z-chunk length.
#define CS_V3F(ID, R, SH, BY, BX, KF, ZC, NW)                                                             \
  Variant {                                                                                               \
    ID, 8, R, SH, KF, BY, BX, NW, Cfg3F<R, SH, BY, BX, KF, ZC, NW>::LDS,                                  \
        Cfg3F<R, SH, BY, BX, KF, ZC, NW>::PX, ZC, &launch_k3f<R, SH, BY, BX, KF, ZC, NW>,                 \
        &prepare_k3f<R, SH, BY, BX, KF, ZC, NW>,                                                          \
        "k3f_f64<r" #R "," #SH "," #BY "x" #BX ",kf" #KF ",zc" #ZC ",w" #NW ">",                          \
        Cfg3F<R, SH, BY, BX, KF, ZC, NW>::CS                                                              \
  }

static const Variant kVariants3D2[] = {
    // single-step: the output planes staged in LDS and written row by row (ids + 300)
    // are the first choice for the box and star classes
    CS_V3D2O(41404, 1, SHAPE3_BOX, 16, 128, 32, 8, 1, ",os1"), CS_V3D2O(41414, 1, SHAPE3_STAR, 16, 128, 32, 8, 1, ",os1"),
    CS_V3D2O(41401, 1, SHAPE3_BOX, 32, 64, 32, 8, 1, ",os1"),  CS_V3D2O(41411, 1, SHAPE3_STAR, 32, 64, 32, 8, 1, ",os1"),
    CS_V3D2O(41402, 1, SHAPE3_BOX, 16, 64, 32, 4, 1, ",os1"),  CS_V3D2O(41412, 1, SHAPE3_STAR, 16, 64, 32, 4, 1, ",os1"),
    CS_V3D2(41101, 1, SHAPE3_BOX, 32, 64, 32, 8),   CS_V3D2(41111, 1, SHAPE3_STAR, 32, 64, 32, 8),
    CS_V3D2(41121, 1, SHAPE3_DIAMOND, 32, 64, 32, 8), CS_V3D2(41102, 1, SHAPE3_BOX, 16, 64, 32, 4),
    CS_V3D2(41112, 1, SHAPE3_STAR, 16, 64, 32, 4),  CS_V3D2(41103, 1, SHAPE3_BOX, 32, 64, 64, 8),
    CS_V3D2(41113, 1, SHAPE3_STAR, 32, 64, 64, 8),  CS_V3D2(41104, 1, SHAPE3_BOX, 16, 128, 32, 8),
    CS_V3D2(41114, 1, SHAPE3_STAR, 16, 128, 32, 8),
    // fused (ids 42000 + shape * 10 + KF, then tile alternatives)
    CS_V3F(42002, 1, SHAPE3_BOX, 16, 64, 2, 64, 8),  CS_V3F(42012, 1, SHAPE3_STAR, 16, 64, 2, 64, 8),
    CS_V3F(42003, 1, SHAPE3_BOX, 16, 64, 3, 64, 8),  CS_V3F(42013, 1, SHAPE3_STAR, 16, 64, 3, 64, 8),
    CS_V3F(42102, 1, SHAPE3_BOX, 16, 64, 2, 64, 4),  CS_V3F(42112, 1, SHAPE3_STAR, 16, 64, 2, 64, 4),
    CS_V3F(42202, 1, SHAPE3_BOX, 8, 128, 2, 64, 4),  CS_V3F(42212, 1, SHAPE3_STAR, 8, 128, 2, 64, 4),
};

const Variant *variants_3d2_f64(int *count) {
  *count = (int)(sizeof(kVariants3D2) / sizeof(kVariants3D2[0]));
  return kVariants3D2;
}

int bfrag_ksteps_3d2_f64(int R, int shape) { return dtable3_ksteps(make_dtable3(R, shape)); }

// B fragments: plane dz's k-steps in table order (pair steps, then singles); lane group
// k = lane >> 4 holds footprint cell t, output column n = lane & 3 is moved by
// v_n = (n >> 1, n & 1), so B[t][n] = W[dz][t - v_n].
void build_bfrag_3d2_f64(int R, int shape, const double *w, double *bfrag) {
  const DTable3 t3 = make_dtable3(R, shape);
  const int D = 2 * R + 1;
  int ks = 0;
  for (int dz = -R; dz <= R; ++dz) {
    const DTable &t = t3.t[dz + R];
    auto put = [&](int kk, int k, int ty, int tx, int live) {
      for (int lane = 16 * k; lane < 16 * k + 16; ++lane) {
        const int n = lane & 3;
        const int dy = ty - (n >> 1), dx = tx - (n & 1);
        double v = 0.0;
        if (live && dy >= -R && dy <= R && dx >= -R && dx <= R) v = w[((dz + R) * D + (dy + R)) * D + (dx + R)];
        bfrag[kk * 64 + lane] = v;
      }
    };
    for (int ps = 0; ps < t.np; ++ps)
      for (int k = 0; k < 4; ++k)
        for (int h = 0; h < 2; ++h) put(ks + 2 * ps + h, k, t.pty[ps][k], t.ptx[ps][k] + h, t.plive[ps][k]);
    for (int ss = 0; ss < t.ns; ++ss)
      for (int k = 0; k < 4; ++k) put(ks + 2 * t.np + ss, k, t.sty[ss][k], t.stx[ss][k], t.slive[ss][k]);
    ks += dtable_ksteps(t);
  }
}

}  // namespace cs
