// k3d_split.hip -- split-precision 3-D ConvStencil kernels (v_mfma_f32_16x16x32 on 16 x 16 blocks).
//
// Computes the same radius-1 3-D stencils as the FP64 kernels (k3d2_f64.hip; reference
// gpu_box_3d1r_kernel, src/3d/gpu_box.cu:186-238, and gpu_star_3d1r_step3_kernel,
// src/3d/gpu_star.cu:178-236, which are FP64 only) with the arithmetic contract of the
// 2-D dilated split form (k2d_split2_body.h):
//   * the state is FP64 in HBM; an input plane entering LDS is rounded to FP32 once and
//     split into slices (three BF16, or two FP16 with three weight slices); products
//     accumulate in FP32 (level 0 and the kept lower levels in two accumulators); a fused
//     level recombines in FP32 and re-splits, the last level recombines in FP64 and
//     writes FP64;
//   * an output plane is a sum over dz of 2-D products, each with the chunk table of its
//     plane (cs_ktable.h, CTable3) and its own A fragments (the weights' slices, in
//     registers for the whole kernel); every MFMA reads 16 bytes per lane from one plane;
//   * KF time levels stream through z together as in k3f_f64 (KF = 1 is the unfused
//     kernel): level l keeps a ring of 2R+1 planes of slices, input plane p + 1 is loaded
//     into registers while plane p is computed, and a fused launch is exact for a zero
//     halo only (the plan fuses only then).
#include <vector>

#include "cs_wcomp.h"
#include "k2d_split2_body.h"

namespace cs {

typedef uint32_t z4 __attribute__((ext_vector_type(4)));  // 16 bytes of zeros

template <class Tr_, int R_, int SH_, int BY_, int BX_, int KF_, int ZC_, int NW_>
struct CfgS3 {
  using Tr = Tr_;
  using T = typename Tr::T;
  static constexpr int R = R_, SH = SH_, BY = BY_, BX = BX_, KF = KF_, ZC = ZC_, NW = NW_;
  static constexpr int NSU = Tr::NSU, NSW = Tr::NSW, NP = 2 * R + 1, ESZ = (int)sizeof(T);
  static constexpr CTable3 CT = CTab3<R, SH>::T;
  // MFMAs of plane dz, and the index of its first one in the per-lane arrays
  static constexpr int nins(int dz) { return CT.t[dz + R].n / 4; }
  static constexpr int ibase(int dz) {
    int n = 0;
    for (int d = -R; d < dz; ++d) n += nins(d);
    return n;
  }
  static constexpr int NINS = ibase(R + 1);
  static constexpr int cmin() {
    int m = 1 << 20;
    for (int d = -R; d <= R; ++d) m = CT.t[d + R].cmin < m ? CT.t[d + R].cmin : m;
    return m;
  }
  static constexpr int cmax() {
    int m = -(1 << 20);
    for (int d = -R; d <= R; ++d) m = CT.t[d + R].cmax > m ? CT.t[d + R].cmax : m;
    return m;
  }
  static constexpr int tmin() {
    int m = 1 << 20;
    for (int d = -R; d <= R; ++d) m = CT.t[d + R].tmin < m ? CT.t[d + R].tmin : m;
    return m;
  }
  static constexpr int tmax() {
    int m = -(1 << 20);
    for (int d = -R; d <= R; ++d) m = CT.t[d + R].tmax > m ? CT.t[d + R].tmax : m;
    return m;
  }
  static constexpr int EY = BY + 2 * KF * R;
  static constexpr int EX = BX + 2 * KF * R;
  // level l (1..KF): rows [LO, LO + H), columns [LO, HI) of the level-0 tile
  static constexpr int LO(int l) { return R * l; }
  static constexpr int H(int l) { return BY + 2 * R * (KF - l); }
  static constexpr int HI(int l) { return LO(l) + BX + 2 * R * (KF - l); }
  // tile column c is stored at c + CS; blocks start at columns x with x + CS + c0 = 0
  // (mod 8), the first at most 7 columns left of the level's region
  static constexpr int lox(int l, int cs) { return LO(l) - (((LO(l) + cs + CT.c0) % 8) + 8) % 8; }
  static constexpr int nbx(int l, int cs) { return (HI(l) - lox(l, cs) + 15) / 16; }
  static constexpr int find_cs() {
    for (int cs = 0; cs < 64; ++cs) {
      bool ok = true;
      for (int l = 1; l <= KF; ++l) ok = ok && lox(l, cs) + cmin() + cs >= 0;
      if (ok) return cs;
    }
    return 64;
  }
  static constexpr int CS = find_cs();
  static constexpr int LOX(int l) { return lox(l, CS); }
  static constexpr int NBX(int l) { return nbx(l, CS); }
  static constexpr int NBY(int l) { return (H(l) + 15) / 16; }
  static constexpr int NB(int l) { return NBY(l) * NBX(l); }
  static constexpr int PERS(int l) { return (NB(l) + NW - 1) / NW; }
  static constexpr int max_col() {  // largest tile column read (chunk + 7) or written (block + 15)
    int m = EX - 1;
    for (int l = 1; l <= KF; ++l) {
      const int xl = LOX(l) + 16 * (NBX(l) - 1);
      m = xl + 8 + cmax() + 7 > m ? xl + 8 + cmax() + 7 : m;
      m = xl + 15 > m ? xl + 15 : m;
    }
    return m;
  }
  static constexpr int PX0 = CS + max_col() + 1;
  // a multiple of 8 elements with the fewest LDS bank conflicts that keeps the most
  // workgroups per CU (cs_ktable.h, lds_cost)
  static constexpr int PX = best_pitch_ctable3(CT, PX0, ESZ, (long)KF * NP * NSU * EY * ESZ, 8 * ESZ);
  static constexpr int PLANE = EY * PX;   // elements of one slice plane
  static constexpr int SLOT = NSU * PLANE;  // one ring slot: all slices of a plane
  static constexpr int RING = NP * SLOT;
  static constexpr int SPARE = KF * RING;  // scratch cells after the rings
  static constexpr size_t LDS = ((size_t)KF * RING + 8) * ESZ;
  static constexpr int NT = NW * 64;
  static constexpr int QPR = PX / 4;                // quads (4 adjacent cells) per row
  static constexpr int NQ = EY * QPR;               // quads per plane
  static constexpr int PFQ = (NQ + NT - 1) / NT;    // quads per thread per input plane
  static constexpr void origin(int l, int blk, int &y, int &x) {
    const int ba = blk / NBX(l), bb = blk - ba * NBX(l);
    y = LO(l) + (16 * ba < H(l) - 16 ? 16 * ba : H(l) - 16);
    x = LOX(l) + 16 * bb;
  }
  static_assert(BY % 16 == 0 && BX % 16 == 0, "tile must be a multiple of 16");
  static_assert(CS < 64 && (CS + LOX(1) + CT.c0) % 8 == 0, "block columns must align the chunk reads");
  static_assert(PX % 8 == 0, "16-byte reads need a pitch that is a multiple of 8 elements");
  static_assert(LO(1) + tmin() >= 0 && LO(1) + H(1) - 16 + 14 + tmax() < EY, "block reads outside the plane");
  static_assert(4 * PFQ <= 64, "one validity bit per load");
  static_assert(LDS <= 160 * 1024, "rings do not fit gfx950 LDS");
};

struct Split3Ctx {
  int tid, wave, orow, ocol;
  bool full;  // the tile with its halo lies inside the domain interior
  int64_t Y0, X0, ny, nx, nz, ay0, ax0, ext1, ext2, org0, Z0, Z1, base;
};

// Adds the products of planes DZ .. R of one block (src: level ring; sl: byte offsets of
// the ring slots of planes q - R .. q + R; boff: byte offset of the block origin).
template <class C, int DZ>
__device__ __forceinline__ void split3_planes(const char *src, const int *sl, int boff, const int *off,
                                              const typename C::Tr::V (*w)[C::NINS], f4 &d0, f4 &d1) {
  if constexpr (DZ <= C::R) {
    using Tr = typename C::Tr;
    using V = typename Tr::V;
    constexpr int I0 = C::ibase(DZ), NI = C::nins(DZ), PB = C::PLANE * C::ESZ;
    const char *p = src + sl[DZ + C::R] + boff;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      V u[C::NSU];
#pragma unroll
      for (int s = 0; s < C::NSU; ++s) u[s] = *reinterpret_cast<const V *>(p + off[I0 + i] + s * PB);
      d0 = Tr::mma(w[0][I0 + i], u[0], d0);
#pragma unroll
      for (int s = 0; s < C::NSU; ++s)
#pragma unroll
        for (int q = 0; q < C::NSW; ++q)
          if (keep_product<Tr>(s, q)) d1 = Tr::mma(w[q][I0 + i], u[s], d1);
    }
    split3_planes<C, DZ + 1>(src, sl, boff, off, w, d0, d1);
  }
}

// Level L of iteration p (plane q = p - L R), then the levels above it.
template <class C, int L>
__device__ __forceinline__ void split3_level(const Split3Ctx &t, int64_t p, char *lds, double *out,
                                             const typename C::Tr::V (*w)[C::NINS], const int *off) {
  if constexpr (L <= C::KF) {
    using Tr = typename C::Tr;
    using T = typename C::T;
    using Q = typename Quad<Tr>::V;
    constexpr int R = C::R, NW = C::NW, PX = C::PX, ESZ = C::ESZ;
    constexpr int PB = C::PLANE * ESZ, SB = C::SLOT * ESZ, RB = C::RING * ESZ;
    constexpr float INV = 1.0f / (float)Tr::SCALE;
    const int64_t q = p - (int64_t)L * R;
    if (q >= t.Z0 - (int64_t)(C::KF - L) * R && q <= t.Z1 - 1 + (int64_t)(C::KF - L) * R) {
      if (L < C::KF && (q < 0 || q >= t.nz)) {  // a halo plane of level L: zero slices
        z4 *dst = reinterpret_cast<z4 *>(lds + L * RB + (int)((q - t.base) % C::NP) * SB);
        for (int idx = t.tid; idx < SB / 16; idx += NW * 64) dst[idx] = z4{0u, 0u, 0u, 0u};
      } else {
        int sl[C::NP];
#pragma unroll
        for (int d = 0; d < C::NP; ++d)
          sl[d] = __builtin_amdgcn_readfirstlane((int)((q - R + d - t.base) % C::NP) * SB);
        const char *src = lds + (L - 1) * RB;
        constexpr int NB = C::NB(L), PERS = C::PERS(L);
        f4 d0[PERS], d1[PERS];
#pragma unroll
        for (int j = 0; j < PERS; ++j) {
          int y, x;
          C::origin(L, min(j * NW + t.wave, NB - 1), y, x);
          d0[j] = f4{0.f, 0.f, 0.f, 0.f};
          d1[j] = f4{0.f, 0.f, 0.f, 0.f};
          split3_planes<C, -R>(src, sl, (y * PX + x) * ESZ, off, w, d0[j], d1[j]);
        }
        char *dst = lds + (L < C::KF ? L * RB + (int)((q - t.base) % C::NP) * SB : 0);
#pragma unroll
        for (int j = 0; j < PERS; ++j) {
          const int blk = j * NW + t.wave;
          if (!((NB % NW == 0 || j < PERS - 1) || blk < NB)) continue;
          int y, x;
          C::origin(L, blk, y, x);
          const int ty = y + t.orow, tx0 = x + t.ocol;
          if constexpr (L < C::KF) {
            // cells beyond the level's region are never read by a later level's region;
            // cells outside the domain interior keep their zero
            Q qv[C::NSU];
            const float f[4] = {d0[j][0] + d1[j][0] * INV, d0[j][1] + d1[j][1] * INV, d0[j][2] + d1[j][2] * INV,
                                d0[j][3] + d1[j][3] * INV};
            split4<Tr>(f, qv);
            char *cell = dst + (ty * PX + tx0 + C::CS) * ESZ;
            const bool row_in = t.full || (t.Y0 + ty >= 0 && t.Y0 + ty < t.ny);
            if (t.full || (row_in && t.X0 + tx0 >= 0 && t.X0 + tx0 + 3 < t.nx)) {
#pragma unroll
              for (int k = 0; k < C::NSU; ++k) *reinterpret_cast<Q *>(cell + k * PB) = qv[k];
            } else if (row_in) {
#pragma unroll
              for (int v = 0; v < 4; ++v)
                if (t.X0 + tx0 + v >= 0 && t.X0 + tx0 + v < t.nx)
#pragma unroll
                  for (int k = 0; k < C::NSU; ++k) reinterpret_cast<T *>(cell + k * PB)[v] = qv[k][v];
            }
          } else {
            double *o = out + ((t.org0 + q) * t.ext1 + t.ay0 + ty) * t.ext2 + t.ax0 + tx0;
            const bool row_in = t.full || (t.Y0 + ty >= 0 && t.Y0 + ty < t.ny);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int tx = tx0 + v;
              const bool in = row_in && tx >= C::LO(L) && tx < C::HI(L) &&
                              (t.full || (t.X0 + tx >= 0 && t.X0 + tx < t.nx));
              if (in) o[v] = (double)d0[j][v] + (double)d1[j][v] * (double)INV;
            }
          }
        }
      }
    }
    lds_barrier();  // level L of plane q is complete before level L + 1 reads it
    split3_level<C, L + 1>(t, p, lds, out, w, off);
  }
}

template <class Tr, int R, int SH, int BY, int BX, int KF, int ZC, int NW>
__global__ __launch_bounds__(NW * 64) void k3d_split(const double *__restrict__ in, double *__restrict__ out,
                                                     const typename Tr::T *__restrict__ bfrag, Geo g) {
  using C = CfgS3<Tr, R, SH, BY, BX, KF, ZC, NW>;
  using V = typename Tr::V;
  using Q = typename Quad<Tr>::V;
  constexpr int PX = C::PX, NT = C::NT, ESZ = C::ESZ, PB = C::PLANE * ESZ, SB = C::SLOT * ESZ;
  extern __shared__ __align__(16) unsigned char smem[];
  char *lds = reinterpret_cast<char *>(smem);
  const int64_t tiles_yx = g.tiles[1] * g.tiles[2];
  const int64_t ntiles = g.tiles[0] * tiles_yx;
  const int64_t per = (ntiles + 7) / 8;
  const int64_t tile = (blockIdx.x % 8) * per + blockIdx.x / 8;  // XCD-contiguous tile bands
  if (tile >= ntiles) return;
  const int64_t zc = tile / tiles_yx, rem = tile % tiles_yx;
  Split3Ctx t;
  t.tid = threadIdx.x;
  const int lane = t.tid & 63, grp = lane >> 4, n = lane & 15;
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
  // D[m][n], m = 4 grp + v: cell (2 (n & 7) + (grp >> 1), 8 (n >> 3) + 4 (grp & 1) + v)
  t.orow = 2 * (n & 7) + (grp >> 1);
  t.ocol = 8 * (n >> 3) + 4 * (grp & 1);

  V w[C::NSW][C::NINS];
#pragma unroll
  for (int q = 0; q < C::NSW; ++q)
#pragma unroll
    for (int i = 0; i < C::NINS; ++i)
      w[q][i] = *reinterpret_cast<const V *>(bfrag + ((size_t)(q * C::NINS + i) * 64 + lane) * 8);
  // B reads: chunk 4 i + grp of plane dz's instruction i, anchor u_n = (2 (n & 7), 8 (n >> 3))
  int off[C::NINS];
#pragma unroll
  for (int dz = -R; dz <= R; ++dz) {
    const CTable &ct = C::CT.t[dz + R];
#pragma unroll
    for (int i = 0; i < C::nins(dz); ++i)
      off[C::ibase(dz) + i] = ((2 * (n & 7) + ct.ty[4 * i + grp]) * PX + 8 * (n >> 3) + ct.tx[4 * i + grp] + C::CS) * ESZ;
  }
  // the rings of levels >= 1 start at zero
  {
    z4 *z = reinterpret_cast<z4 *>(lds + C::RING * ESZ);
    for (int idx = t.tid; idx < (KF - 1) * C::RING * ESZ / 16; idx += NT) z[idx] = z4{0u, 0u, 0u, 0u};
  }

  // input plane p (interior z) as quads of 4 adjacent storage columns, zero outside the
  // array and in pad columns (the loaded values are not touched until store(): a select
  // here would wait for them)
  double pf[C::PFQ][4];
  uint64_t okm = 0;
  auto issue = [&](int64_t p) {
    const bool plane_ok = t.org0 + p >= 0 && t.org0 + p < g.ext[0];
    okm = 0;
#pragma unroll
    for (int i = 0; i < C::PFQ; ++i) {
      const int qi = t.tid + i * NT;
      const int r = qi / C::QPR, c4 = 4 * (qi - r * C::QPR);
      const bool row_ok = plane_ok && qi < C::NQ && (all_in || (t.ay0 + r >= 0 && t.ay0 + r < t.ext1));
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int c = c4 + e - C::CS;
        const bool ok = row_ok && c >= 0 && c < C::EX && (all_in || (t.ax0 + c >= 0 && t.ax0 + c < t.ext2));
        const int64_t o = ((t.org0 + p) * t.ext1 + t.ay0 + r) * t.ext2 + t.ax0 + c;
        pf[i][e] = *(ok ? in + o : in);
        okm |= (uint64_t)ok << (4 * i + e);
      }
    }
  };
  auto store = [&](int64_t p) {
    char *dst = lds + (int)((p - t.base) % C::NP) * SB;
#pragma unroll
    for (int i = 0; i < C::PFQ; ++i) {
      const int qi = t.tid + i * NT;
      const int r = qi / C::QPR, c4 = 4 * (qi - r * C::QPR);
      float f[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) f[e] = (okm >> (4 * i + e)) & 1 ? (float)pf[i][e] : 0.0f;
      Q qv[C::NSU];
      split4<Tr>(f, qv);
      const bool in_plane = i < C::PFQ - 1 || qi < C::NQ;
#pragma unroll
      for (int k = 0; k < C::NSU; ++k)
        *reinterpret_cast<Q *>(in_plane ? dst + k * PB + (r * PX + c4) * ESZ : lds + C::SPARE * ESZ) = qv[k];
    }
  };

  const int64_t p0 = t.base, p1 = t.Z1 - 1 + (int64_t)KF * R;
  issue(p0);
  for (int64_t p = p0; p <= p1; ++p) {
    store(p);
    lds_barrier();  // input plane p is in ring 0 (and, at p0, the zeroed rings are visible)
    if (p < p1) issue(p + 1);
    split3_level<C, 1>(t, p, lds, out, w, off);
  }
}

template <class Tr, int R, int SH, int BY, int BX, int KF, int ZC, int NW>
static hipError_t prepare_ks3() {
  using C = CfgS3<Tr, R, SH, BY, BX, KF, ZC, NW>;
  return hipFuncSetAttribute((const void *)k3d_split<Tr, R, SH, BY, BX, KF, ZC, NW>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::LDS);
}

template <class Tr, int R, int SH, int BY, int BX, int KF, int ZC, int NW>
static hipError_t launch_ks3(const double *in, double *out, const double *bfrag, const Geo &g0, hipStream_t s) {
  using C = CfgS3<Tr, R, SH, BY, BX, KF, ZC, NW>;
  Geo g = g0;
  g.tiles[0] = (g.n[0] + ZC - 1) / ZC;
  const int64_t ntiles = g.tiles[0] * g.tiles[1] * g.tiles[2];
  const int64_t grid = ((ntiles + 7) / 8) * 8;
  hipLaunchKernelGGL((k3d_split<Tr, R, SH, BY, BX, KF, ZC, NW>), dim3((unsigned)grid), dim3(NW * 64), C::LDS, s, in,
                     out, reinterpret_cast<const typename Tr::T *>(bfrag), g);
  return hipGetLastError();
}

// Variant::mf holds the precision (1 BF16x3, 4 FP16 w3); st = 2 (dilated form).
#define CS_VS3(ID, P, TR, R, SH, BY, BX, KF, ZC, NW)                                                         \
  Variant {                                                                                                  \
    ID, P, R, SH, KF, BY, BX, NW, CfgS3<TR, R, SH, BY, BX, KF, ZC, NW>::LDS,                                 \
        CfgS3<TR, R, SH, BY, BX, KF, ZC, NW>::PX, 2, &launch_ks3<TR, R, SH, BY, BX, KF, ZC, NW>,             \
        &prepare_ks3<TR, R, SH, BY, BX, KF, ZC, NW>,                                                         \
        "k3d_split<" #TR ",r" #R "," #SH "," #BY "x" #BX ",kf" #KF ",zc" #ZC ",w" #NW ">",                   \
        CfgS3<TR, R, SH, BY, BX, KF, ZC, NW>::CS                                                             \
  }

// ids 90000 + precision * 1000 + shape * 100 + KF * 10 + tile alternative
static const Variant kVariants3DSplit[] = {
    CS_VS3(91010, 1, TrBF16, 1, SHAPE3_BOX, 32, 64, 1, 32, 8),
    CS_VS3(91020, 1, TrBF16, 1, SHAPE3_BOX, 16, 64, 2, 64, 8),
    CS_VS3(91021, 1, TrBF16, 1, SHAPE3_BOX, 16, 64, 2, 64, 4),
    CS_VS3(91110, 1, TrBF16, 1, SHAPE3_STAR, 32, 64, 1, 32, 8),
    CS_VS3(91120, 1, TrBF16, 1, SHAPE3_STAR, 16, 64, 2, 64, 8),
    CS_VS3(91121, 1, TrBF16, 1, SHAPE3_STAR, 16, 64, 2, 64, 4),
    CS_VS3(94010, 4, TrF16W3, 1, SHAPE3_BOX, 32, 64, 1, 32, 8),
    CS_VS3(94020, 4, TrF16W3, 1, SHAPE3_BOX, 16, 64, 2, 64, 8),
    CS_VS3(94021, 4, TrF16W3, 1, SHAPE3_BOX, 16, 64, 2, 64, 4),
    CS_VS3(94110, 4, TrF16W3, 1, SHAPE3_STAR, 32, 64, 1, 32, 8),
    CS_VS3(94120, 4, TrF16W3, 1, SHAPE3_STAR, 16, 64, 2, 64, 8),
    CS_VS3(94121, 4, TrF16W3, 1, SHAPE3_STAR, 16, 64, 2, 64, 4),
};

const Variant *variants_3d_split(int *count) {
  *count = (int)(sizeof(kVariants3DSplit) / sizeof(kVariants3DSplit[0]));
  return kVariants3DSplit;
}

// A fragments: weight slice q, instruction i (plane dz's instructions in order), lane l,
// element j is W[dz][t - v_m] with m = l & 15, chunk 4 i' + (l >> 4) of plane dz's table
// (i' = i - first instruction of dz) and t = (ty, tx0 + j).
template <class Tr>
static size_t build_split3(int R, int shape, const double *w, unsigned char *out, size_t cap) {
  using T = typename Tr::T;
  const CTable3 c3 = make_ctable3(R, shape);
  int nins = 0;
  for (int dz = -R; dz <= R; ++dz) nins += c3.t[dz + R].n / 4;
  const size_t bytes = (size_t)Tr::NSW * nins * 64 * 8 * sizeof(T);
  if (bytes > cap) return 0;
  T *o = reinterpret_cast<T *>(out);
  const int D = 2 * R + 1, N = D * D * D;
  if (N > 343) return 0;
  // slices of every tap; BF16x3 moves its last slice by ulps so that the weights' errors
  // cancel in the mass and the first moments (cs_wcomp.h)
  std::vector<T> tap((size_t)3 * N);
  for (int t = 0; t < N; ++t) {
    T sl[3];
    split_weight<Tr>(w[t], sl);
    for (int q = 0; q < Tr::NSW; ++q) tap[(size_t)q * N + t] = sl[q];
  }
  if constexpr (Tr::NSW == Tr::NSU && Tr::NSW == 3) {
    std::vector<double> fixed((size_t)N);
    std::vector<int> offv((size_t)3 * N);
    for (int t = 0; t < N; ++t) {
      fixed[(size_t)t] = (double)(float)tap[(size_t)t] + (double)(float)tap[(size_t)N + t] / Tr::SCALE;
      offv[(size_t)3 * t] = t / (D * D) - R;
      offv[(size_t)3 * t + 1] = (t / D) % D - R;
      offv[(size_t)3 * t + 2] = t % D - R;
    }
    compensate_low_slice<T>(N, w, reinterpret_cast<const int(*)[3]>(offv.data()), fixed.data(), &tap[(size_t)2 * N],
                            Tr::SCALE * Tr::SCALE);
  }
  for (int q = 0; q < Tr::NSW; ++q) {
    int i = 0;
    for (int dz = -R; dz <= R; ++dz) {
      const CTable &ct = c3.t[dz + R];
      for (int li = 0; li < ct.n / 4; ++li, ++i)
        for (int lane = 0; lane < 64; ++lane)
          for (int j = 0; j < 8; ++j) {
            const int m = lane & 15, c = 4 * li + (lane >> 4);
            const int dy = ct.ty[c] - (m >> 3), dx = ct.tx[c] + j - (m & 7);
            T v = (T)0.0f;
            if (ct.live[c] && dy >= -R && dy <= R && dx >= -R && dx <= R)
              v = tap[(size_t)q * N + ((dz + R) * D + (dy + R)) * D + (dx + R)];
            o[((size_t)(q * nins + i) * 64 + lane) * 8 + j] = v;
          }
    }
  }
  return bytes;
}

size_t build_bfrag_split3(int prec, int R, int shape, const double *w, void *out, size_t cap) {
  unsigned char *o = reinterpret_cast<unsigned char *>(out);
  if (prec == 1) return build_split3<TrBF16>(R, shape, w, o, cap);
  if (prec == 4) return build_split3<TrF16W3>(R, shape, w, o, cap);
  return 0;
}

}  // namespace cs
