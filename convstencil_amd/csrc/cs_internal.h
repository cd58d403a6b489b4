// cs_internal.h -- types shared by the host runtime and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include "cs_ktable.h"

namespace cs {

// Workgroup barrier for data that waves exchange through LDS only: waits for this wave's
// LDS (and scalar) operations, then for the other waves.  __syncthreads() also waits for
// every outstanding global load, which would stall each fused step behind loads issued
// ahead for the next tile.  The memory clobber keeps the compiler from moving memory
// accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Lane shifts of a double across the wave (v_mov_b32_dpp wave_shr:1 / wave_shl:1, bound_ctrl:
// the lane without a source receives zero).  lane_from_lower: lane l gets lane l - 1's value;
// lane_from_upper: lane l gets lane l + 1's value.
__device__ __forceinline__ double lane_from_lower(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane_from_upper(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u2 __attribute__((ext_vector_type(2)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

// Raw buffer resource over `bytes` bytes from p (no stride, no format; word 3 as for the
// gfx9 family).  The hardware drops a buffer store whose offset lies past `bytes`, so a
// store that only some rows or lanes want can still be issued by every lane on every row:
// no branch, and the compiler's waits count the stores exactly.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_over(double *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, (int)bytes, 0x00020000);
}
constexpr int kBufferDrop = 0x40000000;  // a store offset past every range used

__device__ __forceinline__ void buffer_store2(__amdgpu_buffer_rsrc_t r, int off, double a, double b) {
  u4 v;
  v.x = (unsigned)__double2loint(a);
  v.y = (unsigned)__double2hiint(a);
  v.z = (unsigned)__double2loint(b);
  v.w = (unsigned)__double2hiint(b);
  __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, 0);
}

__device__ __forceinline__ void buffer_store1(__amdgpu_buffer_rsrc_t r, int off, double a) {
  u2 v;
  v.x = (unsigned)__double2loint(a);
  v.y = (unsigned)__double2hiint(a);
  __builtin_amdgcn_raw_buffer_store_b64(v, r, off, 0, 0);
}

// Geometry of one padded grid as the kernels see it.  All coordinates slowest first.
struct Geo {
  int64_t ext[3];    // padded extents
  int64_t org[3];    // interior origin
  int64_t n[3];      // interior sizes
  int64_t tiles[3];  // tiles per axis for this launch
};

// Prefetch of a tile into L2 for the persistent kernels: one 8-byte load per thread and
// pass touches every 128-byte line of the tile's rows (EY rows, EX columns from
// (ay0, ax0), clamped to the array).  The loaded values go to pf[] and must be consumed
// (prefetch_consume) after the work they should overlap; nothing waits for them before.
template <class C>
struct Prefetch {
  static constexpr int NT = C::NW * 64;
  static constexpr int LPR = (C::EX + 15) / 16 + 1;  // lines per row (a row may straddle one more)
  static constexpr int RPP = NT / LPR;
  static constexpr int N = (C::EY + RPP - 1) / RPP;
};

template <class C, class Ctx>
__device__ __forceinline__ void prefetch_tile(const double *__restrict__ in, const Geo &g, const Ctx &t,
                                              double (&pf)[Prefetch<C>::N]) {
  using P = Prefetch<C>;
  const int r0 = t.tid / P::LPR, k = t.tid - r0 * P::LPR;
  int64_t ax = t.ax0 + 16 * k < t.ax0 + C::EX - 1 ? t.ax0 + 16 * k : t.ax0 + C::EX - 1;
  ax = ax < 0 ? 0 : (ax > t.pitch - 1 ? t.pitch - 1 : ax);
#pragma unroll
  for (int i = 0; i < P::N; ++i) {
    int64_t ay = t.ay0 + r0 + (int64_t)P::RPP * i;
    ay = ay < 0 ? 0 : (ay > g.ext[1] - 1 ? g.ext[1] - 1 : ay);
    pf[i] = in[ay * t.pitch + ax];
  }
}

template <class C>
__device__ __forceinline__ void prefetch_consume(const double (&pf)[Prefetch<C>::N]) {
#pragma unroll
  for (int i = 0; i < Prefetch<C>::N; ++i) asm volatile("" ::"v"(pf[i]));
}

// Decomposition of one launch of the register-streaming 2-D kernels (k2d_reg.hip) into
// waves: each wave sweeps rows [y0, y1) of one strip of output columns.  Strips 0 and
// strips - 1 (the edge strips, which run the checked path) are cut into `we` chunks of
// `he` rows; every other strip into `wi` chunks: `he` rows at the top, `he` at the bottom
// (the checked path) and chunks of `hm` rows between them; hm = 0 marks strips of at most
// 2 he rows, cut like the edge strips.  Waves: the edge strips' chunks first, then the other
// strips' chunk by chunk.
struct RegSplit {
  int64_t strips, he, hm, we, wi;
};

__host__ __device__ inline int64_t reg_waves(const RegSplit &sp) {
  const int64_t ne = sp.strips >= 2 ? 2 : 1;
  return ne * sp.we + (sp.strips - ne) * sp.wi;
}

// Strip and rows [y0, y1) of wave wid; false for a wave without work.
__host__ __device__ inline bool reg_chunk(const RegSplit &sp, int64_t ny, int64_t wid, int64_t *strip, int64_t *y0,
                                          int64_t *y1) {
  const int64_t ne = sp.strips >= 2 ? 2 : 1;
  int64_t a, b;
  if (wid < ne * sp.we) {
    *strip = wid < sp.we ? 0 : sp.strips - 1;
    a = (wid % sp.we) * sp.he;
    b = a + sp.he;
  } else {
    // consecutive waves take the same chunk of neighbouring strips, so that the columns
    // that neighbouring strips share are loaded at about the same time on one CU (L2 hits)
    const int64_t ni = sp.strips - ne, u = wid - ne * sp.we;
    if (ni <= 0 || u >= ni * sp.wi) return false;
    const int64_t c = u / ni;
    *strip = 1 + u % ni;
    if (sp.hm == 0) {  // a short strip: chunks of he rows from the top
      a = c * sp.he;
      b = a + sp.he;
    } else if (c == 0) {
      a = 0;
      b = sp.he;
    } else if (c == sp.wi - 1) {
      a = ny - sp.he;
      b = ny;
    } else {
      a = sp.he + (c - 1) * sp.hm;
      b = a + sp.hm < ny - sp.he ? a + sp.hm : ny - sp.he;
    }
  }
  if (a >= ny || b <= a) return false;
  *y0 = a;
  *y1 = b < ny ? b : ny;
  return true;
}

// Time of a wave on the checked path (edge strips, first and last chunk of a strip)
// relative to the plain path, per row, for strips of 64 px columns.  Measured at 10240^2
// (box2d1r depth 8, GStencil/s): 2 columns per lane (whose checked path spills at 3 waves
// per SIMD) 1920-1928 at 2.5, 1868 at 1.75, 1752 at 1.25; 4 columns per lane 1948-1950 at
// 2.5, 1979-1985 at 1.75, 1888 at 1.25; star2d3r (2 columns) 968 at 2.5, 980 at 1.75.
inline double reg_edge_cost(int px, int R) { return px >= 4 || R > 1 ? 1.75 : 2.5; }

// Iterations per unrolled block of a radius-R sweep with a load ring of `ring` rows (the
// 2R + 1 partial-sum slots and the ring both repeat), and the rows a plain-path chunk needs
// above it and below it inside the grid.
__host__ __device__ constexpr int reg_gcd(int a, int b) { return b == 0 ? a : reg_gcd(b, a % b); }
__host__ __device__ constexpr int reg_period(int R, int ring) { return (2 * R + 1) / reg_gcd(2 * R + 1, ring) * ring; }
__host__ __device__ constexpr int reg_least(int R, int kf, int ring) { return R * kf + reg_period(R, ring) + ring - 2; }

// The decomposition for a grid of ny x nx interior cells, `ox` output columns per strip,
// h = R kf warm-up rows on either side and `slots` resident waves: the smallest per-wave cost
// (rows plus the 2 h warm-up rows, checked rows counted `edge_cost` times) at which every wave is resident at
// once, so that no wave waits for a second round; checked chunks are at least `least` rows.
inline RegSplit reg_split(int64_t ny, int64_t nx, int ox, int h, int64_t slots, double edge_cost, int least) {
  RegSplit sp{};
  sp.strips = (nx + ox - 1) / ox;
  const int64_t warm = 2 * (int64_t)h;
  auto make = [&](int64_t cost) {
    RegSplit t = sp;
    t.hm = cost - warm;
    t.he = (int64_t)((double)cost / edge_cost) - warm;
    if (t.he < least) t.he = least;  // plain-path chunks start and end `least` rows inside
    if (t.hm < 1) t.hm = 1;
    t.we = (ny + t.he - 1) / t.he;
    if (ny <= 2 * t.he) {
      t.wi = t.we;
      t.hm = 0;
    } else {
      t.wi = 2 + (ny - 2 * t.he + t.hm - 1) / t.hm;
    }
    return t;
  };
  int64_t lo = warm + 1, hi = ny + warm;
  if (reg_waves(make(hi)) > slots) return make(hi);  // more strips than slots: one chunk per strip
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (reg_waves(make(mid)) <= slots)
      hi = mid;
    else
      lo = mid + 1;
  }
  return make(lo);
}

typedef hipError_t (*LaunchFn)(const double *in, double *out, const double *bfrag, const Geo &g, hipStream_t s);

// One launch of the register-streaming 3-D kernels (k3d_reg.hip): strips x y-tiles x
// z-chunks of zc planes; wave w takes y-tile w mod ytiles (neighbouring tiles share two
// input rows and run together), then strip, then z-chunk.  As many waves as are resident
// at once, with z-chunks of at least 8 planes.
struct Reg3Split {
  int64_t strips, ytiles, zchunks, zc;
};

inline Reg3Split reg3_split(int64_t nz, int64_t ny, int64_t nx, int ox, int by, int64_t slots) {
  Reg3Split sp{};
  sp.strips = (nx + ox - 1) / ox;
  sp.ytiles = (ny + by - 1) / by;
  int64_t zch = slots / (sp.strips * sp.ytiles);
  const int64_t most = nz / 8;
  if (zch > most) zch = most;
  if (zch < 1) zch = 1;
  sp.zc = (nz + zch - 1) / zch;
  sp.zchunks = (nz + sp.zc - 1) / sp.zc;
  return sp;
}

// Several grids in one launch of a register-streaming 2-D kernel (the batched workload):
// wave w sweeps rows [y0, y1) of strip `strip` of grid items[w].grid, from that grid's
// buffer cur ^ flip into the other one.
struct RegGrid {
  double *buf[2];
  const double *wts;  // dense weights
  Geo geo;
  int cur;  // buffer holding the grid's state at the start of the run
  int pad;
};
struct RegItem {
  int grid, strip, y0, y1;
};
struct RegMulti {
  const RegGrid *grids;
  const RegItem *items;  // null: a single-grid launch
  int nitems, flip;
};
typedef hipError_t (*LaunchMultiFn)(const RegMulti &mm, hipStream_t s);

// One kernel variant of the FP64 stencil2row families.
struct Variant {
  int id;
  int mf;      // MFMA shape: 16 = v_mfma_f64_16x16x4, 4 = v_mfma_f64_4x4x4; 1, 2 = the
               // --breakdown ladder's FP64 FMA kernels (global, LDS)
  int R, shape, kf, by, bx, nw;
  size_t lds_bytes;
  int px;  // LDS row pitch in doubles
  int st;  // fused-step staging: 0 = two LDS buffers, 1 = one buffer + registers;
           // split-precision: 0 = Toeplitz form, 2 = dilated form
  LaunchFn launch;  // one launch = kf fused steps from `in` to `out`
  hipError_t (*prepare)();  // one-time function attributes (dynamic LDS above 64 KiB)
  const char *name;
  int cs = 0;  // 2-D 4x4x4: LDS storage column of tile column 0
  LaunchMultiFn launch_multi = nullptr;  // register-streaming 2-D: several grids per launch
};

const Variant *variants_1d_f64(int *count);
const Variant *variants_1d_reg(int *count);   // register-resident FP64 1-D (k1d_reg.hip; mf 6)
const Variant *variants_2d_f64(int *count);
const Variant *variants_2d_reg(int *count);     // register-streaming FP64 2-D, radius 1 (k2d_reg.hip; mf 7)
const Variant *variants_2d_stream(int *count);  // pipelined persistent 2-D FP64 (k2d_stream.hip)
const Variant *variants_3d_f64(int *count);  // all 3-D FP64 variants: k3d2 (mf 8) first, then k3d (mf 4)
const Variant *variants_3d_reg(int *count);  // register-streaming FP64 3-D single steps (k3d_reg.hip; mf 9)
const Variant *variants_3d2_f64(int *count);  // 8 x 8 block family only
const Variant *variants_2d_split(int *count);
const Variant *variants_2d_ablation(int *count);  // --breakdown rungs 1-2 (mf 1, 2; forced by id only)
const Variant *variants_1d_ablation(int *count);  // the same for 1-D
const Variant *variants_3d_ablation(int *count);  // the same for 3-D
const Variant *variants_3d_fma(int *count);       // dense FP64 FMA, radius 2-3 (mf 3)
const Variant *variants_3d_split(int *count);  // split-precision 3-D (k3d_split.hip)

// Host-side B-fragment construction for the 2-D FP64 family: bfrag[ks * 64 + lane].
void build_bfrag_2d_f64(int mf, int R, int shape, const double *w_dense, double *bfrag);
int bfrag_ksteps_2d_f64(int mf, int R, int shape);
void build_bfrag_1d_f64(int R, const double *w_dense, double *bfrag);
int bfrag_ksteps_1d_f64(int R);
void build_bfrag_3d_f64(int R, int shape, const double *w_dense, double *bfrag);
void build_bfrag_3d2_f64(int R, int shape, const double *w_dense, double *bfrag);
int bfrag_ksteps_3d2_f64(int R, int shape);
// Split-precision B fragments (prec: 1 BF16x3, 2 FP32x2, 3 FP16x2); returns bytes written, 0 if cap is short.
size_t build_bfrag_split(int prec, int R, const double *w_dense, void *out, size_t cap);
// A fragments of the dilated split form (Variant::st == 2; prec 1 BF16x3 or 3 FP16x2).
size_t build_bfrag_split2(int prec, int R, int shape, const double *w_dense, void *out, size_t cap);
int bfrag_ksteps_3d_f64(int R, int shape);
// A fragments of the split-precision 3-D kernels (prec 1 BF16x3 or 4 FP16 w3).
size_t build_bfrag_split3(int prec, int R, int shape, const double *w_dense, void *out, size_t cap);

// Device-side synthetic input fill (same bits as the host generator).
hipError_t launch_fill_random(double *a, const Geo &g, uint64_t stream_seed, hipStream_t s);
hipError_t launch_zero(double *a, size_t count, hipStream_t s);
hipError_t launch_copy(const double *src, double *dst, size_t count, hipStream_t s);

}  // namespace cs
