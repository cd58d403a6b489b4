// k1d_reg.hip -- register-resident FP64 1-D kernels for gfx950: vector FMAs, the cells of
// neighbouring lanes exchanged by DPP wave shifts; no LDS and no barriers.
//
// Replaces gpu_1d1r_kernel (reference src/1d/gpu_1r.cu:26-75) and
// gpu_star_1d2r_step2_kernel (src/1d/gpu_2r.cu:23-77) for fused runs.  On gfx950 FP64
// vector FMAs and FP64 MFMAs share one peak (128 flops per clock and CU), and a 1-D
// stencil needs 2R + 1 multiply-adds per cell and step; the 4 x 4 x 4 MFMA blocks of
// k1d_f64.hip pay an LDS round trip and a workgroup barrier every step instead.
//
// Each wave owns a segment of 64 P cells, lane l the P consecutive cells l P .. l P + P - 1,
// held in registers for the whole launch.  A step needs the R cells on either side of a
// lane's run: they come from lanes l - 1 and l + 1 through v_mov_b32_dpp wave_shr:1 /
// wave_shl:1 (lanes 0 and 63 receive zeros), so the valid cells shrink by R from each end
// of the segment per step and after KF steps the central O = 64 P - 2 R KF cells are the
// outputs.  Segments overlap by 2 R KF recomputed cells; waves never communicate.  Cells
// outside the interior keep their value every step (the halo is never written); cells
// outside the array load as zero and no interior cell ever reads them.
#include "cs_internal.h"

namespace cs {

template <int R, int P, int KF>
struct Reg1D {
  static constexpr int H = R * KF;          // recomputed cells on each side of a segment
  static constexpr int O = 64 * P - 2 * H;  // outputs per wave
  static_assert(P >= R, "a lane's run must cover the radius");
  static_assert(O > 0, "segment too short for the fusion depth");
};

// One step from a to b: b[j] = sum_d w[d] a[j - R + d], the R cells beyond either end of
// the run from the neighbouring lanes; cells outside the interior keep their value.
template <int R, int P>
__device__ __forceinline__ void step1d(const double (&a)[P], double (&b)[P], const double (&w)[2 * R + 1], bool full,
                                       int64_t c0, int64_t nx) {
  double lo[R], hi[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    lo[r] = lane_from_lower(a[P - R + r]);
    hi[r] = lane_from_upper(a[r]);
  }
#pragma unroll
  for (int j = 0; j < P; ++j) {
    // left-to-right over the taps, as the reference's loop sums them
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d <= 2 * R; ++d) {
      const int k = j - R + d;
      const double v = k < 0 ? lo[k + R] : k >= P ? hi[k - P] : a[k];
      acc = d == 0 ? w[0] * v : __builtin_fma(w[d], v, acc);
    }
    b[j] = acc;
  }
  if (!full) {
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const int64_t c = c0 + j;
      b[j] = c >= 0 && c < nx ? b[j] : a[j];
    }
  }
}

template <int R, int P, int KF, int NW>
__global__ __launch_bounds__(NW * 64) void k1d_reg(const double *__restrict__ in, double *__restrict__ out,
                                                   const double *__restrict__ bfrag, Geo g) {
  using C = Reg1D<R, P, KF>;
  const int lane = threadIdx.x & 63;
  const int64_t seg = (int64_t)blockIdx.x * NW + (threadIdx.x >> 6);
  const int64_t nx = g.n[2], ext = g.ext[2], org = g.org[2];
  if (seg * C::O >= nx) return;  // wave-uniform: the whole segment lies past the interior
  const int64_t X0 = seg * C::O - C::H;  // interior coordinate of the segment's cell 0
  const int64_t c0 = X0 + (int64_t)lane * P;

  // the 4 x 4 x 4 B fragment's entry (k-row f, column 0) is w[f] (build_bfrag_1d_f64)
  double w[2 * R + 1];
#pragma unroll
  for (int f = 0; f <= 2 * R; ++f) w[f] = bfrag[(f >> 2) * 64 + ((f & 3) << 4)];

  double x[P], y[P];
  if (org + X0 >= 0 && org + X0 + 64 * P <= ext) {
    const double *src = in + org + c0;
#pragma unroll
    for (int j = 0; j < P; ++j) x[j] = src[j];
  } else {
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const int64_t a = org + c0 + j;
      const bool ok = a >= 0 && a < ext;
      const double v = in[ok ? a : 0];
      x[j] = ok ? v : 0.0;
    }
  }

  // steps in pairs x -> y -> x, so that no step copies its result back
  const bool full = X0 >= 0 && X0 + 64 * P <= nx;  // wave-uniform: every cell is interior
  if (KF & 1) {
    step1d<R, P>(x, y, w, full, c0, nx);
#pragma unroll
    for (int j = 0; j < P; ++j) x[j] = y[j];
  }
  for (int s = 0; s < KF / 2; ++s) {
    step1d<R, P>(x, y, w, full, c0, nx);
    step1d<R, P>(y, x, w, full, c0, nx);
  }

#pragma unroll
  for (int j = 0; j < P; ++j) {
    const int d = lane * P + j;
    const int64_t c = c0 + j;  // >= 0 wherever d >= H
    if (d >= C::H && d < C::H + C::O && c < nx) out[org + c] = x[j];
  }
}

static hipError_t prepare_reg1d() { return hipSuccess; }

template <int R, int P, int KF, int NW>
static hipError_t launch_reg1d(const double *in, double *out, const double *bfrag, const Geo &g, hipStream_t s) {
  hipLaunchKernelGGL((k1d_reg<R, P, KF, NW>), dim3((unsigned)g.tiles[2]), dim3(NW * 64), 0, s, in, out, bfrag, g);
  return hipGetLastError();
}

// Variant::mf 6 marks this family; bx is the outputs per workgroup, px the cells per lane.
#define CS_VR1D(ID, R, P, KF, NW)                                                                          \
  Variant {                                                                                                \
    ID, 6, R, SHAPE_BOX, KF, 1, NW * Reg1D<R, P, KF>::O, NW, 0, P, 0, &launch_reg1d<R, P, KF, NW>,         \
        &prepare_reg1d, "k1d_reg<r" #R ",p" #P ",kf" #KF ",w" #NW ">"                                    \
  }

// ids 600000 + R * 10000 + P * 100 + KF
static const Variant kVariantsReg1D[] = {
    CS_VR1D(611864, 1, 18, 64, 4), CS_VR1D(611832, 1, 18, 32, 4), CS_VR1D(611816, 1, 18, 16, 4),
    CS_VR1D(611808, 1, 18, 8, 4),  CS_VR1D(611804, 1, 18, 4, 4),  CS_VR1D(611802, 1, 18, 2, 4),
    CS_VR1D(611801, 1, 18, 1, 4),  CS_VR1D(611032, 1, 10, 32, 4), CS_VR1D(612664, 1, 26, 64, 4),
    // shorter segments: more waves per SIMD (FP64 issue) for more recomputed cells
    CS_VR1D(611064, 1, 10, 64, 4), CS_VR1D(611264, 1, 12, 64, 4), CS_VR1D(610864, 1, 8, 64, 4),
    CS_VR1D(610832, 1, 8, 32, 4),
    CS_VR1D(621832, 2, 18, 32, 4), CS_VR1D(621816, 2, 18, 16, 4), CS_VR1D(621808, 2, 18, 8, 4),
    CS_VR1D(621804, 2, 18, 4, 4),  CS_VR1D(621802, 2, 18, 2, 4),  CS_VR1D(621801, 2, 18, 1, 4),
    CS_VR1D(631816, 3, 18, 16, 4), CS_VR1D(631808, 3, 18, 8, 4),  CS_VR1D(631804, 3, 18, 4, 4),
    CS_VR1D(631802, 3, 18, 2, 4),  CS_VR1D(631801, 3, 18, 1, 4),  CS_VR1D(641816, 4, 18, 16, 4),
    CS_VR1D(641808, 4, 18, 8, 4),  CS_VR1D(641804, 4, 18, 4, 4),  CS_VR1D(641802, 4, 18, 2, 4),
    CS_VR1D(641801, 4, 18, 1, 4),
};

const Variant *variants_1d_reg(int *count) {
  *count = (int)(sizeof(kVariantsReg1D) / sizeof(kVariantsReg1D[0]));
  return kVariantsReg1D;
}

}  // namespace cs
