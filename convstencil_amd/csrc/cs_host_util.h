// cs_host_util.h -- host-side helpers shared by the plan and batch runtimes.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/convstencil.h"
#include "cs_internal.h"

namespace cs {
namespace host {

#define CS_HIP(call)                                                                                        \
  do {                                                                                                      \
    hipError_t e_ = (call);                                                                                 \
    if (e_ != hipSuccess) {                                                                                 \
      fprintf(stderr, "convstencil: HIP error %d (%s) at %s:%d: %s\n", (int)e_, hipGetErrorString(e_), __FILE__, \
              __LINE__, #call);                                                                             \
      return CS_ERR_DEVICE;                                                                                 \
    }                                                                                                       \
  } while (0)

#define CS_TRY(expr)                  \
  do {                                \
    cs_status s_ = (expr);            \
    if (s_ != CS_OK) return s_;       \
  } while (0)

inline bool layout_valid(const cs_layout *L, int R) {
  if (!L || L->ndim < 1 || L->ndim > 3) return false;
  for (int d = 0; d < 3; ++d) {
    const bool active = d >= 3 - L->ndim;
    if (!active) {
      if (L->ext[d] != 1 || L->n[d] != 1 || L->org[d] != 0) return false;
      continue;
    }
    if (L->ext[d] <= 0 || L->n[d] < 0 || L->org[d] < R) return false;
    if (L->org[d] + L->n[d] + R > L->ext[d]) return false;
  }
  const double cells = (double)L->ext[0] * (double)L->ext[1] * (double)L->ext[2];
  return cells < 9.0e18 / 8.0;
}

inline int64_t cell_count(const cs_layout *L) { return L->ext[0] * L->ext[1] * L->ext[2]; }

// 2-D shape class from the dense weights: a star has no off-axis taps.
inline int classify_2d(int R, const double *w) {
  const int D = 2 * R + 1;
  for (int dy = -R; dy <= R; ++dy)
    for (int dx = -R; dx <= R; ++dx)
      if (dy != 0 && dx != 0 && w[(dy + R) * D + (dx + R)] != 0.0) return SHAPE_BOX;
  return SHAPE_STAR;
}

// 3-D shape class: star (taps on the axes), diamond (|dz|+|dy|+|dx| <= R) or box.
inline int classify_3d(int R, const double *w) {
  const int D = 2 * R + 1;
  bool star = true, diamond = true;
  for (int dz = -R; dz <= R; ++dz)
    for (int dy = -R; dy <= R; ++dy)
      for (int dx = -R; dx <= R; ++dx) {
        if (w[((dz + R) * D + (dy + R)) * D + (dx + R)] == 0.0) continue;
        const int nz = (dz != 0) + (dy != 0) + (dx != 0);
        if (nz > 1) star = false;
        if (std::abs(dz) + std::abs(dy) + std::abs(dx) > R) diamond = false;
      }
  return star ? SHAPE3_STAR : diamond ? SHAPE3_DIAMOND : SHAPE3_BOX;
}

// True when every halo cell of the padded host array is zero.
inline bool halo_is_zero(const cs_layout *L, const double *a) {
  for (int64_t z = 0; z < L->ext[0]; ++z)
    for (int64_t y = 0; y < L->ext[1]; ++y) {
      const bool row_in = z >= L->org[0] && z < L->org[0] + L->n[0] && y >= L->org[1] && y < L->org[1] + L->n[1];
      const double *row = a + (z * L->ext[1] + y) * L->ext[2];
      for (int64_t x = 0; x < L->ext[2]; ++x) {
        if (row_in && x >= L->org[2] && x < L->org[2] + L->n[2]) {
          x = L->org[2] + L->n[2] - 1;
          continue;
        }
        if (row[x] != 0.0) return false;
      }
    }
  return true;
}

// Largest magnitude in a[0 .. n); +inf if any value is not finite.
inline double abs_max(const double *a, size_t n) {
  double m = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double v = std::fabs(a[i]);
    if (!(v <= m)) m = std::isfinite(v) ? v : HUGE_VAL;
  }
  return m;
}

inline double abs_sum(const double *w, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += std::fabs(w[i]);
  return s;
}

// FP16 split forms hold every value as FP16 slices, so the state must stay below 2^15 in
// magnitude.  One step maps a state bounded by a to one bounded by a * sum|w|, so `steps`
// steps from a bound amax end below amax * max(1, sum|w|)^steps (*bound_out).  Other
// precisions always pass.
inline bool fp16_range_ok(int precision, double amax, double wabs, int64_t steps, double *bound_out) {
  double b = amax;
  if (wabs > 1.0 && b > 0.0) b *= std::pow(wabs, (double)steps);  // inf when it overflows
  *bound_out = b;
  if (precision != CS_PREC_FP16X2 && precision != CS_PREC_FP16W3) return true;
  return b < 32768.0;
}

inline cs_status check_device(int *dev_out) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return CS_ERR_NO_DEVICE;
  int dev = 0;
  CS_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  CS_HIP(hipGetDeviceProperties(&prop, dev));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    fprintf(stderr, "convstencil: device %d is %s, this library is built for gfx950\n", dev, prop.gcnArchName);
    return CS_ERR_NO_DEVICE;
  }
  *dev_out = dev;
  return CS_OK;
}

}  // namespace host
}  // namespace cs
