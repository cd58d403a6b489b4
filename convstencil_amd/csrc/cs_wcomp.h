// cs_wcomp.h -- host side: weight slices whose rounding errors cancel in the low moments.
//
// A split form represents each weight w_k by slices s0 + s1 / S + ... of a narrow type.
// The error d_k = w_k - (slices) is the same at every time step, so over T steps the
// moments that every step repeats grow with T: the sum of d (mass added or removed per
// step) and its first moments sum d * (dz, dy, dx) (a drift).  Higher moments change the
// diffusion per step and do not accumulate that way.  The last slice of each tap is
// moved by whole ulps of its type (at most kMaxMove either way from the rounded value;
// zero taps stay zero) to make |sum d| + sum over axes |sum d * offset| as small as the
// ulp grid allows: a local search over single moves and pair moves (one tap up, another
// down), taking the best improvement each round.  Pair moves trade the two taps' ulps
// against each other, which reaches residuals far below either ulp and leaves local
// minima that single moves cannot.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace cs {

// Adjacent representable value of a 16-bit floating type (FP16 or BF16) in direction dir.
template <class T>
inline T step16(T v, int dir) {
  static_assert(sizeof(T) == 2, "16-bit types only");
  uint16_t b;
  std::memcpy(&b, &v, 2);
  const bool neg = (b & 0x8000u) != 0, zero = (b & 0x7fffu) == 0;
  if (zero) {
    b = (uint16_t)(dir > 0 ? 0x0001u : 0x8001u);  // smallest subnormal of the direction's sign
  } else if (neg == (dir < 0)) {
    b = (uint16_t)(b + 1);  // away from zero
  } else {
    b = (uint16_t)(b - 1);  // towards zero
  }
  T r;
  std::memcpy(&r, &b, 2);
  return r;
}

// taps: n weights w[k] with offsets off[k][0..2] (dz, dy, dx); fixed[k] = the value of the
// slices that stay as rounded; low[k] = the last slice (moved in place); the last slice
// carries weight 1 / scale.  Returns the final objective.
template <class T>
inline double compensate_low_slice(int n, const double *w, const int (*off)[3], const double *fixed, T *low,
                                   double scale) {
  constexpr int kMaxMove = 4, kSpan = 2 * kMaxMove + 1;
  std::vector<double> dv((size_t)n * kSpan);  // d_k for move a in [-kMaxMove, kMaxMove]
  std::vector<T> lv((size_t)n * kSpan);
  std::vector<int> mv((size_t)n, 0);
  for (int k = 0; k < n; ++k) {
    T up = low[k], dn = low[k];
    lv[(size_t)k * kSpan + kMaxMove] = low[k];
    for (int a = 1; a <= kMaxMove; ++a) {
      up = step16(up, +1);
      dn = step16(dn, -1);
      lv[(size_t)k * kSpan + kMaxMove + a] = up;
      lv[(size_t)k * kSpan + kMaxMove - a] = dn;
    }
    for (int a = 0; a < kSpan; ++a) {
      const double l = (double)(float)lv[(size_t)k * kSpan + a];
      dv[(size_t)k * kSpan + a] = (w[k] == 0.0 || !std::isfinite(l)) ? NAN : w[k] - (fixed[k] + l / scale);
    }
    if (w[k] == 0.0) dv[(size_t)k * kSpan + kMaxMove] = 0.0;  // zero taps stay zero
  }
  double r[4] = {0, 0, 0, 0};
  auto d_of = [&](int k, int a) { return dv[(size_t)k * kSpan + kMaxMove + a]; };
  for (int k = 0; k < n; ++k) {
    const double d = d_of(k, 0);
    r[0] += d;
    for (int c = 0; c < 3; ++c) r[c + 1] += d * off[k][c];
  }
  auto cost = [](const double *q) { return std::fabs(q[0]) + std::fabs(q[1]) + std::fabs(q[2]) + std::fabs(q[3]); };
  double cur = cost(r);
  for (int iter = 0; iter < 64 * n && cur > 0.0; ++iter) {
    double best = cur, bq[4] = {0, 0, 0, 0};
    int bk1 = -1, ba1 = 0, bk2 = -1, ba2 = 0;
    auto delta = [&](int k, int na, double *q) {  // residual change when tap k moves to na
      if (na < -kMaxMove || na > kMaxMove || w[k] == 0.0) return false;
      const double nd = d_of(k, na);
      if (!std::isfinite(nd)) return false;
      const double dd = nd - d_of(k, mv[(size_t)k]);
      q[0] += dd;
      for (int c = 0; c < 3; ++c) q[c + 1] += dd * off[k][c];
      return true;
    };
    for (int k1 = 0; k1 < n; ++k1)
      for (int s1 = -1; s1 <= 1; s1 += 2) {
        double q1[4] = {r[0], r[1], r[2], r[3]};
        if (!delta(k1, mv[(size_t)k1] + s1, q1)) continue;
        if (cost(q1) < best) {
          best = cost(q1);
          std::memcpy(bq, q1, sizeof bq);
          bk1 = k1, ba1 = mv[(size_t)k1] + s1, bk2 = -1;
        }
        for (int k2 = k1 + 1; k2 < n; ++k2)
          for (int s2 = -1; s2 <= 1; s2 += 2) {
            double q2[4] = {q1[0], q1[1], q1[2], q1[3]};
            if (!delta(k2, mv[(size_t)k2] + s2, q2)) continue;
            if (cost(q2) < best) {
              best = cost(q2);
              std::memcpy(bq, q2, sizeof bq);
              bk1 = k1, ba1 = mv[(size_t)k1] + s1, bk2 = k2, ba2 = mv[(size_t)k2] + s2;
            }
          }
      }
    if (bk1 < 0) break;
    mv[(size_t)bk1] = ba1;
    if (bk2 >= 0) mv[(size_t)bk2] = ba2;
    std::memcpy(r, bq, sizeof r);
    cur = best;
  }
  for (int k = 0; k < n; ++k) low[k] = lv[(size_t)k * kSpan + kMaxMove + mv[(size_t)k]];
  return cur;
}

}  // namespace cs
