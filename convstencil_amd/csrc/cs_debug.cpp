// cs_debug.cpp -- host-only table introspection (include/convstencil_debug.h).
#include "../../include/convstencil_debug.h"

#include <cstring>

#include "cs_fused_weights.h"
#include "cs_host_util.h"
#include "cs_internal.h"

extern "C" {

static const cs::Variant *family(int32_t ndim, int *n) {
  *n = 0;
  if (ndim == 4) return cs::variants_2d_split(n);
  if (ndim == 5) return cs::variants_3d_split(n);
  if (ndim == 1) return cs::variants_1d_f64(n);
  if (ndim == 2) return cs::variants_2d_f64(n);
  if (ndim == 3) return cs::variants_3d_f64(n);
  if (ndim == 6) return cs::variants_2d_stream(n);
  if (ndim == 7) return cs::variants_1d_reg(n);
  if (ndim == 8) return cs::variants_2d_reg(n);
  if (ndim == 9) return cs::variants_3d_reg(n);
  return nullptr;
}

int32_t cs_debug_variant_count(int32_t ndim) {
  int n = 0;
  (void)family(ndim, &n);
  return n;
}

int32_t cs_debug_variant_info(int32_t ndim, int32_t i, int64_t *info, char *name, int32_t len) {
  int n = 0;
  const cs::Variant *v = family(ndim, &n);
  if (i < 0 || i >= n || !info) return -1;
  const cs::Variant &x = v[i];
  const int64_t vals[12] = {x.id, x.mf, x.R,  x.shape, x.kf, x.by, x.bx, x.nw, (int64_t)x.lds_bytes, x.px, x.st, x.cs};
  memcpy(info, vals, sizeof(vals));
  if (name && len > 0) {
    strncpy(name, x.name, (size_t)len - 1);
    name[len - 1] = 0;
  }
  return 0;
}

int32_t cs_debug_ktable(int32_t mf, int32_t radius, int32_t shape, int32_t *dy, int32_t *c0, int32_t *live,
                        int32_t cap) {
  const cs::KTable t = mf == 16 ? cs::make_ktable(radius, shape) : cs::make_ftable(radius, shape);
  if (t.n > cap) return -1;
  for (int i = 0; i < t.n; ++i) {
    dy[i] = t.dy[i];
    c0[i] = t.c0[i];
    live[i] = t.live[i];
  }
  return t.n;
}

int32_t cs_debug_dtable(int32_t radius, int32_t shape, int32_t *hdr, int32_t *ty, int32_t *tx, int32_t *live,
                        int32_t cap) {
  const cs::DTable t = cs::make_dtable(radius, shape);
  const int n = 4 * (t.np + t.ns);
  if (n > cap || !hdr) return -1;
  hdr[0] = t.np;
  hdr[1] = t.ns;
  hdr[2] = t.pp;
  hdr[3] = t.map;
  for (int ps = 0; ps < t.np; ++ps)
    for (int q = 0; q < 4; ++q) {
      ty[4 * ps + q] = t.pty[ps][q];
      tx[4 * ps + q] = t.ptx[ps][q];
      live[4 * ps + q] = t.plive[ps][q];
    }
  for (int ss = 0; ss < t.ns; ++ss)
    for (int q = 0; q < 4; ++q) {
      ty[4 * (t.np + ss) + q] = t.sty[ss][q];
      tx[4 * (t.np + ss) + q] = t.stx[ss][q];
      live[4 * (t.np + ss) + q] = t.slive[ss][q];
    }
  return n;
}

int32_t cs_debug_bfrag_2d(int32_t mf, int32_t radius, int32_t shape, const double *w, double *bfrag, int32_t cap) {
  const int nks = cs::bfrag_ksteps_2d_f64(mf, radius, shape);
  if (nks * 64 > cap) return -1;
  cs::build_bfrag_2d_f64(mf, radius, shape, w, bfrag);
  return nks;
}

int32_t cs_debug_bfrag_1d(int32_t radius, const double *w, double *bfrag, int32_t cap) {
  const int nks = cs::bfrag_ksteps_1d_f64(radius);
  if (nks * 64 > cap) return -1;
  cs::build_bfrag_1d_f64(radius, w, bfrag);
  return nks;
}

int32_t cs_debug_ktable3d(int32_t radius, int32_t shape, int32_t *dz, int32_t *dy, int32_t *c, int32_t *live,
                          int32_t cap) {
  const cs::KTable3D t = cs::make_ktable3d(radius, shape);
  if (t.n > cap) return -1;
  for (int i = 0; i < t.n; ++i) {
    dz[i] = t.dz[i];
    for (int k = 0; k < 4; ++k) {
      dy[4 * i + k] = t.dy[i][k];
      c[4 * i + k] = t.c[i][k];
      live[4 * i + k] = t.live[i][k];
    }
  }
  return t.n;
}

int32_t cs_debug_bfrag_3d(int32_t radius, int32_t shape, const double *w, double *bfrag, int32_t cap) {
  const int nks = cs::bfrag_ksteps_3d_f64(radius, shape);
  if (nks * 64 > cap) return -1;
  cs::build_bfrag_3d_f64(radius, shape, w, bfrag);
  return nks;
}

int32_t cs_debug_fp16_range(int32_t precision, double amax, double wabs, int64_t steps, double *bound) {
  double b = 0.0;
  const bool ok = cs::host::fp16_range_ok(precision, amax, wabs, steps, &b);
  if (bound) *bound = b;
  return ok ? 1 : 0;
}

double cs_debug_abs_max(const double *a, int64_t n) { return cs::host::abs_max(a, (size_t)n); }

void cs_debug_fused_2d(int32_t exact, const double *p9, double *out49) {
  if (exact)
    cs::fused_2d_exact(p9, out49);
  else
    cs::fused_2d_reference(p9, out49);
}

int64_t cs_debug_bfrag_split(int32_t precision, int32_t radius, const double *w, void *out, int64_t cap) {
  return (int64_t)cs::build_bfrag_split(precision, radius, w, out, (size_t)cap);
}

int64_t cs_debug_bfrag_split2(int32_t precision, int32_t radius, int32_t shape, const double *w, void *out,
                              int64_t cap) {
  return (int64_t)cs::build_bfrag_split2(precision, radius, shape, w, out, (size_t)cap);
}

int32_t cs_debug_ctable(int32_t radius, int32_t shape, int32_t *c0, int32_t *ty, int32_t *tx, int32_t *live,
                        int32_t cap) {
  const cs::CTable t = cs::make_ctable(radius, shape);
  if (t.n > cap || !c0) return -1;
  *c0 = t.c0;
  for (int i = 0; i < t.n; ++i) {
    ty[i] = t.ty[i];
    tx[i] = t.tx[i];
    live[i] = t.live[i];
  }
  return t.n;
}

int32_t cs_debug_reg2d_split(int64_t ny, int64_t nx, int32_t ox, int32_t radius, int32_t kf, int32_t ring,
                             int64_t slots, int64_t *split) {
  if (ny < 1 || nx < 1 || ox < 1 || radius < 1 || kf < 1 || ring < 2 || slots < 1 || !split) return -1;
  const int px = (ox + 2 * radius * kf) / 64;  // strips of 64 px columns
  const cs::RegSplit sp = cs::reg_split(ny, nx, ox, radius * kf, slots, cs::reg_edge_cost(px, radius),
                                        cs::reg_least(radius, kf, ring));
  const int64_t v[5] = {sp.strips, sp.he, sp.hm, sp.we, sp.wi};
  memcpy(split, v, sizeof(v));
  return 0;
}

int32_t cs_debug_reg3d_split(int64_t nz, int64_t ny, int64_t nx, int32_t ox, int32_t by, int64_t slots,
                             int64_t *split) {
  if (nz < 1 || ny < 1 || nx < 1 || ox < 1 || by < 1 || slots < 1 || !split) return -1;
  const cs::Reg3Split sp = cs::reg3_split(nz, ny, nx, ox, by, slots);
  const int64_t v[4] = {sp.strips, sp.ytiles, sp.zchunks, sp.zc};
  memcpy(split, v, sizeof(v));
  return 0;
}

int32_t cs_debug_reg2d_chunk(const int64_t *split, int64_t ny, int64_t wid, int64_t *rows) {
  if (!split || !rows) return -1;
  const cs::RegSplit sp{split[0], split[1], split[2], split[3], split[4]};
  int64_t strip = 0, y0 = 0, y1 = 0;
  if (!cs::reg_chunk(sp, ny, wid, &strip, &y0, &y1)) return 0;
  rows[0] = strip;
  rows[1] = y0;
  rows[2] = y1;
  return 1;
}

int32_t cs_debug_ctable_map(int32_t radius, int32_t shape) { return cs::make_ctable(radius, shape).map; }

int32_t cs_debug_ctable3(int32_t radius, int32_t shape, int32_t dz, int32_t *c0, int32_t *ty, int32_t *tx,
                         int32_t *live, int32_t cap) {
  if (radius < 1 || radius > 3 || dz < -radius || dz > radius || !c0) return -1;
  const cs::CTable3 t3 = cs::make_ctable3(radius, shape);
  const cs::CTable &t = t3.t[dz + radius];
  if (t.n > cap) return -1;
  *c0 = t3.c0;
  for (int i = 0; i < t.n; ++i) {
    ty[i] = t.ty[i];
    tx[i] = t.tx[i];
    live[i] = t.live[i];
  }
  return t.n;
}

int64_t cs_debug_bfrag_split3(int32_t precision, int32_t radius, int32_t shape, const double *w, void *out,
                              int64_t cap) {
  return (int64_t)cs::build_bfrag_split3(precision, radius, shape, w, out, (size_t)cap);
}

int32_t cs_debug_dtable3(int32_t radius, int32_t shape, int32_t dz, int32_t *hdr, int32_t *ty, int32_t *tx,
                         int32_t *live, int32_t cap) {
  if (radius < 1 || radius > 3 || dz < -radius || dz > radius || !hdr) return -1;
  const cs::DTable3 t3 = cs::make_dtable3(radius, shape);
  const cs::DTable &t = t3.t[dz + radius];
  const int n = 4 * (t.np + t.ns);
  if (n > cap) return -1;
  hdr[0] = t.np;
  hdr[1] = t.ns;
  hdr[2] = t.pp;
  for (int ps = 0; ps < t.np; ++ps)
    for (int q = 0; q < 4; ++q) {
      ty[4 * ps + q] = t.pty[ps][q];
      tx[4 * ps + q] = t.ptx[ps][q];
      live[4 * ps + q] = t.plive[ps][q];
    }
  for (int ss = 0; ss < t.ns; ++ss)
    for (int q = 0; q < 4; ++q) {
      ty[4 * (t.np + ss) + q] = t.sty[ss][q];
      tx[4 * (t.np + ss) + q] = t.stx[ss][q];
      live[4 * (t.np + ss) + q] = t.slive[ss][q];
    }
  return n;
}

int32_t cs_debug_bfrag_3d2(int32_t radius, int32_t shape, const double *w, double *bfrag, int32_t cap) {
  const int nks = cs::bfrag_ksteps_3d2_f64(radius, shape);
  if (nks * 64 > cap) return -1;
  cs::build_bfrag_3d2_f64(radius, shape, w, bfrag);
  return nks;
}

int32_t cs_debug_copy_bench(int64_t count, int32_t reps, double *ms_per_copy) {
  // Plain device-to-device copy of `count` doubles (one element per thread), `reps`
  // times: the reference point for HBM byte counters (FETCH_SIZE / WRITE_SIZE).
  double *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, (size_t)count * sizeof(double)) != hipSuccess) return 3;
  if (hipMalloc(&b, (size_t)count * sizeof(double)) != hipSuccess) {
    (void)hipFree(a);
    return 3;
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipError_t e = cs::launch_zero(a, (size_t)count, nullptr);
  if (e == hipSuccess) e = cs::launch_copy(a, b, (size_t)count, nullptr);
  (void)hipEventRecord(e0, nullptr);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = cs::launch_copy(a, b, (size_t)count, nullptr);
  (void)hipEventRecord(e1, nullptr);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  if (ms_per_copy) *ms_per_copy = reps > 0 ? ms / reps : 0.0;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return e == hipSuccess ? 0 : 4;
}

}  // extern "C"
