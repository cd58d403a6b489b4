// cs_compat.cpp -- the reference's host launchers, by name, on the plan API.
//
// Each function keeps the reference's argument meaning, padded host layout,
// weight layout and stdout lines, and like the reference times the launch loop
// with the host clock around the launches plus one device synchronisation
// (src/2d/gpu.cu:316-329).  `times` counts launches; each launch applies the
// given (already time-fused) weights once, exactly as the reference kernels do.
//
// Halo semantics follow the reference's buffers (src/2d/gpu.cu:273-277): buffer 0
// holds the whole input including its halo, buffer 1 starts as zeros, step t
// reads buffer t % 2.  When the input halo is not zero the two halos differ, so
// the plan runs one step per launch (no in-kernel fusion) to keep that
// alternation exact.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/convstencil.h"

namespace {

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  long long ms() const {
    return std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
  }
  double secs() const {
    return std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count() / 1e6;
  }
};

int fail(const char *what, cs_status s) {
  fprintf(stderr, "convstencil: %s failed: %s\n", what, cs_status_string(s));
  fflush(stderr);
  return (int)s;
}

// Runs `times` single-launch steps on `dev_in` (padded, `cells` doubles) and copies
// the final padded array to `dev_out`.  Prints nothing; returns elapsed times.
int run_compat(const cs_layout &L, int radius, const std::vector<double> &dense, const double *in, double *out,
               int times, long long *ms, double *secs, int variant = 0) {
  cs_stencil st{L.ndim, radius, dense.data()};
  cs_options opt{};
  opt.precision = CS_PREC_FP64;
  opt.tile_variant = variant;
  cs_plan *p = nullptr;
  cs_status s = cs_plan_create(&p, &L, &st, &opt);
  if (s != CS_OK) return fail("plan creation", s);
  s = cs_plan_upload(p, in);
  if (s != CS_OK) {
    cs_plan_destroy(p);
    return fail("upload", s);
  }
  Timer t;
  if (times > 0) s = cs_plan_run(p, times, nullptr);
  if (s == CS_OK) s = cs_plan_sync(p);
  *ms = t.ms();
  *secs = t.secs();
  if (s == CS_OK) s = cs_plan_download(p, out);
  cs_plan_destroy(p);
  if (s != CS_OK) return fail("time loop", s);
  return 0;
}

cs_layout layout_1d(int64_t ext, int64_t org, int64_t n) {
  cs_layout L{};
  L.ndim = 1;
  L.ext[0] = L.ext[1] = 1;
  L.n[0] = L.n[1] = 1;
  L.ext[2] = ext;
  L.org[2] = org;
  L.n[2] = n;
  return L;
}

cs_layout layout_2d(int64_t rows, int64_t cols, int64_t oy, int64_t ox, int64_t m, int64_t n) {
  cs_layout L{};
  L.ndim = 2;
  L.ext[0] = 1;
  L.n[0] = 1;
  L.ext[1] = rows;
  L.ext[2] = cols;
  L.org[1] = oy;
  L.org[2] = ox;
  L.n[1] = m;
  L.n[2] = n;
  return L;
}

cs_layout layout_3d(int64_t h, int64_t m, int64_t n, int64_t halo) {
  cs_layout L{};
  L.ndim = 3;
  L.ext[0] = h + 2 * halo;
  L.ext[1] = m + 2 * halo;
  L.ext[2] = n + 2 * halo;
  L.org[0] = L.org[1] = L.org[2] = halo;
  L.n[0] = h;
  L.n[1] = m;
  L.n[2] = n;
  return L;
}

bool bad_sizes(int times, std::initializer_list<int> sizes) {
  if (times < 0) return true;
  for (int v : sizes)
    if (v < 0) return true;
  return false;
}

}  // namespace

extern "C" {

// src/1d/gpu_1r.cu:227-301: device array of n + 7, interior [4, n + 4), 7 weights,
// host output = device array from index 1.
int gpu_1d1r(const double *in, double *out, const double *params, int times, int input_n) {
  if (!in || !out || !params || bad_sizes(times, {input_n})) return fail("gpu_1d1r", CS_ERR_INVALID);
  const int64_t ext = (int64_t)input_n + 7;
  const cs_layout L = layout_1d(ext, 4, input_n);
  std::vector<double> w(params, params + 7), fin((size_t)ext);
  long long ms = 0;
  double secs = 0;
  int rc = run_compat(L, 3, w, in, fin.data(), times, &ms, &secs);
  if (rc) return rc;
  memcpy(out, fin.data() + 1, (size_t)(ext - 1) * sizeof(double));
  printf("ConvStencil(1D): \nTime = %lld[ms]\n", ms);
  printf("GStencil/s = %f\n", ((double)input_n * times * 3) / secs / 1e9);
  fflush(stdout);
  return 0;
}

// The 1-D --breakdown ladder (src/1d/gpu_1r.cu:305-651, run by src/1d/main.cu:224-228):
// gpu_1d1r's launch loop and result on one rung's kernel (k1d_ablation.hip, k1d_f64.hip):
// 1 FP64 FMA from HBM/L2, 2 FP64 FMA from an LDS tile, 3 the MFMA kernel on the reference's
// 1024-output tiling, 4 the production single-step kernel.
static int run_breakdown_1d(int rung, int variant, const double *in, double *out, const double *params, int times,
                            int input_n) {
  if (!in || !out || !params || bad_sizes(times, {input_n})) return fail("gpu_1d1r_breakdown", CS_ERR_INVALID);
  const int64_t ext = (int64_t)input_n + 7;
  const cs_layout L = layout_1d(ext, 4, input_n);
  std::vector<double> w(params, params + 7), fin((size_t)ext);
  long long ms = 0;
  double secs = 0;
  int rc = run_compat(L, 3, w, in, fin.data(), times, &ms, &secs, variant);
  if (rc) return rc;
  memcpy(out, fin.data() + 1, (size_t)(ext - 1) * sizeof(double));
  printf("Experiment - Breakdown(1D) %d: \nTime = %lld[ms]\n", rung, ms);
  printf("GStencil/s = %f\n\n", ((double)input_n * times * 3) / secs / 1e9);
  fflush(stdout);
  return 0;
}

int gpu_1d1r_breakdown1(const double *in, double *out, const double *params, int times, int input_n) {
  return run_breakdown_1d(1, 913000, in, out, params, times, input_n);
}
int gpu_1d1r_breakdown2(const double *in, double *out, const double *params, int times, int input_n) {
  return run_breakdown_1d(2, 923000, in, out, params, times, input_n);
}
int gpu_1d1r_breakdown3(const double *in, double *out, const double *params, int times, int input_n) {
  return run_breakdown_1d(3, 933001, in, out, params, times, input_n);
}
int gpu_1d1r_breakdown4(const double *in, double *out, const double *params, int times, int input_n) {
  return run_breakdown_1d(4, 13001, in, out, params, times, input_n);
}

// src/1d/gpu_2r.cu:85-159: device array of n + 8, interior [4, n + 4), 9 weights,
// host output = the first n + 7 device elements.
int gpu_star_1d2r_step2(const double *in, double *out, const double *params, int times, int input_n) {
  if (!in || !out || !params || bad_sizes(times, {input_n})) return fail("gpu_star_1d2r_step2", CS_ERR_INVALID);
  const int64_t ext = (int64_t)input_n + 8;
  const cs_layout L = layout_1d(ext, 4, input_n);
  std::vector<double> w(params, params + 9), fin((size_t)ext);
  long long ms = 0;
  double secs = 0;
  int rc = run_compat(L, 4, w, in, fin.data(), times, &ms, &secs);
  if (rc) return rc;
  memcpy(out, fin.data(), (size_t)(ext - 1) * sizeof(double));
  printf("Time = %lld[ms]\n", ms);
  printf("GStencil/s = %f\n", ((double)input_n * times * 2) / secs / 1e9);
  fflush(stdout);
  return 0;
}

static int run_2d(const double *in, double *out, const double *params, int times, int m, int n, double steps_per) {
  if (!in || !out || !params || bad_sizes(times, {m, n})) return fail("gpu_box_2d", CS_ERR_INVALID);
  // rows = m + 2*3, cols = n + 2*3 + 2; interior rows [3, m+3), cols [4, n+4)  (src/2d/gpu.cu:269-271)
  const cs_layout L = layout_2d((int64_t)m + 6, (int64_t)n + 8, 3, 4, m, n);
  std::vector<double> w(params, params + 49);
  long long ms = 0;
  double secs = 0;
  int rc = run_compat(L, 3, w, in, out, times, &ms, &secs);
  if (rc) return rc;
  printf("ConvStencil(2D): \nTime = %lld[ms]\n", ms);
  printf("GStencil/s = %f\n", ((double)m * n * times * steps_per) / secs / 1e9);
  fflush(stdout);
  return 0;
}

// src/2d/gpu.cu:245-333 (GStencil/s counts 3 steps per launch, :329)
int gpu_box_2d1r(const double *in, double *out, const double *params, int times, int input_m, int input_n) {
  return run_2d(in, out, params, times, input_m, input_n, 3.0);
}

// src/2d/gpu.cu:336-425 (1 step per launch, :420)
int gpu_box_2d3r(const double *in, double *out, const double *params, int times, int input_m, int input_n) {
  return run_2d(in, out, params, times, input_m, input_n, 1.0);
}

// The --breakdown ladder (src/2d/gpu.cu:434-801, run by src/2d/main.cu:349-357): the same
// launch loop as gpu_box_2d1r on one rung's kernel, printed as the reference prints its
// rungs.  Rung kernels (k2d_ablation.hip, k2d_f64.hip): 1 FP64 FMA from HBM/L2, 2 FP64 FMA
// from an LDS tile, 3 the dilated MFMA tables with 128-byte aligned LDS rows, 4 the
// production single-step kernel (the bank model's pitch, the output staged in LDS).
static int run_breakdown(int rung, int variant, const double *in, double *out, const double *params, int times, int m,
                         int n) {
  if (!in || !out || !params || bad_sizes(times, {m, n})) return fail("gpu_box_2d1r_breakdown", CS_ERR_INVALID);
  const cs_layout L = layout_2d((int64_t)m + 6, (int64_t)n + 8, 3, 4, m, n);
  std::vector<double> w(params, params + 49);
  long long ms = 0;
  double secs = 0;
  int rc = run_compat(L, 3, w, in, out, times, &ms, &secs, variant);
  if (rc) return rc;
  printf("Experiment - Breakdown(2D) %d: \nTime = %lld[ms]\n", rung, ms);
  printf("GStencil/s = %f\n\n", ((double)m * n * times * 3) / secs / 1e9);
  fflush(stdout);
  return 0;
}

int gpu_box_2d1r_breakdown1(const double *in, double *out, const double *params, int times, int input_m,
                            int input_n) {
  return run_breakdown(1, 913000, in, out, params, times, input_m, input_n);
}
int gpu_box_2d1r_breakdown2(const double *in, double *out, const double *params, int times, int input_m,
                            int input_n) {
  return run_breakdown(2, 923000, in, out, params, times, input_m, input_n);
}
int gpu_box_2d1r_breakdown3(const double *in, double *out, const double *params, int times, int input_m,
                            int input_n) {
  return run_breakdown(3, 933101, in, out, params, times, input_m, input_n);
}
int gpu_box_2d1r_breakdown4(const double *in, double *out, const double *params, int times, int input_m,
                            int input_n) {
  return run_breakdown(4, 713101, in, out, params, times, input_m, input_n);
}

// 196 box weights: plane k = 3 - |dz| holds the 7x7 (dy, dx) slice (src/3d/main.cu:115-467).
static std::vector<double> dense_box3d(const double *p) {
  std::vector<double> w(343, 0.0);
  for (int dz = -3; dz <= 3; ++dz) {
    const int k = 3 - (dz < 0 ? -dz : dz);
    for (int i = 0; i < 49; ++i) w[(dz + 3) * 49 + i] = p[k * 49 + i];
  }
  return w;
}

// 44 star weights: planes |dz| = 3, 2, 1, 0 are diamonds of radius 0, 1, 2, 3 with
// 1, 5, 13, 25 taps in row-major order at offsets 0, 1, 6, 19 (src/3d/main.cu:42-113).
static std::vector<double> dense_star3d(const double *p) {
  std::vector<double> w(343, 0.0);
  const int base[4] = {0, 1, 6, 19};
  for (int dz = -3; dz <= 3; ++dz) {
    const int adz = dz < 0 ? -dz : dz;
    const int r = 3 - adz;
    int idx = base[r];
    for (int dy = -r; dy <= r; ++dy) {
      const int span = r - (dy < 0 ? -dy : dy);
      for (int dx = -span; dx <= span; ++dx) w[((dz + 3) * 7 + (dy + 3)) * 7 + (dx + 3)] = p[idx++];
    }
  }
  return w;
}

static int run_3d(const char *name, const std::vector<double> &w, const double *in, double *out, int times, int h,
                  int m, int n, bool header) {
  const cs_layout L = layout_3d(h, m, n, 3);
  long long ms = 0;
  double secs = 0;
  int rc = run_compat(L, 3, w, in, out, times, &ms, &secs);
  if (rc) return rc;
  (void)name;
  if (header) printf("ConvStencil(3D): \n");
  printf("Time = %lld[ms]\n", ms);
  printf("%g\n", secs);
  printf("GStencil/s = %f\n", ((double)m * n * h * times * 3) / secs / 1e9);
  fflush(stdout);
  return 0;
}

// src/3d/gpu_box.cu:446-536
int gpu_box_3d1r(const double *in, double *out, const double *params, int times, int input_h, int input_m,
                 int input_n) {
  if (!in || !out || !params || bad_sizes(times, {input_h, input_m, input_n}))
    return fail("gpu_box_3d1r", CS_ERR_INVALID);
  return run_3d("box", dense_box3d(params), in, out, times, input_h, input_m, input_n, true);
}

// The 3-D --breakdown ladder (src/3d/gpu_box.cu:538-904, run by src/3d/main.cu:623-628):
// gpu_box_3d1r's launch loop and result on one rung's kernel (k3d_ablation.hip,
// k3d_f64.hip, k3d_fma.hip): 1 FP64 FMA from HBM/L2, 2 FP64 FMA from a ring of LDS planes,
// 3 the MFMA strip kernel with 128-byte aligned LDS rows, 4 the production kernel
// (register-blocked FMAs over streamed planes).  The reference prints the elapsed seconds
// on a line of their own here too (:618-622).
static int run_breakdown_3d(int rung, int variant, const double *in, double *out, const double *params, int times,
                            int h, int m, int n) {
  if (!in || !out || !params || bad_sizes(times, {h, m, n})) return fail("gpu_box_3d1r_breakdown", CS_ERR_INVALID);
  const cs_layout L = layout_3d(h, m, n, 3);
  long long ms = 0;
  double secs = 0;
  int rc = run_compat(L, 3, dense_box3d(params), in, out, times, &ms, &secs, variant);
  if (rc) return rc;
  printf("Experiment - Breakdown(3D) %d: \nTime = %lld[ms]\n%g\n", rung, ms, secs);
  printf("GStencil/s = %f\n\n", ((double)m * n * h * times * 3) / secs / 1e9);
  fflush(stdout);
  return 0;
}

int gpu_box_3d1r_breakdown1(const double *in, double *out, const double *params, int times, int input_h, int input_m,
                            int input_n) {
  return run_breakdown_3d(1, 913000, in, out, params, times, input_h, input_m, input_n);
}
int gpu_box_3d1r_breakdown2(const double *in, double *out, const double *params, int times, int input_h, int input_m,
                            int input_n) {
  return run_breakdown_3d(2, 923000, in, out, params, times, input_h, input_m, input_n);
}
int gpu_box_3d1r_breakdown3(const double *in, double *out, const double *params, int times, int input_h, int input_m,
                            int input_n) {
  return run_breakdown_3d(3, 933001, in, out, params, times, input_h, input_m, input_n);
}
int gpu_box_3d1r_breakdown4(const double *in, double *out, const double *params, int times, int input_h, int input_m,
                            int input_n) {
  return run_breakdown_3d(4, 35302, in, out, params, times, input_h, input_m, input_n);
}

// src/3d/gpu_star.cu:238-332
int gpu_star_3d1r(const double *in, double *out, const double *params, int times, int input_h, int input_m,
                  int input_n) {
  if (!in || !out || !params || bad_sizes(times, {input_h, input_m, input_n}))
    return fail("gpu_star_3d1r", CS_ERR_INVALID);
  return run_3d("star", dense_star3d(params), in, out, times, input_h, input_m, input_n, false);
}

}  // extern "C"
