// miopen_stencil.cpp -- the vendor library as a stencil engine: MIOpen convolution as the
// baseline beside ConvStencil (the counterpart of the reference's cuDNN programs,
// src/cudnn/conv_box2d9p.cu:16-172, conv_box2d49p.cu, conv_1d3p.cu, conv_1d5p.cu,
// conv_box3d27p.cu).
//
// Usage: miopen_stencil shape sizes... T [--seed=S] [--fp32]
//   shape: 1d1r (3 points), 1d2r (5), box2d1r (3 x 3), box2d3r (7 x 7), star2d3r (7 x 7 with
//   zeros off the axes), box3d1r (3 x 3 x 3), star3d1r (3 x 3 x 3 with zeros off the axes).
// One single-channel cross-correlation per time step with zero padding of the radius -- the
// same zero-halo stencil as ConvStencil's physical mode, on the same seeded grid and
// weights -- ping-ponging two buffers as the reference's programs do.  FP64 first; when
// MIOpen has no FP64 solver for the problem it says so and runs FP32 (or use --fp32).
// Prints the reference programs' "Time = N[ms]" and "GStencil/s = X" lines (timed with HIP
// events around T steps after one untimed step), the algorithm, and the relative L2 and max
// error of the MIOpen result against ConvStencil's FP64 plan after the same T steps.
#include <hip/hip_runtime.h>
#include <miopen/miopen.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/convstencil.h"

namespace {

struct Shape {
  const char *name;
  int ndim, r;
  bool star;
};
const Shape kShapes[] = {{"1d1r", 1, 1, false},    {"1d2r", 1, 2, false},    {"box2d1r", 2, 1, false},
                         {"box2d3r", 2, 3, false}, {"star2d3r", 2, 3, true}, {"box3d1r", 3, 1, false},
                         {"star3d1r", 3, 1, true}};

int usage() {
  printf("Usage: miopen_stencil shape sizes... time_steps [--seed=S] [--fp32]\n"
         "  shapes: 1d1r 1d2r (n), box2d1r box2d3r star2d3r (m n), box3d1r star3d1r (h m n)\n");
  return 1;
}

#define MIO(x)                                                                       \
  do {                                                                               \
    const miopenStatus_t st_ = (x);                                                  \
    if (st_ != miopenStatusSuccess) {                                                \
      fprintf(stderr, "miopen_stencil: %s failed (status %d)\n", #x, (int)st_);      \
      return 1;                                                                      \
    }                                                                                \
  } while (0)
#define HIPC(x)                                                                       \
  do {                                                                                \
    const hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                           \
      fprintf(stderr, "miopen_stencil: %s failed: %s\n", #x, hipGetErrorString(e_));  \
      return 1;                                                                       \
    }                                                                                 \
  } while (0)

// One precision's run: T steps of the convolution on `x0` (interior cells, slowest first);
// the result after T steps goes to `result` (as doubles).
template <class T>
int run(miopenDataType_t dt, const std::vector<int> &dims, int r, const std::vector<double> &dense,
        const std::vector<double> &x0, int steps, std::vector<double> &result, double *ms, const char **algo,
        bool *unsupported) {
  *unsupported = false;
  miopenHandle_t h;
  MIO(miopenCreate(&h));
  // MIOpen takes 4-D and 5-D tensors: a 1-D grid is a 2-D one of a single row, its filter a
  // single row of taps
  const bool one_d = dims.size() == 1;
  const int sd = one_d ? 2 : (int)dims.size(), D = 2 * r + 1;
  std::vector<int> xd = {1, 1}, wd = {1, 1}, pad(sd, r), one(sd, 1);
  if (one_d) {
    xd.push_back(1);
    wd.push_back(1);
    pad[0] = 0;
  }
  for (int d : dims) xd.push_back(d);
  for (int i = one_d ? 1 : 0; i < sd; ++i) wd.push_back(D);
  miopenTensorDescriptor_t xdesc, wdesc, ydesc;
  miopenConvolutionDescriptor_t cdesc;
  MIO(miopenCreateTensorDescriptor(&xdesc));
  MIO(miopenCreateTensorDescriptor(&wdesc));
  MIO(miopenCreateTensorDescriptor(&ydesc));
  MIO(miopenCreateConvolutionDescriptor(&cdesc));
  MIO(miopenSetTensorDescriptor(xdesc, dt, 2 + sd, xd.data(), nullptr));
  MIO(miopenSetTensorDescriptor(wdesc, dt, 2 + sd, wd.data(), nullptr));
  MIO(miopenInitConvolutionNdDescriptor(cdesc, sd, pad.data(), one.data(), one.data(), miopenConvolution));
  int nd = 0;
  std::vector<int> yd(2 + sd);
  MIO(miopenGetConvolutionNdForwardOutputDim(cdesc, xdesc, wdesc, &nd, yd.data()));
  MIO(miopenSetTensorDescriptor(ydesc, dt, nd, yd.data(), nullptr));
  size_t cells = 1;
  for (int d : dims) cells *= (size_t)d;
  std::vector<T> hx(cells), hw(dense.size());
  for (size_t i = 0; i < cells; ++i) hx[i] = (T)x0[i];
  for (size_t i = 0; i < dense.size(); ++i) hw[i] = (T)dense[i];
  T *buf[2] = {nullptr, nullptr}, *w = nullptr;
  HIPC(hipMalloc(&buf[0], cells * sizeof(T)));
  HIPC(hipMalloc(&buf[1], cells * sizeof(T)));
  HIPC(hipMalloc(&w, hw.size() * sizeof(T)));
  HIPC(hipMemcpy(buf[0], hx.data(), cells * sizeof(T), hipMemcpyHostToDevice));
  HIPC(hipMemcpy(w, hw.data(), hw.size() * sizeof(T), hipMemcpyHostToDevice));
  size_t ws = 0;
  MIO(miopenConvolutionForwardGetWorkSpaceSize(h, wdesc, xdesc, cdesc, ydesc, &ws));
  void *work = nullptr;
  if (ws) HIPC(hipMalloc(&work, ws));
  miopenConvAlgoPerf_t perf[4];
  int found = 0;
  const miopenStatus_t fs = miopenFindConvolutionForwardAlgorithm(h, xdesc, buf[0], wdesc, w, cdesc, ydesc, buf[1], 4,
                                                                  &found, perf, work, ws, false);
  if (fs != miopenStatusSuccess || found == 0) {
    *unsupported = true;  // no solver for this data type and problem
  } else {
    const miopenConvFwdAlgorithm_t a = perf[0].fwd_algo;
    *algo = a == miopenConvolutionFwdAlgoGEMM       ? "GEMM"
            : a == miopenConvolutionFwdAlgoDirect   ? "Direct"
            : a == miopenConvolutionFwdAlgoFFT      ? "FFT"
            : a == miopenConvolutionFwdAlgoWinograd ? "Winograd"
            : a == miopenConvolutionFwdAlgoImplicitGEMM ? "ImplicitGEMM"
                                                        : "other";
    const T alpha = 1, beta = 0;
    // one untimed step from a copy of the input, then T timed steps from the input
    MIO(miopenConvolutionForward(h, &alpha, xdesc, buf[0], wdesc, w, cdesc, a, &beta, ydesc, buf[1], work, ws));
    HIPC(hipMemcpy(buf[0], hx.data(), cells * sizeof(T), hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    HIPC(hipEventCreate(&e0));
    HIPC(hipEventCreate(&e1));
    HIPC(hipDeviceSynchronize());
    HIPC(hipEventRecord(e0, nullptr));
    for (int t = 0; t < steps; ++t)
      MIO(miopenConvolutionForward(h, &alpha, xdesc, buf[t % 2], wdesc, w, cdesc, a, &beta, ydesc, buf[(t + 1) % 2],
                                   work, ws));
    HIPC(hipEventRecord(e1, nullptr));
    HIPC(hipEventSynchronize(e1));
    float fms = 0;
    HIPC(hipEventElapsedTime(&fms, e0, e1));
    *ms = fms;
    std::vector<T> out(cells);
    HIPC(hipMemcpy(out.data(), buf[steps % 2], cells * sizeof(T), hipMemcpyDeviceToHost));
    result.assign(out.begin(), out.end());
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipFree(buf[0]);
  (void)hipFree(buf[1]);
  (void)hipFree(w);
  if (work) (void)hipFree(work);
  miopenDestroyTensorDescriptor(xdesc);
  miopenDestroyTensorDescriptor(wdesc);
  miopenDestroyTensorDescriptor(ydesc);
  miopenDestroyConvolutionDescriptor(cdesc);
  miopenDestroy(h);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  unsigned long long seed = 1234;
  bool fp32 = false;
  std::vector<std::string> pos;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a.rfind("--seed=", 0) == 0)
      seed = std::strtoull(a.c_str() + 7, nullptr, 10);
    else if (a == "--fp32")
      fp32 = true;
    else
      pos.push_back(a);
  }
  if (pos.empty()) return usage();
  const Shape *sh = nullptr;
  for (const Shape &s : kShapes)
    if (pos[0] == s.name) sh = &s;
  if (!sh || (int)pos.size() != sh->ndim + 2) return usage();
  std::vector<int> dims;
  for (int k = 0; k < sh->ndim; ++k) dims.push_back(std::atoi(pos[1 + k].c_str()));
  const int steps = std::atoi(pos.back().c_str());
  for (int d : dims)
    if (d < 1) return usage();
  if (steps < 1) return usage();

  // the seeded grid and weights of ConvStencil's physical mode: interior of a zero-halo layout
  const int r = sh->r, D = 2 * r + 1;
  cs_layout L{};
  L.ndim = sh->ndim;
  for (int d = 0; d < 3; ++d) {
    L.ext[d] = 1;
    L.n[d] = 1;
    L.org[d] = 0;
  }
  for (int k = 0; k < sh->ndim; ++k) {
    const int d = 3 - sh->ndim + k;
    L.n[d] = dims[k];
    L.org[d] = r;
    L.ext[d] = dims[k] + 2 * r;
  }
  const size_t ext = (size_t)(L.ext[0] * L.ext[1] * L.ext[2]), cells = (size_t)(L.n[0] * L.n[1] * L.n[2]);
  std::vector<double> padded(ext), x0(cells);
  cs_fill_random_host(padded.data(), &L, seed);
  std::vector<size_t> imap(cells);  // interior cell i -> padded index
  {
    size_t i = 0;
    for (int64_t z = 0; z < L.n[0]; ++z)
      for (int64_t y = 0; y < L.n[1]; ++y)
        for (int64_t x = 0; x < L.n[2]; ++x, ++i)
          imap[i] = (size_t)(((L.org[0] + z) * L.ext[1] + L.org[1] + y) * L.ext[2] + L.org[2] + x);
    for (size_t k = 0; k < cells; ++k) x0[k] = padded[imap[k]];
  }
  int taps = 0;
  std::vector<int> tapidx;
  const int nw = sh->ndim == 1 ? D : sh->ndim == 2 ? D * D : D * D * D;
  for (int t = 0; t < nw; ++t) {
    const int dx = t % D - r, dy = sh->ndim >= 2 ? (t / D) % D - r : 0, dz = sh->ndim == 3 ? t / (D * D) - r : 0;
    const bool on = !sh->star || (dz != 0) + (dy != 0) + (dx != 0) <= 1;
    if (on) tapidx.push_back(t), ++taps;
  }
  std::vector<double> wt((size_t)taps), dense((size_t)nw, 0.0);
  cs_fill_weights_host(wt.data(), taps, seed);
  for (int k = 0; k < taps; ++k) dense[(size_t)tapidx[k]] = wt[(size_t)k];

  // ConvStencil's FP64 result after the same steps (the reference for the error figures)
  std::vector<double> mine(ext);
  cs_stencil st{sh->ndim, r, dense.data()};
  cs_options opt{};
  cs_report rep{};
  if (cs_run(&L, &st, padded.data(), mine.data(), steps, &opt, &rep) != CS_OK) {
    fprintf(stderr, "miopen_stencil: the ConvStencil reference run failed\n");
    return 1;
  }

  std::vector<double> result;
  double ms = 0;
  const char *algo = "none";
  bool unsupported = false;
  const char *dtn = "fp64";
  if (!fp32) {
    if (run<double>(miopenDouble, dims, r, dense, x0, steps, result, &ms, &algo, &unsupported)) return 1;
    if (unsupported) printf("note: MIOpen has no FP64 solver for this convolution; running FP32\n");
  }
  if (fp32 || unsupported) {
    dtn = "fp32";
    if (run<float>(miopenFloat, dims, r, dense, x0, steps, result, &ms, &algo, &unsupported)) return 1;
    if (unsupported) {
      fprintf(stderr, "miopen_stencil: MIOpen has no solver for this convolution\n");
      return 1;
    }
  }
  double num = 0, den = 0, emax = 0, vmax = 0;
  for (size_t k = 0; k < cells; ++k) {
    const double c = mine[imap[k]], e = result[k] - c;
    num += e * e;
    den += c * c;
    emax = std::fmax(emax, std::fabs(e));
    vmax = std::fmax(vmax, std::fabs(c));
  }
  printf("MIOpen(%dD) %s %s, algorithm %s: \n", sh->ndim, sh->name, dtn, algo);
  printf("Time = %lld[ms]\n", (long long)ms);
  printf("GStencil/s = %f\n", (double)cells * steps / (ms * 1e-3) / 1e9);
  printf("rel_l2 vs ConvStencil fp64 = %.3e, rel_max = %.3e\n", den > 0 ? std::sqrt(num / den) : 0.0,
         vmax > 0 ? emax / vmax : 0.0);
  return 0;
}
