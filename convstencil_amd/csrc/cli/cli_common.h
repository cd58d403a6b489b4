// cli_common.h -- shared pieces of the convstencil_{1d,2d,3d} command lines.
//
// The positional grammar, prompts, messages and exit codes are the reference's
// (src/1d/main.cu:49-100, src/2d/main.cu:117-181, src/3d/main.cu:469-522).  Flags of
// the form --name=value are additions; they are removed from argv before the
// reference grammar is applied, so `prog shape sizes... T [--custom]` behaves as in
// the reference:
//   --steps=T        physical-stencil mode: the shape's own stencil (1d1r = 3 points,
//                    box2d1r = 9, ...) with a zero halo of its radius, seeded input,
//                    T physical time steps, fused in the kernels.  --custom then reads
//                    the shape's taps (3, 5, 5, 9, 13, 49, 7 or 27 values).
//   --seed=S         seed of the synthetic grid and weights (physical mode)
//   --fusion=K       fusion depth (0 = library default)
//   --graph=0|1      capture the time loop in a hipGraph (default 1)
//   --autotune=0|1|2 physical mode: 1 = the autotuner's cached choice (timed on a miss),
//                    2 = time the variants again (default 0: library defaults)
//   --precision=P    physical mode: fp64 (default), bf16x3, fp32x2, fp16x2 or fp16w3 (2-D);
//                    fp64, bf16x3 or fp16w3 (3-D)
//   --json           also print a JSON record (physical mode)
//   --exact-fusion   2-D reference mode: compose the 3x3 weights exactly instead of
//                    using the reference's 49-entry table
//   --write-output[=PATH]  write the final padded array in the reference's text format
//                    (its WRITE_OUTPUT build, src/2d/main.cu:410-415; default output_gpu.txt)
//   --save-f64=PATH  write the final padded array as an exact binary grid (cs_grid_save_f64)
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/convstencil.h"

namespace cli {

struct Ext {
  long long steps = -1;
  unsigned long long seed = 1234;
  int fusion = 0;
  int graph = 1;
  int autotune = 0;
  int precision = CS_PREC_FP64;
  bool json = false;
  bool exact_fusion = false;
  std::string write_output;  // text dump path ("" = none)
  std::string save_f64;      // binary grid path ("" = none)
};

// Removes the additive flags from argv; returns false (after printing a message)
// on a malformed one.
inline bool strip_flags(int &argc, char **argv, Ext &e) {
  int w = 1;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto val = [&](const char *k) -> const char * {
      const size_t n = strlen(k);
      return a.compare(0, n, k) == 0 ? argv[i] + n : nullptr;
    };
    try {
      if (const char *v = val("--steps=")) {
        e.steps = std::stoll(v);
      } else if (const char *v2 = val("--seed=")) {
        e.seed = std::stoull(v2);
      } else if (const char *v3 = val("--fusion=")) {
        e.fusion = std::stoi(v3);
      } else if (const char *v4 = val("--graph=")) {
        e.graph = std::stoi(v4);
      } else if (const char *v5 = val("--autotune=")) {
        e.autotune = std::stoi(v5);
      } else if (const char *v6 = val("--precision=")) {
        const std::string pn = v6;
        if (pn == "fp64") e.precision = CS_PREC_FP64;
        else if (pn == "bf16x3") e.precision = CS_PREC_BF16X3;
        else if (pn == "fp32x2") e.precision = CS_PREC_FP32X2;
        else if (pn == "fp16x2") e.precision = CS_PREC_FP16X2;
        else if (pn == "fp16w3") e.precision = CS_PREC_FP16W3;
        else {
          std::cerr << "Invalid argument: unknown precision " << pn << ".\n";
          return false;
        }
      } else if (a == "--json") {
        e.json = true;
      } else if (a == "--exact-fusion") {
        e.exact_fusion = true;
      } else if (a == "--write-output") {
        e.write_output = "output_gpu.txt";
      } else if (const char *v7 = val("--write-output=")) {
        e.write_output = v7;
      } else if (const char *v8 = val("--save-f64=")) {
        e.save_f64 = v8;
      } else {
        argv[w++] = argv[i];
      }
    } catch (const std::exception &) {
      std::cerr << "Invalid argument: cannot convert the parameter(s) to integer.\n";
      return false;
    }
  }
  argc = w;
  argv[argc] = nullptr;
  return true;
}

// Writes the files requested by --write-output / --save-f64 for a final padded array;
// false (after a message) if one cannot be written.
inline bool write_grid(const Ext &e, const cs_layout &L, const double *host) {
  if (!e.write_output.empty()) {
    printf("Writing %s\n", e.write_output.c_str());
    fflush(stdout);
    if (cs_grid_save_txt(e.write_output.c_str(), &L, host) != CS_OK) {
      fprintf(stderr, "convstencil: cannot write %s\n", e.write_output.c_str());
      return false;
    }
  }
  if (!e.save_f64.empty() && cs_grid_save_f64(e.save_f64.c_str(), &L, host) != CS_OK) {
    fprintf(stderr, "convstencil: cannot write %s\n", e.save_f64.c_str());
    return false;
  }
  return true;
}

inline bool wants_output(const Ext &e) { return !e.write_output.empty() || !e.save_f64.empty(); }

// Layout of the reference's padded host arrays (slowest first): ext, interior origin, n.
inline cs_layout ref_layout(int ndim, const long long *ext, const long long *org, const long long *n) {
  cs_layout L{};
  L.ndim = ndim;
  for (int d = 0; d < 3; ++d) {
    L.ext[d] = 1;
    L.org[d] = 0;
    L.n[d] = 1;
  }
  for (int k = 0; k < ndim; ++k) {
    L.ext[3 - ndim + k] = ext[k];
    L.org[3 - ndim + k] = org[k];
    L.n[3 - ndim + k] = n[k];
  }
  return L;
}

// Reads n doubles from stdin after the reference's prompt; false if one is missing.
inline bool read_params(int n, double *out) {
  printf("Please enter %d parameters:\n", n);
  fflush(stdout);
  for (int i = 0; i < n; ++i)
    if (scanf("%lf", &out[i]) != 1) return false;
  return true;
}

// Physical-stencil mode: dense weights of radius r in ndim dims, interior sizes
// (slowest first), zero halo of width r.
inline int run_physical(const char *tag, int ndim, int r, const std::vector<double> &dense,
                        const std::vector<long long> &sizes, const Ext &e) {
  cs_layout L{};
  L.ndim = ndim;
  for (int d = 0; d < 3; ++d) {
    L.ext[d] = 1;
    L.n[d] = 1;
    L.org[d] = 0;
  }
  for (int k = 0; k < ndim; ++k) {
    const int d = 3 - ndim + k;
    L.n[d] = sizes[k];
    L.org[d] = r;
    L.ext[d] = sizes[k] + 2 * r;
  }
  cs_stencil st{ndim, r, dense.data()};
  cs_options opt{};
  opt.fusion_depth = e.fusion;
  opt.use_graph = e.graph;
  opt.autotune = e.autotune;
  opt.precision = e.precision;
  cs_plan *p = nullptr;
  cs_status s = cs_plan_create(&p, &L, &st, &opt);
  if (s != CS_OK) {
    fprintf(stderr, "convstencil: %s\n", cs_status_string(s));
    return 1;
  }
  s = cs_plan_fill_random(p, e.seed);
  cs_report rep{};
  if (s == CS_OK) s = cs_plan_time(p, e.steps, 1, 1, &rep);
  char desc[512] = {0};
  cs_plan_describe(p, desc, sizeof desc);
  std::vector<double> final_grid;
  if (s == CS_OK && wants_output(e)) {  // the grid after exactly `steps` steps from the seed
    final_grid.resize((size_t)(L.ext[0] * L.ext[1] * L.ext[2]));
    s = cs_plan_fill_random(p, e.seed);
    cs_report r2{};
    if (s == CS_OK) s = cs_plan_run(p, e.steps, &r2);
    if (s == CS_OK) s = cs_plan_download(p, final_grid.data());
  }
  cs_plan_destroy(p);
  if (s != CS_OK) {
    fprintf(stderr, "convstencil: %s\n", cs_status_string(s));
    return 1;
  }
  printf("ConvStencil(%s): \n", tag);
  printf("Time = %lld[ms]\n", (long long)rep.elapsed_ms);
  printf("GStencil/s = %f\n", rep.gstencil_per_s);
  if (e.json)
    printf("{\"mode\": \"physical\", \"steps\": %lld, \"seed\": %llu, \"elapsed_ms\": %.6f, \"gstencil_per_s\": %.4f, "
           "\"launches\": %lld, \"fusion_depth\": %d, \"plan\": \"%s\"}\n",
           e.steps, e.seed, rep.elapsed_ms, rep.gstencil_per_s, (long long)rep.launches, rep.fusion_depth, desc);
  fflush(stdout);
  if (wants_output(e) && !write_grid(e, L, final_grid.data())) return 1;
  return 0;
}

// Seeded physical weights (uniform in (0,1), normalised) unless --custom supplied them.
inline std::vector<double> physical_taps(int count, bool custom, unsigned long long seed, bool &ok) {
  std::vector<double> w((size_t)count);
  ok = true;
  if (custom)
    ok = read_params(count, w.data());
  else
    cs_fill_weights_host(w.data(), count, seed);
  return w;
}

}  // namespace cli
