// convstencil_1d -- drop-in for the reference's 1-D driver (src/1d/main.cu:61-286).
#include <cstdlib>

#include "../cs_fused_weights.h"
#include "cli_common.h"

static const char *kShapeStr[2] = {"1d1r", "1d2r"};

static void print_help() {
  const char *msg =
      "Program name: convstencil_1d\n"
      "Usage: convstencil_2 shape input_size time_iteration_size [Options]\n"
      "Shape: 1d1r or 1d2r\n"
      "Options:\n"
      "  --help    Display this help message and exit\n"
      "  --custom  If you want to use costum parameters, please use this option and input your parameters like 0.2 "
      "0.2 0.2 0.2 0.2 if the shape is star2d1r\n";
  printf("%s\n", msg);
}

int main(int argc, char *argv[]) {
  cli::Ext e;
  if (!cli::strip_flags(argc, argv, e)) return 1;
  if (argc < 4) {
    print_help();
    return 1;
  }
  const std::string arg1 = argv[1];
  int shape;
  if (arg1 == "1d1r")
    shape = 0;
  else if (arg1 == "1d2r")
    shape = 1;
  else {
    print_help();
    return 1;
  }
  int n = 0, times = 0;
  try {
    n = std::stoi(argv[2]);
    times = std::stoi(argv[3]);
  } catch (const std::invalid_argument &) {
    std::cerr << "Invalid argument: cannot convert the parameter(s) to integer.\n";
    return 1;
  } catch (const std::out_of_range &) {
    std::cerr << "Argument out of range: the parameter(s) is(are) too large.\n";
    return 1;
  }
  const bool custom = argc == 5 && std::string(argv[4]) == "--custom";
  const bool breakdown = argc == 5 && std::string(argv[4]) == "--breakdown";

  if (e.steps >= 0) {  // physical-stencil mode: 3-point (1d1r) or 5-point (1d2r)
    const int r = shape == 0 ? 1 : 2;
    bool ok = true;
    std::vector<double> taps = cli::physical_taps(2 * r + 1, custom, e.seed, ok);
    if (!ok) return 1;
    printf("INFO: shape = %s,  n = %d, steps = %lld (physical)\n", kShapeStr[shape], n, e.steps);
    return cli::run_physical("1D", 1, r, taps, {n}, e);
  }

  // reference mode (src/1d/main.cu:102-170): default weights 1..7 / 1..9
  double p1[7], p2[9];
  for (int i = 0; i < 7; ++i) p1[i] = i + 1;
  for (int i = 0; i < 9; ++i) p2[i] = i + 1;
  if (custom) {
    const int num = shape == 0 ? 3 : 5;
    double v[5];
    if (!cli::read_params(num, v)) return 1;
    if (shape == 0)
      cs::fused_1d1r(v, p1);
    else
      cs::fused_1d2r(v, p2);
  }
  // the ladder runs for 1d1r only, as in the reference (src/1d/main.cu:221-233)
  const bool ladder = breakdown && shape == 0;
  if (breakdown && !ladder) printf("note: the --breakdown ladder exists for 1d1r; running the main kernel\n");
  const int halo = shape == 0 ? 3 : 4;
  printf("INFO: shape = %s,  n = %d, times = %d\n", kShapeStr[shape], n, times);
  const size_t cols = (size_t)n + 2 * halo;
  double *input = (double *)malloc((cols + 1) * sizeof(double));
  double *output = (double *)calloc(cols + 1, sizeof(double));
  if (!input || !output) {
    fprintf(stderr, "convstencil: host allocation failed\n");
    return 1;
  }
  for (size_t i = 0; i < cols + 1; ++i) input[i] = (double)(rand() % 10000);
  int rc = 0;
  if (ladder) {
    int (*const rungs[4])(const double *, double *, const double *, int, int) = {
        gpu_1d1r_breakdown1, gpu_1d1r_breakdown2, gpu_1d1r_breakdown3, gpu_1d1r_breakdown4};
    for (int k = 0; k < 4 && rc == 0; ++k) rc = rungs[k](input, output, p1, times, n);
  }
  if (rc == 0) rc = shape == 0 ? gpu_1d1r(input, output, p1, times, n) : gpu_star_1d2r_step2(input, output, p2, times, n);
  if (rc == 0 && cli::wants_output(e)) {  // the reference dumps the first n + 2 halo cells (src/1d/main.cu:275)
    const long long ext[1] = {(long long)cols}, org[1] = {halo}, nn[1] = {n};
    if (!cli::write_grid(e, cli::ref_layout(1, ext, org, nn), output)) rc = 1;
  }
  free(output);
  free(input);
  return rc ? 1 : 0;
}
