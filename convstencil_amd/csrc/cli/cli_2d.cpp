// convstencil_2d -- drop-in for the reference's 2-D driver (src/2d/main.cu:129-423).
#include <cstdlib>

#include "../cs_fused_weights.h"
#include "cli_common.h"

static const char *kShapeStr[4] = {"star_2d1r", "box_2d1r", "star_2d3r", "box_2d3r"};
enum Shape { star_2d1r, box_2d1r, star_2d3r, box_2d3r };

static void print_help() {
  const char *msg =
      "Program name: convstencil_2d\n"
      "Usage: convstencil_2d shape input_size_of_first_dimension input_size_of_second_dimension "
      "time_iteration_size [Options]\n"
      "Shape: box2d1r or star2d1r or box2d3r or star2d3r\n"
      "Options:\n"
      "  --help    Display this help message and exit\n"
      "  --custom  If you want to use costum parameters, please use this option and input your parameters like 0.2 "
      "0.2 0.2 0.2 0.2 if the shape is star2d1r\n";
  printf("%s\n", msg);
}

int main(int argc, char *argv[]) {
  cli::Ext e;
  if (!cli::strip_flags(argc, argv, e)) return 1;
  if (argc < 5) {
    print_help();
    return 1;
  }
  const std::string arg1 = argv[1];
  Shape shape;
  if (arg1 == "box2d1r")
    shape = box_2d1r;
  else if (arg1 == "star2d1r")
    shape = star_2d1r;
  else if (arg1 == "star2d3r")
    shape = star_2d3r;
  else if (arg1 == "box2d3r")
    shape = box_2d3r;
  else {
    print_help();
    return 1;
  }
  int m = 0, n = 0, times = 0;
  try {
    m = std::stoi(argv[2]);
    n = std::stoi(argv[3]);
    times = std::stoi(argv[4]);
  } catch (const std::invalid_argument &) {
    std::cerr << "Invalid argument: cannot convert the parameter(s) to integer.\n";
    return 1;
  } catch (const std::out_of_range &) {
    std::cerr << "Argument out of range: the parameter(s) is(are) too large.\n";
    return 1;
  }
  const bool custom = argc == 6 && std::string(argv[5]) == "--custom";
  const bool breakdown = argc == 6 && std::string(argv[5]) == "--breakdown";

  if (e.steps >= 0) {  // physical-stencil mode
    const int r = (shape == box_2d1r || shape == star_2d1r) ? 1 : 3;
    const bool star = shape == star_2d1r || shape == star_2d3r;
    const int D = 2 * r + 1;
    const int count = star ? 4 * r + 1 : D * D;
    bool ok = true;
    std::vector<double> taps = cli::physical_taps(count, custom, e.seed, ok);
    if (!ok) return 1;
    std::vector<double> dense((size_t)D * D, 0.0);
    int k = 0;
    for (int dy = -r; dy <= r; ++dy)
      for (int dx = -r; dx <= r; ++dx)
        if (!star || dy == 0 || dx == 0) dense[(size_t)(dy + r) * D + (dx + r)] = taps[(size_t)k++];
    printf("INFO: shape = %s, m = %d, n = %d, steps = %lld (physical)\n", kShapeStr[shape], m, n, e.steps);
    return cli::run_physical("2D", 2, r, dense, {m, n}, e);
  }

  // reference mode: pre-fused 7x7 weights applied `times` times (src/2d/main.cu:183-303)
  double p1r[9] = {0.0};
  if (custom) {
    const int num = shape == star_2d1r ? 5 : 9;
    double v[9];
    if (!cli::read_params(num, v)) return 1;
    if (num == 9) {
      for (int i = 0; i < 9; ++i) p1r[i] = v[i];
    } else {
      p1r[1] = v[0];
      p1r[3] = v[1];
      p1r[4] = v[2];
      p1r[5] = v[3];
      p1r[7] = v[4];
    }
  }
  const bool ladder = breakdown && (shape == box_2d1r || shape == star_2d1r);
  if (breakdown && !ladder) printf("note: the --breakdown ladder exists for box2d1r and star2d1r; running the main kernel\n");
  double box[49], star[49] = {0.0};
  if (e.exact_fusion)
    cs::fused_2d_exact(p1r, box);
  else
    cs::fused_2d_reference(p1r, box);
  // star shapes run with the all-zero weight set, as in the reference (src/2d/main.cu:226,291-298)
  const double *param = (shape == box_2d1r || shape == box_2d3r) ? box : star;

  printf("INFO: shape = %s, m = %d, n = %d, times = %d\n", kShapeStr[shape], m, n, times);
  const size_t rows = (size_t)m + 6, cols = (size_t)n + 8;
  double *matrix = (double *)malloc(rows * cols * sizeof(double));
  double *output = (double *)malloc(rows * cols * sizeof(double));
  if (!matrix || !output) {
    fprintf(stderr, "convstencil: host allocation failed\n");
    return 1;
  }
  for (size_t i = 0; i < rows * cols; ++i) matrix[i] = (double)(rand() % 100);
  int rc = 0;
  if (ladder) {  // the reference's order (src/2d/main.cu:349-357): rungs 1-4, then the main launcher
    int (*const rungs[4])(const double *, double *, const double *, int, int, int) = {
        gpu_box_2d1r_breakdown1, gpu_box_2d1r_breakdown2, gpu_box_2d1r_breakdown3, gpu_box_2d1r_breakdown4};
    for (int k = 0; k < 4 && rc == 0; ++k) rc = rungs[k](matrix, output, param, times, m, n);
  }
  if (rc == 0)
    rc = (shape == box_2d1r || shape == star_2d1r) ? gpu_box_2d1r(matrix, output, param, times, m, n)
                                                   : gpu_box_2d3r(matrix, output, param, times, m, n);
  if (rc == 0 && cli::wants_output(e)) {
    const long long ext[2] = {(long long)rows, (long long)cols}, org[2] = {3, 4}, nn[2] = {m, n};
    if (!cli::write_grid(e, cli::ref_layout(2, ext, org, nn), output)) rc = 1;
  }
  free(output);
  free(matrix);
  return rc ? 1 : 0;
}
