// convstencil_3d -- drop-in for the reference's 3-D driver (src/3d/main.cu:481-692).
#include <cstdlib>

#include "cli_common.h"

static const char *kShapeStr[2] = {"box_3d1r", "star_3d1r"};

static void print_help() {
  const char *msg =
      "Program name: convstencil_2d\n"
      "Usage: convstencil_3d shape input_size_of_first_dimension input_size_of_second_dimension "
      "input_size_of_third_dimension time_iteration_size [Options]\n"
      "Shape: box3d1r or star3d1r\n"
      "Options:\n"
      "  --help    Display this help message and exit\n"
      "  --custom  If you want to use costum parameters, please use this option and input your parameters like 0.2 "
      "0.2 0.2 0.2 0.2 if the shape is star2d1r\n";
  printf("%s\n", msg);
}

int main(int argc, char *argv[]) {
  cli::Ext e;
  if (!cli::strip_flags(argc, argv, e)) return 1;
  if (argc < 6) {
    print_help();
    return 1;
  }
  const std::string arg1 = argv[1];
  int shape;
  if (arg1 == "box3d1r")
    shape = 0;
  else if (arg1 == "star3d1r")
    shape = 1;
  else {
    print_help();
    return 1;
  }
  int h = 0, m = 0, n = 0, times = 0;
  try {
    h = std::stoi(argv[2]);
    m = std::stoi(argv[3]);
    n = std::stoi(argv[4]);
    times = std::stoi(argv[5]);
  } catch (const std::invalid_argument &) {
    std::cerr << "Invalid argument: cannot convert the parameter(s) to integer.\n";
    return 1;
  } catch (const std::out_of_range &) {
    std::cerr << "Argument out of range: the parameter(s) is(are) too large.\n";
    return 1;
  }
  const bool custom = argc == 7 && std::string(argv[6]) == "--custom";
  const bool breakdown = argc == 7 && std::string(argv[6]) == "--breakdown";

  if (e.steps >= 0) {  // physical-stencil mode: 27-point box or 7-point star, radius 1
    const bool star = shape == 1;
    bool ok = true;
    std::vector<double> taps = cli::physical_taps(star ? 7 : 27, custom, e.seed, ok);
    if (!ok) return 1;
    std::vector<double> dense(27, 0.0);
    int k = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx)
          if (!star || (dz != 0) + (dy != 0) + (dx != 0) <= 1) dense[(size_t)((dz + 1) * 3 + dy + 1) * 3 + dx + 1] =
              taps[(size_t)k++];
    printf("INFO: shape = %s, h = %d, m = %d, n = %d, steps = %lld (physical)\n", kShapeStr[shape], h, m, n, e.steps);
    return cli::run_physical("3D", 3, 1, dense, {h, m, n}, e);
  }

  // reference mode (src/3d/main.cu:526-586): fixed defaults; --custom values are read and not used
  double box[196], star[44];
  for (int i = 0; i < 196; ++i) box[i] = (double)i / 100;
  for (int i = 0; i < 44; ++i) star[i] = 1;
  if (custom) {
    double v[27];
    if (!cli::read_params(shape == 0 ? 27 : 7, v)) return 1;
  }
  // the ladder runs for box3d1r only, as in the reference (src/3d/main.cu:620-637)
  const bool ladder = breakdown && shape == 0;
  if (breakdown && !ladder) printf("note: the --breakdown ladder exists for box3d1r; running the main kernel\n");
  printf("INFO: shape = %s, h = %d, m = %d, n = %d, times = %d\n", kShapeStr[shape], h, m, n, times);
  const size_t cells = ((size_t)h + 6) * ((size_t)m + 6) * ((size_t)n + 6);
  double *matrix = (double *)malloc(cells * sizeof(double));
  double *output = (double *)malloc(cells * sizeof(double));
  if (!matrix || !output) {
    fprintf(stderr, "convstencil: host allocation failed\n");
    return 1;
  }
  for (size_t i = 0; i < cells; ++i) matrix[i] = (double)(rand() % 100);
  int rc = 0;
  if (ladder) {
    int (*const rungs[4])(const double *, double *, const double *, int, int, int, int) = {
        gpu_box_3d1r_breakdown1, gpu_box_3d1r_breakdown2, gpu_box_3d1r_breakdown3, gpu_box_3d1r_breakdown4};
    for (int k = 0; k < 4 && rc == 0; ++k) rc = rungs[k](matrix, output, box, times, h, m, n);
  }
  if (rc == 0)
    rc = shape == 0 ? gpu_box_3d1r(matrix, output, box, times, h, m, n)
                    : gpu_star_3d1r(matrix, output, star, times, h, m, n);
  if (rc == 0 && cli::wants_output(e)) {
    const long long ext[3] = {(long long)h + 6, (long long)m + 6, (long long)n + 6}, org[3] = {3, 3, 3},
                    nn[3] = {h, m, n};
    if (!cli::write_grid(e, cli::ref_layout(3, ext, org, nn), output)) rc = 1;
  }
  free(output);
  free(matrix);
  return rc ? 1 : 0;
}
