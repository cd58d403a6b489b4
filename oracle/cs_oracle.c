/*
 * cs_oracle.c -- CPU oracle for the ConvStencil time loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libconvstencil, the CLIs,
 * the Python package) links, loads or calls this file.  Only tests/, the
 * smoke() check and bench.py's CPU-baseline leg use it, as the checker.
 *
 * What it restates (reference paths relative to the ConvStencil repository):
 *   - naive_1d            src/1d/main.cu:38-47
 *   - naive_box2d1r       src/2d/main.cu:57-115
 *   - naive_star3d1r_step3 / naive_box3d1r_step3   src/3d/main.cu:42-467
 *   - the ping-pong driver of the disabled CHECK_ERROR blocks
 *     (src/2d/main.cu:375-394, src/1d/main.cu:239-258, src/3d/main.cu:642-661)
 *     together with the device-side buffers of the GPU launchers
 *     (src/2d/gpu.cu:272-277, 318-320): buffer 0 holds the input, buffer 1 starts
 *     as zeros, step t reads buffer t%2 and writes only the interior of buffer
 *     (t+1)%2, so halo cells are never written.
 *
 * Semantics kept from those loops:
 *   - every interior point is one expression summed left to right over the taps in
 *     row-major offset order (z outer, then y, then x); the reference's naive_1d
 *     starts from 0 and adds w*x, which is bit-identical to starting from w0*x0;
 *   - plain IEEE double arithmetic, no FMA contraction (built with
 *     -ffp-contract=off), so the result is independent of thread count;
 *   - generalised from the reference's fixed shapes to any tap list in 1-3 D.
 *
 * The synthetic-input generator (SplitMix64, Steele, Lea & Flood 2014) is the
 * same definition the product's host code uses; tests pin both against each
 * other and against fixture values.
 *
 * Parity pins for this file live in tests/test_oracle.py: independent numpy
 * brute force on small grids, impulse response == repeated self-convolution
 * (numpy.convolve / scipy.signal.convolve), constant-field invariance away from
 * the boundary, separable box == successive 1-D convolutions, and golden vectors
 * under tests/golden/.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CSO_MAX_TAPS 512

uint64_t cso_splitmix64(uint64_t *state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* Stream seed: stream 0 = grid data, 1 = weights.  One SplitMix64 draw from
 * (seed * 2 + stream) decorrelates neighbouring seeds. */
uint64_t cso_stream_seed(uint64_t seed, uint64_t stream) {
  uint64_t s = seed * 2ULL + stream;
  return cso_splitmix64(&s);
}

/* Interior filled in row-major interior order with (x>>11)*2^-53*2-1 in [-1,1);
 * every other cell of the padded array is set to 0. */
void cso_fill_data(double *a, const int64_t ext[3], const int64_t org[3], const int64_t n[3], uint64_t seed) {
  uint64_t st = cso_stream_seed(seed, 0);
  size_t total = (size_t)(ext[0] * ext[1] * ext[2]);
  memset(a, 0, total * sizeof(double));
  for (int64_t z = 0; z < n[0]; ++z)
    for (int64_t y = 0; y < n[1]; ++y)
      for (int64_t x = 0; x < n[2]; ++x) {
        uint64_t r = cso_splitmix64(&st);
        double u = (double)(r >> 11) * 0x1.0p-53;
        a[((org[0] + z) * ext[1] + (org[1] + y)) * ext[2] + (org[2] + x)] = u * 2.0 - 1.0;
      }
}

/* count weights uniform in (0,1): ((x>>11)+0.5)*2^-53, normalised by their
 * left-to-right double sum. */
void cso_fill_weights(double *w, int count, uint64_t seed) {
  uint64_t st = cso_stream_seed(seed, 1);
  double s = 0.0;
  for (int i = 0; i < count; ++i) {
    uint64_t r = cso_splitmix64(&st);
    w[i] = ((double)(r >> 11) + 0.5) * 0x1.0p-53;
    s = s + w[i];
  }
  for (int i = 0; i < count; ++i) w[i] = w[i] / s;
}

/* One time step over the interior.  offs holds ntaps triples (dz, dy, dx) in the
 * order the taps are summed.  Each row is accumulated tap by tap into a row
 * buffer; for every point this is the same left-to-right sum as the reference's
 * single expression (acc = w0*x0; acc = acc + w1*x1; ...), written so that the
 * compiler can vectorise along x. */
static void step_rows(const double *in, double *out, const int64_t ext[3], const int64_t org[3], const int64_t n[3],
                      int ntaps, const int64_t *lin, const double *w, double *acc, int64_t z, int64_t y) {
  const int64_t base = ((org[0] + z) * ext[1] + (org[1] + y)) * ext[2] + org[2];
  const int64_t nx = n[2];
  const double *src = in + base + lin[0];
  const double w0 = w[0];
  for (int64_t x = 0; x < nx; ++x) acc[x] = w0 * src[x];
  for (int t = 1; t < ntaps; ++t) {
    const double *s = in + base + lin[t];
    const double wt = w[t];
    for (int64_t x = 0; x < nx; ++x) acc[x] = acc[x] + wt * s[x];
  }
  memcpy(out + base, acc, (size_t)nx * sizeof(double));
}

void cso_step(const double *in, double *out, const int64_t ext[3], const int64_t org[3], const int64_t n[3],
              int ntaps, const int32_t *offs, const double *w, int nthreads) {
  int64_t lin[CSO_MAX_TAPS];
  if (ntaps <= 0 || ntaps > CSO_MAX_TAPS) return;
  for (int t = 0; t < ntaps; ++t)
    lin[t] = ((int64_t)offs[3 * t] * ext[1] + offs[3 * t + 1]) * ext[2] + offs[3 * t + 2];
  const int64_t rows = n[0] * n[1];
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
  {
    double *acc = (double *)malloc((size_t)(n[2] > 0 ? n[2] : 1) * sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int64_t r = 0; r < rows; ++r) step_rows(in, out, ext, org, n, ntaps, lin, w, acc, r / n[1], r % n[1]);
    free(acc);
  }
  (void)nthreads;
}

/* steps time steps ping-ponging between a (holding the input) and b (whose halo
 * is whatever the caller put there; the reference zeroes it).  Returns 0 if the
 * final state is in a, 1 if it is in b. */
int cso_run(double *a, double *b, const int64_t ext[3], const int64_t org[3], const int64_t n[3], int ntaps,
            const int32_t *offs, const double *w, int64_t steps, int nthreads) {
  double *buf[2] = {a, b};
  for (int64_t t = 0; t < steps; ++t) cso_step(buf[t & 1], buf[(t + 1) & 1], ext, org, n, ntaps, offs, w, nthreads);
  return (int)(steps & 1);
}

int cso_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
