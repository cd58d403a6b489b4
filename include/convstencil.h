/*
 * convstencil.h -- C ABI of the MI355X (gfx950) ConvStencil library.
 *
 * Two layers:
 *
 * 1. Drop-in launchers with the reference's names and argument meaning.  The
 *    reference declares them (with C++ linkage and a void return) in
 *      src/1d/1d_utils.h:75,85   gpu_1d1r, gpu_star_1d2r_step2
 *      src/2d/2d_utils.h:77-79   gpu_box_2d1r, gpu_box_2d3r
 *      src/3d/3d_utils.h:75-77   gpu_box_3d1r, gpu_star_3d1r
 *    and implements them in src/1d/gpu_1r.cu:227-301, src/1d/gpu_2r.cu:85-159,
 *    src/2d/gpu.cu:245-333,336-425, src/3d/gpu_box.cu:446-536,
 *    src/3d/gpu_star.cu:238-332.  Same host-pointer arguments, same padded
 *    layouts, same weight layouts, same stdout lines ("ConvStencil(2D): ",
 *    "Time = N[ms]", "GStencil/s = X").  `times` counts launches as in the
 *    reference (each launch applies the given, already time-fused, weights once).
 *    The --breakdown ladders' launchers (src/1d/1d_utils.h:77-83,
 *    src/2d/2d_utils.h:81-87, src/3d/3d_utils.h:79-85) are declared beside them.
 *    Where the reference prints a CUDA error and exits, these print the error and
 *    return a non-zero cs_status; the CLIs turn that into exit code 1.
 *
 * 2. A plan API (cs_plan_*) for device-resident time loops: many steps on grids
 *    that stay in HBM, optional hipGraph capture, split precision, temporal
 *    fusion, device-pointer access for benchmarking.  This has no single
 *    counterpart in the reference; it generalises the launchers' bodies
 *    (B-matrix build, ping-pong buffers, time loop) so that they can be reused.
 *
 * All functions are thread-compatible: distinct plans may be used from distinct
 * threads; a single plan must not be used concurrently.
 */
#ifndef CONVSTENCIL_H
#define CONVSTENCIL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cs_status {
  CS_OK = 0,
  CS_ERR_INVALID = 1,     /* bad argument: null pointer, bad size, halo narrower than the radius ... */
  CS_ERR_UNSUPPORTED = 2, /* valid request this build has no kernel for */
  CS_ERR_ALLOC = 3,       /* device or host allocation failed */
  CS_ERR_DEVICE = 4,      /* a HIP call failed */
  CS_ERR_NO_DEVICE = 5    /* no gfx950 device visible */
} cs_status;

typedef enum cs_precision {
  CS_PREC_FP64 = 0,   /* plain FP64: stencil2row on FP64 MFMA (v_mfma_f64_4x4x4 / 16x16x4) */
  CS_PREC_BF16X3 = 1, /* 2-D: state/weights split into 3 BF16 slices, BF16 MFMA, FP32 accumulate per level
                         (6 of 9 slice products), FP64 recombine; state FP64 in HBM */
  CS_PREC_FP32X2 = 2, /* 2-D: FP32 hi/lo slices, FP32 MFMA (3 products), FP64 recombine */
  CS_PREC_FP16X2 = 3, /* 2-D: 2 FP16 slices rescaled by 2^11 (3 products), FP32 accumulate, FP64
                         recombine; values must stay below 2^15 in magnitude */
  CS_PREC_FP16W3 = 4  /* 2-D: weights in 3 FP16 slices, state in 2 (4 products: the weights' error no
                         longer biases every step), FP32 accumulate, FP64 recombine; values below 2^15 */
} cs_precision;
/* FP16 range guard (CS_PREC_FP16X2 / CS_PREC_FP16W3, plans and batches): the library keeps a
 * bound on the state's magnitude -- 0 for a new plan, max |cell| (halo included) of an
 * uploaded array, 1 after a random fill -- and multiplies it by max(1, sum |w|) per step.  A
 * run whose bound would reach 2^15 is refused with CS_ERR_INVALID before anything is
 * launched.  Values written into the device buffers directly (cs_plan_device_buffers) are
 * outside the guard. */

/* Padded array layout.  Dimensions are listed slowest first; a 2-D grid has
 * ext[0] = n[0] = 1, org[0] = 0 and a 1-D grid also has ext[1] = n[1] = 1.
 * Element (z, y, x) of the padded array is at ((z * ext[1]) + y) * ext[2] + x.
 * The interior is [org[d], org[d] + n[d]) on every axis; every other cell is
 * halo, read by the stencil and never written. */
typedef struct cs_layout {
  int32_t ndim;   /* 1, 2 or 3 */
  int32_t reserved;
  int64_t ext[3];
  int64_t org[3];
  int64_t n[3];
} cs_layout;

/* A stencil of radius r in ndim dimensions is given by its dense weight cube
 * w[(2r+1)^ndim] in row-major (z, y, x) order, tap (dz, dy, dx) at index
 * ((dz + r) * (2r+1) + (dy + r)) * (2r+1) + (dx + r); taps that are not part of
 * the stencil are 0.  Radius 1..3 is supported. */
typedef struct cs_stencil {
  int32_t ndim;
  int32_t radius;
  const double *weights;
} cs_stencil;

typedef struct cs_options {
  int32_t precision;     /* cs_precision */
  int32_t fusion_depth;  /* time steps per launch; 0 = choose (autotuner / defaults) */
  int32_t use_graph;     /* 1 = capture the launch sequence of cs_plan_run in a hipGraph */
  int32_t tile_variant;  /* 0 = choose; otherwise a kernel-variant id from cs_plan_describe */
  int32_t verbose;
  int32_t autotune;      /* 0 = defaults; 1 = the autotuner's cached choice, timing every fitting
                            variant on a miss; 2 = time them again.  Ignored with tile_variant. */
} cs_options;

typedef struct cs_report {
  double elapsed_ms;       /* device time of the time loop (HIP events on the plan stream) */
  double gstencil_per_s;   /* interior points * steps / elapsed / 1e9 */
  int64_t launches;        /* kernel launches issued */
  int32_t fusion_depth;    /* executed steps per launch */
  int32_t variant;         /* kernel variant used */
} cs_report;

typedef struct cs_plan cs_plan;

/* ---- library ---- */
const char *cs_version(void);
const char *cs_status_string(cs_status s);
/* Fills *count with the number of visible devices whose gfx name is gfx950. */
cs_status cs_device_count(int32_t *count);

/* ---- plans: state lives in two device buffers (ping-pong) ---- */
cs_status cs_plan_create(cs_plan **plan, const cs_layout *layout, const cs_stencil *stencil, const cs_options *opt);
cs_status cs_plan_destroy(cs_plan *plan);
/* Copies the padded host array into buffer 0 and zeroes buffer 1 (the reference's
 * cudaMemcpy + cudaMemset pair, src/2d/gpu.cu:273-277); resets the step count. */
cs_status cs_plan_upload(cs_plan *plan, const double *host_in);
/* Fills buffer 0 with the seeded synthetic grid on the device (interior uniform in
 * [-1,1) from SplitMix64, halo 0; same bits as the host generator) and zeroes buffer 1. */
cs_status cs_plan_fill_random(cs_plan *plan, uint64_t seed);
/* Advances the state by `steps` physical steps (asynchronous on the plan stream). */
cs_status cs_plan_run(cs_plan *plan, int64_t steps, cs_report *report);
/* Runs `steps` once untimed, then `reps` timed repetitions of `steps` steps; reports
 * the mean device time per repetition in report->elapsed_ms. */
cs_status cs_plan_time(cs_plan *plan, int64_t steps, int32_t warmup, int32_t reps, cs_report *report);
cs_status cs_plan_sync(cs_plan *plan);
/* Copies the whole padded array holding the current state to the host. */
cs_status cs_plan_download(cs_plan *plan, double *host_out);
/* Device pointer of the buffer holding the current state / the other buffer. */
cs_status cs_plan_device_buffers(cs_plan *plan, double **current, double **other);
/* Human-readable description of the kernel choice (variant, tile, fusion depth). */
cs_status cs_plan_describe(cs_plan *plan, char *buf, int32_t len);

/* ---- batches: many independent 2-D grids advanced by one persistent launch ----
 * No counterpart in the reference (one grid per process).  Every grid has its own
 * layout and stencil; supported stencils are the radius-1 box (box2d1r) and the
 * radius-3 star (star2d3r), in FP64, CS_PREC_BF16X3 and CS_PREC_FP16X2.  cs_batch_run advances every grid by
 * `steps` steps with a single persistent kernel launch (plus a small reset kernel),
 * optionally captured in a hipGraph (opt->use_graph). */
typedef struct cs_batch cs_batch;
cs_status cs_batch_create(cs_batch **batch, int32_t count, const cs_layout *layouts, const cs_stencil *stencils,
                          const cs_options *opt);
cs_status cs_batch_destroy(cs_batch *batch);
cs_status cs_batch_upload(cs_batch *batch, int32_t grid, const double *host_in);
/* Seeds grid i with seeds[i] (as cs_plan_fill_random). */
cs_status cs_batch_fill_random(cs_batch *batch, const uint64_t *seeds);
/* FP64 batches whose grids all carry a uniform halo advance by multi-grid launches of the
 * register-streaming kernels; the others by one persistent launch, which needs every
 * workgroup resident at once and whose grid-wide barriers wait a bounded time.  A barrier that times out sets the batch's sticky error word: that run,
 * and every later cs_batch_run / cs_batch_time / cs_batch_download of this batch, returns
 * CS_ERR_DEVICE (the grids' contents are undefined from then on). */
cs_status cs_batch_run(cs_batch *batch, int64_t steps, cs_report *report);
cs_status cs_batch_time(cs_batch *batch, int64_t steps, int32_t warmup, int32_t reps, cs_report *report);
cs_status cs_batch_download(cs_batch *batch, int32_t grid, double *host_out);
cs_status cs_batch_describe(cs_batch *batch, char *buf, int32_t len);

/* ---- autotuner cache ----
 * No counterpart in the reference (fixed tiles).  A plan created with autotune times the
 * fitting variants for 24 steps on its own buffers (before any data is uploaded) and keeps
 * the fastest; choices are cached per key "d<ndim>_r<radius>_s<shape>_p<precision>_n<nz>x
 * <ny>x<nx>_g<graph>" in memory and in a text file: $CONVSTENCIL_TUNE_CACHE (empty: memory
 * only), else $XDG_CACHE_HOME/convstencil_amd/tune.txt, else $HOME/.cache/convstencil_amd/tune.txt. */
const char *cs_tune_cache_path(void);
/* Forgets the in-memory cache; the next use re-reads the environment and the file. */
cs_status cs_tune_cache_reset(void);
/* CS_OK and the cached variant (and its time per step, if ms_per_step is not NULL), or
 * CS_ERR_UNSUPPORTED if the key has no entry; keys contain no white space. */
cs_status cs_tune_lookup(const char *key, int32_t *variant, double *ms_per_step);
cs_status cs_tune_store(const char *key, int32_t variant, double ms_per_step);
/* The cache key of a plan with these arguments (host only; no device needed). */
cs_status cs_tune_key(const cs_layout *layout, const cs_stencil *stencil, const cs_options *opt, char *buf,
                      int32_t len);

/* ---- allocation-free device run ----
 * No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): advances
 * a 2-D FP64 grid held in two caller-owned device buffers by `steps` steps, ping-ponging
 * between them on `stream` (a hipStream_t; NULL = the default stream), with the plan's
 * kernels and default fusion depth (opt: fusion_depth and tile_variant as for a plan;
 * precision FP64 only, use_graph and autotune 0).  buf0 holds the initial state; both
 * buffers must carry the same halo (fused launches read only their input's halo), else
 * pass fusion_depth 1 for the reference's alternating-halo single steps.  *final_buffer
 * receives 0 or 1, the buffer that holds the result.  Nothing is allocated: the kernels'
 * weight operands (dense weights and/or B fragments, radius <= 3: at most 1024 doubles)
 * live in a fixed device array of the library, one per GPU, that every call on that GPU
 * shares; calls that overlap on one GPU race on it.  Larger radii return CS_ERR_UNSUPPORTED.  The call only enqueues kernels:
 * synchronise the stream before reading the result. */
cs_status cs_run_2d_noalloc(const cs_layout *layout, const cs_stencil *stencil, double *buf0, double *buf1,
                            int64_t steps, const cs_options *opt, void *stream, int32_t *final_buffer);

/* ---- one-shot host-to-host run (what the reference launchers do, minus printing) ---- */
cs_status cs_run(const cs_layout *layout, const cs_stencil *stencil, const double *host_in, double *host_out,
                 int64_t steps, const cs_options *opt, cs_report *report);

/* ---- synthetic input generator (host side; bit-identical to cs_plan_fill_random) ---- */
void cs_fill_random_host(double *a, const cs_layout *layout, uint64_t seed);
void cs_fill_weights_host(double *w, int32_t count, uint64_t seed);

/* ---- grid files (host only; no device needed) ---- */
/* The reference's text dump of a whole padded array (save_to_txt, src/1d/main.cu:21-36,
 * src/2d/main.cu:36-55, src/3d/main.cu:22-40; written as output_gpu.txt by its drivers
 * built with WRITE_OUTPUT, src/2d/main.cu:410-415): "%.0f" values, 1-D "index value"
 * lines, 2-D tab-terminated rows, 3-D planes of rows each followed by "\nNext\n". */
cs_status cs_grid_save_txt(const char *path, const cs_layout *layout, const double *host);
/* Binary FP64 grid (no counterpart in the reference): a 96-byte header ("CSGRID64", version
 * 1, ndim, ext[3], org[3], n[3]) and the padded array, exact.  cs_grid_load_f64 fills
 * *layout and, if host is not NULL, reads the array into host (capacity in doubles);
 * CS_ERR_INVALID for a missing, foreign, truncated or too large file. */
cs_status cs_grid_save_f64(const char *path, const cs_layout *layout, const double *host);
cs_status cs_grid_load_f64(const char *path, cs_layout *layout, double *host, int64_t capacity);

/* ---- drop-in launchers (reference names; see the header comment) ---- */
/* 1d1r: host arrays of input_n + 7 doubles; interior at [4, input_n + 4) of `in`;
 * 7 fused weights; out[j] = final[j + 1] for j < input_n + 6  (src/1d/gpu_1r.cu:227-301). */
int gpu_1d1r(const double *in, double *out, const double *params, int times, int input_n);
/* 1d2r: host arrays of input_n + 8 doubles; interior [4, input_n + 4); 9 fused weights;
 * out[j] = final[j] for j < input_n + 7  (src/1d/gpu_2r.cu:85-159). */
int gpu_star_1d2r_step2(const double *in, double *out, const double *params, int times, int input_n);
/* The 1-D --breakdown ladder (src/1d/1d_utils.h:77-83, src/1d/gpu_1r.cu:305-651): gpu_1d1r's
 * arguments and result on one rung's kernel each -- 1 FP64 FMA from HBM/L2, 2 FP64 FMA from an
 * LDS tile, 3 the MFMA kernel on the reference's 1024-output tiling, 4 the production
 * single-step kernel -- printing "Experiment - Breakdown(1D) K: ", "Time = N[ms]", "GStencil/s = X". */
int gpu_1d1r_breakdown1(const double *in, double *out, const double *params, int times, int input_n);
int gpu_1d1r_breakdown2(const double *in, double *out, const double *params, int times, int input_n);
int gpu_1d1r_breakdown3(const double *in, double *out, const double *params, int times, int input_n);
int gpu_1d1r_breakdown4(const double *in, double *out, const double *params, int times, int input_n);
/* 2D: host arrays of (input_m + 6) x (input_n + 8); interior rows [3, m+3), cols [4, n+4);
 * 49 weights row-major over (dy, dx) in [-3,3]^2  (src/2d/gpu.cu:245-333).  Prints with the
 * 3-steps-per-launch GStencil/s formula of src/2d/gpu.cu:329. */
int gpu_box_2d1r(const double *in, double *out, const double *params, int times, int input_m, int input_n);
/* Same computation, 1 step per launch in the GStencil/s line (src/2d/gpu.cu:336-425). */
int gpu_box_2d3r(const double *in, double *out, const double *params, int times, int input_m, int input_n);
/* The --breakdown ladder (src/2d/2d_utils.h:81-87, src/2d/gpu.cu:434-801): gpu_box_2d1r's
 * arguments and result on one rung's kernel each -- 1 FP64 FMA from HBM/L2, 2 FP64 FMA from an
 * LDS tile, 3 the dilated MFMA tables with 128-byte aligned LDS rows, 4 the production
 * single-step kernel (the bank model's pitch, the output staged in LDS) -- printing "Experiment - Breakdown(2D) K: ", "Time = N[ms]", "GStencil/s = X". */
int gpu_box_2d1r_breakdown1(const double *in, double *out, const double *params, int times, int input_m, int input_n);
int gpu_box_2d1r_breakdown2(const double *in, double *out, const double *params, int times, int input_m, int input_n);
int gpu_box_2d1r_breakdown3(const double *in, double *out, const double *params, int times, int input_m, int input_n);
int gpu_box_2d1r_breakdown4(const double *in, double *out, const double *params, int times, int input_m, int input_n);
/* 3D box: (h+6)(m+6)(n+6) arrays, interior [3, .+3); 196 weights = 4 planes of 7x7 for
 * |dz| = 3, 2, 1, 0 (z-symmetric 343-point stencil)  (src/3d/gpu_box.cu:446-536). */
int gpu_box_3d1r(const double *in, double *out, const double *params, int times, int input_h, int input_m,
                 int input_n);
/* The 3-D --breakdown ladder (src/3d/3d_utils.h:79-85, src/3d/gpu_box.cu:538-904):
 * gpu_box_3d1r's arguments and result on one rung's kernel each -- 1 FP64 FMA from HBM/L2, 2
 * FP64 FMA from a ring of LDS planes, 3 the MFMA strip kernel with 128-byte aligned LDS rows,
 * 4 the production kernel (register-blocked FMAs over streamed planes) -- printing "Experiment - Breakdown(3D) K: ", "Time = N[ms]", the
 * seconds, "GStencil/s = X". */
int gpu_box_3d1r_breakdown1(const double *in, double *out, const double *params, int times, int input_h, int input_m,
                            int input_n);
int gpu_box_3d1r_breakdown2(const double *in, double *out, const double *params, int times, int input_h, int input_m,
                            int input_n);
int gpu_box_3d1r_breakdown3(const double *in, double *out, const double *params, int times, int input_h, int input_m,
                            int input_n);
int gpu_box_3d1r_breakdown4(const double *in, double *out, const double *params, int times, int input_h, int input_m,
                            int input_n);
/* 3D star: same layout; 44 weights for the 63-point octahedron |dz|+|dy|+|dx| <= 3, planes
 * |dz| = 3, 2, 1, 0 with 1, 5, 13, 25 taps, z-symmetric  (src/3d/gpu_star.cu:238-332). */
int gpu_star_3d1r(const double *in, double *out, const double *params, int times, int input_h, int input_m,
                  int input_n);

#ifdef __cplusplus
}
#endif
#endif /* CONVSTENCIL_H */
