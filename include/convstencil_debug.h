/*
 * convstencil_debug.h -- host-only introspection of libconvstencil's kernel tables.
 *
 * No counterpart in the reference.  These functions touch no device and exist so
 * that CPU tests can check the kernel variants' tables (k-step lists, B
 * fragments, LDS sizes) and emulate a kernel's index arithmetic before anything
 * runs on a GPU.
 */
#ifndef CONVSTENCIL_DEBUG_H
#define CONVSTENCIL_DEBUG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of compiled kernel variants of a family: 1, 2, 3 = FP64 kernels for 1-, 2-, 3-D grids;
 * 4 = 2-D and 5 = 3-D split-precision kernels (there info[1] holds the cs_precision; the 3-D
 * split kernels' z-chunk length is in their name, "zc<n>"); 6 = the pipelined persistent 2-D
 * FP64 kernels (two LDS tile buffers, producer waves "p<n>" in the name; info[8] = both
 * buffers' bytes); 7 = the register-resident 1-D FP64 kernels (info[1] = 6, info[6] = outputs
 * per workgroup, info[9] = cells per lane); 8 = the register-streaming 2-D FP64 kernels
 * (info[1] = 7, info[6] = output columns per wave, info[9] = columns per lane, info[10] =
 * rows in the load ring); 9 = the register-streaming 3-D FP64 kernels, single-step (info[4] = 1)
 * and two fused steps (info[4] = 2; strips of 64 x columns-per-lane columns of which the outer
 * two on either side are recomputed) (info[1] = 9, info[5] = rows per wave, info[6] = output
 * columns per wave, info[7] = waves per workgroup, info[9] = columns per lane, info[10] =
 * planes in the load ring). */
int32_t cs_debug_variant_count(int32_t ndim);
/* info[0..11] = id, mfma shape (16 or 4), radius, shape class (0 box, 1 star), fusion depth,
 * tile rows, tile cols (1-D: tile length), waves, LDS bytes, LDS row pitch / buffer length
 * (doubles), staging of fused steps (0 = two LDS buffers, 1 = one buffer + registers; for 3-D
 * variants: the z-chunk length), LDS storage column of tile column 0 (2-D 4x4x4 variants, else 0).
 * Returns 0, or -1 if i is out of range.  name receives the variant's name. */
int32_t cs_debug_variant_info(int32_t ndim, int32_t i, int64_t *info, char *name, int32_t len);
/* k-step table of the 16x16x4 kernels (mf = 16) or the flattened 4x4x4 table of the 1-D kernels'
 * layout (mf = 4): dy[], c0[], live[] (cap entries each); returns the table length, or -1. */
int32_t cs_debug_ktable(int32_t mf, int32_t radius, int32_t shape, int32_t *dy, int32_t *c0, int32_t *live,
                        int32_t cap);
/* Dilated 4x4x4 table of the 2-D FP64 kernels (8 x 8 output blocks): hdr[4] = {pair steps np, single
 * steps ns, pair column parity, lane map}; entry e = 4 * step + q (pair steps first, then single
 * steps) is the footprint cell (ty[e], tx[e]) read by lane group q (a pair also reads tx + 1),
 * live[e] = 0 for a zero-weight slot.  Lane map 1: lane l (block b = (l >> 2) & 3, A row i = l & 3)
 * reads at row 2 i + ty, column 2 ((b + i) mod 4) + tx of the block, and holds the result of D row
 * l >> 4 at the same kind of sub-block; map 0: row 2 b, column 2 i.  Returns 4 * (np + ns), or -1
 * if cap is too small. */
int32_t cs_debug_dtable(int32_t radius, int32_t shape, int32_t *hdr, int32_t *ty, int32_t *tx, int32_t *live,
                        int32_t cap);
/* B fragments bfrag[ks * 64 + lane] for dense weights w[(2r+1)^2]; returns the k-step count,
 * or -1 if cap (doubles) is too small. */
int32_t cs_debug_bfrag_2d(int32_t mf, int32_t radius, int32_t shape, const double *w, double *bfrag, int32_t cap);
/* Same for the 1-D family (v_mfma_f64_4x4x4, weights w[2r+1]). */
int32_t cs_debug_bfrag_1d(int32_t radius, const double *w, double *bfrag, int32_t cap);

/* 3-D k-step table (grouped by dz): dz[ks], and dy/c/live[4 * ks + k] for lane group k; returns
 * the k-step count or -1 if cap (k-steps) is too small.  shape: 0 box, 1 star, 2 diamond. */
int32_t cs_debug_ktable3d(int32_t radius, int32_t shape, int32_t *dz, int32_t *dy, int32_t *c, int32_t *live,
                          int32_t cap);
/* 3-D B fragments bfrag[ks * 64 + lane] for dense weights w[(2r+1)^3]. */
int32_t cs_debug_bfrag_3d(int32_t radius, int32_t shape, const double *w, double *bfrag, int32_t cap);

/* The 2-D CLI's fused weights: the 49 coefficients of the cube of a row-major 3x3 stencil, either
 * exactly (exact != 0) or as the reference's hand-expanded table computes them
 * (src/2d/main.cu:233-281). */
void cs_debug_fused_2d(int32_t exact, const double *p9, double *out49);

/* Split-precision B fragments as raw slice elements (bf16 / fp32 / fp16 by precision), layout
 * [weight slice][instruction][lane][element]; returns the byte count, or 0 if cap is too small. */
int64_t cs_debug_bfrag_split(int32_t precision, int32_t radius, const double *w, void *out, int64_t cap);
/* Dilated split form (variants with staging 2): A fragments of the weights' slices, layout
 * [weight slice][instruction][lane][8 elements], for shape class 0 box / 1 star; returns the byte
 * count, or 0 if cap is too small or the precision has no dilated form (FP32x2). */
int64_t cs_debug_bfrag_split2(int32_t precision, int32_t radius, int32_t shape, const double *w, void *out,
                              int64_t cap);
/* Chunk table of the dilated split form: *c0 = chunk start residue (mod 8); chunk e reads footprint
 * row ty[e], columns tx[e] .. tx[e] + 7 (live[e] = 0: zero-weight padding).  Returns the chunk count
 * (4 per MFMA), or -1 if cap is too small. */
int32_t cs_debug_ctable(int32_t radius, int32_t shape, int32_t *c0, int32_t *ty, int32_t *tx, int32_t *live,
                        int32_t cap);

/* Anchor map of that chunk table's 16 x 16 blocks: D column n (lane l: n = l & 15) is the 2 x 8
 * cell group at row pair n & 7 and column octet (n >> 3) ^ ((n >> 2) & 1) for map 1, n >> 3 for
 * map 0 (the 3-D split kernels); D row m = 4 (l >> 4) + v is its cell (m >> 3, m & 7). */
int32_t cs_debug_ctable_map(int32_t radius, int32_t shape);

/* Wave decomposition of one launch of the register-streaming 2-D kernels (family 8): for
 * ny x nx interior cells, `ox` output columns per strip, the stencil's radius, fusion depth
 * kf, a load ring of `ring` rows and `slots` resident waves, split[0..4] = strips, he, hm, we,
 * wi (rows per checked chunk, rows per plain chunk, chunks per edge strip, chunks per other
 * strip); waves = the edge strips' chunks first.  Returns 0, or -1 for invalid arguments. */
int32_t cs_debug_reg2d_split(int64_t ny, int64_t nx, int32_t ox, int32_t radius, int32_t kf, int32_t ring,
                             int64_t slots, int64_t *split);

/* Wave `wid` of that decomposition: rows[0..2] = strip, first row, end row; returns 1, or 0 for
 * a wave without work. */
int32_t cs_debug_reg2d_chunk(const int64_t *split, int64_t ny, int64_t wid, int64_t *rows);

/* Decomposition of one launch of the register-streaming 3-D kernels (family 9) for nz x ny x
 * nx interior cells, `ox` output columns and `by` rows per wave and `slots` resident waves:
 * split[0..3] = strips, y-tiles, z-chunks, planes per z-chunk; wave w takes y-tile w mod
 * y-tiles, then strip, then z-chunk.  Returns 0, or -1 for invalid arguments. */
int32_t cs_debug_reg3d_split(int64_t nz, int64_t ny, int64_t nx, int32_t ox, int32_t by, int64_t slots,
                             int64_t *split);

/* Dilated table of plane dz of the 3-D 8 x 8 block kernels (variants with mfma shape 8), in the
 * layout of cs_debug_dtable; all planes share the pair parity hdr[2].  Returns 4 * (np + ns) or -1. */
int32_t cs_debug_dtable3(int32_t radius, int32_t shape, int32_t dz, int32_t *hdr, int32_t *ty, int32_t *tx,
                         int32_t *live, int32_t cap);
/* B fragments of the 3-D 8 x 8 block kernels, planes dz = -r .. r in order; returns the k-step count. */
int32_t cs_debug_bfrag_3d2(int32_t radius, int32_t shape, const double *w, double *bfrag, int32_t cap);

/* Chunk table of plane dz of the split-precision 3-D kernels (family 5 of cs_debug_variant_info),
 * in the layout of cs_debug_ctable; every plane shares *c0.  Returns the plane's chunk count (4 per
 * MFMA) or -1. */
int32_t cs_debug_ctable3(int32_t radius, int32_t shape, int32_t dz, int32_t *c0, int32_t *ty, int32_t *tx,
                         int32_t *live, int32_t cap);
/* A fragments of the split-precision 3-D kernels, layout [weight slice][instruction][lane][8
 * elements] with the instructions of planes dz = -r .. r in order; precision 1 (BF16x3) or 4 (FP16
 * with three weight slices).  Returns the byte count, or 0. */
int64_t cs_debug_bfrag_split3(int32_t precision, int32_t radius, int32_t shape, const double *w, void *out,
                              int64_t cap);

/* The FP16 range guard of plans and batches (CS_PREC_FP16X2 / CS_PREC_FP16W3 need |values| < 2^15):
 * from a bound amax on the state's magnitude and the weights' sum of magnitudes wabs, *bound
 * receives the bound after `steps` steps (amax * max(1, wabs)^steps); returns 1 if a run of
 * that many steps is allowed for `precision`, 0 if it is refused. */
int32_t cs_debug_fp16_range(int32_t precision, double amax, double wabs, int64_t steps, double *bound);
/* The guard's bound of an uploaded array: max |a[i]|, +inf if any value is not finite. */
double cs_debug_abs_max(const double *a, int64_t n);
/* The batch's error latch: feeds the device error word's successive values flags[0..n) to a
 * host-only batch object and stores the status each read returns (0 = CS_OK, else
 * CS_ERR_DEVICE: once a barrier has timed out, every later call fails).  Returns 0. */
int32_t cs_debug_batch_latch(const uint32_t *flags, int32_t n, int32_t *status);

/* Times `reps` plain device copies of `count` doubles (needs a GPU); the kernel is the
 * reference for HBM byte counters.  Returns 0 on success. */
int32_t cs_debug_copy_bench(int64_t count, int32_t reps, double *ms_per_copy);

/* Runs the FP64 MFMA lane-layout probe on the current device (needs a GPU).  which = 0:
 * v_mfma_f64_16x16x4, host_out[64 * 64 * 4] = D[lane][item] for A one-hot at lane p (p-major);
 * which = 1: v_mfma_f64_4x4x4, host_out[64 * 64].  B lane q holds 1 + q/64.  Returns 0 on success. */
int32_t cs_debug_mfma_layout(int32_t which, double *host_out);

#ifdef __cplusplus
}
#endif
#endif /* CONVSTENCIL_DEBUG_H */
