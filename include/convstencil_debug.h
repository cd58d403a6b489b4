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

/* Number of compiled 2-D FP64 kernel variants. */
int32_t cs_debug_variant_count(void);
/* info[0..9] = id, mfma shape (16 or 4), radius, shape class (0 box, 1 star), fusion depth,
 * tile rows, tile cols, waves, LDS bytes, LDS row pitch (doubles).  Returns 0, or -1 if i is out
 * of range.  name receives the variant's name (truncated to len). */
int32_t cs_debug_variant_info(int32_t i, int64_t *info, char *name, int32_t len);
/* k-step table of an MFMA shape: dy[], c0[], live[] (cap entries each); returns the table length
 * (k-steps for 16x16x4, flattened entries for 4x4x4), or -1 if cap is too small. */
int32_t cs_debug_ktable(int32_t mf, int32_t radius, int32_t shape, int32_t *dy, int32_t *c0, int32_t *live,
                        int32_t cap);
/* B fragments bfrag[ks * 64 + lane] for dense weights w[(2r+1)^2]; returns the k-step count,
 * or -1 if cap (doubles) is too small. */
int32_t cs_debug_bfrag_2d(int32_t mf, int32_t radius, int32_t shape, const double *w, double *bfrag, int32_t cap);

/* Runs the FP64 MFMA lane-layout probe on the current device (needs a GPU).  which = 0:
 * v_mfma_f64_16x16x4, host_out[64 * 64 * 4] = D[lane][item] for A one-hot at lane p (p-major);
 * which = 1: v_mfma_f64_4x4x4, host_out[64 * 64].  B lane q holds 1 + q/64.  Returns 0 on success. */
int32_t cs_debug_mfma_layout(int32_t which, double *host_out);

#ifdef __cplusplus
}
#endif
#endif /* CONVSTENCIL_DEBUG_H */
