// cs_batch.h -- device-side tables of the persistent batched kernel.
#pragma once
#include "cs_internal.h"

namespace cs {

// Spin iterations (each ~128 clocks of s_sleep) a barrier waits before it gives up
// and flags an error: about 1 s at 2.4 GHz, far beyond any legitimate wait.
static constexpr unsigned kBatchSpinLimit = 1u << 24;

struct BatchGrid {
  double *buf[2];        // ping-pong buffers of this grid
  const double *bfrag;   // B fragments (FP64 doubles, or split slices reinterpreted)
  Geo geo;               // layout and 64 x 64 tile counts
};

struct BatchItem {
  int grid;  // index into the grid table
  int tile;  // tile index within the grid
  int kind;  // tile body (see k2d_batch.hip)
  int src;   // buffer holding this block's input
};

size_t batch_lds_bytes(int prec);
hipError_t batch_prepare(int prec, int *max_wg);
// sync: nsync words (barrier counter, then one work counter per step block), cleared by
// every launch; error: the batch's sticky error word (set by a barrier that timed out,
// never cleared by a launch).
hipError_t batch_launch(int prec, const BatchGrid *grids, const BatchItem *items, const int *block_start,
                        int nblocks, unsigned *sync, int nsync, unsigned *error, unsigned nwg, hipStream_t s);
int batch_kind(int prec, int R, int shape, int kf);
int batch_main_kf(int prec, int R);

}  // namespace cs
