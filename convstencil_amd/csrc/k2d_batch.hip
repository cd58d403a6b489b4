// k2d_batch.hip -- many independent 2-D grids advanced by one persistent launch.
//
// The batched workload (hundreds of grids of uneven sizes and mixed stencils) has
// no counterpart in the reference, which runs one grid per process.  Design:
//   * the host flattens every grid's time loop into "step blocks": in block b each
//     grid still running advances by its own fusion depth (or by one step for a
//     remainder), so block b's work is the list of (grid, tile, kernel kind,
//     input buffer) items of all grids; grids are independent, so the only
//     ordering needed is block b before block b + 1 for the same grid;
//   * one persistent launch: as many workgroups as fit on the device at once
//     (occupancy API) pull items of the current block from an atomic counter
//     (dynamic load balancing over uneven grids), then meet at a grid-wide
//     barrier before the next block;
//   * every wait is bounded: a barrier that does not complete within its spin
//     limit sets an error flag and lets the workgroup continue, so the kernel
//     always terminates; the host reports the flag as an error;
//   * tiles are computed by the same tile bodies as the single-grid kernels
//     (k2d_f64_body.h, k2d_split_body.h).
#include <type_traits>

#include "cs_batch.h"
#include "k2d_f64_body.h"
#include "k2d_split2_body.h"

namespace cs {

// Grid-wide barrier number `index` (0, 1, ...): the counter only grows, the
// target is (index + 1) * nwg.  Release/acquire at agent scope make the tiles
// written before the barrier visible to every XCD after it.
__device__ __forceinline__ void grid_barrier(unsigned *counter, unsigned *error, unsigned target) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    unsigned spins = 0;
    while (__hip_atomic_load(counter, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (++spins > kBatchSpinLimit) {
        __hip_atomic_store(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    __threadfence();
  }
  __syncthreads();
}

// Zeroes the barrier counter and the per-block work counters (a kernel rather than a
// memset so that the sequence can be graph-captured).  The error flag has a word of its
// own that this never clears: a barrier that timed out in any launch stays reported.
__global__ void batch_reset(unsigned *sync, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) sync[i] = 0u;
}

// Kinds (BatchItem::kind): the tile bodies compiled into the batch kernels -- the
// batched workload's two stencils (box2d1r, star2d3r), fused and single-step.
// Every extra body adds to the one kernel's register allocation.
//   FP64:  0 box r1 kf4, 1 box r1 kf1, 2 star r3 kf2, 3 star r3 kf1
//   BF16x3 and FP16x2 with compensated weights (dilated split forms): the same four kinds
template <int PREC>
__device__ __forceinline__ void run_item(const BatchGrid &G, const BatchItem &m, unsigned char *smem, int tid) {
  const double *in = G.buf[m.src];
  double *out = G.buf[m.src ^ 1];
  double *lds = reinterpret_cast<double *>(smem);
  if constexpr (PREC == 0) {
    switch (m.kind) {
      case 0: tile_f64<4, 1, SHAPE_BOX, 4, 64, 64, 8, 1, 64>(in, out, G.bfrag, G.geo, m.tile, lds, tid); break;
      case 1: tile_f64<4, 1, SHAPE_BOX, 1, 64, 64, 8, 1, 64>(in, out, G.bfrag, G.geo, m.tile, lds, tid); break;
      case 2: tile_f64<4, 3, SHAPE_STAR, 2, 64, 64, 8, 1, 64>(in, out, G.bfrag, G.geo, m.tile, lds, tid); break;
      case 3: tile_f64<4, 3, SHAPE_STAR, 1, 64, 64, 8, 1, 64>(in, out, G.bfrag, G.geo, m.tile, lds, tid); break;
      default: break;
    }
  } else {
    using Tr = typename std::conditional<PREC == 1, TrBF16, TrF16>::type;
    const typename Tr::T *bf = reinterpret_cast<const typename Tr::T *>(G.bfrag);
    switch (m.kind) {
      case 0: tile_split2<Tr, 1, SHAPE_BOX, 4, 64, 64, 8, 8>(in, out, bf, G.geo, m.tile, smem, tid); break;
      case 1: tile_split2<Tr, 1, SHAPE_BOX, 1, 64, 64, 8, 8>(in, out, bf, G.geo, m.tile, smem, tid); break;
      case 2: tile_split2<Tr, 3, SHAPE_STAR, 2, 64, 64, 8, 8>(in, out, bf, G.geo, m.tile, smem, tid); break;
      case 3: tile_split2<Tr, 3, SHAPE_STAR, 1, 64, 64, 8, 8>(in, out, bf, G.geo, m.tile, smem, tid); break;
      default: break;
    }
  }
}

template <int PREC>
__global__ __launch_bounds__(512, 4) void k2d_batch(const BatchGrid *__restrict__ grids,
                                                 const BatchItem *__restrict__ items,
                                                 const int *__restrict__ block_start, int nblocks, unsigned *sync,
                                                 unsigned *error, unsigned nwg) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int next;
  unsigned *barrier = sync, *work = sync + 1;
  for (int b = 0; b < nblocks; ++b) {
    const int beg = block_start[b], end = block_start[b + 1];
    for (;;) {
      if (threadIdx.x == 0) next = beg + (int)atomicAdd(&work[b], 1u);
      __syncthreads();
      // uniform over the workgroup; readfirstlane lets the compiler keep the item,
      // the grid's geometry and everything derived from them in scalar registers
      const int it = __builtin_amdgcn_readfirstlane(next);
      __syncthreads();  // every thread has read `next` before thread 0 writes it again
      if (it >= end) break;
      BatchItem m = items[it];
      m.grid = __builtin_amdgcn_readfirstlane(m.grid);
      m.tile = __builtin_amdgcn_readfirstlane(m.tile);
      m.kind = __builtin_amdgcn_readfirstlane(m.kind);
      m.src = __builtin_amdgcn_readfirstlane(m.src);
      // an opaque copy of the thread index: keeps the compiler from hoisting each tile
      // body's per-lane set-up out of this loop, which would hold the set-up of every
      // kind in registers at once
      int tid;
      asm volatile("v_mov_b32 %0, %1" : "=v"(tid) : "v"((int)threadIdx.x));
      run_item<PREC>(grids[m.grid], m, smem, tid);
      __syncthreads();  // LDS is reused by the next item
    }
    grid_barrier(barrier, error, (unsigned)(b + 1) * nwg);
  }
}

template <int PREC>
static size_t batch_lds() {
  if constexpr (PREC == 0) {
    size_t m = 0;
    m = std::max(m, Cfg2D<4, 1, SHAPE_BOX, 4, 64, 64, 8, 1>::LDS);
    m = std::max(m, Cfg2D<4, 3, SHAPE_STAR, 2, 64, 64, 8, 1>::LDS);
    return m;
  } else {
    using Tr = typename std::conditional<PREC == 1, TrBF16, TrF16>::type;
    return std::max(CfgS2<Tr, 1, SHAPE_BOX, 4, 64, 64, 8>::LDS, CfgS2<Tr, 3, SHAPE_STAR, 2, 64, 64, 8>::LDS);
  }
}

// prec: 0 FP64, 1 BF16x3, 2 FP16x2
size_t batch_lds_bytes(int prec) { return prec == 0 ? batch_lds<0>() : prec == 1 ? batch_lds<1>() : batch_lds<2>(); }

static const void *batch_fn(int prec) {
  return prec == 0 ? (const void *)k2d_batch<0> : prec == 1 ? (const void *)k2d_batch<1> : (const void *)k2d_batch<2>;
}

hipError_t batch_prepare(int prec, int *max_wg) {
  const size_t lds = batch_lds_bytes(prec);
  const void *fn = batch_fn(prec);
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  int per_cu = 0, dev = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 512, lds);
  if (e != hipSuccess) return e;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  int cus = 0;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  *max_wg = per_cu * cus;
  return hipSuccess;
}

hipError_t batch_launch(int prec, const BatchGrid *grids, const BatchItem *items, const int *block_start,
                        int nblocks, unsigned *sync, int nsync, unsigned *error, unsigned nwg, hipStream_t s) {
  hipLaunchKernelGGL(batch_reset, dim3(1), dim3(256), 0, s, sync, nsync);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t lds = batch_lds_bytes(prec);
  if (prec == 0)
    hipLaunchKernelGGL(k2d_batch<0>, dim3(nwg), dim3(512), lds, s, grids, items, block_start, nblocks, sync, error, nwg);
  else if (prec == 1)
    hipLaunchKernelGGL(k2d_batch<1>, dim3(nwg), dim3(512), lds, s, grids, items, block_start, nblocks, sync, error, nwg);
  else
    hipLaunchKernelGGL(k2d_batch<2>, dim3(nwg), dim3(512), lds, s, grids, items, block_start, nblocks, sync, error, nwg);
  return hipGetLastError();
}

int batch_kind(int prec, int R, int shape, int kf) {
  if (prec == 0) {
    if (R == 1 && shape == SHAPE_BOX) return kf == 4 ? 0 : kf == 1 ? 1 : -1;
    if (R == 3 && shape == SHAPE_STAR) return kf == 2 ? 2 : kf == 1 ? 3 : -1;
    return -1;
  }
  if (R == 1 && shape == SHAPE_BOX) return kf == 4 ? 0 : kf == 1 ? 1 : -1;
  if (R == 3 && shape == SHAPE_STAR) return kf == 2 ? 2 : kf == 1 ? 3 : -1;
  return -1;
}

int batch_main_kf(int prec, int R) {
  (void)prec;
  return R == 1 ? 4 : 2;
}

}  // namespace cs
