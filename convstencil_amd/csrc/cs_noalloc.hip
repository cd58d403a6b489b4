// cs_noalloc.hip -- the allocation-free entry point (cs_run_2d_noalloc, include/convstencil.h).
//
// A plan (cs_plan_*) owns its device memory: the two grid buffers and the kernels' weight
// operands (the B fragments of the dilated v_mfma_f64_4x4x4 tables, or the dense weights
// of the register-streaming kernels), which it allocates with hipMalloc.  Callers that
// cannot allocate -- code under HIP graph capture, tight serving loops, harnesses that
// forbid it -- pass both grid buffers themselves; the weight operands go to `scratch`, a
// fixed device array of this library (each GPU holds one copy until the process ends).
// They are built on the host (the same tables as a plan's) in compact form, 16 values per
// k-step or the dense weights, and written into scratch by a small kernel whose argument
// carries them, so no host-to-device copy and no allocation is made: the whole call is
// kernel launches on the caller's stream.
#include <atomic>

#include "cs_host_util.h"
#include "cs_internal.h"

namespace cs {

cs_status pick_fp64_variants(int ndim, int R, int shape, int fusion_depth, int tile_variant, int64_t cells,
                             const Variant **vmain, const Variant **vone, const Variant **vrem, int vrem_cap);

namespace {

// B fragments of one 2-D FP64 plan: at most 16 k-steps (radius-3 box) x 64 lanes.
constexpr int kScratchDoubles = 16 * 64;
constexpr int kMaxDevices = 64;

// Compact fragments: lane l of k-step ks holds value (ks, l >> 4, l & 3) -- the dilated
// tables give the four lane groups one tap each and the lanes of a group differ only in
// their output column l & 3.
struct CompactFrags {
  double v[16 * 16];
};

}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

// dense == 0: expand nks k-steps of compact fragments; dense == 1: copy c.v[0 .. 63] (the
// dense weights, zero-padded) to the first 64 entries
__global__ void expand_frags(CompactFrags c, int nks, int dense, double *dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (dense) {
    if (i < 64) dst[i] = c.v[i];
  } else if (i < nks * 64) {
    const int ks = i >> 6, lane = i & 63;
    dst[i] = c.v[ks * 16 + (lane >> 4) * 4 + (lane & 3)];
  }
}

namespace {

// scratch's device address on each GPU (hipGetSymbolAddress, once per device with that
// device current; the host-side name `scratch` is not a device address)
std::atomic<double *> g_scratch[kMaxDevices];

cs_status scratch_ptr(int dev, double **out) {
  if (dev < 0 || dev >= kMaxDevices) return CS_ERR_UNSUPPORTED;
  double *p = g_scratch[dev].load(std::memory_order_acquire);
  if (!p) {
    void *q = nullptr;
    CS_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
    p = static_cast<double *>(q);
    g_scratch[dev].store(p, std::memory_order_release);  // racing first calls store the same address
  }
  *out = p;
  return CS_OK;
}

}  // namespace
}  // namespace cs

using namespace cs::host;

extern "C" cs_status cs_run_2d_noalloc(const cs_layout *layout, const cs_stencil *st, double *buf0, double *buf1,
                                       int64_t steps, const cs_options *opt, void *stream, int32_t *final_buffer) {
  if (!layout || !st || !st->weights || !buf0 || !buf1 || buf0 == buf1 || steps < 0) return CS_ERR_INVALID;
  if (layout->ndim != 2 || st->ndim != 2 || st->radius < 1) return CS_ERR_INVALID;
  if (!layout_valid(layout, st->radius)) return CS_ERR_INVALID;
  const int R = st->radius;
  if (R > 3) {
    fprintf(stderr, "convstencil: cs_run_2d_noalloc: radius %d exceeds the scratch array (radius 3 at most)\n", R);
    return CS_ERR_UNSUPPORTED;
  }
  cs_options o{};
  if (opt) o = *opt;
  if (o.precision != CS_PREC_FP64 || o.use_graph || o.autotune) return CS_ERR_UNSUPPORTED;
  if (final_buffer) *final_buffer = 0;
  int dev = 0;
  CS_TRY(check_device(&dev));
  const int shape = classify_2d(R, st->weights);
  const cs::Variant *vmain = nullptr, *vone = nullptr, *vrem[32] = {};
  CS_TRY(cs::pick_fp64_variants(2, R, shape, o.fusion_depth, o.tile_variant, layout->n[1] * layout->n[2], &vmain,
                                &vone, vrem, 32));
  // the dilated tables (mf 4: B fragments) and the register-streaming kernels (mf 7: dense
  // weights, their single steps on the mf 4 kernel as in a plan): scratch holds the dense
  // weights first when the main variant is mf 7, then the B fragments when an mf 4 kernel runs
  const bool dense = vmain->mf == 7, frags4 = vmain->mf == 4 || vone->mf == 4;
  if (vmain->mf != 4 && !dense) return CS_ERR_UNSUPPORTED;
  const int D = 2 * R + 1;
  const int nks = frags4 ? cs::bfrag_ksteps_2d_f64(4, R, vmain->shape) : 0;
  const int64_t off4 = dense ? 64 : 0;  // doubles before the B fragments
  if (nks > 16 || off4 + nks * 64 > cs::kScratchDoubles) {
    fprintf(stderr, "convstencil: cs_run_2d_noalloc: %d k-steps exceed the scratch array (%d doubles)\n", nks,
            cs::kScratchDoubles);
    return CS_ERR_UNSUPPORTED;
  }
  // weight operands on the host (fixed arrays: nothing is allocated), compacted
  double wd[49];
  for (int i = 0; i < D * D; ++i) wd[i] = st->weights[i];
  cs::CompactFrags cd{}, cf{};
  if (dense)
    for (int i = 0; i < D * D; ++i) cd.v[i] = wd[i];
  if (frags4) {
    double full[16 * 64];
    cs::build_bfrag_2d_f64(4, R, vmain->shape, wd, full);
    for (int ks = 0; ks < nks; ++ks)
      for (int lane = 0; lane < 64; ++lane) {
        const double v = full[ks * 64 + lane], e = cf.v[ks * 16 + (lane >> 4) * 4 + (lane & 3)];
        if ((lane & 12) == 0)
          cf.v[ks * 16 + (lane >> 4) * 4 + (lane & 3)] = v;
        else if (v != e)
          return CS_ERR_DEVICE;  // the tables' lane structure changed: never expand wrongly
      }
  }
  double *frags = nullptr;
  CS_TRY(cs::scratch_ptr(dev, &frags));
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (dense) hipLaunchKernelGGL(cs::expand_frags, dim3(1), dim3(256), 0, s, cd, 1, 1, frags);
  if (frags4)
    hipLaunchKernelGGL(cs::expand_frags, dim3((nks * 64 + 255) / 256), dim3(256), 0, s, cf, nks, 0, frags + off4);
  CS_HIP(hipGetLastError());
  for (const cs::Variant *u : {vmain, vone}) CS_HIP(u->prepare());
  // the reference's ping-pong: fused launches are exact when both buffers carry the same
  // halo (the caller's contract); fusion_depth 1 gives the alternating-halo single steps
  if (layout->n[1] == 0 || layout->n[2] == 0) {
    if (final_buffer) *final_buffer = (int32_t)(steps & 1);
    return CS_OK;
  }
  cs::Geo g{};
  for (int d = 0; d < 3; ++d) {
    g.ext[d] = layout->ext[d];
    g.org[d] = layout->org[d];
    g.n[d] = layout->n[d];
  }
  g.tiles[0] = 1;
  double *b[2] = {buf0, buf1};
  int cur = 0;
  int64_t left = steps;
  while (left > 0) {
    const cs::Variant *v = vone;
    if (left >= vmain->kf) {
      v = vmain;
    } else {
      for (int r = 1; r < vmain->kf; ++r)
        if (vrem[r] && vrem[r]->kf <= left && vrem[r]->kf > v->kf) v = vrem[r];
    }
    if (v != vmain && v != vone) CS_HIP(v->prepare());
    g.tiles[1] = (layout->n[1] + v->by - 1) / v->by;
    g.tiles[2] = (layout->n[2] + v->bx - 1) / v->bx;
    CS_HIP(v->launch(b[cur], b[cur ^ 1], v->mf == 4 ? frags + off4 : frags, g, s));
    cur ^= 1;
    left -= v->kf;
  }
  if (final_buffer) *final_buffer = cur;
  return CS_OK;
}
