// cs_batch.cpp -- host side of the batched workload (cs_batch_* in convstencil.h).
//
// FP64 batches whose grids all carry a uniform halo run on the register-streaming 2-D
// kernels (k2d_reg.hip): launch j of the run advances every grid whose launch sequence
// has a j-th launch, all grids of one kernel variant in one multi-grid launch (the kernel
// boundary orders the steps; no device-wide barrier).  Every other batch runs the
// persistent kernel: `steps` time steps of every grid become step blocks of work items
// (see k2d_batch.hip), the item table is uploaded once per step count, and the reset
// kernel plus the persistent kernel are launched.  Both are optionally captured in a
// hipGraph.
#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/convstencil_debug.h"
#include "cs_batch.h"
#include "cs_host_util.h"
#include "cs_rng.h"

using namespace cs::host;

namespace {

// One multi-grid launch of a register-streaming variant.
struct RegLaunch {
  const cs::Variant *v = nullptr;
  cs::RegItem *items = nullptr;
  int nitems = 0, flip = 0;
  int chain = 0;  // the launches of one stencil class form a chain (one stream)
};

struct RegSeq {
  std::vector<RegLaunch> launches;
  std::vector<int> parity_flip;  // per grid: its launches (mod 2)
  int chains = 1;
  hipGraphExec_t graph = nullptr;
};

struct Seq {
  cs::BatchItem *items = nullptr;
  int *block_start = nullptr;
  unsigned *sync = nullptr;
  int nblocks = 0, nsync = 0;
  int64_t nitems = 0;
  std::vector<int> parity_flip;  // per grid: number of launches' buffer swaps (mod 2)
  hipGraphExec_t graph = nullptr;
  std::vector<char> uniform;     // per-grid fusion eligibility the sequence was built for
};

}  // namespace

struct cs_batch {
  int count = 0;
  int prec = 0;  // 0 FP64, 1 BF16x3, 2 FP16x2 (compensated weights)
  cs_options opt{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<cs_layout> layouts;
  std::vector<int> R, shape, kind_main, kind_one, kf;
  std::vector<double *> buf0, buf1, bfrag;
  std::vector<int> cur;        // buffer holding each grid's state
  std::vector<char> uniform;   // both halos equal -> fused blocks are exact
  std::vector<double> amax, wabs;  // FP16 range guard: bound on |cell| and sum |w| per grid
  cs::BatchGrid *d_grids = nullptr;
  std::vector<cs::BatchGrid> h_grids;  // host copy; must outlive the asynchronous upload
  int max_wg = 0;
  std::map<int64_t, Seq> seqs;  // keyed by steps (rebuilt when the per-grid state changes)
  int64_t launches = 0;
  unsigned *d_error = nullptr;  // sticky device error word: set by a barrier that timed out
  bool failed = false;          // latched once the device word has been seen set
  // register-streaming path (FP64, every grid's stencil has the family)
  bool reg = false;
  std::vector<double *> wdense;         // dense weights per grid
  cs::RegGrid *d_rgrids = nullptr;
  std::vector<cs::RegGrid> h_rgrids;    // host copy; must outlive the asynchronous upload
  std::map<int64_t, RegSeq> rseqs;      // keyed by steps
  // grids of different stencil classes are independent: their chains of launches run on
  // side streams (default priority) forked from and joined to the batch's stream
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> events;
};

namespace cs {
void reg_sequence(int R, int shape, int64_t steps, std::vector<const Variant *> *seq);
}

namespace {

void free_seq(Seq &s) {
  if (s.graph) (void)hipGraphExecDestroy(s.graph);
  if (s.items) (void)hipFree(s.items);
  if (s.block_start) (void)hipFree(s.block_start);
  if (s.sync) (void)hipFree(s.sync);
  s = Seq{};
}

// Maps each grid's buffer roles (0 = holds the state now) to its buffers; copied in
// stream order, so a previous launch still reading the table is not disturbed.
cs_status upload_grid_table(cs_batch *b) {
  std::vector<cs::BatchGrid> &g = b->h_grids;
  g.assign((size_t)b->count, cs::BatchGrid{});
  for (int i = 0; i < b->count; ++i) {
    // BatchItem::src indexes buf[]: buf[0] is the buffer that holds the state now
    g[(size_t)i].buf[0] = b->cur[(size_t)i] == 0 ? b->buf0[(size_t)i] : b->buf1[(size_t)i];
    g[(size_t)i].buf[1] = b->cur[(size_t)i] == 0 ? b->buf1[(size_t)i] : b->buf0[(size_t)i];
    g[(size_t)i].bfrag = b->bfrag[(size_t)i];
    cs::Geo &geo = g[(size_t)i].geo;
    const cs_layout &L = b->layouts[(size_t)i];
    for (int d = 0; d < 3; ++d) {
      geo.ext[d] = L.ext[d];
      geo.org[d] = L.org[d];
      geo.n[d] = L.n[d];
    }
    geo.tiles[0] = 1;
    geo.tiles[1] = (L.n[1] + 63) / 64;
    geo.tiles[2] = (L.n[2] + 63) / 64;
  }
  CS_HIP(hipMemcpyAsync(b->d_grids, g.data(), g.size() * sizeof(cs::BatchGrid), hipMemcpyHostToDevice, b->stream));
  return CS_OK;
}

// Builds the step blocks for `steps` steps from the current per-grid state.
cs_status build_seq(cs_batch *b, int64_t steps, Seq *s) {
  std::vector<std::vector<cs::BatchItem>> blocks;
  s->parity_flip.assign((size_t)b->count, 0);
  for (int g = 0; g < b->count; ++g) {
    const int kf = b->uniform[(size_t)g] ? b->kf[(size_t)g] : 1;
    const int64_t ntiles = ((b->layouts[(size_t)g].n[1] + 63) / 64) * ((b->layouts[(size_t)g].n[2] + 63) / 64);
    int64_t done = 0;
    int blk = 0;
    while (done < steps) {
      const bool full = steps - done >= kf;
      const int kind = (full && kf > 1) ? b->kind_main[(size_t)g] : b->kind_one[(size_t)g];
      if ((int)blocks.size() <= blk) blocks.resize((size_t)blk + 1);
      for (int64_t t = 0; t < ntiles; ++t) blocks[(size_t)blk].push_back(cs::BatchItem{g, (int)t, kind, blk & 1});
      done += (full && kf > 1) ? kf : 1;
      ++blk;
    }
    s->parity_flip[(size_t)g] = blk & 1;
  }
  std::vector<cs::BatchItem> flat;
  std::vector<int> start;
  for (auto &v : blocks) {
    start.push_back((int)flat.size());
    flat.insert(flat.end(), v.begin(), v.end());
  }
  start.push_back((int)flat.size());
  if (flat.size() > (size_t)0x7fffffff) return CS_ERR_UNSUPPORTED;
  s->nblocks = (int)blocks.size();
  s->nitems = (int64_t)flat.size();
  s->nsync = s->nblocks + 1;  // barrier counter + one work counter per step block
  if (!flat.empty()) {
    CS_HIP(hipMalloc(&s->items, flat.size() * sizeof(cs::BatchItem)));
    CS_HIP(hipMemcpy(s->items, flat.data(), flat.size() * sizeof(cs::BatchItem), hipMemcpyHostToDevice));
  }
  CS_HIP(hipMalloc(&s->block_start, start.size() * sizeof(int)));
  CS_HIP(hipMemcpy(s->block_start, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice));
  CS_HIP(hipMalloc(&s->sync, (size_t)s->nsync * sizeof(unsigned)));
  s->uniform = b->uniform;
  return CS_OK;
}

void free_rseq(RegSeq &s) {
  if (s.graph) (void)hipGraphExecDestroy(s.graph);
  for (RegLaunch &l : s.launches)
    if (l.items) (void)hipFree(l.items);
  s = RegSeq{};
}

// Waves of grid g for variant v: strips of v's output columns cut into chunks of about 64
// rows (the checked chunks shorter), as the single-grid launch cuts them.  Measured on the
// batch of 256 grids (FP64, GStencil/s): 224-row chunks 513, 128 622, 64 662, 40 639 --
// many short waves balance the launches' tails better than the warm-up rows they add.
void reg_items(const cs_batch *b, int g, const cs::Variant *v, std::vector<cs::RegItem> *out) {
  const cs_layout &L = b->layouts[(size_t)g];
  const int64_t ny = L.n[1], nx = L.n[2];
  if (ny <= 0 || nx <= 0) return;
  const int ox = v->bx, h = v->R * v->kf;
  const int64_t strips = (nx + ox - 1) / ox;
  const int64_t slots = strips * std::max<int64_t>(1, (ny + 63) / 64);
  const cs::RegSplit sp = cs::reg_split(ny, nx, ox, h, slots, cs::reg_edge_cost(v->px, v->R),
                                         cs::reg_least(v->R, v->kf, v->st));
  const int64_t waves = cs::reg_waves(sp);
  for (int64_t w = 0; w < waves; ++w) {
    int64_t strip = 0, y0 = 0, y1 = 0;
    if (cs::reg_chunk(sp, ny, w, &strip, &y0, &y1)) out->push_back(cs::RegItem{g, (int)strip, (int)y0, (int)y1});
  }
}

// The register path's launches for `steps` steps: launch j groups, per variant, every grid
// whose sequence has a j-th launch.
cs_status build_rseq(cs_batch *b, int64_t steps, RegSeq *s) {
  std::vector<std::vector<const cs::Variant *>> seq((size_t)b->count);
  std::vector<int> chain((size_t)b->count);
  std::vector<std::pair<int, int>> classes;  // (R, shape) in order of first appearance
  size_t nj = 0;
  for (int g = 0; g < b->count; ++g) {
    const std::pair<int, int> c(b->R[(size_t)g], b->shape[(size_t)g]);
    size_t k = 0;
    while (k < classes.size() && classes[k] != c) ++k;
    if (k == classes.size()) classes.push_back(c);
    chain[(size_t)g] = (int)k;
    cs::reg_sequence(b->R[(size_t)g], b->shape[(size_t)g], steps, &seq[(size_t)g]);
    if (seq[(size_t)g].empty()) return CS_ERR_UNSUPPORTED;
    nj = std::max(nj, seq[(size_t)g].size());
    s->parity_flip.push_back((int)(seq[(size_t)g].size() & 1));
  }
  std::map<std::pair<int, const cs::Variant *>, std::vector<cs::RegItem>> cache;
  for (size_t j = 0; j < nj; ++j) {
    std::map<const cs::Variant *, std::vector<cs::RegItem>> groups;
    std::map<const cs::Variant *, int> group_chain;
    for (int g = 0; g < b->count; ++g) {
      if (j >= seq[(size_t)g].size()) continue;
      const cs::Variant *v = seq[(size_t)g][j];
      group_chain[v] = chain[(size_t)g];  // a variant serves one stencil class
      auto key = std::make_pair(g, v);
      auto c = cache.find(key);
      if (c == cache.end()) {
        std::vector<cs::RegItem> it;
        reg_items(b, g, v, &it);
        c = cache.emplace(key, it).first;
      }
      std::vector<cs::RegItem> &dst = groups[v];
      dst.insert(dst.end(), c->second.begin(), c->second.end());
    }
    for (auto &kv : groups) {
      if (kv.second.empty()) continue;
      if (kv.second.size() > (size_t)0x7fffffff) return CS_ERR_UNSUPPORTED;
      RegLaunch l;
      l.v = kv.first;
      l.nitems = (int)kv.second.size();
      l.flip = (int)(j & 1);
      l.chain = group_chain[kv.first];
      CS_HIP(hipMalloc(&l.items, kv.second.size() * sizeof(cs::RegItem)));
      s->launches.push_back(l);  // freed with the sequence from here on
      CS_HIP(hipMemcpy(l.items, kv.second.data(), kv.second.size() * sizeof(cs::RegItem), hipMemcpyHostToDevice));
    }
  }
  s->chains = (int)classes.size();
  return CS_OK;
}

// Side streams and the fork / join events for `chains` chains (created once).
cs_status ensure_side(cs_batch *b, int chains) {
  while ((int)b->side.size() < chains - 1) {
    hipStream_t t = nullptr;
    CS_HIP(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
    b->side.push_back(t);
  }
  while ((int)b->events.size() < chains) {
    hipEvent_t e = nullptr;
    CS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->events.push_back(e);
  }
  return CS_OK;
}

// Chain 0 on the batch's stream, chain c on side stream c - 1: forked after the grid table
// upload, joined back before the run ends (also under graph capture).
cs_status issue_reg(cs_batch *b, RegSeq &s) {
  const int nc = s.chains;
  if (nc > 1) {
    CS_HIP(hipEventRecord(b->events[0], b->stream));
    for (int c = 1; c < nc; ++c) CS_HIP(hipStreamWaitEvent(b->side[(size_t)c - 1], b->events[0], 0));
  }
  for (const RegLaunch &l : s.launches) {
    const hipStream_t st = l.chain == 0 ? b->stream : b->side[(size_t)l.chain - 1];
    CS_HIP(l.v->launch_multi(cs::RegMulti{b->d_rgrids, l.items, l.nitems, l.flip}, st));
  }
  for (int c = 1; c < nc; ++c) {
    CS_HIP(hipEventRecord(b->events[(size_t)c], b->side[(size_t)c - 1]));
    CS_HIP(hipStreamWaitEvent(b->stream, b->events[(size_t)c], 0));
  }
  return CS_OK;
}

cs_status run_reg(cs_batch *b, int64_t steps) {
  auto it = b->rseqs.find(steps);
  if (it == b->rseqs.end()) {
    RegSeq s;
    cs_status st = build_rseq(b, steps, &s);
    if (st != CS_OK) {
      free_rseq(s);
      return st;
    }
    it = b->rseqs.emplace(steps, s).first;
  }
  RegSeq &s = it->second;
  CS_TRY(ensure_side(b, s.chains));
  // grid table: buffers, weights, layout and the buffer that holds each grid's state
  b->h_rgrids.assign((size_t)b->count, cs::RegGrid{});
  for (int g = 0; g < b->count; ++g) {
    cs::RegGrid &G = b->h_rgrids[(size_t)g];
    G.buf[0] = b->buf0[(size_t)g];
    G.buf[1] = b->buf1[(size_t)g];
    G.wts = b->wdense[(size_t)g];
    const cs_layout &L = b->layouts[(size_t)g];
    for (int d = 0; d < 3; ++d) {
      G.geo.ext[d] = L.ext[d];
      G.geo.org[d] = L.org[d];
      G.geo.n[d] = L.n[d];
    }
    G.geo.tiles[0] = G.geo.tiles[1] = G.geo.tiles[2] = 1;
    G.cur = b->cur[(size_t)g];
  }
  CS_HIP(hipMemcpyAsync(b->d_rgrids, b->h_rgrids.data(), b->h_rgrids.size() * sizeof(cs::RegGrid),
                        hipMemcpyHostToDevice, b->stream));
  if (b->opt.use_graph) {
    if (!s.graph) {
      hipGraph_t graph = nullptr;
      CS_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      cs_status st = issue_reg(b, s);
      hipError_t e = hipStreamEndCapture(b->stream, &graph);
      if (st != CS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
      }
      CS_HIP(e);
      hipError_t ei = hipGraphInstantiate(&s.graph, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      CS_HIP(ei);
    }
    CS_HIP(hipGraphLaunch(s.graph, b->stream));
  } else {
    CS_TRY(issue_reg(b, s));
  }
  // the host table must stay unchanged until the copy above has run
  CS_HIP(hipStreamSynchronize(b->stream));
  b->launches += (int64_t)s.launches.size();
  for (int g = 0; g < b->count; ++g) b->cur[(size_t)g] ^= s.parity_flip[(size_t)g];
  return CS_OK;
}

bool reg_path(const cs_batch *b) {
  if (b->prec != 0 || !b->reg) return false;
  for (char u : b->uniform)
    if (!u) return false;
  return true;
}

cs_status issue(cs_batch *b, Seq &s) {
  const unsigned nwg = (unsigned)b->max_wg;
  CS_HIP(cs::batch_launch(b->prec, b->d_grids, s.items, s.block_start, s.nblocks, s.sync, s.nsync, b->d_error, nwg,
                          b->stream));
  return CS_OK;
}

cs_status check_error_flag(cs_batch *b);

cs_status run(cs_batch *b, int64_t steps) {
  if (steps < 0) return CS_ERR_INVALID;
  if (b->failed) return CS_ERR_DEVICE;
  if (steps == 0) return CS_OK;
  std::vector<double> bound((size_t)b->count);
  const int precision = b->prec == 2 ? CS_PREC_FP16X2 : b->prec == 1 ? CS_PREC_BF16X3 : CS_PREC_FP64;
  for (int g = 0; g < b->count; ++g)
    if (!fp16_range_ok(precision, b->amax[(size_t)g], b->wabs[(size_t)g], steps, &bound[(size_t)g])) {
      fprintf(stderr,
              "convstencil: FP16 split precision needs |values| < 2^15; grid %d may reach %.3g within %lld step(s)\n",
              g, bound[(size_t)g], (long long)steps);
      return CS_ERR_INVALID;
    }
  CS_HIP(hipSetDevice(b->device));
  if (reg_path(b)) {
    CS_TRY(run_reg(b, steps));
    b->amax = bound;
    return check_error_flag(b);
  }
  auto it = b->seqs.find(steps);
  if (it != b->seqs.end() && it->second.uniform != b->uniform) {
    free_seq(it->second);
    b->seqs.erase(it);
    it = b->seqs.end();
  }
  if (it == b->seqs.end()) {
    Seq s;
    cs_status st = build_seq(b, steps, &s);
    if (st != CS_OK) {
      free_seq(s);
      return st;
    }
    it = b->seqs.emplace(steps, s).first;
  }
  Seq &s = it->second;
  CS_TRY(upload_grid_table(b));
  if (b->opt.use_graph) {
    if (!s.graph) {
      hipGraph_t graph = nullptr;
      CS_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      cs_status st = issue(b, s);
      hipError_t e = hipStreamEndCapture(b->stream, &graph);
      if (st != CS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
      }
      CS_HIP(e);
      hipError_t ei = hipGraphInstantiate(&s.graph, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      CS_HIP(ei);
    }
    CS_HIP(hipGraphLaunch(s.graph, b->stream));
  } else {
    CS_TRY(issue(b, s));
  }
  // the host table must stay unchanged until the copy above has run
  CS_HIP(hipStreamSynchronize(b->stream));
  b->launches += 1;
  for (int g = 0; g < b->count; ++g) b->cur[(size_t)g] ^= s.parity_flip[(size_t)g];
  b->amax = bound;
  return check_error_flag(b);
}

// Latches the batch's failure: once a barrier has timed out (the grids' contents are then
// undefined), every later run, timing and download of this batch returns CS_ERR_DEVICE.
cs_status latch_error(cs_batch *b, unsigned flag) {
  if (flag && !b->failed) fprintf(stderr, "convstencil: batch barrier did not complete within its spin limit\n");
  if (flag) b->failed = true;
  return b->failed ? CS_ERR_DEVICE : CS_OK;
}

// Reads the sticky device word (the stream has been synchronised by the caller).
cs_status check_error_flag(cs_batch *b) {
  if (b->failed) return CS_ERR_DEVICE;
  unsigned flag = 0;
  CS_HIP(hipMemcpy(&flag, b->d_error, sizeof(unsigned), hipMemcpyDeviceToHost));
  return latch_error(b, flag);
}

}  // namespace

extern "C" {

cs_status cs_batch_create(cs_batch **out, int32_t count, const cs_layout *layouts, const cs_stencil *stencils,
                          const cs_options *opt) {
  if (!out || count <= 0 || !layouts || !stencils) return CS_ERR_INVALID;
  *out = nullptr;
  const int prec = opt ? opt->precision : CS_PREC_FP64;
  if (prec != CS_PREC_FP64 && prec != CS_PREC_BF16X3 && prec != CS_PREC_FP16X2) return CS_ERR_UNSUPPORTED;
  for (int i = 0; i < count; ++i) {
    const cs_layout &L = layouts[i];
    const cs_stencil &st = stencils[i];
    if (L.ndim != 2 || st.ndim != 2 || !st.weights || (st.radius != 1 && st.radius != 3)) return CS_ERR_UNSUPPORTED;
    if (!layout_valid(&L, st.radius)) return CS_ERR_INVALID;
  }
  int dev = 0;
  CS_TRY(check_device(&dev));
  cs_batch *b = new cs_batch();
  b->count = count;
  b->prec = prec == CS_PREC_FP64 ? 0 : prec == CS_PREC_BF16X3 ? 1 : 2;
  if (opt) b->opt = *opt;
  b->device = dev;
  b->layouts.assign(layouts, layouts + count);
  cs_status s = CS_OK;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) s = CS_ERR_DEVICE;
  for (int i = 0; i < count && s == CS_OK; ++i) {
    const cs_stencil &st = stencils[i];
    const int R = st.radius;
    const int D = 2 * R + 1;
    std::vector<double> w(st.weights, st.weights + D * D);
    const int shape = classify_2d(R, w.data());
    const int kf = cs::batch_main_kf(b->prec, R);
    const int km = cs::batch_kind(b->prec, R, shape, kf), k1 = cs::batch_kind(b->prec, R, shape, 1);
    if (km < 0 || k1 < 0) {
      s = CS_ERR_UNSUPPORTED;
      break;
    }
    b->R.push_back(R);
    b->shape.push_back(shape);
    b->kf.push_back(kf);
    b->kind_main.push_back(km);
    b->kind_one.push_back(k1);
    // B fragments
    double *d = nullptr;
    if (b->prec == 0) {
      const int nks = cs::bfrag_ksteps_2d_f64(4, R, shape);
      std::vector<double> h((size_t)nks * 64);
      cs::build_bfrag_2d_f64(4, R, shape, w.data(), h.data());
      if (hipMalloc(&d, h.size() * sizeof(double)) != hipSuccess ||
          hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        s = CS_ERR_ALLOC;
    } else {
      std::vector<unsigned char> h(1 << 20);
      const size_t bytes = cs::build_bfrag_split2(b->prec == 1 ? CS_PREC_BF16X3 : CS_PREC_FP16X2, R, shape, w.data(),
                                                   h.data(), h.size());
      if (bytes == 0 || hipMalloc(&d, bytes) != hipSuccess ||
          hipMemcpy(d, h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
        s = CS_ERR_ALLOC;
    }
    b->bfrag.push_back(d);
    // dense weights for the register-streaming path (FP64, when the family has the stencil)
    double *wd = nullptr;
    if (s == CS_OK && b->prec == 0) {
      std::vector<double> h(64, 0.0);
      std::copy(w.begin(), w.end(), h.begin());
      if (hipMalloc(&wd, h.size() * sizeof(double)) != hipSuccess ||
          hipMemcpy(wd, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        s = CS_ERR_ALLOC;
    }
    b->wdense.push_back(wd);
    double *p0 = nullptr, *p1 = nullptr;
    const size_t bytes = (size_t)cell_count(&layouts[i]) * sizeof(double);
    if (s == CS_OK && (hipMalloc(&p0, bytes) != hipSuccess || hipMalloc(&p1, bytes) != hipSuccess)) {
      (void)hipGetLastError();
      s = CS_ERR_ALLOC;
    }
    b->buf0.push_back(p0);
    b->buf1.push_back(p1);
    if (s == CS_OK && (cs::launch_zero(p0, (size_t)cell_count(&layouts[i]), b->stream) != hipSuccess ||
                       cs::launch_zero(p1, (size_t)cell_count(&layouts[i]), b->stream) != hipSuccess))
      s = CS_ERR_DEVICE;
    b->cur.push_back(0);
    b->uniform.push_back(1);
    b->amax.push_back(0.0);  // both buffers are zeroed below
    b->wabs.push_back(abs_sum(w.data(), w.size()));
  }
  if (s == CS_OK && hipMalloc(&b->d_grids, (size_t)count * sizeof(cs::BatchGrid)) != hipSuccess) s = CS_ERR_ALLOC;
  if (s == CS_OK && b->prec == 0) {
    b->reg = true;
    std::vector<const cs::Variant *> one;
    for (int i = 0; i < count && b->reg; ++i) {
      cs::reg_sequence(b->R[(size_t)i], b->shape[(size_t)i], 1, &one);
      b->reg = !one.empty();
    }
    if (b->reg && hipMalloc(&b->d_rgrids, (size_t)count * sizeof(cs::RegGrid)) != hipSuccess) s = CS_ERR_ALLOC;
  }
  if (s == CS_OK && (hipMalloc(&b->d_error, sizeof(unsigned)) != hipSuccess ||
                     hipMemset(b->d_error, 0, sizeof(unsigned)) != hipSuccess))
    s = CS_ERR_ALLOC;
  if (s == CS_OK && hipStreamSynchronize(b->stream) != hipSuccess) s = CS_ERR_DEVICE;
  if (s == CS_OK && cs::batch_prepare(b->prec, &b->max_wg) != hipSuccess) s = CS_ERR_DEVICE;
  if (s == CS_OK && b->max_wg <= 0) s = CS_ERR_UNSUPPORTED;
  if (s != CS_OK) {
    cs_batch_destroy(b);
    return s;
  }
  *out = b;
  return CS_OK;
}

cs_status cs_batch_destroy(cs_batch *b) {
  if (!b) return CS_ERR_INVALID;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto &kv : b->seqs) free_seq(kv.second);
  for (auto &kv : b->rseqs) free_rseq(kv.second);
  for (double *p : b->wdense)
    if (p) (void)hipFree(p);
  if (b->d_rgrids) (void)hipFree(b->d_rgrids);
  for (hipStream_t t : b->side)
    if (t) (void)hipStreamDestroy(t);
  for (hipEvent_t e : b->events)
    if (e) (void)hipEventDestroy(e);
  for (double *p : b->buf0)
    if (p) (void)hipFree(p);
  for (double *p : b->buf1)
    if (p) (void)hipFree(p);
  for (double *p : b->bfrag)
    if (p) (void)hipFree(p);
  if (b->d_grids) (void)hipFree(b->d_grids);
  if (b->d_error) (void)hipFree(b->d_error);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return CS_OK;
}

cs_status cs_batch_upload(cs_batch *b, int32_t grid, const double *host_in) {
  if (!b || !host_in || grid < 0 || grid >= b->count) return CS_ERR_INVALID;
  CS_HIP(hipSetDevice(b->device));
  const cs_layout &L = b->layouts[(size_t)grid];
  const size_t n = (size_t)cell_count(&L);
  CS_HIP(hipMemcpy(b->buf0[(size_t)grid], host_in, n * sizeof(double), hipMemcpyHostToDevice));
  CS_HIP(cs::launch_zero(b->buf1[(size_t)grid], n, b->stream));
  CS_HIP(hipStreamSynchronize(b->stream));
  b->cur[(size_t)grid] = 0;
  b->uniform[(size_t)grid] = halo_is_zero(&L, host_in) ? 1 : 0;
  b->amax[(size_t)grid] = abs_max(host_in, n);
  return CS_OK;
}

cs_status cs_batch_fill_random(cs_batch *b, const uint64_t *seeds) {
  if (!b || !seeds) return CS_ERR_INVALID;
  CS_HIP(hipSetDevice(b->device));
  for (int i = 0; i < b->count; ++i) {
    const cs_layout &L = b->layouts[(size_t)i];
    cs::Geo g{};
    for (int d = 0; d < 3; ++d) {
      g.ext[d] = L.ext[d];
      g.org[d] = L.org[d];
      g.n[d] = L.n[d];
    }
    CS_HIP(cs::launch_fill_random(b->buf0[(size_t)i], g, cs::stream_seed(seeds[i], 0), b->stream));
    CS_HIP(cs::launch_zero(b->buf1[(size_t)i], (size_t)cell_count(&L), b->stream));
    b->cur[(size_t)i] = 0;
    b->uniform[(size_t)i] = 1;
    b->amax[(size_t)i] = 1.0;  // uniform in [-1, 1) with a zero halo
  }
  CS_HIP(hipStreamSynchronize(b->stream));
  return CS_OK;
}

cs_status cs_batch_run(cs_batch *b, int64_t steps, cs_report *rep) {
  if (!b) return CS_ERR_INVALID;
  const int64_t l0 = b->launches;
  CS_TRY(run(b, steps));
  if (rep) {
    memset(rep, 0, sizeof(*rep));
    rep->launches = b->launches - l0;
  }
  return CS_OK;
}

cs_status cs_batch_time(cs_batch *b, int64_t steps, int32_t warmup, int32_t reps, cs_report *rep) {
  if (!b || reps <= 0 || warmup < 0 || steps <= 0) return CS_ERR_INVALID;
  CS_HIP(hipSetDevice(b->device));
  for (int i = 0; i < warmup; ++i) CS_TRY(run(b, steps));
  CS_HIP(hipStreamSynchronize(b->stream));
  hipEvent_t e0, e1;
  CS_HIP(hipEventCreate(&e0));
  CS_HIP(hipEventCreate(&e1));
  const int64_t l0 = b->launches;
  CS_HIP(hipEventRecord(e0, b->stream));
  for (int i = 0; i < reps; ++i) CS_TRY(run(b, steps));
  CS_HIP(hipEventRecord(e1, b->stream));
  CS_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  CS_HIP(hipEventElapsedTime(&ms, e0, e1));
  CS_HIP(hipEventDestroy(e0));
  CS_HIP(hipEventDestroy(e1));
  CS_TRY(check_error_flag(b));
  if (rep) {
    memset(rep, 0, sizeof(*rep));
    rep->elapsed_ms = ms / reps;
    double pts = 0;
    for (const cs_layout &L : b->layouts) pts += (double)L.n[1] * (double)L.n[2];
    rep->gstencil_per_s = rep->elapsed_ms > 0 && pts > 0 ? pts * (double)steps / (rep->elapsed_ms * 1e-3) / 1e9 : 0.0;
    rep->launches = (b->launches - l0) / reps;
  }
  return CS_OK;
}

cs_status cs_batch_download(cs_batch *b, int32_t grid, double *host_out) {
  if (!b || !host_out || grid < 0 || grid >= b->count) return CS_ERR_INVALID;
  CS_HIP(hipSetDevice(b->device));
  CS_HIP(hipStreamSynchronize(b->stream));
  CS_TRY(check_error_flag(b));
  const double *src = b->cur[(size_t)grid] == 0 ? b->buf0[(size_t)grid] : b->buf1[(size_t)grid];
  CS_HIP(hipMemcpy(host_out, src, (size_t)cell_count(&b->layouts[(size_t)grid]) * sizeof(double),
                   hipMemcpyDeviceToHost));
  return CS_OK;
}

cs_status cs_batch_describe(cs_batch *b, char *buf, int32_t len) {
  if (!b || !buf || len <= 0) return CS_ERR_INVALID;
  int64_t tiles = 0;
  for (const cs_layout &L : b->layouts) tiles += ((L.n[1] + 63) / 64) * ((L.n[2] + 63) / 64);
  snprintf(buf, (size_t)len, "batch grids=%d precision=%s path=%s tiles=%lld workgroups=%d lds=%zu", b->count,
           b->prec == 0 ? "fp64" : b->prec == 1 ? "bf16x3" : "fp16x2",
           reg_path(b) ? "register-streaming" : "persistent", (long long)tiles, b->max_wg,
           cs::batch_lds_bytes(b->prec));
  return CS_OK;
}

int32_t cs_debug_batch_latch(const uint32_t *flags, int32_t n, int32_t *status) {
  cs_batch b;  // host-only: no device state is touched
  for (int32_t i = 0; i < n; ++i) status[i] = (int32_t)latch_error(&b, flags[i]);
  return 0;
}

}  // extern "C"
