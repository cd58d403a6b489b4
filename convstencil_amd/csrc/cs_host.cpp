// cs_host.cpp -- host runtime and C ABI of libconvstencil (see include/convstencil.h).
//
// The reference's host launchers (src/2d/gpu.cu:245-333 and siblings) each build
// the B matrices, allocate two device buffers, copy the input, run the time loop
// of single-launch steps and copy the result back.  Here that body is a plan:
// buffers and B fragments are built once, the time loop is a sequence of fused
// launches (optionally captured in a hipGraph), and the reference-named
// launchers at the bottom of this file are thin wrappers that keep the
// reference's layouts, weight layouts and stdout lines.
#include "../../include/convstencil.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "cs_host_util.h"
#include "cs_internal.h"
#include "cs_rng.h"

using cs::Geo;
using cs::Variant;

using namespace cs::host;

struct cs_plan {
  cs_layout L{};
  int ndim = 0, R = 0, shape = 0;
  std::vector<double> w;  // dense weights
  cs_options opt{};
  int device = 0;
  hipStream_t stream = nullptr;
  double *buf[2] = {nullptr, nullptr};
  double *bfrag = nullptr;  // B fragments for the selected MFMA shape (or dense weights)
  int64_t frag_off = 0;     // when the register kernels' dense weights come first: where the
                            // MFMA single-step kernel's B fragments start (doubles)
  int cur = 0;              // buffer holding the current state
  int64_t steps_done = 0;
  bool halo_uniform = true;  // both buffers carry the same halo -> fusion is exact
  const Variant *vmain = nullptr;
  std::vector<const Variant *> vrem;  // vrem[r] runs r fused steps (r < vmain->kf), may be null
  const Variant *vone = nullptr;      // single-step variant (the register families use the MFMA one)
  std::map<std::pair<int64_t, int>, hipGraphExec_t> graphs;
  int64_t launches = 0;
  int tuned = 0;  // 0 defaults or a forced variant, 1 cached autotuner choice, 2 timed now
  // FP16 range guard: a bound on |cell| of both buffers (halo included) and the weights'
  // sum of magnitudes, which bounds each step's growth
  double amax = 0.0, wabs = 0.0;
};

namespace {

Geo make_geo(const cs_layout &L, const Variant *v) {
  Geo g{};
  for (int d = 0; d < 3; ++d) {
    g.ext[d] = L.ext[d];
    g.org[d] = L.org[d];
    g.n[d] = L.n[d];
  }
  g.tiles[0] = 1;
  g.tiles[1] = (L.n[1] + v->by - 1) / v->by;
  g.tiles[2] = (L.n[2] + v->bx - 1) / v->bx;
  return g;
}

// Default fusion depth when the caller does not choose one (the autotuner's
// cache, when present, overrides this).
// FP64 MFMA families (measured, GStencil/s per physical step): 1d1r 2^20 kf16 878 (kf32
// 842); box2d1r 10240^2 kf6 925-990 (kf4 861); star2d3r 8192^2 kf3 494 (kf2 454); box3d1r
// 512^3 kf1 260-265 (kf2 219); star3d1r kf2 303 (kf1 272).  Split precision: kf4 (2-D
// radius 1), kf2 (2-D radius 2-3 and 3-D).  The register kernels that pick_main prefers
// for 1-D and 2-D radius 1 choose their own depths there.
int default_kf(int ndim, int R, int top, int shape, bool fp64) {
  if (ndim == 1) return std::min(top, 16);
  if (ndim == 3) return std::min(top, fp64 && shape == cs::SHAPE3_BOX ? 1 : 2);
  if (!fp64) return std::min(top, R == 1 ? 4 : 2);
  return std::min(top, R == 1 ? 6 : 3);
}

std::vector<Variant> joined(const Variant *(*a)(int *), const Variant *(*b)(int *)) {
  int na = 0, nb = 0;
  const Variant *x = a(&na), *y = b(&nb);
  std::vector<Variant> v(x, x + na);
  v.insert(v.end(), y, y + nb);
  return v;
}

std::vector<Variant> joined3(const Variant *(*a)(int *), const Variant *(*b)(int *), const Variant *(*c)(int *)) {
  std::vector<Variant> v = joined(a, b);
  int nc = 0;
  const Variant *z = c(&nc);
  v.insert(v.end(), z, z + nc);
  return v;
}

std::vector<Variant> joined4(const Variant *(*a)(int *), const Variant *(*b)(int *), const Variant *(*c)(int *),
                             const Variant *(*d)(int *)) {
  std::vector<Variant> v = joined3(a, b, c);
  int nd = 0;
  const Variant *z = d(&nd);
  v.insert(v.end(), z, z + nd);
  return v;
}

const Variant *family(int ndim, int *cnt, int prec = CS_PREC_FP64) {
  *cnt = 0;
  if (prec != CS_PREC_FP64)
    return ndim == 2 ? cs::variants_2d_split(cnt) : ndim == 3 ? cs::variants_3d_split(cnt) : nullptr;
  if (ndim < 1 || ndim > 3) return nullptr;
  // the MFMA kernels, then the --breakdown ladder's FMA kernels
  static const std::vector<Variant> all[3] = {
      joined3(cs::variants_1d_reg, cs::variants_1d_f64, cs::variants_1d_ablation),
      joined4(cs::variants_2d_f64, cs::variants_2d_reg, cs::variants_2d_stream, cs::variants_2d_ablation),
      joined4(cs::variants_3d_reg, cs::variants_3d_f64, cs::variants_3d_fma, cs::variants_3d_ablation)};
  *cnt = (int)all[ndim - 1].size();
  return all[ndim - 1].data();
}

// First variant of the family with this (MFMA shape or precision, radius, shape class,
// fusion depth) and, if st >= 0, this Variant::st; or the variant with id prefer_id.
const Variant *find_variant(int ndim, int mf, int R, int shape, int kf, int prefer_id, int prec = CS_PREC_FP64,
                            int st = -1) {
  int cnt = 0;
  const Variant *vs = family(ndim, &cnt, prec);
  const Variant *best = nullptr;
  for (int i = 0; i < cnt; ++i) {
    const Variant &v = vs[i];
    if (prefer_id > 0 && v.id == prefer_id && v.R == R && v.shape == shape) return &v;
    if (v.mf == mf && v.R == R && v.shape == shape && v.kf == kf && (st < 0 || v.st == st) && !best) best = &v;
  }
  return best;
}

int max_kf(int ndim, int mf, int R, int shape, int prec = CS_PREC_FP64, int st = -1) {
  int cnt = 0, m = 0;
  const Variant *vs = family(ndim, &cnt, prec);
  for (int i = 0; i < cnt; ++i)
    if (vs[i].mf == mf && vs[i].R == R && vs[i].shape == shape && (st < 0 || vs[i].st == st))
      m = std::max(m, vs[i].kf);
  return m;
}

// The single-step variant of an FP64 plan whose main variant is v: v itself when it is a
// single-step one (a variant forced by id must be the one that runs), else its family's
// first single-step variant -- except that the register kernels' single steps (alternating
// halos, remainders of one step) run on the LDS-staged MFMA single-step kernel, which keeps
// more bytes in flight: measured (GStencil/s) box2d1r 10240^2 300 against 203, star2d3r
// 8192^2 286 against 253; box2d3r keeps its own (246 against 204).
const Variant *single_step(int nd, int R, int shape, const Variant *v) {
  if (v->kf == 1) return v;
  // the fused 3-D register kernels: the 8 x 8 block single-step kernel (512^3: 258-268
  // GStencil/s against 242 for the register single step)
  if (v->mf == 9 && nd == 3)
    if (const Variant *m = find_variant(nd, 8, R, shape, 1, 0, CS_PREC_FP64)) return m;
  if ((v->mf == 6 || v->mf == 7) && !(nd == 2 && R > 1 && shape == cs::SHAPE_BOX))
    if (const Variant *m = find_variant(nd, 4, R, shape, 1, 0, CS_PREC_FP64)) return m;
  return find_variant(nd, v->mf, R, shape, 1, 0, CS_PREC_FP64);
}

// 2-D grids of at most this many interior cells take depth 4 from the radius-1 box register
// kernels: their launches run short of waves and depth 8's longer tail dominates (2048^2:
// 825 GStencil/s at depth 4 against 635 at depth 8; 3072^2: 1136 against 1183).
constexpr int64_t kRegSmallGrid = 2560LL * 2560LL;

// The main variant of a plan: a forced id, else the preferred family's default depth (for
// 2-D, `cells` = the interior cells of one grid, 0 when no single grid decides).
const Variant *pick_main(int nd, int R, int shape, int prec, int fusion_depth, int tile_variant, int64_t cells = 0) {
  const int mf = prec != CS_PREC_FP64 ? prec : 4;
  const Variant *v = nullptr;
  if (tile_variant > 0) {
    // a forced variant of the stencil's shape class, or a box variant (it computes any
    // stencil of its radius: the taps outside the stencil get zero weights)
    v = find_variant(nd, 0, R, shape, 0, tile_variant, prec);
    if (!v) v = find_variant(nd, 0, R, cs::SHAPE_BOX, 0, tile_variant, prec);
  } else if (prec == CS_PREC_FP64) {
    // 1-D: the register-resident kernels (mf 6) at their deepest compiled fusion, else the
    // 4 x 4 x 4 blocks; 2-D: the register-streaming kernels (mf 7), else the dilated 8 x 8
    // blocks; 3-D: the fused register-streaming kernels (mf 9, radius 1), else the 8 x 8
    // block family (mf 8) where it exists, else the register-blocked FMA kernels (mf 3,
    // radius 2-3), else the 4 x 16 strips.  The 2-D box radius-3 register kernels spill; the
    // dilated tables are faster there (8192^2, 64 steps: 304-308 GStencil/s at depth 2-3
    // against 184-195).
    const bool reg2d = !(nd == 2 && R > 1 && shape == cs::SHAPE_BOX);
    const int fams[4] = {nd == 3 ? 9 : nd == 1 ? 6 : reg2d ? 7 : mf, nd == 3 ? 8 : mf, nd == 3 ? 3 : mf, mf};
    // a requested depth that some family compiles exactly is honoured exactly; single
    // steps of the register families run on the MFMA single-step kernel, as a plan's
    // remainders do (see single_step)
    if (fusion_depth > 0)
      for (int fmf : fams) {
        if (fusion_depth == 1 && (fmf == 6 || fmf == 7 || fmf == 9)) continue;
        if (!v) v = find_variant(nd, fmf, R, shape, fusion_depth, 0, prec);
      }
    for (int fmf : fams) {
      if (v) break;
      const int top = max_kf(nd, fmf, R, shape, prec);
      if (top == 0) continue;
      // register kernels (measured at 10240^2 / 8192^2, 64 steps): 2-D radius 1 depth 8
      // (1980-2001 GStencil/s; depth 12 1890-1908), star2d3r depth 4 (1016-1018; 3: 797,
      // 6: 761); 1-D the deepest compiled
      // the 3-D register kernels fuse two steps for the box and three for the star (512^3:
      // box 433-451 GStencil/s at depth 2 and 3 alike, star 621 at depth 3 against 449-472);
      // their single steps are slower than the 8 x 8 blocks', which a plan takes instead
      if (fmf == 9 && top < 2) continue;
      int kf = fusion_depth > 0 ? std::min(fusion_depth, top)
               : fmf == 6       ? top
               : fmf == 7       ? std::min(top, R == 1 ? (shape == cs::SHAPE_BOX && cells > 0 &&
                                                          cells <= kRegSmallGrid ? 4 : 8)
                                                 : shape == cs::SHAPE_STAR ? 4 : 2)
               : fmf == 9       ? std::min(top, shape == cs::SHAPE3_STAR ? 3 : 2)
                                : default_kf(nd, R, top, shape, true);
      for (; kf >= 1 && !v; --kf) v = find_variant(nd, fmf, R, shape, kf, 0, prec);
      if (v) break;
    }
  }
  return v;
}

cs_status select_variants(cs_plan *p) {
  const int nd = p->ndim;
  const int prec = p->opt.precision;
  if (nd < 1 || nd > 3) return CS_ERR_UNSUPPORTED;
  if (prec != CS_PREC_FP64 && (prec < CS_PREC_BF16X3 || prec > CS_PREC_FP16W3)) return CS_ERR_INVALID;
  if (prec != CS_PREC_FP64 && nd == 1) return CS_ERR_UNSUPPORTED;
  // Split precision (Variant::mf holds the precision): the dilated form (st 2) for the
  // stencil's shape class, else its box variants (zero weights off the stencil), else the
  // Toeplitz form (st 0, box only).  The remainder launches come from the same form.
  const int mf = prec != CS_PREC_FP64 ? prec : 4;
  const Variant *v = nullptr;
  if (p->opt.tile_variant > 0) {
    v = pick_main(nd, p->R, p->shape, prec, p->opt.fusion_depth, p->opt.tile_variant);
    if (!v) return CS_ERR_INVALID;
  } else if (prec == CS_PREC_FP64) {
    v = pick_main(nd, p->R, p->shape, prec, p->opt.fusion_depth, 0, nd == 2 ? p->L.n[1] * p->L.n[2] : 0);
  } else {
    const int tries[3][2] = {{p->shape, 2}, {cs::SHAPE_BOX, 2}, {cs::SHAPE_BOX, 0}};
    for (const auto &tr : tries) {
      const int top = max_kf(nd, mf, p->R, tr[0], prec, tr[1]);
      if (top == 0) continue;
      int kf = p->opt.fusion_depth > 0 ? std::min(p->opt.fusion_depth, top)
                                       : default_kf(nd, p->R, top, tr[0], false);
      for (; kf >= 1 && !v; --kf) v = find_variant(nd, mf, p->R, tr[0], kf, 0, prec, tr[1]);
      if (v) break;
    }
  }
  if (!v) return CS_ERR_UNSUPPORTED;
  p->shape = v->shape;
  const int fam = prec != CS_PREC_FP64 ? v->st : -1;
  p->vmain = v;
  // single-step launches: the chosen variant itself when it is a single-step one (a variant
  // forced by id must be the one that runs), else the family's first single-step variant
  p->vone = prec == CS_PREC_FP64 ? single_step(nd, p->R, p->shape, v)
                                  : v->kf == 1 ? v : find_variant(nd, v->mf, p->R, p->shape, 1, 0, prec, fam);
  if (!p->vone) return CS_ERR_UNSUPPORTED;
  p->vrem.assign(v->kf, nullptr);
  for (int r = 1; r < v->kf; ++r) p->vrem[r] = find_variant(nd, v->mf, p->R, p->shape, r, 0, prec, fam);
  for (const Variant *u : {p->vmain, p->vone}) CS_HIP(u->prepare());
  for (const Variant *u : p->vrem)
    if (u) CS_HIP(u->prepare());
  if (prec != CS_PREC_FP64) {
    std::vector<unsigned char> hb(1 << 20);
    const size_t bytes = nd == 3      ? cs::build_bfrag_split3(prec, p->R, p->shape, p->w.data(), hb.data(), hb.size())
                         : v->st == 2 ? cs::build_bfrag_split2(prec, p->R, p->shape, p->w.data(), hb.data(), hb.size())
                                      : cs::build_bfrag_split(prec, p->R, p->w.data(), hb.data(), hb.size());
    if (bytes == 0) return CS_ERR_UNSUPPORTED;
    if (p->bfrag) (void)hipFree(p->bfrag);
    p->bfrag = nullptr;
    CS_HIP(hipMalloc(&p->bfrag, bytes));
    CS_HIP(hipMemcpy(p->bfrag, hb.data(), bytes, hipMemcpyHostToDevice));
    return CS_OK;
  }
  // the FMA kernels (the ladder's, the 3-D mf 3 family and the register-streaming mf 7 and
  // mf 9 families) read the dense weights (zero-padded to whole 64-entry rows); a plan of
  // those whose single steps run on an MFMA kernel (mf 7: the 2-D 4 x 4 x 4 kernel, mf 9:
  // the 3-D 8 x 8 blocks) appends that kernel's B fragments
  p->frag_off = 0;
  if (v->mf == 1 || v->mf == 2 || v->mf == 3 || v->mf == 7 || v->mf == 9) {
    std::vector<double> h((p->w.size() + 63) / 64 * 64, 0.0);
    std::copy(p->w.begin(), p->w.end(), h.begin());
    if (v->mf == 7 && p->vone->mf == 4) {
      p->frag_off = (int64_t)h.size();
      h.resize(h.size() + (size_t)cs::bfrag_ksteps_2d_f64(4, p->R, p->shape) * 64);
      cs::build_bfrag_2d_f64(4, p->R, p->shape, p->w.data(), h.data() + p->frag_off);
    } else if (v->mf == 9 && p->vone->mf == 8) {
      p->frag_off = (int64_t)h.size();
      h.resize(h.size() + (size_t)cs::bfrag_ksteps_3d2_f64(p->R, p->shape) * 64);
      cs::build_bfrag_3d2_f64(p->R, p->shape, p->w.data(), h.data() + p->frag_off);
    }
    if (p->bfrag) (void)hipFree(p->bfrag);
    p->bfrag = nullptr;
    CS_HIP(hipMalloc(&p->bfrag, h.size() * sizeof(double)));
    CS_HIP(hipMemcpy(p->bfrag, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return CS_OK;
  }
  // B fragments for the family's MFMA shape
  const int nks = nd == 1   ? cs::bfrag_ksteps_1d_f64(p->R)
                  : nd == 2 ? cs::bfrag_ksteps_2d_f64(v->mf, p->R, p->shape)
                  : v->mf == 8 ? cs::bfrag_ksteps_3d2_f64(p->R, p->shape)
                               : cs::bfrag_ksteps_3d_f64(p->R, p->shape);
  std::vector<double> h((size_t)nks * 64);
  if (nd == 1)
    cs::build_bfrag_1d_f64(p->R, p->w.data(), h.data());
  else if (nd == 2)
    cs::build_bfrag_2d_f64(v->mf, p->R, p->shape, p->w.data(), h.data());
  else if (v->mf == 8)
    cs::build_bfrag_3d2_f64(p->R, p->shape, p->w.data(), h.data());
  else
    cs::build_bfrag_3d_f64(p->R, p->shape, p->w.data(), h.data());
  if (p->bfrag) (void)hipFree(p->bfrag);
  p->bfrag = nullptr;
  CS_HIP(hipMalloc(&p->bfrag, h.size() * sizeof(double)));
  CS_HIP(hipMemcpy(p->bfrag, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  return CS_OK;
}

cs_status launch_one(cs_plan *p, const Variant *v) {
  const Geo g = make_geo(p->L, v);
  const double *wts = (v->mf == 4 || v->mf == 8) && p->frag_off > 0 ? p->bfrag + p->frag_off : p->bfrag;
  CS_HIP(v->launch(p->buf[p->cur], p->buf[p->cur ^ 1], wts, g, p->stream));
  p->cur ^= 1;
  p->launches += 1;
  return CS_OK;
}

// Launch sequence for `steps` physical steps.  With uniform halos, fused
// launches are exact: take the main depth as often as it fits, then the largest
// compiled depth that fits the rest.  Otherwise every launch is a single step so
// that each step reads the halo of its own buffer, as the reference's ping-pong
// does.
void plan_sequence(const cs_plan *p, int64_t steps, std::vector<const Variant *> *seq) {
  seq->clear();
  if (steps <= 0) return;
  if (!p->halo_uniform || p->vmain->kf == 1) {
    seq->assign((size_t)steps, p->vone);
    return;
  }
  const int kf = p->vmain->kf;
  seq->assign((size_t)(steps / kf), p->vmain);
  int64_t rem = steps % kf;
  while (rem > 0) {
    const Variant *best = p->vone;
    for (const Variant *u : p->vrem)
      if (u && u->kf <= rem && u->kf > best->kf) best = u;
    seq->push_back(best);
    rem -= best->kf;
  }
}

cs_status issue_steps(cs_plan *p, int64_t steps) {
  std::vector<const Variant *> seq;
  plan_sequence(p, steps, &seq);
  for (const Variant *v : seq) CS_TRY(launch_one(p, v));
  return CS_OK;
}

// Number of launches issue_steps() makes for `steps` steps.
int64_t count_launches(const cs_plan *p, int64_t steps) {
  std::vector<const Variant *> seq;
  plan_sequence(p, steps, &seq);
  return (int64_t)seq.size();
}

cs_status run_steps(cs_plan *p, int64_t steps) {
  if (steps < 0) return CS_ERR_INVALID;
  if (steps == 0) return CS_OK;
  double bound = 0.0;
  if (!fp16_range_ok(p->opt.precision, p->amax, p->wabs, steps, &bound)) {
    fprintf(stderr,
            "convstencil: FP16 split precision needs |values| < 2^15; the state may reach %.3g within %lld "
            "step(s)\n",
            bound, (long long)steps);
    return CS_ERR_INVALID;
  }
  p->amax = bound;
  if (p->L.n[0] == 0 || p->L.n[1] == 0 || p->L.n[2] == 0) {
    // no interior cell: every step only swaps the two buffers (no kernel has work)
    if (steps & 1) p->cur ^= 1;
    p->steps_done += steps;
    return CS_OK;
  }
  CS_HIP(hipSetDevice(p->device));
  if (!p->opt.use_graph) {
    CS_TRY(issue_steps(p, steps));
    p->steps_done += steps;
    return CS_OK;
  }
  // Graphs are keyed by (step count, starting buffer, fusion rule): the captured
  // kernel arguments name the buffers directly.
  const auto key = std::make_pair(steps * 4 + (p->halo_uniform ? 2 : 0), p->cur);
  auto it = p->graphs.find(key);
  if (it == p->graphs.end()) {
    const int cur0 = p->cur;
    const int64_t launches0 = p->launches;
    hipGraph_t graph = nullptr;
    CS_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    cs_status st = issue_steps(p, steps);
    hipError_t e = hipStreamEndCapture(p->stream, &graph);
    p->cur = cur0;
    p->launches = launches0;
    if (st != CS_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return st;
    }
    CS_HIP(e);
    hipGraphExec_t exec = nullptr;
    hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    CS_HIP(ei);
    it = p->graphs.emplace(key, exec).first;
  }
  CS_HIP(hipGraphLaunch(it->second, p->stream));
  const int64_t n = count_launches(p, steps);
  p->cur = (int)((p->cur + n) & 1);
  p->launches += n;
  p->steps_done += steps;
  return CS_OK;
}

void destroy_graphs(cs_plan *p) {
  for (auto &kv : p->graphs) (void)hipGraphExecDestroy(kv.second);
  p->graphs.clear();
}

std::string tune_key(const cs_layout &L, int ndim, int R, int shape, const cs_options &o) {
  char k[160];
  snprintf(k, sizeof k, "d%d_r%d_s%d_p%d_n%lldx%lldx%lld_g%d_f%d", ndim, R, shape, o.precision, (long long)L.n[0],
           (long long)L.n[1], (long long)L.n[2], o.use_graph ? 1 : 0, o.fusion_depth);
  return std::string(k);
}

// Variants a plan may run: same radius and precision, its shape class (split precision:
// also the box variants), and the requested fusion depth if one is set.
std::vector<const Variant *> tune_candidates(const cs_plan *p, int classified_shape) {
  std::vector<const Variant *> out;
  int cnt = 0;
  const Variant *vs = family(p->ndim, &cnt, p->opt.precision);
  for (int i = 0; i < cnt; ++i) {
    const Variant &v = vs[i];
    if (v.R != p->R) continue;
    if (v.shape != classified_shape && !(p->opt.precision != CS_PREC_FP64 && v.shape == cs::SHAPE_BOX)) continue;
    if (p->opt.precision != CS_PREC_FP64 && v.mf != p->opt.precision) continue;
    // the --breakdown ladder's FMA kernels (for split plans mf holds the precision instead)
    if (p->opt.precision == CS_PREC_FP64 && (v.mf == 1 || v.mf == 2)) continue;
    if (p->opt.fusion_depth > 0 && v.kf != p->opt.fusion_depth) continue;
    out.push_back(&v);
  }
  return out;
}

}  // namespace

extern "C" {

const char *cs_version(void) { return "convstencil-mi355x 0.1.0 (gfx950)"; }

const char *cs_status_string(cs_status s) {
  switch (s) {
    case CS_OK: return "ok";
    case CS_ERR_INVALID: return "invalid argument";
    case CS_ERR_UNSUPPORTED: return "unsupported configuration";
    case CS_ERR_ALLOC: return "allocation failed";
    case CS_ERR_DEVICE: return "HIP error";
    case CS_ERR_NO_DEVICE: return "no gfx950 device";
  }
  return "unknown status";
}

cs_status cs_device_count(int32_t *count) {
  if (!count) return CS_ERR_INVALID;
  *count = 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return CS_OK;
  for (int d = 0; d < n; ++d) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d) == hipSuccess && strncmp(prop.gcnArchName, "gfx950", 6) == 0) ++*count;
  }
  return CS_OK;
}

void cs_fill_random_host(double *a, const cs_layout *L, uint64_t seed) {
  if (!a || !L) return;
  const int64_t total = cell_count(L);
  std::fill(a, a + total, 0.0);
  uint64_t st = cs::stream_seed(seed, 0);
  for (int64_t z = 0; z < L->n[0]; ++z)
    for (int64_t y = 0; y < L->n[1]; ++y) {
      double *row = a + ((L->org[0] + z) * L->ext[1] + (L->org[1] + y)) * L->ext[2] + L->org[2];
      for (int64_t x = 0; x < L->n[2]; ++x) row[x] = cs::unit_data(cs::splitmix_next(st));
    }
}

void cs_fill_weights_host(double *w, int32_t count, uint64_t seed) {
  if (!w || count <= 0) return;
  uint64_t st = cs::stream_seed(seed, 1);
  double s = 0.0;
  for (int i = 0; i < count; ++i) {
    w[i] = cs::unit_weight(cs::splitmix_next(st));
    s = s + w[i];
  }
  for (int i = 0; i < count; ++i) w[i] = w[i] / s;
}

}  // extern "C"

namespace {

// Times every candidate variant for at least kTuneSteps steps, a whole number of its
// fused launches (a remainder launch would come from another variant), with the graph
// as the plan asks, on the plan's own buffers and seeded synthetic data, and keeps the
// fastest -- unless it beats the library's default choice by less than kTuneMargin, a
// difference that timings on one device repeat within (box2d1r 10240^2: depth 8 and 12
// within 2%, picked either way run to run); the buffers are zeroed afterwards (the caller
// uploads or fills next).
constexpr int64_t kTuneSteps = 96;
constexpr double kTuneMargin = 0.03;

cs_status autotune_plan(cs_plan *p, int classified) {
  const std::vector<const Variant *> cand = tune_candidates(p, classified);
  if (cand.empty()) return CS_OK;
  const cs_options user = p->opt;
  int best = 0, dflt = 0;
  double best_ms = 1e300, dflt_ms = 1e300;
  p->shape = classified;
  if (select_variants(p) == CS_OK) dflt = p->vmain->id;
  for (const Variant *v : cand) {
    destroy_graphs(p);
    p->opt = user;
    p->opt.tile_variant = v->id;
    p->shape = classified;
    if (select_variants(p) != CS_OK) continue;
    CS_TRY(cs_plan_fill_random(p, 1));
    cs_report r{};
    const int64_t steps = (kTuneSteps + v->kf - 1) / v->kf * v->kf;
    CS_TRY(cs_plan_time(p, steps, 1, 3, &r));
    const double ms = r.elapsed_ms / (double)steps;
    if (p->opt.verbose) fprintf(stderr, "convstencil: autotune %s %.4f ms/step\n", v->name, ms);
    if (ms < best_ms) {
      best_ms = ms;
      best = v->id;
    }
    if (v->id == dflt) dflt_ms = ms;
  }
  if (dflt != 0 && best_ms > dflt_ms * (1.0 - kTuneMargin)) {
    best = dflt;
    best_ms = dflt_ms;
  }
  destroy_graphs(p);
  p->opt = user;
  p->shape = classified;
  if (best == 0) return select_variants(p);
  p->opt.tile_variant = best;
  CS_TRY(select_variants(p));
  p->tuned = 2;
  CS_HIP(cs::launch_zero(p->buf[0], (size_t)cell_count(&p->L), p->stream));
  CS_HIP(cs::launch_zero(p->buf[1], (size_t)cell_count(&p->L), p->stream));
  CS_HIP(hipStreamSynchronize(p->stream));
  p->cur = 0;
  p->steps_done = 0;
  p->halo_uniform = true;
  return cs_tune_store(tune_key(p->L, p->ndim, p->R, classified, user).c_str(), best, best_ms);
}

}  // namespace

namespace cs {

// The FP64 kernel choice of a plan, made without a plan (nothing allocated, no device
// touched): the main variant, the single-step variant and one variant per remainder depth
// 1 .. kf - 1 (vrem[r], null where the family has none), as select_variants picks them for
// a grid of `cells` interior cells.
// The launch sequence of the register-streaming 2-D family for `steps` steps of a radius-R
// stencil of this shape class: the plan's default depth as often as it fits, then the
// family's largest depth that fits the rest; empty when the plan would not take the family.
void reg_sequence(int R, int shape, int64_t steps, std::vector<const Variant *> *seq) {
  seq->clear();
  const Variant *v = pick_main(2, R, shape, CS_PREC_FP64, 0, 0);
  if (!v || v->mf != 7) return;
  for (int64_t done = 0; done < steps;) {
    const int64_t rem = steps - done;
    const Variant *u = nullptr;
    if (rem >= v->kf) {
      u = v;
    } else {
      const int top = max_kf(2, 7, R, shape);  // the largest compiled depth that fits
      for (int kf = (int)std::min<int64_t>(rem, top); kf >= 1 && !u; --kf) u = find_variant(2, 7, R, shape, kf, 0);
    }
    if (!u) {
      seq->clear();
      return;
    }
    seq->push_back(u);
    done += u->kf;
  }
}

cs_status pick_fp64_variants(int ndim, int R, int shape, int fusion_depth, int tile_variant, int64_t cells,
                             const Variant **vmain, const Variant **vone, const Variant **vrem, int vrem_cap) {
  const Variant *v = pick_main(ndim, R, shape, CS_PREC_FP64, fusion_depth, tile_variant, cells);
  if (!v) return tile_variant > 0 ? CS_ERR_INVALID : CS_ERR_UNSUPPORTED;
  if (v->kf > vrem_cap) return CS_ERR_UNSUPPORTED;
  *vmain = v;
  *vone = single_step(ndim, R, v->shape, v);
  if (!*vone) return CS_ERR_UNSUPPORTED;
  for (int r = 0; r < vrem_cap; ++r)
    vrem[r] = r >= 1 && r < v->kf ? find_variant(ndim, v->mf, R, v->shape, r, 0, CS_PREC_FP64) : nullptr;
  return CS_OK;
}

}  // namespace cs

extern "C" {

cs_status cs_tune_key(const cs_layout *layout, const cs_stencil *st, const cs_options *opt, char *buf, int32_t len) {
  if (!layout || !st || !st->weights || !buf || len <= 0 || st->ndim != layout->ndim || st->radius < 1 ||
      st->radius > 4)
    return CS_ERR_INVALID;
  const int R = st->radius;
  const int shape = layout->ndim == 2 ? classify_2d(R, st->weights)
                    : layout->ndim == 3 ? classify_3d(R, st->weights)
                                        : cs::SHAPE_BOX;
  cs_options o{};
  if (opt) o = *opt;
  const std::string k = tune_key(*layout, layout->ndim, R, shape, o);
  if ((int32_t)k.size() >= len) return CS_ERR_INVALID;
  memcpy(buf, k.c_str(), k.size() + 1);
  return CS_OK;
}

cs_status cs_plan_create(cs_plan **out, const cs_layout *layout, const cs_stencil *st, const cs_options *opt) {
  if (!out || !layout || !st || !st->weights) return CS_ERR_INVALID;
  *out = nullptr;
  if (st->ndim != layout->ndim || st->radius < 1) return CS_ERR_INVALID;
  if (!layout_valid(layout, st->radius)) return CS_ERR_INVALID;
  if (st->radius > (st->ndim == 1 ? 4 : 3)) return CS_ERR_UNSUPPORTED;
  int dev = 0;
  CS_TRY(check_device(&dev));
  cs_plan *p = new cs_plan();
  p->L = *layout;
  p->ndim = layout->ndim;
  p->R = st->radius;
  const int D = 2 * p->R + 1;
  const int nw = p->ndim == 1 ? D : p->ndim == 2 ? D * D : D * D * D;
  p->w.assign(st->weights, st->weights + nw);
  if (opt) p->opt = *opt;
  p->device = dev;
  p->shape = p->ndim == 2 ? classify_2d(p->R, p->w.data()) : p->ndim == 3 ? classify_3d(p->R, p->w.data()) : cs::SHAPE_BOX;
  const int classified = p->shape;
  if (p->opt.autotune > 0 && p->opt.tile_variant == 0) {  // a cached choice needs no timing
    int32_t vid = 0;
    if (p->opt.autotune == 1 &&
        cs_tune_lookup(tune_key(p->L, p->ndim, p->R, classified, p->opt).c_str(), &vid, nullptr) == CS_OK) {
      p->opt.tile_variant = vid;
      p->tuned = 1;
    }
  }
  cs_status s = select_variants(p);
  if (s != CS_OK && p->tuned) {  // a stale cache entry: choose again
    p->opt.tile_variant = 0;
    p->tuned = 0;
    p->shape = classified;
    s = select_variants(p);
  }
  if (s != CS_OK) {
    cs_plan_destroy(p);
    return s;
  }
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
    cs_plan_destroy(p);
    return CS_ERR_DEVICE;
  }
  const size_t bytes = (size_t)cell_count(layout) * sizeof(double);
  if (hipMalloc(&p->buf[0], bytes) != hipSuccess || hipMalloc(&p->buf[1], bytes) != hipSuccess) {
    (void)hipGetLastError();
    cs_plan_destroy(p);
    return CS_ERR_ALLOC;
  }
  // both buffers start at zero, so the range guard's bound (0) holds before any upload
  if (cs::launch_zero(p->buf[0], (size_t)cell_count(layout), p->stream) != hipSuccess ||
      cs::launch_zero(p->buf[1], (size_t)cell_count(layout), p->stream) != hipSuccess ||
      hipStreamSynchronize(p->stream) != hipSuccess) {
    cs_plan_destroy(p);
    return CS_ERR_DEVICE;
  }
  p->amax = 0.0;
  p->wabs = abs_sum(p->w.data(), p->w.size());
  if (p->opt.autotune > 0 && p->opt.tile_variant == 0) {
    s = autotune_plan(p, classified);
    if (s != CS_OK) {
      cs_plan_destroy(p);
      return s;
    }
  }
  *out = p;
  return CS_OK;
}

cs_status cs_plan_destroy(cs_plan *p) {
  if (!p) return CS_ERR_INVALID;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  destroy_graphs(p);
  for (double *&b : p->buf)
    if (b) (void)hipFree(b);
  if (p->bfrag) (void)hipFree(p->bfrag);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return CS_OK;
}

cs_status cs_plan_upload(cs_plan *p, const double *host_in) {
  if (!p || !host_in) return CS_ERR_INVALID;
  CS_HIP(hipSetDevice(p->device));
  const size_t n = (size_t)cell_count(&p->L);
  CS_HIP(hipMemcpyAsync(p->buf[0], host_in, n * sizeof(double), hipMemcpyHostToDevice, p->stream));
  CS_HIP(cs::launch_zero(p->buf[1], n, p->stream));
  CS_HIP(hipStreamSynchronize(p->stream));
  p->cur = 0;
  p->steps_done = 0;
  p->halo_uniform = halo_is_zero(&p->L, host_in);
  p->amax = abs_max(host_in, n);
  return CS_OK;
}

cs_status cs_plan_fill_random(cs_plan *p, uint64_t seed) {
  if (!p) return CS_ERR_INVALID;
  CS_HIP(hipSetDevice(p->device));
  Geo g{};
  for (int d = 0; d < 3; ++d) {
    g.ext[d] = p->L.ext[d];
    g.org[d] = p->L.org[d];
    g.n[d] = p->L.n[d];
  }
  CS_HIP(cs::launch_fill_random(p->buf[0], g, cs::stream_seed(seed, 0), p->stream));
  CS_HIP(cs::launch_zero(p->buf[1], (size_t)cell_count(&p->L), p->stream));
  CS_HIP(hipStreamSynchronize(p->stream));
  p->cur = 0;
  p->steps_done = 0;
  p->halo_uniform = true;
  p->amax = 1.0;  // the fill is uniform in [-1, 1) with a zero halo
  return CS_OK;
}

cs_status cs_plan_run(cs_plan *p, int64_t steps, cs_report *rep) {
  if (!p) return CS_ERR_INVALID;
  const int64_t l0 = p->launches;
  CS_TRY(run_steps(p, steps));
  if (rep) {
    memset(rep, 0, sizeof(*rep));
    rep->launches = p->launches - l0;
    rep->fusion_depth = p->halo_uniform ? p->vmain->kf : 1;
    rep->variant = p->vmain->id;
  }
  return CS_OK;
}

cs_status cs_plan_time(cs_plan *p, int64_t steps, int32_t warmup, int32_t reps, cs_report *rep) {
  if (!p || reps <= 0 || warmup < 0 || steps <= 0) return CS_ERR_INVALID;
  CS_HIP(hipSetDevice(p->device));
  // graphs are captured per starting buffer: a run of an odd number of launches alternates
  // the two, so a warm-up captures both before the timed runs
  if (warmup > 0 && p->opt.use_graph && (count_launches(p, steps) & 1)) warmup = std::max(warmup, 2);
  for (int i = 0; i < warmup; ++i) CS_TRY(run_steps(p, steps));
  CS_HIP(hipStreamSynchronize(p->stream));
  hipEvent_t e0, e1;
  CS_HIP(hipEventCreate(&e0));
  CS_HIP(hipEventCreate(&e1));
  const int64_t l0 = p->launches;
  CS_HIP(hipEventRecord(e0, p->stream));
  for (int i = 0; i < reps; ++i) CS_TRY(run_steps(p, steps));
  CS_HIP(hipEventRecord(e1, p->stream));
  CS_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  CS_HIP(hipEventElapsedTime(&ms, e0, e1));
  CS_HIP(hipEventDestroy(e0));
  CS_HIP(hipEventDestroy(e1));
  if (rep) {
    memset(rep, 0, sizeof(*rep));
    rep->elapsed_ms = ms / reps;
    const double pts = (double)p->L.n[0] * (double)p->L.n[1] * (double)p->L.n[2];
    // a grid without interior cells launches nothing: no rate
    rep->gstencil_per_s = rep->elapsed_ms > 0 && pts > 0 ? pts * (double)steps / (rep->elapsed_ms * 1e-3) / 1e9 : 0.0;
    rep->launches = (p->launches - l0) / reps;
    rep->fusion_depth = p->halo_uniform ? p->vmain->kf : 1;
    rep->variant = p->vmain->id;
  }
  return CS_OK;
}

cs_status cs_plan_sync(cs_plan *p) {
  if (!p) return CS_ERR_INVALID;
  CS_HIP(hipStreamSynchronize(p->stream));
  return CS_OK;
}

cs_status cs_plan_download(cs_plan *p, double *host_out) {
  if (!p || !host_out) return CS_ERR_INVALID;
  CS_HIP(hipSetDevice(p->device));
  CS_HIP(hipMemcpyAsync(host_out, p->buf[p->cur], (size_t)cell_count(&p->L) * sizeof(double), hipMemcpyDeviceToHost,
                        p->stream));
  CS_HIP(hipStreamSynchronize(p->stream));
  return CS_OK;
}

cs_status cs_plan_device_buffers(cs_plan *p, double **current, double **other) {
  if (!p) return CS_ERR_INVALID;
  if (current) *current = p->buf[p->cur];
  if (other) *other = p->buf[p->cur ^ 1];
  return CS_OK;
}

cs_status cs_plan_describe(cs_plan *p, char *buf, int32_t len) {
  if (!p || !buf || len <= 0) return CS_ERR_INVALID;
  static const char *how[3] = {"default", "cached", "timed"};
  snprintf(buf, (size_t)len,
           "variant=%d name=%s mf=%d R=%d shape=%s kf=%d tile=%dx%d waves=%d lds=%zu halo_uniform=%d autotune=%s",
           p->vmain->id, p->vmain->name, p->vmain->mf, p->R, p->shape == cs::SHAPE_BOX ? "box" : "star", p->vmain->kf,
           p->vmain->by, p->vmain->bx, p->vmain->nw, p->vmain->lds_bytes, (int)p->halo_uniform, how[p->tuned]);
  return CS_OK;
}

cs_status cs_run(const cs_layout *layout, const cs_stencil *st, const double *host_in, double *host_out, int64_t steps,
                 const cs_options *opt, cs_report *rep) {
  if (!host_in || !host_out || steps < 0) return CS_ERR_INVALID;
  cs_plan *p = nullptr;
  CS_TRY(cs_plan_create(&p, layout, st, opt));
  cs_status s = cs_plan_upload(p, host_in);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s == CS_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) s = CS_ERR_DEVICE;
  if (s == CS_OK && hipEventRecord(e0, p->stream) != hipSuccess) s = CS_ERR_DEVICE;
  const int64_t l0 = p->launches;  // the autotuner's timing launches are not this run's
  if (s == CS_OK) s = run_steps(p, steps);
  if (s == CS_OK && hipEventRecord(e1, p->stream) != hipSuccess) s = CS_ERR_DEVICE;
  if (s == CS_OK) s = cs_plan_download(p, host_out);
  if (s == CS_OK && rep) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    memset(rep, 0, sizeof(*rep));
    rep->elapsed_ms = ms;
    const double pts = (double)layout->n[0] * (double)layout->n[1] * (double)layout->n[2];
    rep->gstencil_per_s = ms > 0 ? pts * (double)steps / (ms * 1e-3) / 1e9 : 0.0;
    rep->launches = p->launches - l0;
    rep->fusion_depth = p->halo_uniform ? p->vmain->kf : 1;
    rep->variant = p->vmain->id;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  cs_plan_destroy(p);
  return s;
}

}  // extern "C"
