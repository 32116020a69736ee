// grn_host.cpp -- the C ABI of libgrn_hip.so: device context, experiment
// upload, model compilation, kernel dispatch and result collection.
//
// Host-side mirror of Src/GPUSimulator (CEvaluatorMorphogens,
// CModelSimulator): one context per device and caller thread, one HIP stream
// per context, experiment planes uploaded once, models compiled on the host
// and evaluated by the fused on-chip kernel.  Batches are grouped by kernel
// class (grid width, gene range), ordered costliest-first and run as one
// persistent launch per class, synchronizing once per batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "grn.h"
#include "grn_internal.hpp"
#include "grn_kernels.h"
#include "grn_model.hpp"
#include "grn_pool.hpp"
#include "grn_program.h"

// The batch buffers of grn_evaluate_batch_noalloc: one fixed array per device
// (each device holds its own copy, at its own address, until the process
// ends).  Layout: the queue table, then programs, work order, maxError bounds
// and results for at most GRN_NOALLOC_MAX_MODELS models, each region aligned
// to 256 bytes.  Host code never uses the array's name as an address: the
// address comes from hipGetSymbolAddress, once per device.
namespace noalloc {
constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
constexpr int kMax = GRN_NOALLOC_MAX_MODELS;
constexpr size_t kBatch = 0;
constexpr size_t kProgs = align256(sizeof(GrnBatch));
constexpr size_t kOrder = kProgs + align256(sizeof(GrnProgram) * kMax);
constexpr size_t kMaxErr = kOrder + align256(sizeof(int) * kMax);
constexpr size_t kOut = kMaxErr + align256(sizeof(double) * kMax);
constexpr size_t kBytes = kOut + align256(sizeof(GrnModelOut) * kMax);
static_assert(kBytes == GRN_NOALLOC_SCRATCH_BYTES, "grn.h states the scratch array's size");
}  // namespace noalloc

static __device__ __align__(256) unsigned char scratch[noalloc::kBytes];

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(GRN_E_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                 \
  } while (0)

// Grow-only device buffer.
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int reserve(size_t bytes) {
    if (bytes <= cap) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max(bytes, (size_t)4096);
    if (hipMalloc(&p, want) != hipSuccess) return fail(GRN_E_NOMEM, "hipMalloc(%zu) failed", want);
    cap = want;
    return 0;
  }
};

struct HostDataset {
  grn_sim_params params;
  int n_exp;
  bool nonneg;
  void *d_init = nullptr;
  void *d_target = nullptr;
};

struct TileGeom {
  int tw = 0, th = 0, sx = 0, sy = 0, halo = 0;
  bool operator==(const TileGeom &o) const {
    return tw == o.tw && th == o.th && sx == o.sx && sy == o.sy && halo == o.halo;
  }
};

struct KernelChoice {
  const GrnKernelEntry *k = nullptr;
  int block = 0;
  size_t lds = 0;
  TileGeom tile;       // halo-tiled kernels only
  long slot_floats = 0;  // device-memory state per workgroup (halo-tiled)
};

// A size class of the autotuner: grid, genes in use, target genes, guard.
struct TuneKey {
  int W, H, G, n_tgt;
  bool guard;
  bool operator<(const TuneKey &o) const {
    return std::tie(W, H, G, n_tgt, guard) < std::tie(o.W, o.H, o.G, o.n_tgt, o.guard);
  }
};

// The device buffers of a batch: the context's own (grow-only, reused by
// every grn_evaluate_batch) or a staged batch's.
struct BatchBufs {
  DevBuf progs, progs64, order, maxerr, out, batch, scratch;
};

// Device addresses a batch's launches read and write: a BatchBufs' buffers,
// or regions of the fixed scratch array of grn_evaluate_batch_noalloc.
struct DevPtrs {
  void *progs, *progs64, *order, *maxerr, *out, *batch, *scratch;
};

DevPtrs ptrs_of(const BatchBufs &B) {
  return DevPtrs{B.progs.p, B.progs64.p, B.order.p, B.maxerr.p, B.out.p, B.batch.p, B.scratch.p};
}

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // a batch's odd kernel classes (default priority, like stream)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int n_cu = 0;
  size_t max_lds = 0;
  std::vector<HostDataset> datasets;
  DevBuf d_datasets, d_fields;
  BatchBufs bufs;  // the buffers of grn_evaluate_batch and its kin
  std::vector<const GrnKernelEntry *> kernels;
  std::map<const void *, size_t> lds_set;  // dynamic-LDS attribute already applied
  int tile_halo = 4;                        // Euler steps per tile load
  int launch_mode = GRN_LAUNCH_EAGER;
  std::map<std::vector<long>, hipGraphExec_t> graphs;  // by launch layout
  std::map<TuneKey, KernelChoice> tuned;               // autotuner's choices
  mutable std::map<std::pair<TuneKey, int>, KernelChoice> defaults;  // default choice per class and path
  grn_stats stats{};
  // Pinned host staging of the program uploads (a batch's largest upload:
  // 2 KB per model).  Every batch that fills it waits for its stream before
  // returning, so the next batch can refill it.
  GrnProgram *h_progs = nullptr;
  size_t h_progs_cap = 0;
  // A batch with a host stage (evaluate_batch_hooked): its results are copied
  // into pinned memory and a host function -- result conversion plus the
  // caller's hook -- runs in stream order after the launches, inside the
  // captured graph in graph mode.  The host function reads `hook`, whose
  // address the graph bakes in and whose fields are set before each launch.
  struct Hook {
    Ctx *c = nullptr;
    const void *plan = nullptr;
    grn_result *out = nullptr;
    int *n_products = nullptr;
    grn::BatchHook fn = nullptr;
    void *user = nullptr;
    int rc = 0;
    std::string err;
    double cell_steps = 0, flops = 0;
  } hook;
  GrnModelOut *h_out = nullptr;
  size_t h_out_cap = 0;
  // grn_evaluate_batch_noalloc: this device's scratch array and host-side
  // staging sized for its capacity at creation
  void *noalloc = nullptr;
  std::vector<GrnProgram> na_progs;
  std::vector<double> na_maxerr;
  std::vector<GrnModelOut> na_res;

  void drop_graphs() {
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
  }

  ~Ctx() {
    drop_graphs();
    for (HostDataset &d : datasets) {
      if (d.d_init) (void)hipFree(d.d_init);
      if (d.d_target) (void)hipFree(d.d_target);
    }
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream) (void)hipStreamDestroy(stream);
    if (h_progs) (void)hipHostFree(h_progs);
    if (h_out) (void)hipHostFree(h_out);
  }
};

void collect_kernels(std::vector<const GrnKernelEntry *> *out) {
  const GrnKernelEntry *(*tables[])(int *) = {grn_field_table_w32, grn_field_table_w32wide, grn_field_table_w32wide9,
                                              grn_field_table_w32wide13, grn_field_table_w8, grn_field_table_w48,
                                              grn_field_table_w64, grn_field_table_wany, grn_field_table_wanyguard,
                                              grn_field_table_tiled512, grn_field_table_tiled,
                                              grn_field_table_tiled512s, grn_field_table_tiled512g,
                                              grn_field_table_tiled1024, grn_field_table_exact};
  for (auto fn : tables) {
    int n = 0;
    const GrnKernelEntry *t = fn(&n);
    for (int i = 0; i < n; ++i) out->push_back(&t[i]);
  }
}

// Window of the halo-tiled kernel for a W x H grid: extended tiles of
// tw x th cells (th a multiple of cpt, at most ntmax threads, tw at most the
// kernel's fixed LDS pitch when it has one, G LDS planes within the LDS
// budget), interior (tw - 2 halo) x (th - 2 halo).  Minimises the cells
// computed per step over all tiles plus a per-tile overhead.
bool tile_geometry(int W, int H, int G, int cpt, int ntmax, int pitch, size_t max_lds, int halo, TileGeom *out) {
  double best = 1e300;
  bool found = false;
  const int th_max = ((H + 2 * halo + cpt - 1) / cpt) * cpt;
  for (int th = cpt; th <= th_max; th += cpt) {
    const int sy = std::min(th - 2 * halo, H);
    if (sy < 1) continue;
    for (int tw = 2 * halo + 1; tw <= W + 2 * halo; ++tw) {
      const int sx = std::min(tw - 2 * halo, W);
      const long threads = (long)tw * (th / cpt);
      if (threads > ntmax || (pitch > 0 && tw > pitch)) break;
      const size_t lds = (size_t)G * ((pitch > 0 ? pitch : tw) + 2) * (th + 2) * sizeof(float);
      if (lds > max_lds) break;
      const int ntx = (W + sx - 1) / sx, nty = (H + sy - 1) / sy;
      const double cost = (double)ntx * nty * ((double)tw * th + 2048.0);
      if (cost < best) {
        best = cost;
        found = true;
        out->tw = tw;
        out->th = th;
        out->sx = sx;
        out->sy = sy;
        out->halo = halo;
      }
    }
  }
  return found;
}

// Every kernel configuration that can run a W x H grid with G genes, in the
// order the default heuristic prefers them: on-chip kernels (exact width,
// then any width) unless the tiled path is forced, then the halo-tiled
// kernels with halo widths 4, 2 and 8 (512-thread instantiations first:
// they keep their state in registers without spilling).  Of two on-chip
// kernels of one pass that launch the same threads (a grid of at most 64
// cells is one wavefront either way), the one with fewer cells per thread
// comes first: the same wavefronts, less serial work per Euler step.
std::vector<KernelChoice> kernel_candidates(const Ctx &c, int W, int H, int G, int n_tgt, bool guard, int path) {
  std::vector<KernelChoice> out;
  if (path == GRN_PATH_EXACT) {  // the FP64 kernel: state in device memory, any grid
    (void)n_tgt;
    (void)guard;
    for (const GrnKernelEntry *k : c.kernels) {
      if (k->wc != -2 || G < k->glo || G > k->ghi) continue;
      KernelChoice kc;
      kc.k = k;
      kc.block = k->ntmax;
      kc.lds = 0;
      kc.slot_floats = 4L * G * W * H;  // two [G][H][W] double plane sets
      out.push_back(kc);
    }
    return out;
  }
  for (int pass = 0; pass < 2 && path != GRN_PATH_TILED; ++pass) {  // exact width first, then any width
    const size_t pass_begin = out.size();
    for (const GrnKernelEntry *k : c.kernels) {
      if (pass == 0 ? k->wc != W : k->wc != 0) continue;
      if (G < k->glo || G > k->ghi || k->guard != guard) continue;
      if (H % k->cpt) continue;
      const int active = W * (H / k->cpt);
      const int block = (active + 63) / 64 * 64;
      if (block > k->ntmax) continue;
      const size_t planes = (size_t)G * (W + 2) * (H + 2) * sizeof(float);
      const size_t stage = (size_t)n_tgt * W * H * sizeof(double);
      const size_t lds = std::max(planes, stage);
      if (lds > c.max_lds) continue;
      KernelChoice kc;
      kc.k = k;
      kc.block = block;
      kc.lds = lds;
      size_t at = out.size();
      for (size_t j = pass_begin; j < out.size(); ++j)
        if (out[j].block == block && out[j].k->cpt > k->cpt) {
          at = j;
          break;
        }
      out.insert(out.begin() + (long)at, kc);
    }
  }
  if (path != GRN_PATH_LDS) {
    for (int halo : {c.tile_halo, 2, 8}) {
      for (const GrnKernelEntry *k : c.kernels) {
        if (k->wc != -1 || G < k->glo || G > k->ghi || k->guard != guard) continue;
        TileGeom t;
        if (!tile_geometry(W, H, G, k->cpt, k->ntmax, k->pitch, c.max_lds, halo, &t)) continue;
        KernelChoice kc;
        kc.k = k;
        kc.tile = t;
        kc.block = (int)(((long)t.tw * (t.th / k->cpt) + 63) / 64 * 64);
        kc.lds = (size_t)G * ((k->pitch > 0 ? k->pitch : t.tw) + 2) * (t.th + 2) * sizeof(float);
        kc.slot_floats = 2L * grn_cell_rec(G) * W * H;  // two sets of per-cell (hi, lo) records
        out.push_back(kc);
      }
    }
  }
  return out;
}

// The kernel for a grid of W x H with G genes: the autotuned choice for the
// size class when there is one (GRN_PATH_AUTO), else the first candidate.
int choose_kernel(const Ctx &c, int W, int H, int G, int n_tgt, bool guard, int path, KernelChoice *out) {
  if (path == GRN_PATH_AUTO) {
    auto it = c.tuned.find(TuneKey{W, H, G, n_tgt, guard});
    if (it != c.tuned.end()) {
      *out = it->second;
      return 0;
    }
  }
  const std::pair<TuneKey, int> dk{TuneKey{W, H, G, n_tgt, guard}, path};
  auto hit = c.defaults.find(dk);
  if (hit != c.defaults.end()) {
    *out = hit->second;
    return 0;
  }
  const std::vector<KernelChoice> cand = kernel_candidates(c, W, H, G, n_tgt, guard, path);
  if (!cand.empty()) {
    *out = cand[0];
    c.defaults[dk] = cand[0];
    return 0;
  }
  *out = KernelChoice();
  return fail(GRN_E_CAPACITY,
              "no on-chip kernel for a %dx%d grid with %d genes (LDS %zu bytes available)", W, H, G,
              c.max_lds);
}

Ctx *as_ctx(void *p) { return static_cast<Ctx *>(p); }

// This device's address of the scratch array, looked up on the first context
// created on the device (with the device current) and kept per device.
std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_addr;

int scratch_address(int device, void **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  if (it == g_scratch_addr.end()) {
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    it = g_scratch_addr.emplace(device, p).first;
  }
  *out = it->second;
  return 0;
}

int to_model(const grn_model *gm, grn::Model *m) {
  if (!grn::model_from_abi(gm, m)) return fail(GRN_E_INVALID, "invalid grn_model");
  return 0;
}

struct Item {
  int dataset;
  GrnProgram prog;
  grn::ProgramInfo info;
  KernelChoice kc;
  double cost;
  double max_err;
};

// Options of a batch beyond the public arguments (autotuner timing runs).
struct BatchOptions {
  const KernelChoice *force = nullptr;  // run every model with this configuration
  int step_cap = 0;                     // > 0: at most this many Euler steps
};

// A compiled and planned batch: everything its launches need but the device
// buffers.  Queue position j runs model idx[j]; group g is queue g.
struct Plan {
  int n = 0;
  std::vector<Item> items;  // input order
  std::vector<GrnProgram64> p64;  // [n] when a model runs on the FP64 kernel (else empty)
  std::vector<int> idx;
  std::vector<int> rank;  // per item: its kernel class's rank (the order of the launches)
  std::vector<std::pair<int, int>> groups;  // [begin, end) in idx order
  struct Launch {
    KernelChoice kc;
    int grid;
    size_t scratch_off;  // floats: the launch's own state slots (launches may run concurrently)
  };
  std::vector<Launch> launches;
  GrnBatch batch{};
  size_t scratch_bytes = 0;
  int step_cap = 0;
};

// Compiles every model, chooses its kernel configuration, groups the models
// by configuration (and grid, for tiled kernels), orders each group
// costliest-first and sizes each group's persistent launch.
int make_plan(Ctx *c, int n, const int *dataset_ids, const grn_model *models, const double *max_errors,
              const double *dt_override, int path, const BatchOptions &opt, Plan *P) {
  if (n < 0 || (n > 0 && !models)) return fail(GRN_E_INVALID, "invalid batch arguments");
  P->n = n;
  P->step_cap = opt.step_cap;
  std::vector<Item> &items = P->items;
  items.resize(n);  // every field of every item is set below (a reused plan keeps its memory)
  for (int i = 0; i < n; ++i) {
    const int ds = dataset_ids ? dataset_ids[i] : 0;
    if (ds < 0 || ds >= (int)c->datasets.size()) return fail(GRN_E_INVALID, "model %d: unknown dataset %d", i, ds);
    items[i].dataset = ds;
  }
  // The path each model takes: GRN_PATH_AUTO sends models of datasets with
  // REF fitness and an early-stop threshold to the FP64 kernel, whose stop
  // step and penalty are the reference's (the FP32 kernels' rounding of the
  // convergence measure can move both); everything else to the FP32 kernels.
  auto model_path = [&](int ds) {
    if (path != GRN_PATH_AUTO) return path;
    const grn_sim_params &sp = c->datasets[ds].params;
    return (sp.fitness_kind == GRN_FITNESS_REF && sp.min_conc_change >= 0) ? (int)GRN_PATH_EXACT : path;
  };
  bool any64 = false;
  for (int i = 0; i < n; ++i) any64 |= model_path(items[i].dataset) == GRN_PATH_EXACT;
  P->p64.clear();
  if (any64) P->p64.resize(n);
  // Compile every model (independent work: host threads for large batches).
  std::vector<int> crc(n, 0);
  std::vector<std::string> cerr(n);
  auto compile_range = [&](int t, int nt) {
    for (int i = t; i < n; i += nt) {
      const HostDataset &D = c->datasets[items[i].dataset];
      grn::Model m;
      if (!grn::model_from_abi(&models[i], &m)) {
        crc[i] = GRN_E_INVALID;
        cerr[i] = "invalid grn_model";
        continue;
      }
      const double dt = (dt_override && dt_override[i] > 0) ? dt_override[i] : D.params.dt;
      const int n_morph = D.params.n_input_morphogens + D.params.n_target_morphogens;
      GrnProgram64 *p64 = model_path(items[i].dataset) == GRN_PATH_EXACT ? &P->p64[i] : nullptr;
      crc[i] = grn::compile_model(m, n_morph, items[i].dataset, (float)dt, &items[i].prog, &items[i].info, &cerr[i],
                                  p64, dt);
      if (!crc[i]) crc[i] = grn::validate_program(items[i].prog, (int)c->datasets.size(), &cerr[i]);
    }
  };
  // (the host pool's threads for large batches, in strided chunks)
  const int n_chunks = n >= 256 ? grn::host_threads() : 1;
  grn::parallel_for(n_chunks, [&](int t) { compile_range(t, n_chunks); });
  for (int i = 0; i < n; ++i) {
    if (crc[i]) return fail(crc[i], "model %d: %s", i, cerr[i].c_str());
    const int ds = items[i].dataset;
    const HostDataset &D = c->datasets[ds];
    // the fast step also needs a nonnegative start and no positive minConc
    if (!(D.nonneg && D.params.min_conc <= 0)) items[i].prog.fast = 0;
    int rc;
    const int W = D.params.dim_x, H = D.params.dim_y;
    if (opt.force) {
      items[i].kc = *opt.force;
    } else {
      rc = choose_kernel(*c, W, H, items[i].prog.G, D.params.n_target_morphogens, !D.nonneg, model_path(ds),
                         &items[i].kc);
      if (rc) return fail(rc, "model %d: %s", i, g_last_error.c_str());
    }
    items[i].cost = (double)W * H * D.params.num_sim_steps * D.n_exp * items[i].info.flops_per_cell_step;
    items[i].max_err = max_errors ? max_errors[i] : std::numeric_limits<double>::infinity();
  }

  // Group by kernel, costliest first inside each group (longest-processing-
  // time order for the persistent work queue).
  std::vector<int> &idx = P->idx;
  idx.resize(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  auto same_class = [&](int a, int b) {
    const KernelChoice &x = items[a].kc, &y = items[b].kc;
    if (x.k != y.k || !(x.tile == y.tile)) return false;
    if (x.k->wc == -1) {  // the tiled window is per grid size
      const grn_sim_params &pa = c->datasets[items[a].dataset].params, &pb = c->datasets[items[b].dataset].params;
      return pa.dim_x == pb.dim_x && pa.dim_y == pb.dim_y;
    }
    return true;
  };
  auto class_key = [&](int a) {
    const KernelChoice &x = items[a].kc;
    const grn_sim_params &pa = c->datasets[items[a].dataset].params;
    return std::make_tuple((const void *)x.k, x.tile.tw, x.tile.th, x.tile.sx, x.tile.sy, x.tile.halo,
                           x.k->wc == -1 ? pa.dim_x : 0, x.k->wc == -1 ? pa.dim_y : 0);
  };
  // each item's class as its rank among the batch's (few) distinct keys, so
  // the sort compares two ints, not two tuples built per comparison
  std::vector<decltype(class_key(0))> keys;
  for (int i = 0; i < n; ++i)
    if (i == 0 || class_key(i) != class_key(i - 1)) keys.push_back(class_key(i));
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  std::vector<int> &rank = P->rank;
  rank.resize(n);
  for (int i = 0; i < n; ++i)
    rank[i] = (i > 0 && class_key(i) == class_key(i - 1))
                  ? rank[i - 1]
                  : (int)(std::lower_bound(keys.begin(), keys.end(), class_key(i)) - keys.begin());
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    if (rank[a] != rank[b]) return rank[a] < rank[b];
    return items[a].cost > items[b].cost;
  });

  // One persistent launch per kernel class.
  std::vector<std::pair<int, int>> &groups = P->groups;
  groups.clear();
  for (int i = 0; i < n;) {
    int j = i;
    while (j < n && same_class(idx[j], idx[i])) ++j;
    groups.push_back({i, j});
    i = j;
  }
  if ((int)groups.size() > GRN_MAX_QUEUES) return fail(GRN_E_CAPACITY, "too many kernel classes in one batch");
  P->batch = GrnBatch{};
  P->batch.n_total = n;
  P->batch.n_queues = (int)groups.size();
  P->launches.clear();
  P->scratch_bytes = 0;
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    const int b = groups[gi].first, e = groups[gi].second;
    P->batch.q[gi] = GrnQueue{0, e - b, b, 0};
    // the class's block size and LDS: the largest over its members
    KernelChoice kc = items[idx[b]].kc;
    for (int j = b; j < e; ++j) {
      kc.block = std::max(kc.block, items[idx[j]].kc.block);
      kc.lds = std::max(kc.lds, items[idx[j]].kc.lds);
      kc.slot_floats = std::max(kc.slot_floats, items[idx[j]].kc.slot_floats);
    }
    auto it = c->lds_set.find(kc.k->fn);
    if (it == c->lds_set.end() || it->second < kc.lds) {
      HIP_TRY(hipFuncSetAttribute(kc.k->fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kc.lds));
      c->lds_set[kc.k->fn] = kc.lds;
    }
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kc.k->fn, kc.block, kc.lds));
    if (per_cu < 1) return fail(GRN_E_CAPACITY, "kernel cannot be resident (block %d, LDS %zu)", kc.block, kc.lds);
    // Eager launches start no more workgroups than the class has models; in
    // graph mode every launch starts the full resident grid (the extra
    // workgroups find the queue empty and exit), so a launch's layout does
    // not change with its class's size and the captured graph is replayed
    // for the next batch -- the next generation of a search.
    const int grid = c->launch_mode == GRN_LAUNCH_GRAPH ? per_cu * c->n_cu : std::min(e - b, per_cu * c->n_cu);
    const size_t off = P->scratch_bytes / sizeof(float);
    P->scratch_bytes += sizeof(float) * (size_t)kc.slot_floats * grid;
    P->launches.push_back({kc, grid, off});
  }
  return 0;
}

// Sizes the buffers and enqueues the uploads of programs, order, maxError
// bounds and the batch table.  The programs go through the context's pinned
// staging; the other host arrays are the plan's own.  The caller waits for
// the stream before the staging or those arrays are reused.
int upload_plan(Ctx *c, const Plan &P, BatchBufs &B, std::vector<double> *maxerr) {
  const int n = P.n;
  int rc;
  if ((rc = B.progs.reserve(sizeof(GrnProgram) * std::max(n, 1)))) return rc;
  if (!P.p64.empty() && (rc = B.progs64.reserve(sizeof(GrnProgram64) * n))) return rc;
  if ((rc = B.order.reserve(sizeof(int) * std::max(n, 1)))) return rc;
  if ((rc = B.maxerr.reserve(sizeof(double) * std::max(n, 1)))) return rc;
  if ((rc = B.out.reserve(sizeof(GrnModelOut) * std::max(n, 1)))) return rc;
  if ((rc = B.batch.reserve(sizeof(GrnBatch)))) return rc;
  if (P.scratch_bytes && (rc = B.scratch.reserve(P.scratch_bytes))) return rc;
  hipStream_t st = c->stream;
  if ((size_t)n > c->h_progs_cap) {
    HIP_TRY(hipStreamSynchronize(st));  // no upload still reads the old staging
    if (c->h_progs) (void)hipHostFree(c->h_progs);
    c->h_progs = nullptr;
    c->h_progs_cap = 0;
    const size_t cap = std::max<size_t>((size_t)n + n / 2, 1024);
    HIP_TRY(hipHostMalloc((void **)&c->h_progs, sizeof(GrnProgram) * cap, hipHostMallocDefault));
    c->h_progs_cap = cap;
  }
  maxerr->resize(n);
  // 2 KB per program: a large batch is copied to the staging buffer by
  // several threads (the copy is bound by host memory bandwidth)
  auto stage_range = [&](int t, int nt) {
    const int b = (int)((long)n * t / nt), e = (int)((long)n * (t + 1) / nt);
    for (int i = b; i < e; ++i) {
      c->h_progs[i] = P.items[i].prog;
      (*maxerr)[i] = P.items[i].max_err;
    }
  };
  const int n_copy = n >= 2048 ? 8 : 1;
  grn::parallel_for(n_copy, [&](int t) { stage_range(t, n_copy); });
  if (n > 0) {
    HIP_TRY(hipMemcpyAsync(B.progs.p, c->h_progs, sizeof(GrnProgram) * n, hipMemcpyHostToDevice, st));
    if (!P.p64.empty())
      HIP_TRY(hipMemcpyAsync(B.progs64.p, P.p64.data(), sizeof(GrnProgram64) * n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(B.order.p, P.idx.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(B.maxerr.p, maxerr->data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  }
  HIP_TRY(hipMemcpyAsync(B.batch.p, &P.batch, sizeof(GrnBatch), hipMemcpyHostToDevice, st));
  return 0;
}

void hook_host_fn(void *p);

// Enqueues the batch: result clearing (which also rewinds the work queues),
// then one persistent kernel per class, eagerly or as a captured graph,
// bracketed by the context's timing events.
int launch_plan(Ctx *c, const Plan &P, const DevPtrs &B, DevBuf *fields, long field_stride, int field_exp,
                bool *replayed, bool allow_graph = true, bool hooked = false) {
  *replayed = false;
  if (P.n == 0) return 0;
  std::vector<GrnLaunch> L(P.launches.size());
  for (size_t gi = 0; gi < P.launches.size(); ++gi) {
    const KernelChoice &kc = P.launches[gi].kc;
    GrnLaunch &l = L[gi];
    l.scratch = (float *)B.scratch + P.launches[gi].scratch_off;
    l.slot_floats = kc.slot_floats;
    l.tile_w = kc.tile.tw;
    l.tile_h = kc.tile.th;
    l.tile_sx = kc.tile.sx;
    l.tile_sy = kc.tile.sy;
    l.tile_halo = kc.tile.halo;
    l.progs = (const GrnProgram *)B.progs;
    l.progs64 = P.p64.empty() ? nullptr : (const GrnProgram64 *)B.progs64;
    l.queue = &((GrnBatch *)B.batch)->q[gi];
    l.datasets = (const GrnDataset *)c->d_datasets.p;
    l.max_err = (const double *)B.maxerr;
    l.out = (GrnModelOut *)B.out;
    l.fields = fields ? (double *)fields->p : nullptr;
    l.field_stride = field_stride;
    l.field_exp = field_exp;
    l.step_cap = P.step_cap;
  }
  hipStream_t st = c->stream;
  // With several kernel classes the odd ones run on the second stream, so
  // two persistent launches share the device and one class's drain overlaps
  // the other's work (fork after the clearing kernel, join before the end
  // event; the same pattern is captured into a graph).
  const bool fork = P.launches.size() > 1;
  auto issue = [&]() -> int {
    HIP_TRY(grn_launch_clear_results((GrnBatch *)B.batch, (GrnModelOut *)B.out, st));
    if (fork) {
      HIP_TRY(hipEventRecord(c->ev_fork, st));
      HIP_TRY(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
    }
    hipError_t e = hipSuccess;
    for (size_t gi = 0; gi < P.launches.size() && e == hipSuccess; ++gi) {
      const Plan::Launch &x = P.launches[gi];
      const GrnProgram *pp = L[gi].progs;
      const GrnDataset *dp = L[gi].datasets;
      const int *op = (const int *)B.order;
      void *args[] = {&L[gi], &pp, &dp, &op};
      e = hipLaunchKernel(x.kc.k->fn, dim3(x.grid), dim3(x.kc.block), args, x.kc.lds, gi % 2 ? c->stream2 : st);
    }
    if (fork) {  // joined also after a failed launch, so a wait on st covers both streams
      HIP_TRY(hipEventRecord(c->ev_join, c->stream2));
      HIP_TRY(hipStreamWaitEvent(st, c->ev_join, 0));
    }
    if (e != hipSuccess) return fail(GRN_E_DEVICE, "hipLaunchKernel failed: %s", hipGetErrorString(e));
    if (hooked) {  // the host stage, in stream order: results to pinned memory, then the host function
      HIP_TRY(hipMemcpyAsync(c->h_out, B.out, sizeof(GrnModelOut) * P.n, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipLaunchHostFunc(st, hook_host_fn, &c->hook));
    }
    return 0;
  };
  if (c->launch_mode == GRN_LAUNCH_GRAPH && allow_graph) {
    // the layout: everything a captured launch bakes in
    std::vector<long> key = {(long)B.batch, (long)B.out, (long)B.progs,
                             (long)(P.p64.empty() ? nullptr : B.progs64), (long)B.order,
                             (long)B.maxerr, (long)c->d_datasets.p, (long)B.scratch,
                             (long)(fields ? fields->p : nullptr), field_stride, field_exp, P.step_cap,
                             hooked ? (long)c->h_out : 0L, hooked ? (long)P.n : -1L};
    for (const Plan::Launch &x : P.launches) {
      key.insert(key.end(), {(long)x.kc.k->fn, x.grid, x.kc.block, (long)x.kc.lds, x.kc.slot_floats,
                             x.kc.tile.tw, x.kc.tile.th, x.kc.tile.sx, x.kc.tile.sy, x.kc.tile.halo});
    }
    auto it = c->graphs.find(key);
    if (it == c->graphs.end()) {
      if (c->graphs.size() >= 32) c->drop_graphs();
      HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      const int rc_issue = issue();
      hipGraph_t g = nullptr;
      const hipError_t e_end = hipStreamEndCapture(st, &g);
      if (rc_issue) {
        if (g) (void)hipGraphDestroy(g);
        return rc_issue;
      }
      if (e_end != hipSuccess) return fail(GRN_E_DEVICE, "hipStreamEndCapture failed: %s", hipGetErrorString(e_end));
      hipGraphExec_t ge = nullptr;
      const hipError_t e_inst = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e_inst != hipSuccess) return fail(GRN_E_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(e_inst));
      it = c->graphs.emplace(key, ge).first;
    } else {
      *replayed = true;
    }
    HIP_TRY(hipEventRecord(c->ev0, st));
    HIP_TRY(hipGraphLaunch(it->second, st));
    HIP_TRY(hipEventRecord(c->ev1, st));
  } else {
    HIP_TRY(hipEventRecord(c->ev0, st));
    const int rc_issue = issue();
    if (rc_issue) return rc_issue;
    HIP_TRY(hipEventRecord(c->ev1, st));
  }
  return 0;
}

// Waits for the batch, turns the kernels' outputs into results
// (CEvaluatorMorphogens::evaluate, cevaluatormorphogens.cpp:239-242) and
// records the context's statistics.
// Turns the kernels' outputs into results (CEvaluatorMorphogens::evaluate,
// cevaluatormorphogens.cpp:239-242): the maxError sentinel and ceilS.  Pure
// host code with no HIP call, so it also runs as a graph's host function;
// it reports failures through its return value and *err.
int convert_results(const Ctx *c, const Plan &P, const GrnModelOut *res, grn_result *out, int *n_products_out,
                    double *cell_steps_out, double *flops_out, std::string *err) {
  const int n = P.n;
  double cell_steps = 0, flops = 0;
  for (int i = 0; i < n; ++i) {
    const Item &it = P.items[i];
    const HostDataset &D = c->datasets[it.dataset];
    const GrnModelOut &o = res[i];
    if (!o.done) {
      *err = "model " + std::to_string(i) + ": the kernel wrote no result";
      return GRN_E_RESULT;
    }
    grn_result &r = out[i];
    r.raw_error = o.err;
    r.change = o.change;
    r.steps = o.steps;
    r.experiments = o.exps;
    r.n_products = it.prog.G;
    r.reserved = 0;
    if (o.err > it.max_err)
      r.fitness = 100;
    else if (D.params.error_precision > 0)
      r.fitness = grn_ceil_s(o.err, D.params.error_precision);
    else
      r.fitness = o.err;
    const double cs = (double)D.params.dim_x * D.params.dim_y * o.steps;
    cell_steps += cs;
    flops += cs * it.info.flops_per_cell_step;
    if (n_products_out) n_products_out[i] = it.prog.G;
  }
  *cell_steps_out = cell_steps;
  *flops_out = flops;
  return 0;
}

// The host stage of a hooked batch (see Ctx::hook).
void hook_host_fn(void *p) {
  Ctx::Hook *h = static_cast<Ctx::Hook *>(p);
  const Plan &P = *static_cast<const Plan *>(h->plan);
  h->err.clear();
  h->rc = convert_results(h->c, P, h->c->h_out, h->out, h->n_products, &h->cell_steps, &h->flops, &h->err);
  if (!h->rc && h->fn) h->rc = h->fn(h->user, h->out, P.n, &h->err);
}

// Waits for the batch and records the context's statistics.  Without a
// host stage the results are copied back and converted here; with one
// (hooked) the launches already copied and converted them.
int collect_plan(Ctx *c, const Plan &P, const DevPtrs &B, bool replayed, grn_result *out, int *n_products_out,
                 std::chrono::steady_clock::time_point t_start, GrnModelOut *host_res = nullptr,
                 bool hooked = false) {
  const int n = P.n;
  hipStream_t st = c->stream;
  std::vector<GrnModelOut> res_own;
  if (!hooked && !host_res) {
    res_own.resize(n);
    host_res = res_own.data();
  }
  if (!hooked && n > 0)
    HIP_TRY(hipMemcpyAsync(host_res, B.out, sizeof(GrnModelOut) * n, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  float ms = 0.f;
  if (n > 0) HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  double cell_steps = 0, flops = 0;
  if (hooked) {
    if (c->hook.rc) return fail(c->hook.rc, "%s", c->hook.err.c_str());
    cell_steps = c->hook.cell_steps;
    flops = c->hook.flops;
  } else {
    std::string err;
    const int rc = convert_results(c, P, host_res, out, n_products_out, &cell_steps, &flops, &err);
    if (rc) return fail(rc, "%s", err.c_str());
  }
  c->stats.kernel_ms = ms;
  c->stats.cell_steps = cell_steps;
  c->stats.alg_flops = flops;
  c->stats.n_launches = (int)P.launches.size();
  c->stats.n_models = n;
  c->stats.launch_mode = c->launch_mode;
  c->stats.graph_replayed = replayed ? 1 : 0;
  c->stats.host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  return 0;
}

// Compiles, uploads, launches and collects one batch on the context's own
// buffers.  fields (optional) receives each model's final field of
// experiment field_exp.
int run_batch(Ctx *c, int n, const int *dataset_ids, const grn_model *models,
              const double *max_errors, const double *dt_override, int path, grn_result *out,
              double *fields_host, long field_stride, int field_exp, int *n_products_out,
              const BatchOptions &opt = BatchOptions(), grn::BatchHook hook = nullptr, void *hook_user = nullptr) {
  const auto t_start = std::chrono::steady_clock::now();
  if (n < 0 || (n > 0 && (!models || !out))) return fail(GRN_E_INVALID, "invalid batch arguments");
  // one plan per host thread, reused: a batch's items hold 2 KB programs, and
  // fresh memory for them each batch costs page faults and zeroing
  static thread_local Plan plan_cache;
  Plan &P = plan_cache;
  int rc = make_plan(c, n, dataset_ids, models, max_errors, dt_override, path, opt, &P);
  if (rc) return rc;
  if (fields_host) {
    if ((rc = c->d_fields.reserve(sizeof(double) * (size_t)field_stride * std::max(n, 1)))) return rc;
    if (P.groups.size() > 1) return fail(GRN_E_INVALID, "field dumps take one kernel class per call");
  }
  const bool hooked = hook != nullptr && n > 0;
  if (hooked) {
    if (fields_host) return fail(GRN_E_INVALID, "a hooked batch does not dump fields");
    if ((size_t)n > c->h_out_cap) {
      HIP_TRY(hipStreamSynchronize(c->stream));  // no copy still writes the old buffer
      if (c->h_out) (void)hipHostFree(c->h_out);
      c->h_out = nullptr;
      c->h_out_cap = 0;
      c->drop_graphs();  // captured copies target the old buffer
      const size_t cap = std::max<size_t>((size_t)n + n / 4, 1024);
      HIP_TRY(hipHostMalloc((void **)&c->h_out, sizeof(GrnModelOut) * cap, hipHostMallocDefault));
      c->h_out_cap = cap;
    }
    c->hook.c = c;
    c->hook.plan = &P;
    c->hook.out = out;
    c->hook.n_products = n_products_out;
    c->hook.fn = hook;
    c->hook.user = hook_user;
    c->hook.rc = 0;
  }
  std::vector<double> maxerr;
  bool replayed = false;
  if ((rc = upload_plan(c, P, c->bufs, &maxerr)) ||
      (rc = launch_plan(c, P, ptrs_of(c->bufs), fields_host ? &c->d_fields : nullptr, field_stride, field_exp,
                        &replayed, true, hooked))) {
    (void)hipStreamSynchronize(c->stream);  // the uploads read host buffers of this frame
    return rc;
  }
  if (fields_host && n > 0) {
    // d_fields rows are indexed by queue position; return them in input order.
    std::vector<double> tmp((size_t)field_stride * n);
    HIP_TRY(hipMemcpyAsync(tmp.data(), c->d_fields.p, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    // The kernel writes a model's field at its queue position inside its
    // launch; map the rows back to input order.
    for (size_t gi = 0; gi < P.groups.size(); ++gi)
      for (int j = P.groups[gi].first; j < P.groups[gi].second; ++j) {
        const int row = j - P.groups[gi].first;
        std::memcpy(fields_host + (size_t)field_stride * P.idx[j], tmp.data() + (size_t)field_stride * row,
                    sizeof(double) * field_stride);
      }
  }
  return collect_plan(c, P, ptrs_of(c->bufs), replayed, out, n_products_out, t_start, nullptr, hooked);
}

// A batch compiled, planned and uploaded once (grn_stage_batch), launched by
// grn_run_staged with nothing crossing the bus but the results.
struct Staged {
  Ctx *c = nullptr;
  Plan P;
  BatchBufs B;
  std::vector<double> maxerr;  // a source of the uploads
};

// Live staged batches.  grn_destroy frees the ones its context still owns,
// and the staged-batch calls accept only handles found here, so a handle
// used after its context (or itself) was destroyed is rejected instead of
// reading freed memory.
std::mutex g_staged_mu;
std::set<Staged *> g_staged;

bool staged_live(Staged *S) {
  std::lock_guard<std::mutex> lk(g_staged_mu);
  return g_staged.count(S) != 0;
}

}  // namespace

namespace grn {

int evaluate_batch_hooked(void *ctx, int n, const int *dataset_ids, const grn_model *models,
                          const double *max_errors, const double *dt_override, int path, grn_result *out,
                          BatchHook hook, void *user) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(GRN_E_INVALID, "null context");
  if (path < GRN_PATH_AUTO || path > GRN_PATH_EXACT) return fail(GRN_E_INVALID, "unknown path %d", path);
  HIP_TRY(hipSetDevice(c->device));
  return run_batch(c, n, dataset_ids, models, max_errors, dt_override, path, out, nullptr, 0, 0, nullptr,
                   BatchOptions(), hook, user);
}

int set_last_error(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

}  // namespace grn

extern "C" {

int grn_abi_version(void) { return GRN_ABI_VERSION; }

const char *grn_last_error(void) { return g_last_error.c_str(); }

double grn_ceil_s(double n, int d) {
  // MathAlgo::ceilS (Src/Common/mathalgo.h:55-62)
  if (n == 0.0) return 0.0;
  const double factor = std::pow(10.0, d - std::ceil(std::log10(std::fabs(n))));
  return std::ceil(n * factor) / factor;
}

int grn_device_count(int *n) {
  if (!n) return fail(GRN_E_INVALID, "null output");
  int k = 0;
  HIP_TRY(hipGetDeviceCount(&k));
  *n = k;
  return 0;
}

int grn_create(int device, void **out) {
  if (!out) return fail(GRN_E_INVALID, "null output");
  *out = nullptr;
  int count = 0;
  HIP_TRY(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(GRN_E_INVALID, "device %d of %d", device, count);
  HIP_TRY(hipSetDevice(device));
  std::unique_ptr<Ctx> c(new Ctx);
  c->device = device;
  HIP_TRY(hipStreamCreate(&c->stream));
  HIP_TRY(hipStreamCreate(&c->stream2));
  HIP_TRY(hipEventCreate(&c->ev0));
  HIP_TRY(hipEventCreate(&c->ev1));
  HIP_TRY(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  HIP_TRY(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device));
  int lds = 0;
  HIP_TRY(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  // Static LDS of the kernels (queue slot, per-wave partials) stays below 1 KB.
  c->max_lds = lds > 1024 ? (size_t)lds - 1024 : 0;
  collect_kernels(&c->kernels);
  int rc = scratch_address(device, &c->noalloc);
  if (rc) return rc;
  c->na_progs.resize(noalloc::kMax);
  c->na_maxerr.resize(noalloc::kMax);
  c->na_res.resize(noalloc::kMax);
  *out = c.release();
  return 0;
}

void grn_destroy(void *ctx) {
  Ctx *c = as_ctx(ctx);
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  (void)hipStreamSynchronize(c->stream2);
  std::vector<Staged *> mine;
  {
    std::lock_guard<std::mutex> lk(g_staged_mu);
    for (auto it = g_staged.begin(); it != g_staged.end();) {
      if ((*it)->c == c) {
        mine.push_back(*it);
        it = g_staged.erase(it);
      } else {
        ++it;
      }
    }
  }
  for (Staged *S : mine) delete S;
  delete c;
}

int grn_add_dataset(void *ctx, const grn_sim_params *p, int n_exp, const double *init,
                    const double *target, int *dataset_id) {
  Ctx *c = as_ctx(ctx);
  if (!c || !p || !dataset_id) return fail(GRN_E_INVALID, "null argument");
  if (p->dim_x < 1 || p->dim_y < 1 || p->n_input_morphogens < 0 || p->n_target_morphogens < 0 ||
      n_exp < 1 || p->num_sim_steps < 0 || p->fitness_kind < 0 || p->fitness_kind > 1)
    return fail(GRN_E_INVALID, "invalid simulation parameters");
  if (p->n_target_morphogens == 0 && p->fitness_kind == GRN_FITNESS_MSE)
    return fail(GRN_E_INVALID, "MSE fitness needs at least one target plane");
  const size_t N = (size_t)p->dim_x * p->dim_y;
  const size_t n_init = (size_t)n_exp * p->n_input_morphogens * N;
  const size_t n_tgt = (size_t)n_exp * p->n_target_morphogens * N;
  if ((n_init && !init) || (n_tgt && !target)) return fail(GRN_E_INVALID, "missing pattern data");
  HIP_TRY(hipSetDevice(c->device));
  HostDataset D;
  D.params = *p;
  D.n_exp = n_exp;
  bool nonneg = p->min_conc >= 0;
  for (size_t i = 0; i < n_init && nonneg; ++i) nonneg = init[i] >= 0;
  D.nonneg = nonneg;
  HIP_TRY(hipMalloc(&D.d_init, std::max(n_init, (size_t)1) * sizeof(double)));
  HIP_TRY(hipMalloc(&D.d_target, std::max(n_tgt, (size_t)1) * sizeof(double)));
  if (n_init) HIP_TRY(hipMemcpyAsync(D.d_init, init, n_init * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (n_tgt) HIP_TRY(hipMemcpyAsync(D.d_target, target, n_tgt * sizeof(double), hipMemcpyHostToDevice, c->stream));
  c->datasets.push_back(D);

  std::vector<GrnDataset> dev(c->datasets.size());
  for (size_t i = 0; i < c->datasets.size(); ++i) {
    const HostDataset &h = c->datasets[i];
    GrnDataset &g = dev[i];
    std::memset(&g, 0, sizeof g);
    g.W = h.params.dim_x;
    g.H = h.params.dim_y;
    g.n_in = h.params.n_input_morphogens;
    g.n_tgt = h.params.n_target_morphogens;
    g.n_exp = h.n_exp;
    g.steps = h.params.num_sim_steps;
    g.kind = h.params.fitness_kind;
    g.nonneg = h.nonneg ? 1 : 0;
    g.min_conc = (float)h.params.min_conc;
    g.min_conc_change = (float)h.params.min_conc_change;
    g.min_conc_d = h.params.min_conc;
    g.min_conc_change_d = h.params.min_conc_change;
    g.thr = h.params.dist_error_threshold;
    g.init = (const double *)h.d_init;
    g.target = (const double *)h.d_target;
  }
  int rc = c->d_datasets.reserve(sizeof(GrnDataset) * dev.size());
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(c->d_datasets.p, dev.data(), sizeof(GrnDataset) * dev.size(), hipMemcpyHostToDevice,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  *dataset_id = (int)c->datasets.size() - 1;
  return 0;
}

int grn_decode_rgb(int width, int height, const unsigned int *rgb, int n_planes, double *out) {
  if (width < 1 || height < 1 || n_planes < 0 || n_planes > 3 || (n_planes && (!rgb || !out)))
    return fail(GRN_E_INVALID, "invalid image arguments (%dx%d, %d planes)", width, height, n_planes);
  // SimState::loadMorphologyImage (simstate.cpp:71-90): Qt's y axis grows
  // down, so scan line j is row height - 1 - j of the planes.
  const size_t plane = (size_t)width * height;
  for (int j = 0; j < height; ++j) {
    const unsigned int *line = rgb + (size_t)j * width;
    double *row = out + (size_t)(height - 1 - j) * width;
    for (int k = 0; k < n_planes; ++k) {
      const int shift = 16 - 8 * k;
      double *dst = row + k * plane;
      for (int i = 0; i < width; ++i) dst[i] = (double)((line[i] >> shift) & 0xffu) / 255.0;
    }
  }
  return 0;
}

int grn_add_dataset_rgb(void *ctx, const grn_sim_params *p, int n_exp, const unsigned int *init_images,
                        const unsigned int *target_images, int *dataset_id) {
  if (!ctx || !p || !dataset_id || n_exp < 1) return fail(GRN_E_INVALID, "invalid arguments");
  if (p->n_input_morphogens < 0 || p->n_input_morphogens > 3 || p->n_target_morphogens < 0 ||
      p->n_target_morphogens > 3)
    return fail(GRN_E_INVALID, "an RGB image holds at most 3 morphogen planes");
  if (p->dim_x < 1 || p->dim_y < 1) return fail(GRN_E_INVALID, "invalid image size");
  const size_t N = (size_t)p->dim_x * p->dim_y;
  const int n_in = p->n_input_morphogens, n_tgt = p->n_target_morphogens;
  if ((n_in && !init_images) || (n_tgt && !target_images)) return fail(GRN_E_INVALID, "missing images");
  std::vector<double> init((size_t)n_exp * n_in * N), target((size_t)n_exp * n_tgt * N);
  for (int e = 0; e < n_exp; ++e) {
    int rc;
    if (n_in && (rc = grn_decode_rgb(p->dim_x, p->dim_y, init_images + e * N, n_in, init.data() + e * n_in * N)))
      return rc;
    if (n_tgt &&
        (rc = grn_decode_rgb(p->dim_x, p->dim_y, target_images + e * N, n_tgt, target.data() + e * n_tgt * N)))
      return rc;
  }
  return grn_add_dataset(ctx, p, n_exp, init.data(), target.data(), dataset_id);
}

int grn_evaluate(void *ctx, int dataset_id, const grn_model *model, double max_error, grn_result *out) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(GRN_E_INVALID, "null context");
  HIP_TRY(hipSetDevice(c->device));
  return run_batch(c, 1, &dataset_id, model, &max_error, nullptr, GRN_PATH_AUTO, out, nullptr, 0, 0, nullptr);
}

int grn_evaluate_batch(void *ctx, int n_models, const int *dataset_ids, const grn_model *models,
                       const double *max_errors, const double *dt_override, int path, grn_result *out) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(GRN_E_INVALID, "null context");
  if (path < GRN_PATH_AUTO || path > GRN_PATH_EXACT) return fail(GRN_E_INVALID, "unknown path %d", path);
  HIP_TRY(hipSetDevice(c->device));
  return run_batch(c, n_models, dataset_ids, models, max_errors, dt_override, path, out, nullptr, 0, 0, nullptr);
}

int grn_evaluate_batch_noalloc(void *ctx, int n_models, const int *dataset_ids, const grn_model *models,
                               const double *max_errors, const double *dt_override, grn_result *out) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(GRN_E_INVALID, "null context");
  if (n_models < 0 || (n_models > 0 && (!models || !out))) return fail(GRN_E_INVALID, "invalid batch arguments");
  if (n_models > noalloc::kMax)
    return fail(GRN_E_CAPACITY, "grn_evaluate_batch_noalloc takes at most %d models, got %d", noalloc::kMax, n_models);
  HIP_TRY(hipSetDevice(c->device));
  const auto t_start = std::chrono::steady_clock::now();
  static thread_local Plan plan;  // host containers keep their capacity from call to call
  int rc = make_plan(c, n_models, dataset_ids, models, max_errors, dt_override, GRN_PATH_LDS, BatchOptions(), &plan);
  if (rc) return rc;
  if (plan.scratch_bytes || !plan.p64.empty())
    return fail(GRN_E_CAPACITY, "a model needs device memory beyond the fixed scratch array");
  unsigned char *base = static_cast<unsigned char *>(c->noalloc);
  const DevPtrs D{base + noalloc::kProgs, nullptr,       base + noalloc::kOrder, base + noalloc::kMaxErr,
                  base + noalloc::kOut,   base + noalloc::kBatch, nullptr};
  const int n = n_models;
  for (int i = 0; i < n; ++i) {
    c->na_progs[i] = plan.items[i].prog;
    c->na_maxerr[i] = plan.items[i].max_err;
  }
  hipStream_t st = c->stream;
  if (n > 0) {
    HIP_TRY(hipMemcpyAsync(D.progs, c->na_progs.data(), sizeof(GrnProgram) * n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(D.order, plan.idx.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(D.maxerr, c->na_maxerr.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
  }
  HIP_TRY(hipMemcpyAsync(D.batch, &plan.batch, sizeof(GrnBatch), hipMemcpyHostToDevice, st));
  bool replayed = false;
  if ((rc = launch_plan(c, plan, D, nullptr, 0, 0, &replayed, false))) {
    (void)hipStreamSynchronize(st);  // the uploads read host arrays
    return rc;
  }
  return collect_plan(c, plan, D, replayed, out, nullptr, t_start, c->na_res.data());
}

int grn_stage_batch(void *ctx, int n_models, const int *dataset_ids, const grn_model *models,
                    const double *max_errors, const double *dt_override, int path, void **staged) {
  Ctx *c = as_ctx(ctx);
  if (!c || !staged) return fail(GRN_E_INVALID, "null argument");
  *staged = nullptr;
  if (path < GRN_PATH_AUTO || path > GRN_PATH_EXACT) return fail(GRN_E_INVALID, "unknown path %d", path);
  if (n_models < 1 || !models) return fail(GRN_E_INVALID, "a staged batch needs at least one model");
  HIP_TRY(hipSetDevice(c->device));
  std::unique_ptr<Staged> S(new Staged);
  S->c = c;
  int rc = make_plan(c, n_models, dataset_ids, models, max_errors, dt_override, path, BatchOptions(), &S->P);
  if (!rc) rc = upload_plan(c, S->P, S->B, &S->maxerr);
  const hipError_t e = hipStreamSynchronize(c->stream);  // the uploads read S's and the context's host arrays
  if (rc) return rc;
  if (e != hipSuccess) return fail(GRN_E_DEVICE, "hipStreamSynchronize failed: %s", hipGetErrorString(e));
  {
    std::lock_guard<std::mutex> lk(g_staged_mu);
    g_staged.insert(S.get());
  }
  *staged = S.release();
  return 0;
}

int grn_run_staged(void *ctx, void *staged, grn_result *out) {
  Ctx *c = as_ctx(ctx);
  Staged *S = static_cast<Staged *>(staged);
  if (!c || !S || !out) return fail(GRN_E_INVALID, "null argument");
  if (!staged_live(S)) return fail(GRN_E_INVALID, "the staged batch was destroyed");
  if (S->c != c) return fail(GRN_E_INVALID, "the batch was staged on another context");
  HIP_TRY(hipSetDevice(c->device));
  const auto t_start = std::chrono::steady_clock::now();
  bool replayed = false;
  int rc = launch_plan(c, S->P, ptrs_of(S->B), nullptr, 0, 0, &replayed);
  if (rc) {
    (void)hipStreamSynchronize(c->stream);
    return rc;
  }
  return collect_plan(c, S->P, ptrs_of(S->B), replayed, out, nullptr, t_start);
}

void grn_staged_destroy(void *staged) {
  Staged *S = static_cast<Staged *>(staged);
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(g_staged_mu);
    if (!g_staged.erase(S)) return;  // already freed (by itself or by grn_destroy of its context)
  }
  (void)hipSetDevice(S->c->device);
  (void)hipStreamSynchronize(S->c->stream);
  delete S;
}

int grn_simulate_fields(void *ctx, int dataset_id, const grn_model *model, double dt_override, int experiment,
                        int path, double *fields, long capacity, int *n_products) {
  Ctx *c = as_ctx(ctx);
  if (!c || !fields) return fail(GRN_E_INVALID, "null argument");
  if (path < GRN_PATH_AUTO || path > GRN_PATH_EXACT) return fail(GRN_E_INVALID, "unknown path %d", path);
  if (dataset_id < 0 || dataset_id >= (int)c->datasets.size()) return fail(GRN_E_INVALID, "unknown dataset");
  const grn_sim_params &p = c->datasets[dataset_id].params;
  if (experiment < 0 || experiment >= c->datasets[dataset_id].n_exp) return fail(GRN_E_INVALID, "bad experiment");
  const long stride = (long)GRN_PMAX_GENES * p.dim_x * p.dim_y;
  HIP_TRY(hipSetDevice(c->device));
  std::vector<double> buf(stride);
  grn_result r;
  int g = 0;
  const double dt = dt_override;
  const double inf = std::numeric_limits<double>::infinity();
  int rc = run_batch(c, 1, &dataset_id, model, &inf, &dt, path, &r, buf.data(), stride, experiment, &g);
  if (rc) return rc;
  const long need = (long)g * p.dim_x * p.dim_y;
  if (capacity < need) return fail(GRN_E_INVALID, "fields buffer holds %ld of %ld values", capacity, need);
  std::memcpy(fields, buf.data(), sizeof(double) * need);
  if (n_products) *n_products = g;
  return 0;
}

int grn_test_performance(void *ctx, int dataset_id, const char *model_text, int n, double expected,
                         double *max_dif, double *ms) {
  Ctx *c = as_ctx(ctx);
  if (!c || !model_text || n < 1) return fail(GRN_E_INVALID, "invalid arguments");
  grn::Model m;
  std::string err;
  if (!grn::Model::parse(model_text, &m, &err)) return fail(GRN_E_MODEL, "%s", err.c_str());
  std::vector<grn_gene> genes;
  std::vector<grn_link> links;
  for (const grn::Gene &g : m.genes) genes.push_back(grn_gene{g.label, g.init, g.lim, g.deg, g.dif});
  for (const grn::Link &l : m.links)
    links.push_back(grn_link{l.regulator, l.regulated, l.dis_const, l.hill_coef, l.is_and ? 1 : 0});
  const grn_model gm{(int)genes.size(), genes.data(), (int)links.size(), links.data()};
  std::vector<grn_model> models(n, gm);
  std::vector<int> ds(n, dataset_id);
  std::vector<double> me(n, 1000.0);  // Individual without parents (individual.cpp:23-24)
  std::vector<grn_result> res(n);
  HIP_TRY(hipSetDevice(c->device));
  int rc = run_batch(c, n, ds.data(), models.data(), me.data(), nullptr, GRN_PATH_AUTO, res.data(), nullptr, 0, 0,
                     nullptr);
  if (rc) return rc;
  double worst = 0;
  for (const grn_result &r : res) worst = std::max(worst, std::fabs(r.fitness - expected));
  if (max_dif) *max_dif = worst;
  if (ms) *ms = c->stats.kernel_ms;
  if (worst != 0)
    return fail(GRN_E_WRONG, "WRONG ERROR: largest difference %.17g from the expected %.17g", worst, expected);
  return 0;
}

int grn_set_launch_mode(void *ctx, int mode) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(GRN_E_INVALID, "null context");
  if (mode != GRN_LAUNCH_EAGER && mode != GRN_LAUNCH_GRAPH) return fail(GRN_E_INVALID, "unknown launch mode %d", mode);
  c->launch_mode = mode;
  return 0;
}

int grn_autotune(void *ctx, int n, const int *dataset_ids, const grn_model *models, int sample_models,
                 int step_cap, grn_tune_class *report, int cap, int *n_classes) {
  Ctx *c = as_ctx(ctx);
  if (!c || n < 0 || (n && !models) || sample_models < 1 || step_cap < 0 || cap < 0 || (cap && !report))
    return fail(GRN_E_INVALID, "invalid autotune arguments");
  HIP_TRY(hipSetDevice(c->device));
  std::map<TuneKey, std::vector<int>> classes;
  for (int i = 0; i < n; ++i) {
    const int ds = dataset_ids ? dataset_ids[i] : 0;
    if (ds < 0 || ds >= (int)c->datasets.size()) return fail(GRN_E_INVALID, "model %d: unknown dataset %d", i, ds);
    const HostDataset &D = c->datasets[ds];
    grn::Model m;
    int rc = to_model(&models[i], &m);
    if (rc) return rc;
    GrnProgram prog;
    grn::ProgramInfo info;
    std::string err;
    const int n_morph = D.params.n_input_morphogens + D.params.n_target_morphogens;
    rc = grn::compile_model(m, n_morph, ds, (float)D.params.dt, &prog, &info, &err);
    if (rc) return fail(rc, "model %d: %s", i, err.c_str());
    classes[TuneKey{D.params.dim_x, D.params.dim_y, prog.G, D.params.n_target_morphogens, !D.nonneg}].push_back(i);
  }
  const int mode = c->launch_mode;
  c->launch_mode = GRN_LAUNCH_EAGER;  // timing runs are one-off layouts
  int k = 0;
  for (const auto &cls : classes) {
    const TuneKey &key = cls.first;
    const std::vector<int> &idx = cls.second;
    std::vector<KernelChoice> cand = kernel_candidates(*c, key.W, key.H, key.G, key.n_tgt, key.guard, GRN_PATH_AUTO);
    if ((int)cand.size() > GRN_TUNE_MAX_CANDIDATES) cand.resize(GRN_TUNE_MAX_CANDIDATES);
    const int ns = std::min((int)idx.size(), sample_models);
    std::vector<grn_model> sm(ns);
    std::vector<int> sd(ns);
    for (int j = 0; j < ns; ++j) {
      sm[j] = models[idx[j]];
      sd[j] = dataset_ids ? dataset_ids[idx[j]] : 0;
    }
    std::vector<grn_result> res(ns);
    grn_tune_class rep{};
    rep.dim_x = key.W;
    rep.dim_y = key.H;
    rep.genes = key.G;
    rep.n_target_morphogens = key.n_tgt;
    rep.guard = key.guard;
    rep.n_models = (int)idx.size();
    rep.n_timed = ns;
    rep.n_candidates = (int)cand.size();
    int best = -1;
    for (size_t j = 0; j < cand.size(); ++j) {
      BatchOptions o;
      o.force = &cand[j];
      o.step_cap = step_cap;
      grn_tune_candidate &tc = rep.cand[j];
      const KernelChoice &kc = cand[j];
      tc.kind = kc.k->wc == -1 ? GRN_TUNE_TILED : (kc.k->wc == 0 ? GRN_TUNE_ONCHIP_ANY : GRN_TUNE_ONCHIP_EXACT);
      tc.cells_per_thread = kc.k->cpt;
      tc.block = kc.block;
      tc.genes_lo = kc.k->glo;
      tc.genes_hi = kc.k->ghi;
      tc.tile_w = kc.tile.tw;
      tc.tile_h = kc.tile.th;
      tc.tile_sx = kc.tile.sx;
      tc.tile_sy = kc.tile.sy;
      tc.halo = kc.tile.halo;
      // one untimed run (code load, buffer growth), then the timed one
      int rc = run_batch(c, ns, sd.data(), sm.data(), nullptr, nullptr, GRN_PATH_AUTO, res.data(), nullptr, 0, 0,
                         nullptr, o);
      if (!rc) rc = run_batch(c, ns, sd.data(), sm.data(), nullptr, nullptr, GRN_PATH_AUTO, res.data(), nullptr, 0,
                              0, nullptr, o);
      if (rc) {
        if (rc == GRN_E_DEVICE) {
          c->launch_mode = mode;
          return rc;
        }
        tc.ms = -1;
        continue;
      }
      tc.ms = c->stats.kernel_ms;
      tc.cell_steps_per_s = c->stats.cell_steps / (tc.ms * 1e-3);
      if (best < 0 || tc.ms < rep.cand[best].ms) best = (int)j;
    }
    c->launch_mode = mode;
    if (best < 0) return fail(GRN_E_CAPACITY, "no candidate ran the %dx%d class with %d genes", key.W, key.H, key.G);
    c->launch_mode = GRN_LAUNCH_EAGER;
    rep.chosen = best;
    c->tuned[key] = cand[best];
    if (k < cap) report[k] = rep;
    ++k;
  }
  c->launch_mode = mode;
  if (n_classes) *n_classes = k;
  return 0;
}

int grn_apply_tuning(void *ctx, const grn_tune_class *classes, int n) {
  Ctx *c = as_ctx(ctx);
  if (!c || n < 0 || (n && !classes)) return fail(GRN_E_INVALID, "invalid arguments");
  std::vector<std::pair<TuneKey, KernelChoice>> picks;
  for (int i = 0; i < n; ++i) {
    const grn_tune_class &r = classes[i];
    if (r.chosen < 0 || r.chosen >= r.n_candidates || r.chosen >= GRN_TUNE_MAX_CANDIDATES)
      return fail(GRN_E_INVALID, "class %d: chosen candidate out of range", i);
    const grn_tune_candidate &want = r.cand[r.chosen];
    const TuneKey key{r.dim_x, r.dim_y, r.genes, r.n_target_morphogens, r.guard != 0};
    bool found = false;
    for (const KernelChoice &kc : kernel_candidates(*c, key.W, key.H, key.G, key.n_tgt, key.guard, GRN_PATH_AUTO)) {
      const int kind = kc.k->wc == -1 ? GRN_TUNE_TILED : (kc.k->wc == 0 ? GRN_TUNE_ONCHIP_ANY : GRN_TUNE_ONCHIP_EXACT);
      if (kind == want.kind && kc.k->cpt == want.cells_per_thread && kc.block == want.block &&
          kc.k->glo == want.genes_lo && kc.k->ghi == want.genes_hi && kc.tile.tw == want.tile_w &&
          kc.tile.th == want.tile_h && kc.tile.sx == want.tile_sx && kc.tile.sy == want.tile_sy &&
          kc.tile.halo == want.halo) {
        picks.push_back({key, kc});
        found = true;
        break;
      }
    }
    if (!found)
      return fail(GRN_E_INVALID, "class %d (%dx%d, %d genes): its chosen configuration is not available", i, r.dim_x,
                  r.dim_y, r.genes);
  }
  for (auto &p : picks) c->tuned[p.first] = p.second;
  return 0;
}

int grn_clear_tuning(void *ctx) {
  Ctx *c = as_ctx(ctx);
  if (!c) return fail(GRN_E_INVALID, "null context");
  c->tuned.clear();
  return 0;
}

int grn_get_stats(void *ctx, grn_stats *out) {
  Ctx *c = as_ctx(ctx);
  if (!c || !out) return fail(GRN_E_INVALID, "null argument");
  *out = c->stats;
  return 0;
}

int grn_model_parse(const char *text, grn_gene *genes, int gene_cap, int *n_genes, grn_link *links,
                    int link_cap, int *n_links) {
  if (!text || !n_genes || !n_links) return fail(GRN_E_INVALID, "null argument");
  grn::Model m;
  std::string err;
  if (!grn::Model::parse(text, &m, &err)) return fail(GRN_E_MODEL, "%s", err.c_str());
  *n_genes = (int)m.genes.size();
  *n_links = (int)m.links.size();
  if ((int)m.genes.size() > gene_cap || (int)m.links.size() > link_cap)
    return fail(GRN_E_CAPACITY, "model has %zu genes and %zu links", m.genes.size(), m.links.size());
  for (size_t i = 0; i < m.genes.size(); ++i) {
    const grn::Gene &g = m.genes[i];
    genes[i] = grn_gene{g.label, g.init, g.lim, g.deg, g.dif};
  }
  for (size_t i = 0; i < m.links.size(); ++i) {
    const grn::Link &l = m.links[i];
    links[i] = grn_link{l.regulator, l.regulated, l.dis_const, l.hill_coef, l.is_and ? 1 : 0};
  }
  return 0;
}

int grn_model_format(const grn_model *model, int precision, char *buf, long cap, long *len) {
  grn::Model m;
  int rc = to_model(model, &m);
  if (rc) return rc;
  const std::string s = m.to_string(precision > 0 ? precision : 6);
  if (len) *len = (long)s.size();
  if (!buf || cap < (long)s.size() + 1) return fail(GRN_E_CAPACITY, "buffer holds %ld of %zu bytes", cap, s.size() + 1);
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return 0;
}

int grn_model_info(const grn_model *model, int n_morphogens, int *genes_in_use, int *hill_links,
                   double *flops_per_cell_step) {
  grn::Model m;
  int rc = to_model(model, &m);
  if (rc) return rc;
  GrnProgram prog;
  grn::ProgramInfo info;
  std::string err;
  rc = grn::compile_model(m, n_morphogens, 0, 0.f, &prog, &info, &err);
  if (rc) return fail(rc, "%s", err.c_str());
  if (genes_in_use) *genes_in_use = info.G;
  if (hill_links) *hill_links = info.L;
  if (flops_per_cell_step) *flops_per_cell_step = info.flops_per_cell_step;
  return 0;
}

}  // extern "C"
