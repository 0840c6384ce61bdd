// C ABI of batched assembly (hfx_batch_*), declared in include/hfx_hip.h.
//
// Many independent 2-D HDG problems of mixed order are assembled and
// condensed by one persistent launch per order present (at most five),
// running concurrently on side streams forked from and joined back to the
// batch stream, followed by each problem's trace scatter; the whole sequence
// is captured once as a HIP graph and replayed.  Each order's cells are cut
// into items of `tile` cells, sorted largest first and dealt round-robin to
// the workgroups; (waves per workgroup, tile) per order come from an
// autotuner that times the candidate launches on the problems themselves.
// (A single kernel switching on the order was measured 1.3x slower than
// launching the problems one by one: it carries the register budget of the
// highest order for every cell.)
// The reference assembles one problem at a time on the host
// (HDGSolver::assemble, src/solver/HDGSolver.cpp:166-675).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "capi_internal.h"

namespace hfx {
void set_last_error(const std::string& msg);
}

struct hfx_batch {
  static constexpr int NORD = 6;  // orders 1..5
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side[NORD] = {};       // one stream per order: the order kernels run concurrently
  hipEvent_t fork = nullptr, join[NORD] = {};
  std::vector<hfx_hdg*> probs;
  int waves[NORD] = {0, 2, 2, 2, 2, 2}, tile[NORD] = {0, 64, 64, 64, 64, 64};
  bool items_dirty = true;
  std::vector<bool> fused;  // per problem: fused scatter in the argument table
  std::vector<unsigned long long> versions;  // per problem: args_version of the table
  std::vector<hfx::BatchItem> items[NORD];
  hfx::DevBuf<hfx::ElemArgs> dargs;
  hfx::DevBuf<hfx::BatchItem> ditems[NORD];
  int grid[NORD] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool concurrent = true;        // order kernels on side streams (chosen by the tuner)
  bool graph_concurrent = true;  // the same inside the captured graph (timed separately)

  void set_device() const { HFX_HIP(hipSetDevice(device)); }

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }

  bool has(int p) const { return !items[p].empty(); }

  // relative cost of one cell of order p: the condensation's dense work
  // grows like the cube of the local system size
  static double cell_cost(int p) {
    const int nN = (p + 1) * (p + 2) / 2, n = 3 * nN + 3 * (p + 1);
    return (double)n * n * n;
  }

  void build_order(int p) {
    items[p].clear();
    for (size_t k = 0; k < probs.size(); ++k) {
      hfx_hdg* h = probs[k];
      if (h->order != p) continue;
      for (int e0 = 0; e0 < h->mesh.nCells; e0 += tile[p]) {
        hfx::BatchItem it;
        it.prob = (int)k;
        it.order = p;
        it.e0 = e0;
        it.e1 = std::min(h->mesh.nCells, e0 + tile[p]);
        items[p].push_back(it);
      }
    }
    // largest items first, dealt round-robin to the workgroups
    std::stable_sort(items[p].begin(), items[p].end(),
                     [](const hfx::BatchItem& x, const hfx::BatchItem& y) { return (x.e1 - x.e0) > (y.e1 - y.e0); });
    if (items[p].empty()) return;
    hfx::prepare_affine_batch(p, waves[p]);
    ditems[p].upload(items[p], stream);
    // persistent grid: as many workgroups as fit on the 256 CUs at once
    const size_t lds = hfx::affine_batch_lds_bytes(p, waves[p]);
    const int per_cu = std::max(1, (int)((160 * 1024) / (lds + 1024)));
    grid[p] = std::min((int)items[p].size(), 256 * std::min(per_cu, 8));
  }

  void build_items() {
    for (size_t i = 0; i < probs.size() && !items_dirty; ++i)
      if (i >= versions.size() || versions[i] != probs[i]->args_version) items_dirty = true;
    if (!items_dirty) return;
    std::vector<hfx::ElemArgs> args;
    fused.clear();
    versions.clear();
    for (hfx_hdg* h : probs) {
      if (h->ts_scheme == HFX_TS_FEULER || h->ts_scheme == HFX_TS_RK)
        throw std::runtime_error("batched assembly supports implicit Euler only among the time schemes");
      args.push_back(h->affine_args(false));
      fused.push_back(h->fuse_scatter);
      versions.push_back(h->args_version);
    }
    dargs.upload(args, stream);
    for (int p = 1; p < NORD; ++p) build_order(p);
    HFX_HIP(hipStreamSynchronize(stream));
    items_dirty = false;
    drop_graph();
  }

  void launch_order(int p, hipStream_t s) {
    hfx::launch_affine_batch(p, waves[p], dargs.p, ditems[p].p, (int)items[p].size(), grid[p], s);
  }

  // every order's kernel, concurrently on the side streams (joined back) or
  // one after another on the batch stream
  void run_kernels() {
    build_items();
    if (!concurrent) {
      for (int p = 1; p < NORD; ++p)
        if (has(p)) launch_order(p, stream);
      return;
    }
    HFX_HIP(hipEventRecord(fork, stream));
    for (int p = 1; p < NORD; ++p) {
      if (!has(p)) continue;
      HFX_HIP(hipStreamWaitEvent(side[p], fork, 0));
      launch_order(p, side[p]);
      HFX_HIP(hipEventRecord(join[p], side[p]));
      HFX_HIP(hipStreamWaitEvent(stream, join[p], 0));
    }
  }

  void run_all() {
    run_kernels();
    for (size_t i = 0; i < probs.size(); ++i) probs[i]->scatter_on(stream, fused[i]);
  }

  void capture() {
    if (exec) return;
    build_items();
    const bool eager = concurrent;
    concurrent = graph_concurrent;
    HFX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
    try {
      run_all();
    } catch (...) {
      concurrent = eager;
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    concurrent = eager;
    HFX_HIP(hipStreamEndCapture(stream, &graph));
    HFX_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }

  void sync_problems() {
    for (hfx_hdg* h : probs) HFX_HIP(hipStreamSynchronize(h->stream));
  }

  void mark_assembled() {
    for (hfx_hdg* h : probs) {
      h->assembled = true;
      h->solved = false;
      h->last_fused = h->fuse_scatter;  // build_items ran with the current flags
    }
  }
};

#define B_TRY try {
#define B_END                            \
  }                                      \
  catch (const std::exception& e) {      \
    hfx::set_last_error(e.what());       \
    return -1;                           \
  }                                      \
  return 0;

namespace {
void need(bool c, const char* m) {
  if (!c) throw std::runtime_error(std::string("batched assembly: ") + m);
}
}  // namespace

extern "C" int hfx_batch_create(int device, hfx_batch** out) {
  B_TRY
  need(out != nullptr, "null output pointer");
  int nd = 0;
  need(hipGetDeviceCount(&nd) == hipSuccess && device >= 0 && device < nd, "no such GPU");
  auto b = std::unique_ptr<hfx_batch>(new hfx_batch());
  b->device = device;
  b->set_device();
  HFX_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  for (int p = 1; p < hfx_batch::NORD; ++p) {
    HFX_HIP(hipStreamCreateWithFlags(&b->side[p], hipStreamNonBlocking));
    HFX_HIP(hipEventCreateWithFlags(&b->join[p], hipEventDisableTiming));
  }
  HFX_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  *out = b.release();
  B_END
}

extern "C" void hfx_batch_free(hfx_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  b->drop_graph();
  for (int p = 1; p < hfx_batch::NORD; ++p) {
    if (b->side[p]) {
      (void)hipStreamSynchronize(b->side[p]);
      (void)hipStreamDestroy(b->side[p]);
    }
    if (b->join[p]) (void)hipEventDestroy(b->join[p]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  (void)hipStreamDestroy(b->stream);
  delete b;
}

extern "C" int hfx_batch_add(hfx_batch* b, hfx_hdg* h) {
  B_TRY
  need(b && h, "null argument");
  need(h->device == b->device, "the problem lives on another GPU");
  need(h->has_mesh && h->has_fields, "mesh and fields must be set before the problem joins a batch");
  need(h->dim == 2 && h->order >= 1 && h->order <= 5, "the batched kernel takes 2-D problems of order 1..5");
  need(h->affine && !(h->hasD && h->ndD == 4), "the batched kernel needs straight-sided cells and scalar diffusion");
  need(std::find(b->probs.begin(), b->probs.end(), h) == b->probs.end(), "problem already in the batch");
  b->probs.push_back(h);
  b->items_dirty = true;
  b->drop_graph();
  B_END
}

extern "C" int hfx_batch_size(hfx_batch* b, int* n, long long* cells) {
  B_TRY
  need(b != nullptr, "null handle");
  if (n) *n = (int)b->probs.size();
  if (cells) {
    long long c = 0;
    for (hfx_hdg* h : b->probs) c += h->mesh.nCells;
    *cells = c;
  }
  B_END
}

extern "C" int hfx_batch_set_config(hfx_batch* b, int order, int waves, int tile) {
  B_TRY
  need(b != nullptr, "null handle");
  need(order >= 0 && order <= 5, "order must be 1..5, or 0 for all orders");
  need(waves == 1 || waves == 2 || waves == 4, "waves per workgroup must be 1, 2 or 4");
  need(tile >= 1 && tile <= (1 << 20), "tile must be in [1, 2^20] cells");
  for (int p = 1; p < hfx_batch::NORD; ++p) {
    if (order != 0 && p != order) continue;
    need(hfx::affine_batch_lds_bytes(p, waves) <= 160 * 1024, "this wave count needs more than 160 KB of LDS");
    b->waves[p] = waves;
    b->tile[p] = tile;
  }
  b->items_dirty = true;
  b->drop_graph();
  B_END
}

// out[4 * (p - 1) + 0..3] = waves, tile, work items, workgroups of order p (1..5);
// out[20] bit 0: the orders run concurrently, bit 1: they do in the graph
extern "C" int hfx_batch_get_config(hfx_batch* b, int* out) {
  B_TRY
  need(b && out, "null argument");
  b->set_device();
  b->build_items();
  for (int p = 1; p < hfx_batch::NORD; ++p) {
    out[4 * (p - 1) + 0] = b->waves[p];
    out[4 * (p - 1) + 1] = b->tile[p];
    out[4 * (p - 1) + 2] = (int)b->items[p].size();
    out[4 * (p - 1) + 3] = b->has(p) ? b->grid[p] : 0;
  }
  out[20] = (b->concurrent ? 1 : 0) | (b->graph_concurrent ? 2 : 0);
  B_END
}

// Time every (waves, tile) candidate of each order's kernel on this batch
// (alone on the batch stream) and keep the fastest per order, then choose
// between running the orders concurrently or one after another, eagerly and
// (timed on its own: a graph's parallel branches schedule differently) in
// the replayed graph; best_ms returns the tuned eager launch of all orders.
extern "C" int hfx_batch_tune(hfx_batch* b, int reps, double* best_ms) {
  B_TRY
  need(b != nullptr && !b->probs.empty(), "empty batch");
  need(reps >= 1 && reps <= 100, "reps must be in [1, 100]");
  b->set_device();
  b->sync_problems();
  b->build_items();
  hipEvent_t e0, e1;
  HFX_HIP(hipEventCreate(&e0));
  HFX_HIP(hipEventCreate(&e1));
  auto timed = [&](auto&& fn) {
    fn();  // warm-up
    HFX_HIP(hipEventRecord(e0, b->stream));
    for (int r = 0; r < reps; ++r) fn();
    HFX_HIP(hipEventRecord(e1, b->stream));
    HFX_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    HFX_HIP(hipEventElapsedTime(&ms, e0, e1));
    return (double)ms / reps;
  };
  for (int p = 1; p < hfx_batch::NORD; ++p) {
    if (!b->has(p)) continue;
    double best = 1e300;
    int bw = b->waves[p], bt = b->tile[p];
    for (int w : {1, 2, 4}) {
      if (hfx::affine_batch_lds_bytes(p, w) > 160 * 1024) continue;
      for (int t : {8, 16, 32, 64, 128, 256}) {
        if (t < w) continue;
        b->waves[p] = w;
        b->tile[p] = t;
        b->build_order(p);
        HFX_HIP(hipStreamSynchronize(b->stream));
        const double m = timed([&] { b->launch_order(p, b->stream); });
        if (m < best) {
          best = m;
          bw = w;
          bt = t;
        }
      }
    }
    b->waves[p] = bw;
    b->tile[p] = bt;
    b->build_order(p);
    HFX_HIP(hipStreamSynchronize(b->stream));
  }
  b->drop_graph();
  b->concurrent = false;
  const double serial = timed([&] { b->run_kernels(); });
  b->concurrent = true;
  const double conc = timed([&] { b->run_kernels(); });
  b->concurrent = conc <= serial;
  const double all = std::min(conc, serial);
  double gms[2];
  for (int c = 0; c < 2; ++c) {
    b->drop_graph();
    b->graph_concurrent = c == 1;
    b->capture();
    gms[c] = timed([&] { HFX_HIP(hipGraphLaunch(b->exec, b->stream)); });
  }
  b->drop_graph();
  b->graph_concurrent = gms[1] <= gms[0];
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (best_ms) *best_ms = all;
  B_END
}

extern "C" int hfx_batch_assemble(hfx_batch* b, int use_graph) {
  B_TRY
  need(b != nullptr && !b->probs.empty(), "empty batch");
  b->set_device();
  b->sync_problems();
  b->build_items();
  if (use_graph) {
    b->capture();
    HFX_HIP(hipGraphLaunch(b->exec, b->stream));
  } else {
    b->run_all();
  }
  HFX_HIP(hipStreamSynchronize(b->stream));
  b->mark_assembled();
  B_END
}

// what: 0 batched element kernel only, 1 batched kernel + scatters (eager),
// 2 the same as a replayed HIP graph, 3 the problems one after another with
// their own per-problem kernels (element + scatter) on the batch stream.
extern "C" int hfx_batch_time(hfx_batch* b, int what, int warmup, int reps, double* ms) {
  B_TRY
  need(b != nullptr && !b->probs.empty(), "empty batch");
  need(what >= 0 && what <= 3 && warmup >= 0 && reps >= 1 && ms, "bad arguments");
  b->set_device();
  b->sync_problems();
  b->build_items();
  if (what == 2) b->capture();
  hipEvent_t e0, e1;
  HFX_HIP(hipEventCreate(&e0));
  HFX_HIP(hipEventCreate(&e1));
  for (int i = 0; i < warmup + reps; ++i) {
    HFX_HIP(hipEventRecord(e0, b->stream));
    if (what == 0) b->run_kernels();
    else if (what == 1) b->run_all();
    else if (what == 2) HFX_HIP(hipGraphLaunch(b->exec, b->stream));
    else {
      for (hfx_hdg* h : b->probs) {
        hfx::launch_affine(2, h->order, false, h->affine_waves(), h->affine_args(false), h->grid_cap, b->stream);
        h->scatter_on(b->stream, h->fuse_scatter);
      }
    }
    HFX_HIP(hipEventRecord(e1, b->stream));
    HFX_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    HFX_HIP(hipEventElapsedTime(&t, e0, e1));
    if (i >= warmup) ms[i - warmup] = t;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->mark_assembled();
  B_END
}
