// C ABI of batched assembly (hfx_batch_*), declared in include/hfx_hip.h.
//
// Many independent 2-D HDG problems of mixed order are assembled and
// condensed by ONE persistent launch of the batched affine-cell kernel,
// followed by each problem's trace scatter; the whole sequence can be
// captured once as a HIP graph and replayed.  The work list is cut into
// items of `tile` cells, sorted by decreasing estimated cost and dealt
// round-robin to the workgroups; (waves per workgroup, tile) come from an
// autotuner that times the candidate launches on the problems themselves.
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
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<hfx_hdg*> probs;
  int waves = 2, tile = 64;
  bool items_dirty = true;
  std::vector<hfx::BatchItem> items;
  hfx::DevBuf<hfx::ElemArgs> dargs;
  hfx::DevBuf<hfx::BatchItem> ditems;
  int grid = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<double> tune_ms;  // per candidate, from the last tune

  void set_device() const { HFX_HIP(hipSetDevice(device)); }

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }

  // cost of one cell of order p relative to p = 1: the condensation's dense
  // work grows like the cube of the local system size
  static double cell_cost(int p) {
    const int nN = (p + 1) * (p + 2) / 2, n = 3 * nN + 3 * (p + 1);
    return (double)n * n * n;
  }

  void build_items() {
    if (!items_dirty) return;
    std::vector<hfx::ElemArgs> args;
    items.clear();
    for (size_t k = 0; k < probs.size(); ++k) {
      hfx_hdg* h = probs[k];
      hfx::ElemArgs a = h->elem_args();
      a.tabs = h->tabs_aff.p;
      args.push_back(a);
      for (int e0 = 0; e0 < h->mesh.nCells; e0 += tile) {
        hfx::BatchItem it;
        it.prob = (int)k;
        it.order = h->order;
        it.e0 = e0;
        it.e1 = std::min(h->mesh.nCells, e0 + tile);
        items.push_back(it);
      }
    }
    std::stable_sort(items.begin(), items.end(), [](const hfx::BatchItem& x, const hfx::BatchItem& y) {
      return (x.e1 - x.e0) * cell_cost(x.order) > (y.e1 - y.e0) * cell_cost(y.order);
    });
    hfx::prepare_affine_batch(waves);
    dargs.upload(args, stream);
    ditems.upload(items, stream);
    HFX_HIP(hipStreamSynchronize(stream));
    // persistent grid: as many workgroups as fit on the 256 CUs at once
    const size_t lds = hfx::affine_batch_lds_bytes(waves);
    const int per_cu = std::max(1, (int)((160 * 1024) / (lds + 1024)));
    grid = std::min((int)items.size(), 256 * per_cu);
    items_dirty = false;
    drop_graph();
  }

  void run_kernel() {
    build_items();
    hfx::launch_affine_batch(waves, dargs.p, ditems.p, (int)items.size(), grid, stream);
  }

  void run_all() {
    run_kernel();
    for (hfx_hdg* h : probs)
      hfx::launch_scatter(h->A, h->mesh.nFpc, h->S.p, h->S0.p, h->invperm.p, h->is_dir.p, h->g.p, h->rhs.p, stream);
  }

  void sync_problems() {
    for (hfx_hdg* h : probs) HFX_HIP(hipStreamSynchronize(h->stream));
  }

  void mark_assembled() {
    for (hfx_hdg* h : probs) {
      h->assembled = true;
      h->solved = false;
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
  *out = b.release();
  B_END
}

extern "C" void hfx_batch_free(hfx_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  b->drop_graph();
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

extern "C" int hfx_batch_set_config(hfx_batch* b, int waves, int tile) {
  B_TRY
  need(b != nullptr, "null handle");
  need(waves == 1 || waves == 2 || waves == 4, "waves per workgroup must be 1, 2 or 4");
  need(tile >= 1 && tile <= (1 << 20), "tile must be in [1, 2^20] cells");
  need(hfx::affine_batch_lds_bytes(waves) <= 160 * 1024, "this wave count needs more than 160 KB of LDS");
  if (waves != b->waves || tile != b->tile) {
    b->waves = waves;
    b->tile = tile;
    b->items_dirty = true;
    b->drop_graph();
  }
  B_END
}

extern "C" int hfx_batch_get_config(hfx_batch* b, int* out) {
  B_TRY
  need(b && out, "null argument");
  b->set_device();
  b->build_items();
  out[0] = b->waves;
  out[1] = b->tile;
  out[2] = (int)b->items.size();
  out[3] = b->grid;
  B_END
}

// Time every (waves, tile) candidate on the batch itself and keep the fastest.
extern "C" int hfx_batch_tune(hfx_batch* b, int reps, double* best_ms) {
  B_TRY
  need(b != nullptr && !b->probs.empty(), "empty batch");
  need(reps >= 1 && reps <= 100, "reps must be in [1, 100]");
  b->set_device();
  b->sync_problems();
  hipEvent_t e0, e1;
  HFX_HIP(hipEventCreate(&e0));
  HFX_HIP(hipEventCreate(&e1));
  double best = 1e300;
  int bw = b->waves, bt = b->tile;
  b->tune_ms.clear();
  for (int w : {1, 2, 4}) {
    if (hfx::affine_batch_lds_bytes(w) > 160 * 1024) continue;
    for (int t : {8, 16, 32, 64, 128, 256}) {
      if (t < w) continue;
      b->waves = w;
      b->tile = t;
      b->items_dirty = true;
      b->run_kernel();  // warm-up (and item upload)
      HFX_HIP(hipEventRecord(e0, b->stream));
      for (int r = 0; r < reps; ++r) b->run_kernel();
      HFX_HIP(hipEventRecord(e1, b->stream));
      HFX_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      HFX_HIP(hipEventElapsedTime(&ms, e0, e1));
      const double m = ms / reps;
      b->tune_ms.push_back(m);
      if (m < best) {
        best = m;
        bw = w;
        bt = t;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->waves = bw;
  b->tile = bt;
  b->items_dirty = true;
  b->build_items();
  if (best_ms) *best_ms = best;
  B_END
}

extern "C" int hfx_batch_assemble(hfx_batch* b, int use_graph) {
  B_TRY
  need(b != nullptr && !b->probs.empty(), "empty batch");
  b->set_device();
  b->sync_problems();
  b->build_items();
  if (use_graph) {
    if (!b->exec) {
      HFX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
      try {
        b->run_all();
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(b->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HFX_HIP(hipStreamEndCapture(b->stream, &b->graph));
      HFX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
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
  if (what == 2 && !b->exec) {
    HFX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    b->run_all();
    HFX_HIP(hipStreamEndCapture(b->stream, &b->graph));
    HFX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  hipEvent_t e0, e1;
  HFX_HIP(hipEventCreate(&e0));
  HFX_HIP(hipEventCreate(&e1));
  for (int i = 0; i < warmup + reps; ++i) {
    HFX_HIP(hipEventRecord(e0, b->stream));
    if (what == 0) b->run_kernel();
    else if (what == 1) b->run_all();
    else if (what == 2) HFX_HIP(hipGraphLaunch(b->exec, b->stream));
    else {
      for (hfx_hdg* h : b->probs) {
        hfx::ElemArgs a = h->elem_args();
        a.tabs = h->tabs_aff.p;
        hfx::launch_affine(2, h->order, false, h->affine_waves(), a, h->grid_cap, b->stream);
        hfx::launch_scatter(h->A, h->mesh.nFpc, h->S.p, h->S0.p, h->invperm.p, h->is_dir.p, h->g.p, h->rhs.p,
                            b->stream);
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
