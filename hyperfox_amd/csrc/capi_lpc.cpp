// C ABI of the learned trace preconditioner (hfx_lpc_*, hfx_hdg_set_lpc),
// declared in include/hfx_hip.h.  No counterpart in the reference: it
// corrects the face-block Jacobi preconditioner of the trace solve
// (the reference's solve is PetscInterface::solve, PetscInterface.cpp:222-247).
//
// A handle holds the network parameters and Adam state on one GPU and a
// training set of trace systems.  Samples from the same mesh share one graph;
// meshes of different sizes may be mixed.  Each sample keeps its normalised
// FP32 blocks, the normalised solution as the first probe vector and its
// product with Ahat.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "capi_internal.h"

namespace hfx {
void set_last_error(const std::string& msg);
}

namespace {

struct Sample {
  std::shared_ptr<hfx::LpcGraph> g;
  hfx::LpcSystem sys;
  hfx::DevBuf<float> z, v;  // normalised solution probe and Ahat z
};

__global__ void solution_probe_kernel(long long n, int bs, const double* __restrict__ sinv,
                                      const double* __restrict__ lam, float* __restrict__ z) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) z[t] = (float)(lam[t] / sinv[t / bs]);  // S^1/2 lambda
}

__global__ void scale_f32_kernel(long long n, float a, float* __restrict__ z) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) z[t] *= a;
}

__global__ void norm2_f32_kernel(long long n, const float* __restrict__ z, double* __restrict__ out) {
  __shared__ double red[4];
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) s += (double)z[i] * z[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = red[0] + red[1] + red[2] + red[3];
}

struct LaErr : std::runtime_error {
  explicit LaErr(const std::string& m) : std::runtime_error("learned preconditioner: " + m) {}
};

}  // namespace

namespace hfx {
void fill_zero(float* p, long long n, hipStream_t s) {
  if (n > 0) HFX_HIP(hipMemsetAsync(p, 0, (size_t)n * sizeof(float), s));
}
}  // namespace hfx

struct hfx_lpc {
  int device = 0;
  hipStream_t stream = nullptr;
  hfx::LpcLayout L;
  hfx::DevBuf<float> params, grad, am, av;
  hfx::DevBuf<uint16_t> p16;
  bool p16_valid = false;
  int adam_t = 0;
  // graphs of the training set, keyed by the problem's mesh and pattern size
  struct GraphKey {
    unsigned long long mesh_serial;
    int nb, nnzb;
    std::shared_ptr<hfx::LpcGraph> g;
  };
  std::vector<GraphKey> graphs;
  std::vector<std::unique_ptr<Sample>> samples;
  hfx::LpcActs acts;
  hfx::LpcGradWork work;
  hfx::DevBuf<float> zr, vr;
  hfx::DevBuf<double> lossacc, lossp;
  int work_nb = -1, work_E = -1;  // sizes the work buffers are allocated for

  void set_device() const { HFX_HIP(hipSetDevice(device)); }

  // size the activation and gradient buffers for the largest graph
  void ensure_work() {
    int nb = 0, E = 0;
    for (const GraphKey& k : graphs) {
      nb = std::max(nb, k.g->nb);
      E = std::max(E, k.g->E);
    }
    if (nb == work_nb && E == work_E) return;
    acts.alloc(nb, E, L.O);
    work.alloc(nb, E, L.O, 192LL * 64);
    zr.alloc((size_t)nb * L.bs);
    vr.alloc((size_t)nb * L.bs);
    lossacc.alloc(1);
    lossp.alloc(hfx::ceil_div((long long)nb * L.bs, 256));
    work_nb = nb;
    work_E = E;
  }

  // forward + loss (+ gradient into dcn/dce) of sample k; loss accumulated in lossacc
  void forward_loss(int k, int n_random, unsigned long long seed) {
    const Sample& sm = *samples.at(k);
    const hfx::LpcGraph& g = *sm.g;
    hipStream_t s = stream;
    hfx::lpc_forward(L, params.p, g, sm.sys, acts, s);
    const long long nv = (long long)g.nb * L.bs;
    hfx::fill_zero(work.dcn.p, (long long)g.nb * L.O, s);
    hfx::fill_zero(work.dce.p, (long long)std::max(g.E, 1) * L.O, s);
    hfx::lpc_loss(g, sm.sys, acts, sm.z.p, sm.v.p, 1.0f, 1.0f, work.dcn.p, work.dce.p, lossp.p, lossacc.p, s);
    for (int q = 0; q < n_random; ++q) {
      hfx::lpc_rademacher(nv, seed * 1000003ull + (unsigned long long)q * 7919ull + 17ull, zr.p, s);
      hfx::lpc_spmv(g, sm.sys.ahat.p, zr.p, vr.p, s);
      hfx::lpc_loss(g, sm.sys, acts, zr.p, vr.p, 1.0f / (float)nv, 1.0f, work.dcn.p, work.dce.p, lossp.p, lossacc.p, s);
    }
  }
};

#define LPC_TRY try {
#define LPC_END                          \
  }                                      \
  catch (const std::exception& e) {      \
    hfx::set_last_error(e.what());       \
    return -1;                           \
  }                                      \
  return 0;

namespace {
void need(bool c, const char* m) {
  if (!c) throw LaErr(m);
}
}  // namespace

extern "C" int hfx_lpc_create(int device, int block_size, hfx_lpc** out) {
  LPC_TRY
  need(out != nullptr, "null output pointer");
  need(block_size == 2 || block_size == 3 || block_size == 4 || block_size == 5 || block_size == 6 || block_size == 10,
       "block size must be one of 2, 3, 4, 5, 6, 10 (face nodes of 2-D p=1..5 or 3-D p=1..3)");
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || device < 0 || device >= nd) throw LaErr("no such GPU");
  auto h = std::unique_ptr<hfx_lpc>(new hfx_lpc());
  h->device = device;
  h->L = hfx::LpcLayout(block_size);
  h->set_device();
  HFX_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  h->params.alloc(h->L.total);
  h->grad.alloc(h->L.total);
  h->am.alloc(h->L.total);
  h->av.alloc(h->L.total);
  h->p16.alloc(h->L.total);
  hfx::fill_zero(h->params.p, h->L.total, h->stream);
  hfx::fill_zero(h->am.p, h->L.total, h->stream);
  hfx::fill_zero(h->av.p, h->L.total, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  *out = h.release();
  LPC_END
}

extern "C" void hfx_lpc_free(hfx_lpc* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  (void)hipStreamSynchronize(h->stream);
  h->samples.clear();
  (void)hipStreamDestroy(h->stream);
  delete h;
}

extern "C" int hfx_lpc_num_params(hfx_lpc* h, long long* n) {
  LPC_TRY
  need(h && n, "null argument");
  *n = h->L.total;
  LPC_END
}

extern "C" int hfx_lpc_init_params(hfx_lpc* h, unsigned long long seed) {
  LPC_TRY
  need(h != nullptr, "null handle");
  const hfx::LpcLayout& L = h->L;
  std::vector<float> p(L.total, 0.f);
  unsigned long long st = seed ^ 0x243F6A8885A308D3ull;
  auto uni = [&st]() {
    st += 0x9E3779B97F4A7C15ull;
    unsigned long long z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  // He-uniform for the ReLU layers; the decoders start at zero, so the
  // untrained network is exactly block Jacobi.
  auto fill_w = [&](long long off, int fan_in, int fan_out) {
    const double a = std::sqrt(6.0 / fan_in);
    for (long long i = 0; i < (long long)fan_in * fan_out; ++i) p[off + i] = (float)((2.0 * uni() - 1.0) * a);
  };
  const int H = hfx::LPC_H;
  fill_w(L.Wn, L.NI, H);
  fill_w(L.We, L.NE, H);
  for (int l = 0; l < hfx::LPC_LAYERS; ++l) {
    fill_w(L.Wm[l], 3 * H, H);
    fill_w(L.Wu[l], 2 * H, H);
  }
  h->set_device();
  h->params.upload(p.data(), p.size(), h->stream);
  hfx::fill_zero(h->am.p, L.total, h->stream);
  hfx::fill_zero(h->av.p, L.total, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  h->adam_t = 0;
  h->p16_valid = false;
  LPC_END
}

extern "C" int hfx_lpc_set_params(hfx_lpc* h, const float* p, long long n) {
  LPC_TRY
  need(h && p, "null argument");
  need(n == h->L.total, "parameter count mismatch");
  h->set_device();
  h->params.upload(p, n, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  h->p16_valid = false;
  LPC_END
}

extern "C" int hfx_lpc_get_params(hfx_lpc* h, float* p, long long n) {
  LPC_TRY
  need(h && p, "null argument");
  need(n == h->L.total, "parameter count mismatch");
  h->set_device();
  h->params.download(p, n, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  LPC_END
}

extern "C" int hfx_lpc_add_sample(hfx_lpc* h, hfx_hdg* prob) {
  LPC_TRY
  need(h && prob, "null argument");
  need(prob->solved, "the HDG problem must be assembled and solved before it is added as a sample");
  need(prob->device == h->device, "the HDG problem lives on another GPU");
  need(prob->A.bs == h->L.bs, "block size of the problem differs from the network's");
  h->set_device();
  hipStream_t s = h->stream;
  HFX_HIP(hipStreamSynchronize(prob->stream));
  std::shared_ptr<hfx::LpcGraph> g;
  for (const auto& k : h->graphs)
    if (k.mesh_serial == prob->mesh_serial && k.nb == prob->A.nb && k.nnzb == prob->A.nnzb) g = k.g;
  if (!g) {
    std::vector<uint8_t> dir(prob->mesh.nFaces, 0);
    for (int f : prob->mesh.boundary) dir[f] = 1;
    g.reset(new hfx::LpcGraph());
    g->build(prob->pat, dir, s);
    h->graphs.push_back({prob->mesh_serial, prob->A.nb, prob->A.nnzb, g});
  }
  auto sm = std::unique_ptr<Sample>(new Sample());
  sm->g = g;
  sm->sys.build(*g, prob->A.vals.p, s);
  const long long n = (long long)g->nb * h->L.bs;
  sm->z.alloc(n);
  sm->v.alloc(n);
  hipLaunchKernelGGL(solution_probe_kernel, dim3(hfx::ceil_div(n, 256)), dim3(256), 0, s, n, h->L.bs, sm->sys.sinv.p,
                     prob->lambda.p, sm->z.p);
  HFX_HIP(hipGetLastError());
  hfx::DevBuf<double> nrm;
  nrm.alloc(1);
  hipLaunchKernelGGL(norm2_f32_kernel, dim3(1), dim3(256), 0, s, n, sm->z.p, nrm.p);
  HFX_HIP(hipGetLastError());
  double nz2 = 0.0;
  HFX_HIP(hipMemcpyAsync(&nz2, nrm.p, sizeof(double), hipMemcpyDeviceToHost, s));
  HFX_HIP(hipStreamSynchronize(s));
  need(nz2 > 0.0 && std::isfinite(nz2), "the sample's solution is zero or not finite");
  hipLaunchKernelGGL(scale_f32_kernel, dim3(hfx::ceil_div(n, 256)), dim3(256), 0, s, n, (float)(1.0 / std::sqrt(nz2)),
                     sm->z.p);
  HFX_HIP(hipGetLastError());
  hfx::lpc_spmv(*g, sm->sys.ahat.p, sm->z.p, sm->v.p, s);
  HFX_HIP(hipStreamSynchronize(s));
  h->samples.push_back(std::move(sm));
  LPC_END
}

extern "C" int hfx_lpc_num_samples(hfx_lpc* h, int* n) {
  LPC_TRY
  need(h && n, "null argument");
  *n = (int)h->samples.size();
  LPC_END
}

extern "C" int hfx_lpc_clear_samples(hfx_lpc* h) {
  LPC_TRY
  need(h != nullptr, "null handle");
  h->set_device();
  HFX_HIP(hipStreamSynchronize(h->stream));
  h->samples.clear();
  h->graphs.clear();
  LPC_END
}

extern "C" int hfx_lpc_loss_grad(hfx_lpc* h, int sample, int n_random, unsigned long long seed, double* loss,
                                 float* grad) {
  LPC_TRY
  need(h != nullptr, "null handle");
  need(sample >= 0 && sample < (int)h->samples.size(), "sample index out of range");
  need(n_random >= 0, "negative probe count");
  h->set_device();
  h->ensure_work();
  hipStream_t s = h->stream;
  HFX_HIP(hipMemsetAsync(h->lossacc.p, 0, sizeof(double), s));
  h->forward_loss(sample, n_random, seed);
  if (grad) {
    const Sample& sm = *h->samples[sample];
    hfx::lpc_backward(h->L, h->params.p, *sm.g, sm.sys, h->acts, h->work, h->grad.p, s);
    h->grad.download(grad, h->L.total, s);
  }
  if (loss) HFX_HIP(hipMemcpyAsync(loss, h->lossacc.p, sizeof(double), hipMemcpyDeviceToHost, s));
  HFX_HIP(hipStreamSynchronize(s));
  LPC_END
}

extern "C" int hfx_lpc_train(hfx_lpc* h, const int* order, int nsteps, float lr, int n_random,
                             unsigned long long seed, double* losses) {
  LPC_TRY
  need(h && order, "null argument");
  need(nsteps >= 0 && n_random >= 0 && lr > 0.f, "bad training arguments");
  for (int i = 0; i < nsteps; ++i) need(order[i] >= 0 && order[i] < (int)h->samples.size(), "sample index out of range");
  h->set_device();
  h->ensure_work();
  hipStream_t s = h->stream;
  hfx::DevBuf<double> hist;
  hist.alloc(std::max(nsteps, 1));
  HFX_HIP(hipMemsetAsync(hist.p, 0, sizeof(double) * std::max(nsteps, 1), s));
  for (int i = 0; i < nsteps; ++i) {
    const int k = order[i];
    HFX_HIP(hipMemcpyAsync(h->lossacc.p, hist.p + i, sizeof(double), hipMemcpyDeviceToDevice, s));
    h->forward_loss(k, n_random, seed + (unsigned long long)i);
    HFX_HIP(hipMemcpyAsync(hist.p + i, h->lossacc.p, sizeof(double), hipMemcpyDeviceToDevice, s));
    const Sample& sm = *h->samples[k];
    hfx::lpc_backward(h->L, h->params.p, *sm.g, sm.sys, h->acts, h->work, h->grad.p, s);
    ++h->adam_t;
    hfx::lpc_adam(h->L.total, h->params.p, h->grad.p, h->am.p, h->av.p, lr, 0.9f, 0.999f, 1e-8f, h->adam_t, s);
  }
  if (losses && nsteps > 0) HFX_HIP(hipMemcpyAsync(losses, hist.p, sizeof(double) * nsteps, hipMemcpyDeviceToHost, s));
  HFX_HIP(hipStreamSynchronize(s));
  h->p16_valid = false;
  LPC_END
}

extern "C" int hfx_lpc_outputs(hfx_lpc* h, int sample, int bf16, float* cn, float* ce) {
  LPC_TRY
  need(h != nullptr, "null handle");
  need(sample >= 0 && sample < (int)h->samples.size(), "sample index out of range");
  h->set_device();
  h->ensure_work();
  hipStream_t s = h->stream;
  const Sample& sm = *h->samples[sample];
  if (bf16) {
    if (!h->p16_valid) {
      hfx::lpc_params_to_bf16(h->L.total, h->params.p, h->p16.p, s);
      h->p16_valid = true;
    }
    hfx::lpc_forward_bf16(h->L, h->params.p, h->p16.p, *sm.g, sm.sys, h->acts, s);
  } else {
    hfx::lpc_forward(h->L, h->params.p, *sm.g, sm.sys, h->acts, s);
  }
  if (cn) h->acts.cn.download(cn, (size_t)sm.g->nb * h->L.O, s);
  if (ce && sm.g->E > 0) h->acts.ce.download(ce, (size_t)sm.g->E * h->L.O, s);
  HFX_HIP(hipStreamSynchronize(s));
  LPC_END
}

extern "C" int hfx_lpc_graph_sizes(hfx_lpc* h, int sample, long long* out) {
  LPC_TRY
  need(h && out, "null argument");
  need(sample >= 0 && sample < (int)h->samples.size(), "sample index out of range");
  const hfx::LpcGraph& g = *h->samples[sample]->g;
  out[0] = g.nb;
  out[1] = g.nnzb;
  out[2] = g.E;
  out[3] = h->L.bs;
  LPC_END
}

extern "C" int hfx_hdg_set_lpc(hfx_hdg* h, hfx_lpc* lpc, int bf16) {
  LPC_TRY
  need(h != nullptr, "null problem handle");
  if (lpc) {
    need(lpc->device == h->device, "network and problem live on different GPUs");
    need(!h->has_mesh || lpc->L.bs == h->A.bs, "block size of the network differs from the problem's");
  }
  h->lpc = lpc;
  h->lpc_bf16 = bf16 ? 1 : 0;
  LPC_END
}

extern "C" int hfx_hdg_pc_setup_ms(hfx_hdg* h, double* ms) {
  LPC_TRY
  need(h && ms, "null argument");
  *ms = h->last_pc_setup_ms;
  LPC_END
}

namespace hfx {

// Normalise the problem's trace matrix, run the network (BF16 MFMA or FP32)
// and assemble M as FP32 BSR values.
void lpc_prepare(hfx_hdg* h) {
  if (!h->lpc) throw LaErr("no network attached to this problem (hfx_hdg_set_lpc)");
  hfx_lpc* net = h->lpc;
  if (net->L.bs != h->A.bs) throw LaErr("block size of the network differs from the problem's");
  hipStream_t s = h->stream;
  // the network's stream is idle between calls; order its parameter writes first
  HFX_HIP(hipStreamSynchronize(net->stream));
  if (!h->lpc_state) h->lpc_state.reset(new LpcInfer());
  LpcInfer& st = *h->lpc_state;
  const auto t0 = std::chrono::steady_clock::now();
  if (!st.graph_built) {
    std::vector<uint8_t> dir(h->mesh.nFaces, 0);
    for (int f : h->mesh.boundary) dir[f] = 1;
    st.g.build(h->pat, dir, s);
    st.acts.alloc(st.g.nb, st.g.E, net->L.O);
    st.mvals.alloc((size_t)st.g.nnzb * net->L.O);
    st.graph_built = true;
  }
  st.sys.build(st.g, h->A.vals.p, s);
  if (h->lpc_bf16) {
    if (!net->p16_valid) {
      lpc_params_to_bf16(net->L.total, net->params.p, net->p16.p, s);
      net->p16_valid = true;
    }
    lpc_forward_bf16(net->L, net->params.p, net->p16.p, st.g, st.sys, st.acts, s);
  } else {
    lpc_forward(net->L, net->params.p, st.g, st.sys, st.acts, s);
  }
  lpc_build_m(st.g, st.sys, st.acts, st.mvals.p, s);
  HFX_HIP(hipStreamSynchronize(s));
  h->last_pc_setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace hfx
