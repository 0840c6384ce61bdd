// Learned trace preconditioner: a message-passing network over the face-block
// graph of the trace matrix that emits one correction block per matrix block,
// added to the block-Jacobi inverse.  Trained on the GPU with the FP32
// forward / backward / Adam kernels of lpc.hip; inference runs in BF16 on MFMA
// with FP32 accumulation.
//
// Normalisation.  With s_F = mean |diag(A_FF)| and S = diag(s_F) (x) I_b,
// the network sees Ahat = S^-1/2 A S^-1/2 and produces
//   Mhat_FG = [F == G] inv(Ahat_FF) + C_FG      (C from the network)
// and the preconditioner is M = S^-1/2 Mhat S^-1/2, applied as z = M r.
// Dirichlet rows (identity rows of A) keep C = 0, so M's row is exact there.
//
// Network (width H = 64, three layers), row vectors, y = x W + b:
//   h0 = relu([Ahat_FF | dir_F] Wn + bn)          node encoder  (b*b+1 -> H)
//   e0 = relu(Ahat_FG We + be)                    edge encoder  (b*b -> H)
//   layer l: m  = relu([h_F | h_G | e_FG] Wm + bm)    (3H -> H), per edge F->G
//            e  = e + m
//            a_F = sum over edges F->G of m
//            u  = relu([h_F | a_F] Wu + bu)            (2H -> H)
//            h  = h + u
//   C_FF = h Wdn + bdn  (H -> b*b),  C_FG = e Wde + bde  (H -> b*b)
// Training loss per probe z (in the normalised space):
//   L = ||z - Mhat Ahat z||^2 / ||z||^2 , probes = the normalised solution and
//   Rademacher vectors.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "hfx_internal.h"
#include "trace_system.h"

namespace hfx {

constexpr int LPC_H = 64;
constexpr int LPC_LAYERS = 3;

// Parameter layout (FP32, flat), in this order:
//   Wn[NI*H] bn[H] We[NE*H] be[H] {Wm[3H*H] bm[H] Wu[2H*H] bu[H]} x LAYERS
//   Wdn[H*O] bdn[O] Wde[H*O] bde[O]
// with NI = b*b + 1, NE = b*b, O = b*b.
struct LpcLayout {
  int bs = 0, NI = 0, NE = 0, O = 0;
  long long Wn, bn, We, be, Wm[LPC_LAYERS], bm[LPC_LAYERS], Wu[LPC_LAYERS], bu[LPC_LAYERS], Wdn, bdn, Wde, bde;
  long long total = 0;
  explicit LpcLayout(int b = 4);
};

// Graph of one trace-matrix pattern: nodes = faces (block rows), edges = the
// off-diagonal blocks in BSR order (grouped by row).
struct LpcGraph {
  int nb = 0, bs = 0, nnzb = 0, E = 0;
  DevBuf<int> rowptr, col, rowf, diag;       // BSR pattern
  DevBuf<int> erowptr, eblk, esrc, edst, erev;  // edges
  DevBuf<int> blk2edge;                      // nnzb: edge index, -1 for the diagonal block
  DevBuf<float> dirf;                        // nb: 1 on Dirichlet rows
  DevBuf<uint8_t> dir;                       // nb
  void build(const BsrPattern& P, const std::vector<uint8_t>& is_dir_face, hipStream_t s);
};

// Normalised system of one problem (device, FP32 unless noted).
struct LpcSystem {
  DevBuf<float> ahat;   // nnzb * b*b
  DevBuf<float> dinv;   // nb * b*b : inv(Ahat_FF)
  DevBuf<double> sinv;  // nb : s_F^-1/2
  // build from an FP64 BSR matrix with the graph's pattern
  void build(const LpcGraph& g, const double* vals, hipStream_t s);
};

// Activations of one forward pass (sized for a graph).
struct LpcActs {
  int nb = 0, E = 0;
  DevBuf<float> h[LPC_LAYERS + 1], e[LPC_LAYERS + 1], m[LPC_LAYERS], a[LPC_LAYERS], u[LPC_LAYERS];
  DevBuf<float> cn, ce;  // outputs: nb * O, E * O
  void alloc(int nb_, int E_, int O);
};

// Gradient workspace.
struct LpcGradWork {
  DevBuf<float> dh, dh2, de, de2, dP, dX, dXn, dcn, dce, WT, part;
  int parts = 0;
  void alloc(int nb, int E, int O, long long maxKN);
};

// FP32 forward pass; outputs in acts.cn / acts.ce (Dirichlet rows masked).
void lpc_forward(const LpcLayout& L, const float* params, const LpcGraph& g, const LpcSystem& sys, LpcActs& acts,
                 hipStream_t s);

// BF16 MFMA forward pass (inference); params16 = bf16 copy of params (raw bits).
void lpc_forward_bf16(const LpcLayout& L, const float* params, const uint16_t* params16, const LpcGraph& g,
                      const LpcSystem& sys, LpcActs& acts, hipStream_t s);
void lpc_params_to_bf16(long long n, const float* p, uint16_t* q, hipStream_t s);

// Loss of one probe z (nb*b, FP32) with v = Ahat z precomputed; accumulates
// dL/dC into dcn / dce and the loss into loss_acc[0] (device double),
// scaled by weight / ||z||^2 (inv_nz2).  `partials` holds
// lpc_loss_partials(g) doubles.
int lpc_loss_partials(const LpcGraph& g);
void lpc_loss(const LpcGraph& g, const LpcSystem& sys, const LpcActs& acts, const float* z, const float* v,
              float inv_nz2, float weight, float* dcn, float* dce, double* partials, double* loss_acc, hipStream_t s);

// y = Ahat x (FP32).
void lpc_spmv(const LpcGraph& g, const float* vals, const float* x, float* y, hipStream_t s);
// z_i = +-1 from a counter-based hash of (seed, i).
void lpc_rademacher(long long n, unsigned long long seed, float* z, hipStream_t s);

// Full backward pass: grad (same layout as params) = dL/dparams, given dcn/dce.
void lpc_backward(const LpcLayout& L, const float* params, const LpcGraph& g, const LpcSystem& sys,
                  const LpcActs& acts, LpcGradWork& w, float* grad, hipStream_t s);

// Adam step (bias-corrected; t is the 1-based step count).
void lpc_adam(long long n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
              int t, hipStream_t s);

// Preconditioner matrix M = S^-1/2 (inv(Ahat_FF) + C) S^-1/2 as FP32 BSR values.
void lpc_build_m(const LpcGraph& g, const LpcSystem& sys, const LpcActs& acts, float* mvals, hipStream_t s);
// y = M x for FP32 BSR values and FP64 vectors.
void bsr_spmv_f32vals(const LpcGraph& g, const float* mvals, const double* x, double* y, hipStream_t s);

}  // namespace hfx
