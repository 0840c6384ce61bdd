// Kernels of the learned trace preconditioner (see lpc.h for the network and
// the loss).  FP32 training path: LDS-tiled SIMT GEMMs with gathered inputs,
// deterministic two-stage weight-gradient reductions, segmented row sums over
// the BSR edge lists (no atomics anywhere, so a training run is bit
// reproducible), Adam.  Inference: the same network on BF16 MFMA
// (v_mfma_f32_32x32x8bf16_1k) with FP32 accumulation.
#include "lpc.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace hfx {

LpcLayout::LpcLayout(int b) {
  const long long H = LPC_H;
  bs = b;
  NI = b * b + 1;
  NE = b * b;
  O = b * b;
  long long o = 0;
  Wn = o; o += NI * H; bn = o; o += H;
  We = o; o += NE * H; be = o; o += H;
  for (int l = 0; l < LPC_LAYERS; ++l) {
    Wm[l] = o; o += 3 * H * H; bm[l] = o; o += H;
    Wu[l] = o; o += 2 * H * H; bu[l] = o; o += H;
  }
  Wdn = o; o += H * O; bdn = o; o += O;
  Wde = o; o += H * O; bde = o; o += O;
  total = o;
}

void LpcGraph::build(const BsrPattern& P, const std::vector<uint8_t>& is_dir_face, hipStream_t s) {
  nb = P.nb;
  bs = P.bs;
  nnzb = (int)P.col.size();
  std::vector<int> er(nb + 1, 0), eb, es, ed, b2e(nnzb, -1);
  for (int F = 0; F < nb; ++F) {
    for (int k = P.rowptr[F]; k < P.rowptr[F + 1]; ++k)
      if (P.col[k] != F) {
        b2e[k] = (int)eb.size();
        eb.push_back(k);
        es.push_back(F);
        ed.push_back(P.col[k]);
      }
    er[F + 1] = (int)eb.size();
  }
  E = (int)eb.size();
  std::vector<int> rev(E, -1);
  for (int e = 0; e < E; ++e) {
    const int G = ed[e], F = es[e];
    for (int f = er[G]; f < er[G + 1]; ++f)
      if (ed[f] == F) { rev[e] = f; break; }
    if (rev[e] < 0) throw std::runtime_error("learned preconditioner: trace pattern is not structurally symmetric");
  }
  std::vector<float> df(nb);
  std::vector<uint8_t> d8(nb);
  if ((int)is_dir_face.size() != nb) throw std::runtime_error("learned preconditioner: Dirichlet flags size mismatch");
  for (int F = 0; F < nb; ++F) { d8[F] = is_dir_face[F] ? 1 : 0; df[F] = d8[F] ? 1.f : 0.f; }
  rowptr.upload(P.rowptr, s);
  col.upload(P.col, s);
  rowf.upload(P.rowf, s);
  diag.upload(P.diag, s);
  erowptr.upload(er, s);
  eblk.upload(eb, s);
  esrc.upload(es, s);
  edst.upload(ed, s);
  erev.upload(rev, s);
  blk2edge.upload(b2e, s);
  dirf.upload(df, s);
  dir.upload(d8, s);
  HFX_HIP(hipStreamSynchronize(s));
}

void LpcActs::alloc(int nb_, int E_, int O) {
  nb = nb_;
  E = E_;
  const size_t H = LPC_H;
  for (int l = 0; l <= LPC_LAYERS; ++l) {
    h[l].alloc((size_t)nb * H);
    e[l].alloc((size_t)E * H);
  }
  for (int l = 0; l < LPC_LAYERS; ++l) {
    m[l].alloc((size_t)E * H);
    a[l].alloc((size_t)nb * H);
    u[l].alloc((size_t)nb * H);
  }
  cn.alloc((size_t)nb * O);
  ce.alloc((size_t)std::max(E, 1) * O);
}

namespace {

constexpr int WG_PARTS = 256;

void LpcGradWork_alloc(LpcGradWork& w, int nb, int E, int O, long long maxKN) {
  const size_t H = LPC_H, mx = (size_t)std::max(nb, E);
  w.dh.alloc((size_t)nb * H);
  w.dh2.alloc((size_t)nb * H);
  w.de.alloc((size_t)E * H);
  w.de2.alloc((size_t)E * H);
  w.dP.alloc(mx * H);
  w.dX.alloc((size_t)E * 3 * H);
  w.dXn.alloc((size_t)nb * 2 * H);
  w.dcn.alloc((size_t)nb * O);
  w.dce.alloc((size_t)std::max(E, 1) * O);
  w.WT.alloc((size_t)maxKN);
  w.parts = WG_PARTS;
  w.part.alloc((size_t)WG_PARTS * (192 * 64 + 64));
}

// ------------------------------------------------------------ normalisation

__global__ void scale_kernel(int nb, int bs, const int* __restrict__ diag, const double* __restrict__ vals,
                             double* __restrict__ sinv) {
  const int F = blockIdx.x * blockDim.x + threadIdx.x;
  if (F >= nb) return;
  const double* a = vals + (size_t)diag[F] * bs * bs;
  double sm = 0.0;
  for (int i = 0; i < bs; ++i) sm += fabs(a[i * bs + i]);
  sm /= bs;
  sinv[F] = sm > 0.0 ? 1.0 / sqrt(sm) : 1.0;
}

__global__ void normalise_kernel(int nnzb, int bs, const int* __restrict__ rowf, const int* __restrict__ col,
                                 const double* __restrict__ sinv, const double* __restrict__ vals,
                                 float* __restrict__ ahat) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long bb = (long long)bs * bs;
  if (t >= (long long)nnzb * bb) return;
  const int k = (int)(t / bb);
  ahat[t] = (float)(vals[t] * sinv[rowf[k]] * sinv[col[k]]);
}

template <int BS>
__global__ void dinv_kernel(int nb, const int* __restrict__ diag, const double* __restrict__ sinv,
                            const double* __restrict__ vals, float* __restrict__ dinv) {
  const int F = blockIdx.x * blockDim.x + threadIdx.x;
  if (F >= nb) return;
  double A[BS * BS], I[BS * BS];
  const double s2 = sinv[F] * sinv[F];
  const double* src = vals + (size_t)diag[F] * BS * BS;
  for (int i = 0; i < BS * BS; ++i) { A[i] = src[i] * s2; I[i] = 0.0; }
  for (int i = 0; i < BS; ++i) I[i * BS + i] = 1.0;
  for (int k = 0; k < BS; ++k) {
    int p = k;
    for (int r = k + 1; r < BS; ++r)
      if (fabs(A[r * BS + k]) > fabs(A[p * BS + k])) p = r;
    if (p != k)
      for (int c = 0; c < BS; ++c) {
        double t = A[k * BS + c]; A[k * BS + c] = A[p * BS + c]; A[p * BS + c] = t;
        t = I[k * BS + c]; I[k * BS + c] = I[p * BS + c]; I[p * BS + c] = t;
      }
    const double piv = A[k * BS + k];
    const double inv = piv != 0.0 ? 1.0 / piv : 0.0;
    for (int c = 0; c < BS; ++c) { A[k * BS + c] *= inv; I[k * BS + c] *= inv; }
    for (int r = 0; r < BS; ++r) {
      if (r == k) continue;
      const double f = A[r * BS + k];
      if (f == 0.0) continue;
      for (int c = 0; c < BS; ++c) { A[r * BS + c] -= f * A[k * BS + c]; I[r * BS + c] -= f * I[k * BS + c]; }
    }
  }
  float* dst = dinv + (size_t)F * BS * BS;
  for (int i = 0; i < BS * BS; ++i) dst[i] = (float)I[i];
}

// ---------------------------------------------------------------- FP32 GEMM

// One input segment: row r reads p[(idx ? idx[r] : r) * ld + c], c < w.
struct Seg {
  const float* p = nullptr;
  const int* idx = nullptr;
  int w = 0, ld = 0;
};

struct LinArgs {
  int rows = 0, nseg = 1;
  Seg seg[3];
  int K = 0;
  const float* W = nullptr;  // W[k * ldw + n]
  int ldw = 0, ncols = 0;    // ncols <= 64
  const float* bias = nullptr;
  int act = 0;  // 1 = relu
  float* out = nullptr;
  int ldo = 0;
  const float* res = nullptr;  // out2 = res + out
  int ldr = 0;
  float* out2 = nullptr;
  int ldo2 = 0;
  const uint8_t* rowmask = nullptr;  // masked rows produce 0
  const int* maskidx = nullptr;
};

__device__ __forceinline__ float seg_at(const Seg& g, int r, int c) {
  const long long rr = g.idx ? (long long)g.idx[r] : (long long)r;
  return g.p[rr * g.ld + c];
}

template <class A>
__device__ __forceinline__ float gathered(const A& a, int r, int k) {
  if (k < a.seg[0].w) return seg_at(a.seg[0], r, k);
  k -= a.seg[0].w;
  if (a.nseg < 2) return 0.f;
  if (k < a.seg[1].w) return seg_at(a.seg[1], r, k);
  k -= a.seg[1].w;
  if (a.nseg < 3) return 0.f;
  return seg_at(a.seg[2], r, k);
}

// out[r, n] = act(sum_k X[r, k] W[k, n] + bias[n]), 64 x 64 output tiles,
// 4 x 4 per thread, K in chunks of 32 through LDS.
__global__ __launch_bounds__(256) void lin_kernel(LinArgs a) {
  __shared__ __align__(16) float Xs[32][68];
  __shared__ __align__(16) float Ws[32][64];
  const int t = threadIdx.x, tn = t & 15, tm = t >> 4;
  const int ntiles = (a.rows + 63) >> 6;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int r0 = tile << 6;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
    for (int k0 = 0; k0 < a.K; k0 += 32) {
      for (int e = t; e < 2048; e += 256) {
        const int row = e >> 5, kk = e & 31, r = r0 + row, k = k0 + kk;
        Xs[kk][row] = (r < a.rows && k < a.K) ? gathered(a, r, k) : 0.f;
      }
      for (int e = t; e < 2048; e += 256) {
        const int kk = e >> 6, n = e & 63, k = k0 + kk;
        Ws[kk][n] = (k < a.K && n < a.ncols) ? a.W[(long long)k * a.ldw + n] : 0.f;
      }
      __syncthreads();
#pragma unroll 8
      for (int kk = 0; kk < 32; ++kk) {
        const float4 xv = *reinterpret_cast<const float4*>(&Xs[kk][tm * 4]);
        const float4 wv = *reinterpret_cast<const float4*>(&Ws[kk][tn * 4]);
        const float xs[4] = {xv.x, xv.y, xv.z, xv.w}, ws[4] = {wv.x, wv.y, wv.z, wv.w};
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] += xs[i] * ws[j];
      }
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = r0 + tm * 4 + i;
      if (r >= a.rows) continue;
      const bool masked = a.rowmask && a.rowmask[a.maskidx ? a.maskidx[r] : r];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int n = tn * 4 + j;
        if (n >= a.ncols) continue;
        float v = acc[i][j] + (a.bias ? a.bias[n] : 0.f);
        if (a.act) v = fmaxf(v, 0.f);
        if (masked) v = 0.f;
        a.out[(long long)r * a.ldo + n] = v;
        if (a.out2) a.out2[(long long)r * a.ldo2 + n] = a.res[(long long)r * a.ldr + n] + v;
      }
    }
  }
}

// Launch lin_kernel over all output columns in chunks of 64.
void lin(LinArgs a, int N, hipStream_t s) {
  if (a.rows <= 0) return;
  const int ntiles = (a.rows + 63) / 64;
  const int grid = std::min(ntiles, 4096);
  for (int c0 = 0; c0 < N; c0 += 64) {
    LinArgs b = a;
    b.ncols = std::min(64, N - c0);
    b.W = a.W + c0;
    if (a.bias) b.bias = a.bias + c0;
    b.out = a.out + c0;
    if (a.res) b.res = a.res + c0;
    if (a.out2) b.out2 = a.out2 + c0;
    hipLaunchKernelGGL(lin_kernel, dim3(grid), dim3(256), 0, s, b);
    HFX_HIP(hipGetLastError());
  }
}

// ------------------------------------------------------ weight gradients

struct WgArgs {
  int rows = 0, nseg = 1;
  Seg seg[3];
  int K = 0;  // <= 192
  const float* dP = nullptr;
  int ldp = 0, ncols = 0;  // ncols <= 64
  float* part = nullptr;   // parts x (K * ncols + ncols)
  int rows_per_part = 0;
};

// Partial dW = X^T dP and db = sum dP over this block's rows; thread (n, g)
// owns column n and rows k = g, g+4, ... of dW.
__global__ __launch_bounds__(256) void wgrad_kernel(WgArgs a) {
  __shared__ float Xs[32][193];
  __shared__ float Ps[32][64];
  const int t = threadIdx.x, n = t & 63, g = t >> 6;
  float acc[48];
#pragma unroll
  for (int j = 0; j < 48; ++j) acc[j] = 0.f;
  float accb = 0.f;
  const int rbeg = blockIdx.x * a.rows_per_part;
  const int rend = min(a.rows, rbeg + a.rows_per_part);
  for (int c0 = rbeg; c0 < rend; c0 += 32) {
    for (int e = t; e < 32 * 192; e += 256) {
      const int row = e / 192, k = e - row * 192, r = c0 + row;
      Xs[row][k] = (r < rend && k < a.K) ? gathered(a, r, k) : 0.f;
    }
    for (int e = t; e < 32 * 64; e += 256) {
      const int row = e >> 6, nn = e & 63, r = c0 + row;
      Ps[row][nn] = (r < rend && nn < a.ncols) ? a.dP[(long long)r * a.ldp + nn] : 0.f;
    }
    __syncthreads();
    for (int row = 0; row < 32; ++row) {
      const float d = Ps[row][n];
      accb += d;
#pragma unroll
      for (int j = 0; j < 48; ++j) acc[j] += Xs[row][g + 4 * j] * d;
    }
    __syncthreads();
  }
  if (n >= a.ncols) return;
  float* p = a.part + (size_t)blockIdx.x * (a.K * a.ncols + a.ncols);
#pragma unroll
  for (int j = 0; j < 48; ++j) {
    const int k = g + 4 * j;
    if (k < a.K) p[k * a.ncols + n] = acc[j];
  }
  if (g == 0) p[a.K * a.ncols + n] = accb;
}

__global__ void wgrad_reduce_kernel(int parts, int K, int ncols, const float* __restrict__ part, float* __restrict__ gW,
                                    int ldw, float* __restrict__ gb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int tot = K * ncols + ncols;
  if (t >= tot) return;
  const size_t stride = (size_t)tot;
  float acc = 0.f;
  for (int b = 0; b < parts; ++b) acc += part[b * stride + t];
  if (t < K * ncols) gW[(t / ncols) * ldw + (t % ncols)] = acc;
  else gb[t - K * ncols] = acc;
}

void fill(long long n, float v, float* p, hipStream_t s);

// grad W (K x N, row-major) and grad b (N) for the rows of `a`.
void wgrad(WgArgs a, int N, float* gW, float* gb, float* part, hipStream_t s) {
  if (a.K > 192) throw std::runtime_error("learned preconditioner: layer input wider than 192");
  if (a.rows <= 0) {
    for (int k = 0; k < a.K; ++k) fill(N, 0.f, gW + (long long)k * N, s);
    fill(N, 0.f, gb, s);
    return;
  }
  const int parts = std::min(WG_PARTS, (a.rows + 31) / 32);
  int rpp = (a.rows + parts - 1) / parts;
  rpp = (rpp + 31) / 32 * 32;
  const int used = (a.rows + rpp - 1) / rpp;
  for (int c0 = 0; c0 < N; c0 += 64) {
    WgArgs b = a;
    b.ncols = std::min(64, N - c0);
    b.dP = a.dP + c0;
    b.part = part;
    b.rows_per_part = rpp;
    hipLaunchKernelGGL(wgrad_kernel, dim3(used), dim3(256), 0, s, b);
    HFX_HIP(hipGetLastError());
    const int tot = b.K * b.ncols + b.ncols;
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(ceil_div(tot, 256)), dim3(256), 0, s, used, b.K, b.ncols, part,
                       gW + c0, N, gb + c0);
    HFX_HIP(hipGetLastError());
  }
}

// ------------------------------------------------------- small kernels

__global__ void rowsum_kernel(int nb, const int* __restrict__ rowptr, const int* __restrict__ map,
                              const float* __restrict__ src, int lds, int coloff, float* __restrict__ out, int ldo,
                              int accumulate) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nb * LPC_H) return;
  const int F = (int)(t / LPC_H), c = (int)(t % LPC_H);
  float acc = 0.f;
  for (int k = rowptr[F]; k < rowptr[F + 1]; ++k) {
    const long long kk = map ? map[k] : k;
    acc += src[kk * lds + coloff + c];
  }
  float* o = out + (long long)F * ldo + c;
  *o = accumulate ? *o + acc : acc;
}

void rowsum(int nb, const int* rowptr, const int* map, const float* src, int lds, int coloff, float* out, int ldo,
            bool accumulate, hipStream_t s) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(rowsum_kernel, dim3(ceil_div((long long)nb * LPC_H, 256)), dim3(256), 0, s, nb, rowptr, map, src,
                     lds, coloff, out, ldo, accumulate ? 1 : 0);
  HFX_HIP(hipGetLastError());
}

// out = g * (act > 0)   (contiguous rows of 64)
__global__ void mask_mul_kernel(long long n, const float* __restrict__ g, const float* __restrict__ act,
                                float* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = act[t] > 0.f ? g[t] : 0.f;
}

// out[k, c] = (m[k, c] > 0) ? de[k, c] + da[src[k] * ldda + c] : 0
__global__ void gather_add_mask_kernel(int E, const float* __restrict__ de, const float* __restrict__ da, int ldda,
                                       const int* __restrict__ src, const float* __restrict__ m,
                                       float* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)E * LPC_H) return;
  const int k = (int)(t / LPC_H), c = (int)(t % LPC_H);
  out[t] = m[t] > 0.f ? de[t] + da[(long long)src[k] * ldda + c] : 0.f;
}

// out[r, c] = a[r, c] + b[r * ldb + off + c]
__global__ void add_strided_kernel(int rows, const float* __restrict__ a, const float* __restrict__ b, int ldb, int off,
                                   float* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)rows * LPC_H) return;
  const int r = (int)(t / LPC_H), c = (int)(t % LPC_H);
  out[t] = a[t] + b[(long long)r * ldb + off + c];
}

__global__ void transpose_kernel(int K, int N, const float* __restrict__ W, float* __restrict__ WT) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= K * N) return;
  const int k = t / N, n = t % N;
  WT[n * K + k] = W[t];
}

void transpose(int K, int N, const float* W, float* WT, hipStream_t s) {
  hipLaunchKernelGGL(transpose_kernel, dim3(ceil_div(K * N, 256)), dim3(256), 0, s, K, N, W, WT);
  HFX_HIP(hipGetLastError());
}

__global__ void fill_kernel(long long n, float v, float* __restrict__ p) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

void fill(long long n, float v, float* p, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, v, p);
  HFX_HIP(hipGetLastError());
}

// --------------------------------------------------------------- loss

// Thread per (node F, row i): u = (Mhat v)_{F,i}, r = z - u, loss += r^2,
// dC_FG[i][j] += -2 r v_G[j] (times inv_nz2 * weight); Dirichlet rows get no
// gradient (their C is fixed at 0).
template <int BS>
__global__ __launch_bounds__(256) void loss_kernel(int nb, const int* __restrict__ rowptr, const int* __restrict__ col,
                                                   const int* __restrict__ b2e, const uint8_t* __restrict__ dir,
                                                   const float* __restrict__ dinv, const float* __restrict__ cn,
                                                   const float* __restrict__ ce, const float* __restrict__ z,
                                                   const float* __restrict__ v, float scale, float* __restrict__ dcn,
                                                   float* __restrict__ dce, double* __restrict__ lossp) {
  __shared__ double red[4];
  constexpr int O = BS * BS;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double lsq = 0.0;
  if (t < (long long)nb * BS) {
    const int F = (int)(t / BS), i = (int)(t % BS);
    float u = 0.f;
    for (int k = rowptr[F]; k < rowptr[F + 1]; ++k) {
      const int G = col[k], e = b2e[k];
      const float* vg = v + (long long)G * BS;
#pragma unroll
      for (int j = 0; j < BS; ++j) {
        const float mv = e < 0 ? dinv[(long long)F * O + i * BS + j] + cn[(long long)F * O + i * BS + j]
                               : ce[(long long)e * O + i * BS + j];
        u += mv * vg[j];
      }
    }
    const float r = z[t] - u;
    lsq = (double)r * r * scale;
    if (!dir[F]) {
      const float gsc = -2.f * r * scale;
      for (int k = rowptr[F]; k < rowptr[F + 1]; ++k) {
        const int G = col[k], e = b2e[k];
        const float* vg = v + (long long)G * BS;
        float* dst = e < 0 ? dcn + (long long)F * O + i * BS : dce + (long long)e * O + i * BS;
#pragma unroll
        for (int j = 0; j < BS; ++j) dst[j] += gsc * vg[j];
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) lsq += __shfl_xor(lsq, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = lsq;
  __syncthreads();
  if (threadIdx.x == 0) lossp[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void sum_into_kernel(int n, const double* __restrict__ p, double* __restrict__ acc) {
  __shared__ double red[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) acc[0] += red[0] + red[1] + red[2] + red[3];
}

template <int BS>
__global__ void spmv_f32_kernel(int nb, const int* __restrict__ rowptr, const int* __restrict__ col,
                                const float* __restrict__ vals, const float* __restrict__ x, float* __restrict__ y) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nb * BS) return;
  const int F = (int)(t / BS), i = (int)(t % BS);
  float acc = 0.f;
  for (int k = rowptr[F]; k < rowptr[F + 1]; ++k) {
    const float* a = vals + ((long long)k * BS + i) * BS;
    const float* xv = x + (long long)col[k] * BS;
#pragma unroll
    for (int j = 0; j < BS; ++j) acc += a[j] * xv[j];
  }
  y[t] = acc;
}

template <int BS>
__global__ void spmv_f32vals_kernel(int nb, const int* __restrict__ rowptr, const int* __restrict__ col,
                                    const float* __restrict__ vals, const double* __restrict__ x,
                                    double* __restrict__ y) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nb * BS) return;
  const int F = (int)(t / BS), i = (int)(t % BS);
  double acc = 0.0;
  for (int k = rowptr[F]; k < rowptr[F + 1]; ++k) {
    const float* a = vals + ((long long)k * BS + i) * BS;
    const double* xv = x + (long long)col[k] * BS;
#pragma unroll
    for (int j = 0; j < BS; ++j) acc += (double)a[j] * xv[j];
  }
  y[t] = acc;
}

__global__ void build_m_kernel(int nnzb, int bs, const int* __restrict__ rowf, const int* __restrict__ col,
                               const int* __restrict__ b2e, const double* __restrict__ sinv,
                               const float* __restrict__ dinv, const float* __restrict__ cn,
                               const float* __restrict__ ce, float* __restrict__ mv) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int O = bs * bs;
  if (t >= (long long)nnzb * O) return;
  const int k = (int)(t / O), ij = (int)(t % O);
  const int F = rowf[k], G = col[k], e = b2e[k];
  const float c = e < 0 ? dinv[(long long)F * O + ij] + cn[(long long)F * O + ij] : ce[(long long)e * O + ij];
  mv[t] = (float)((double)c * sinv[F] * sinv[G]);
}

__global__ void rademacher_kernel(long long n, unsigned long long seed, float* __restrict__ z) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  unsigned long long x = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(t + 1);
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  z[t] = (x >> 63) ? 1.f : -1.f;
}

__global__ void adam_kernel(long long n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gi = g[t];
  const float mi = b1 * m[t] + (1.f - b1) * gi;
  const float vi = b2 * v[t] + (1.f - b2) * gi * gi;
  m[t] = mi;
  v[t] = vi;
  p[t] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------- BF16 MFMA forward

__device__ __forceinline__ unsigned short f2bf(float f) {
  // round to nearest even
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__global__ void to_bf16_kernel(long long n, const float* __restrict__ p, unsigned short* __restrict__ q) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) q[t] = f2bf(p[t]);
}

struct LinArgs16 {
  LinArgs a;
  const unsigned short* W16 = nullptr;  // bf16 weights, W16[k * ldw + n]
};

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

// out tile: 128 rows x 64 cols per workgroup (4 waves, each 32 rows x 64 cols
// = two 32x32 MFMA accumulators); K in chunks of 32 through LDS as bf16.
// v_mfma_f32_32x32x8bf16_1k operands: lane l supplies A[l % 32][4*(l/32) .. +3]
// and B[4*(l/32) .. +3][l % 32]; accumulator element r of lane l is
// C[8*(r/4) + 4*(l/32) + r%4][l % 32].
__global__ __launch_bounds__(256) void lin_bf16_kernel(LinArgs16 p) {
  const LinArgs& a = p.a;
  __shared__ unsigned short Xs[128][40];  // [row][k], padded
  __shared__ unsigned short Ws[64][40];   // [n][k], padded
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int ntiles = (a.rows + 127) >> 7;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int r0 = tile << 7;
    f32x16_t acc0, acc1;
    for (int i = 0; i < 16; ++i) { acc0[i] = 0.f; acc1[i] = 0.f; }
    for (int k0 = 0; k0 < a.K; k0 += 32) {
      for (int e = t; e < 128 * 32; e += 256) {
        const int row = e >> 5, kk = e & 31, r = r0 + row, k = k0 + kk;
        Xs[row][kk] = f2bf((r < a.rows && k < a.K) ? gathered(a, r, k) : 0.f);
      }
      for (int e = t; e < 64 * 32; e += 256) {
        const int kk = e >> 6, n = e & 63, k = k0 + kk;
        Ws[n][kk] = (k < a.K && n < a.ncols) ? p.W16[(long long)k * a.ldw + n] : (unsigned short)0;
      }
      __syncthreads();
      const int arow = wave * 32 + (lane & 31), kq = (lane >> 5) * 4;
#pragma unroll
      for (int ks = 0; ks < 32; ks += 8) {
        bf16x4_t av, b0, b1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          av[q] = (short)Xs[arow][ks + kq + q];
          b0[q] = (short)Ws[(lane & 31)][ks + kq + q];
          b1[q] = (short)Ws[32 + (lane & 31)][ks + kq + q];
        }
        acc0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, b0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, b1, acc1, 0, 0, 0);
      }
      __syncthreads();
    }
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) {
      const int row = r0 + wave * 32 + 8 * (rr >> 2) + 4 * (lane >> 5) + (rr & 3);
      if (row >= a.rows) continue;
      const bool masked = a.rowmask && a.rowmask[a.maskidx ? a.maskidx[row] : row];
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        const int n = half * 32 + (lane & 31);
        if (n >= a.ncols) continue;
        float v = (half ? acc1[rr] : acc0[rr]) + (a.bias ? a.bias[n] : 0.f);
        if (a.act) v = fmaxf(v, 0.f);
        if (masked) v = 0.f;
        a.out[(long long)row * a.ldo + n] = v;
        if (a.out2) a.out2[(long long)row * a.ldo2 + n] = a.res[(long long)row * a.ldr + n] + v;
      }
    }
  }
}

void lin16(LinArgs a, const unsigned short* W16, int N, hipStream_t s) {
  if (a.rows <= 0) return;
  const int ntiles = (a.rows + 127) / 128;
  const int grid = std::min(ntiles, 4096);
  for (int c0 = 0; c0 < N; c0 += 64) {
    LinArgs16 b;
    b.a = a;
    b.a.ncols = std::min(64, N - c0);
    b.a.W = nullptr;
    b.W16 = W16 + c0;
    if (a.bias) b.a.bias = a.bias + c0;
    b.a.out = a.out + c0;
    if (a.res) b.a.res = a.res + c0;
    if (a.out2) b.a.out2 = a.out2 + c0;
    hipLaunchKernelGGL(lin_bf16_kernel, dim3(grid), dim3(256), 0, s, b);
    HFX_HIP(hipGetLastError());
  }
}

Seg seg(const float* p, const int* idx, int w, int ld) {
  Seg g;
  g.p = p;
  g.idx = idx;
  g.w = w;
  g.ld = ld;
  return g;
}

// Shared body of the two forward passes; `W16` selects the BF16 path.
void forward_impl(const LpcLayout& L, const float* P, const unsigned short* W16, const LpcGraph& g,
                  const LpcSystem& sys, LpcActs& acts, hipStream_t s) {
  const int H = LPC_H;
  auto run = [&](LinArgs a, long long woff, int N) {
    a.W = P + woff;
    if (W16) lin16(a, W16 + woff, N, s);
    else lin(a, N, s);
  };
  {
    LinArgs a;
    a.rows = g.nb;
    a.nseg = 2;
    a.seg[0] = seg(sys.ahat.p, g.diag.p, L.NE, L.NE);
    a.seg[1] = seg(g.dirf.p, nullptr, 1, 1);
    a.K = L.NI;
    a.ldw = H;
    a.bias = P + L.bn;
    a.act = 1;
    a.out = acts.h[0].p;
    a.ldo = H;
    run(a, L.Wn, H);
  }
  if (g.E > 0) {
    LinArgs a;
    a.rows = g.E;
    a.nseg = 1;
    a.seg[0] = seg(sys.ahat.p, g.eblk.p, L.NE, L.NE);
    a.K = L.NE;
    a.ldw = H;
    a.bias = P + L.be;
    a.act = 1;
    a.out = acts.e[0].p;
    a.ldo = H;
    run(a, L.We, H);
  }
  for (int l = 0; l < LPC_LAYERS; ++l) {
    if (g.E > 0) {
      LinArgs a;
      a.rows = g.E;
      a.nseg = 3;
      a.seg[0] = seg(acts.h[l].p, g.esrc.p, H, H);
      a.seg[1] = seg(acts.h[l].p, g.edst.p, H, H);
      a.seg[2] = seg(acts.e[l].p, nullptr, H, H);
      a.K = 3 * H;
      a.ldw = H;
      a.bias = P + L.bm[l];
      a.act = 1;
      a.out = acts.m[l].p;
      a.ldo = H;
      a.res = acts.e[l].p;
      a.ldr = H;
      a.out2 = acts.e[l + 1].p;
      a.ldo2 = H;
      run(a, L.Wm[l], H);
      rowsum(g.nb, g.erowptr.p, nullptr, acts.m[l].p, H, 0, acts.a[l].p, H, false, s);
    } else {
      fill((long long)g.nb * H, 0.f, acts.a[l].p, s);
    }
    LinArgs a;
    a.rows = g.nb;
    a.nseg = 2;
    a.seg[0] = seg(acts.h[l].p, nullptr, H, H);
    a.seg[1] = seg(acts.a[l].p, nullptr, H, H);
    a.K = 2 * H;
    a.ldw = H;
    a.bias = P + L.bu[l];
    a.act = 1;
    a.out = acts.u[l].p;
    a.ldo = H;
    a.res = acts.h[l].p;
    a.ldr = H;
    a.out2 = acts.h[l + 1].p;
    a.ldo2 = H;
    run(a, L.Wu[l], H);
  }
  {
    LinArgs a;
    a.rows = g.nb;
    a.nseg = 1;
    a.seg[0] = seg(acts.h[LPC_LAYERS].p, nullptr, H, H);
    a.K = H;
    a.ldw = L.O;
    a.bias = P + L.bdn;
    a.act = 0;
    a.out = acts.cn.p;
    a.ldo = L.O;
    a.rowmask = g.dir.p;
    run(a, L.Wdn, L.O);
  }
  if (g.E > 0) {
    LinArgs a;
    a.rows = g.E;
    a.nseg = 1;
    a.seg[0] = seg(acts.e[LPC_LAYERS].p, nullptr, H, H);
    a.K = H;
    a.ldw = L.O;
    a.bias = P + L.bde;
    a.act = 0;
    a.out = acts.ce.p;
    a.ldo = L.O;
    a.rowmask = g.dir.p;
    a.maskidx = g.esrc.p;
    run(a, L.Wde, L.O);
  }
}

}  // namespace

void LpcGradWork::alloc(int nb, int E, int O, long long maxKN) { LpcGradWork_alloc(*this, nb, E, O, maxKN); }

void LpcSystem::build(const LpcGraph& g, const double* vals, hipStream_t s) {
  const int bs = g.bs;
  sinv.alloc(g.nb);
  ahat.alloc((size_t)g.nnzb * bs * bs);
  dinv.alloc((size_t)g.nb * bs * bs);
  hipLaunchKernelGGL(scale_kernel, dim3(ceil_div(g.nb, 256)), dim3(256), 0, s, g.nb, bs, g.diag.p, vals, sinv.p);
  HFX_HIP(hipGetLastError());
  const long long nv = (long long)g.nnzb * bs * bs;
  hipLaunchKernelGGL(normalise_kernel, dim3(ceil_div(nv, 256)), dim3(256), 0, s, g.nnzb, bs, g.rowf.p, g.col.p, sinv.p,
                     vals, ahat.p);
  HFX_HIP(hipGetLastError());
#define HFX_BS(B)                                                                                                    \
  case B:                                                                                                            \
    hipLaunchKernelGGL((dinv_kernel<B>), dim3(ceil_div(g.nb, 64)), dim3(64), 0, s, g.nb, g.diag.p, sinv.p, vals, dinv.p); \
    break;
  switch (bs) {
    HFX_BS(2) HFX_BS(3) HFX_BS(4) HFX_BS(5) HFX_BS(6) HFX_BS(10)
    default: throw std::runtime_error("learned preconditioner: unsupported block size");
  }
#undef HFX_BS
  HFX_HIP(hipGetLastError());
}

void lpc_forward(const LpcLayout& L, const float* params, const LpcGraph& g, const LpcSystem& sys, LpcActs& acts,
                 hipStream_t s) {
  forward_impl(L, params, nullptr, g, sys, acts, s);
}

void lpc_forward_bf16(const LpcLayout& L, const float* params, const uint16_t* params16, const LpcGraph& g,
                      const LpcSystem& sys, LpcActs& acts, hipStream_t s) {
  forward_impl(L, params, reinterpret_cast<const unsigned short*>(params16), g, sys, acts, s);
}

void lpc_params_to_bf16(long long n, const float* p, uint16_t* q, hipStream_t s) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, p,
                     reinterpret_cast<unsigned short*>(q));
  HFX_HIP(hipGetLastError());
}

void lpc_spmv(const LpcGraph& g, const float* vals, const float* x, float* y, hipStream_t s) {
  const long long n = (long long)g.nb * g.bs;
#define HFX_BS(B)                                                                                                  \
  case B:                                                                                                          \
    hipLaunchKernelGGL((spmv_f32_kernel<B>), dim3(ceil_div(n, 256)), dim3(256), 0, s, g.nb, g.rowptr.p, g.col.p, vals, \
                       x, y);                                                                                      \
    break;
  switch (g.bs) {
    HFX_BS(2) HFX_BS(3) HFX_BS(4) HFX_BS(5) HFX_BS(6) HFX_BS(10)
    default: throw std::runtime_error("learned preconditioner: unsupported block size");
  }
#undef HFX_BS
  HFX_HIP(hipGetLastError());
}

void bsr_spmv_f32vals(const LpcGraph& g, const float* mvals, const double* x, double* y, hipStream_t s) {
  const long long n = (long long)g.nb * g.bs;
#define HFX_BS(B)                                                                                               \
  case B:                                                                                                       \
    hipLaunchKernelGGL((spmv_f32vals_kernel<B>), dim3(ceil_div(n, 256)), dim3(256), 0, s, g.nb, g.rowptr.p, g.col.p, \
                       mvals, x, y);                                                                            \
    break;
  switch (g.bs) {
    HFX_BS(2) HFX_BS(3) HFX_BS(4) HFX_BS(5) HFX_BS(6) HFX_BS(10)
    default: throw std::runtime_error("learned preconditioner: unsupported block size");
  }
#undef HFX_BS
  HFX_HIP(hipGetLastError());
}

void lpc_rademacher(long long n, unsigned long long seed, float* z, hipStream_t s) {
  hipLaunchKernelGGL(rademacher_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, seed, z);
  HFX_HIP(hipGetLastError());
}

int lpc_loss_partials(const LpcGraph& g) { return ceil_div((long long)g.nb * g.bs, 256); }

void lpc_loss(const LpcGraph& g, const LpcSystem& sys, const LpcActs& acts, const float* z, const float* v,
              float inv_nz2, float weight, float* dcn, float* dce, double* lossp, double* loss_acc, hipStream_t s) {
  const long long n = (long long)g.nb * g.bs;
  const int blocks = ceil_div(n, 256);
  const float scale = inv_nz2 * weight;
#define HFX_BS(B)                                                                                                   \
  case B:                                                                                                           \
    hipLaunchKernelGGL((loss_kernel<B>), dim3(blocks), dim3(256), 0, s, g.nb, g.rowptr.p, g.col.p, g.blk2edge.p,       \
                       g.dir.p, sys.dinv.p, acts.cn.p, acts.ce.p, z, v, scale, dcn, dce, lossp);                  \
    break;
  switch (g.bs) {
    HFX_BS(2) HFX_BS(3) HFX_BS(4) HFX_BS(5) HFX_BS(6) HFX_BS(10)
    default: throw std::runtime_error("learned preconditioner: unsupported block size");
  }
#undef HFX_BS
  HFX_HIP(hipGetLastError());
  hipLaunchKernelGGL(sum_into_kernel, dim3(1), dim3(256), 0, s, blocks, lossp, loss_acc);
  HFX_HIP(hipGetLastError());
}

void lpc_backward(const LpcLayout& L, const float* P, const LpcGraph& g, const LpcSystem& sys, const LpcActs& acts,
                  LpcGradWork& w, float* grad, hipStream_t s) {
  const int H = LPC_H, nb = g.nb, E = g.E, O = L.O;
  float* part = w.part.p;
  // decoders
  {
    WgArgs a;
    a.rows = nb;
    a.seg[0] = seg(acts.h[LPC_LAYERS].p, nullptr, H, H);
    a.K = H;
    a.dP = w.dcn.p;
    a.ldp = O;
    wgrad(a, O, grad + L.Wdn, grad + L.bdn, part, s);
    if (E > 0) {
      WgArgs b;
      b.rows = E;
      b.seg[0] = seg(acts.e[LPC_LAYERS].p, nullptr, H, H);
      b.K = H;
      b.dP = w.dce.p;
      b.ldp = O;
      wgrad(b, O, grad + L.Wde, grad + L.bde, part, s);
    } else {
      fill((long long)H * O + O, 0.f, grad + L.Wde, s);
    }
  }
  float* dh = w.dh.p;
  float* dh2 = w.dh2.p;
  float* de = w.de.p;
  float* de2 = w.de2.p;
  {
    transpose(H, O, P + L.Wdn, w.WT.p, s);
    LinArgs a;
    a.rows = nb;
    a.seg[0] = seg(w.dcn.p, nullptr, O, O);
    a.K = O;
    a.W = w.WT.p;
    a.ldw = H;
    a.out = dh;
    a.ldo = H;
    lin(a, H, s);
    if (E > 0) {
      transpose(H, O, P + L.Wde, w.WT.p, s);
      LinArgs b;
      b.rows = E;
      b.seg[0] = seg(w.dce.p, nullptr, O, O);
      b.K = O;
      b.W = w.WT.p;
      b.ldw = H;
      b.out = de;
      b.ldo = H;
      lin(b, H, s);
    }
  }
  for (int l = LPC_LAYERS - 1; l >= 0; --l) {
    // node update: u = relu([h | a] Wu + bu), h' = h + u
    hipLaunchKernelGGL(mask_mul_kernel, dim3(ceil_div((long long)nb * H, 256)), dim3(256), 0, s, (long long)nb * H, dh,
                       acts.u[l].p, w.dP.p);
    HFX_HIP(hipGetLastError());
    {
      WgArgs a;
      a.rows = nb;
      a.nseg = 2;
      a.seg[0] = seg(acts.h[l].p, nullptr, H, H);
      a.seg[1] = seg(acts.a[l].p, nullptr, H, H);
      a.K = 2 * H;
      a.dP = w.dP.p;
      a.ldp = H;
      wgrad(a, H, grad + L.Wu[l], grad + L.bu[l], part, s);
    }
    transpose(2 * H, H, P + L.Wu[l], w.WT.p, s);
    {
      LinArgs a;
      a.rows = nb;
      a.seg[0] = seg(w.dP.p, nullptr, H, H);
      a.K = H;
      a.W = w.WT.p;
      a.ldw = 2 * H;
      a.out = w.dXn.p;
      a.ldo = 2 * H;
      lin(a, 2 * H, s);
    }
    hipLaunchKernelGGL(add_strided_kernel, dim3(ceil_div((long long)nb * H, 256)), dim3(256), 0, s, nb, dh, w.dXn.p,
                       2 * H, 0, dh2);
    HFX_HIP(hipGetLastError());
    if (E > 0) {
      // message: m = relu([h_F | h_G | e] Wm + bm), e' = e + m, a_F = sum m
      hipLaunchKernelGGL(gather_add_mask_kernel, dim3(ceil_div((long long)E * H, 256)), dim3(256), 0, s, E, de,
                         w.dXn.p + H, 2 * H, g.esrc.p, acts.m[l].p, w.dP.p);
      HFX_HIP(hipGetLastError());
      {
        WgArgs a;
        a.rows = E;
        a.nseg = 3;
        a.seg[0] = seg(acts.h[l].p, g.esrc.p, H, H);
        a.seg[1] = seg(acts.h[l].p, g.edst.p, H, H);
        a.seg[2] = seg(acts.e[l].p, nullptr, H, H);
        a.K = 3 * H;
        a.dP = w.dP.p;
        a.ldp = H;
        wgrad(a, H, grad + L.Wm[l], grad + L.bm[l], part, s);
      }
      transpose(3 * H, H, P + L.Wm[l], w.WT.p, s);
      {
        LinArgs a;
        a.rows = E;
        a.seg[0] = seg(w.dP.p, nullptr, H, H);
        a.K = H;
        a.W = w.WT.p;
        a.ldw = 3 * H;
        a.out = w.dX.p;
        a.ldo = 3 * H;
        lin(a, 3 * H, s);
      }
      rowsum(nb, g.erowptr.p, nullptr, w.dX.p, 3 * H, 0, dh2, H, true, s);
      rowsum(nb, g.erowptr.p, g.erev.p, w.dX.p, 3 * H, H, dh2, H, true, s);
      hipLaunchKernelGGL(add_strided_kernel, dim3(ceil_div((long long)E * H, 256)), dim3(256), 0, s, E, de, w.dX.p,
                         3 * H, 2 * H, de2);
      HFX_HIP(hipGetLastError());
    } else {
      fill(3LL * H * H + H, 0.f, grad + L.Wm[l], s);
    }
    std::swap(dh, dh2);
    std::swap(de, de2);
  }
  // encoders
  if (E > 0) {
    hipLaunchKernelGGL(mask_mul_kernel, dim3(ceil_div((long long)E * H, 256)), dim3(256), 0, s, (long long)E * H, de,
                       acts.e[0].p, w.dP.p);
    HFX_HIP(hipGetLastError());
    WgArgs a;
    a.rows = E;
    a.seg[0] = seg(sys.ahat.p, g.eblk.p, L.NE, L.NE);
    a.K = L.NE;
    a.dP = w.dP.p;
    a.ldp = H;
    wgrad(a, H, grad + L.We, grad + L.be, part, s);
  } else {
    fill((long long)L.NE * H + H, 0.f, grad + L.We, s);
  }
  hipLaunchKernelGGL(mask_mul_kernel, dim3(ceil_div((long long)nb * H, 256)), dim3(256), 0, s, (long long)nb * H, dh,
                     acts.h[0].p, w.dP.p);
  HFX_HIP(hipGetLastError());
  WgArgs a;
  a.rows = nb;
  a.nseg = 2;
  a.seg[0] = seg(sys.ahat.p, g.diag.p, L.NE, L.NE);
  a.seg[1] = seg(g.dirf.p, nullptr, 1, 1);
  a.K = L.NI;
  a.dP = w.dP.p;
  a.ldp = H;
  wgrad(a, H, grad + L.Wn, grad + L.bn, part, s);
}

void lpc_adam(long long n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
              int t, hipStream_t s) {
  const float c1 = 1.f - std::pow(b1, (float)t), c2 = 1.f - std::pow(b2, (float)t);
  hipLaunchKernelGGL(adam_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, p, g, m, v, lr, b1, b2, eps, c1, c2);
  HFX_HIP(hipGetLastError());
}

void lpc_build_m(const LpcGraph& g, const LpcSystem& sys, const LpcActs& acts, float* mvals, hipStream_t s) {
  const long long n = (long long)g.nnzb * g.bs * g.bs;
  hipLaunchKernelGGL(build_m_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, g.nnzb, g.bs, g.rowf.p, g.col.p,
                     g.blk2edge.p, sys.sinv.p, sys.dinv.p, acts.cn.p, acts.ce.p, mvals);
  HFX_HIP(hipGetLastError());
}

}  // namespace hfx
