// Block ILU(0): level-scheduled factorisation and triangular sweeps on the
// GPU.  See ilu.h for the algorithm.
#include <stdexcept>
#include <string>

#include "bsr_row.h"
#include "ilu.h"

#include <algorithm>

namespace hfx {

namespace {

constexpr int FT = 64;   // factorisation: one wavefront per block row
constexpr int AT = 256;  // sweeps: one thread per scalar row

// Factor the block rows rows[0..gridDim.x) of one forward level.  Row i is
// staged in LDS (len x bs^2), updated by the already factored rows of earlier
// levels (lu, dinv; each of their blocks staged in LDS with one coalesced load
// before its product), its diagonal block inverted by Gauss-Jordan with
// partial pivoting (the pivot found by a wavefront argmax: the first row of
// largest magnitude, the diagonal's own row when it is NaN), and written to lu
// (earlier blocks L_ik, diagonal A'_ii, later blocks Uhat_ij) and dinv.
// err[0] receives 1 + row on a zero pivot.
__global__ __launch_bounds__(FT) void ilu_factor_kernel(int bs, const int* __restrict__ rows,
                                                        const int* __restrict__ rowptr, const double* __restrict__ vals,
                                                        const int* __restrict__ lptr, const int* __restrict__ lpos,
                                                        const int* __restrict__ lcol, const int* __restrict__ pptr,
                                                        const int* __restrict__ pslot, const int* __restrict__ ppos,
                                                        const int* __restrict__ uptr, const int* __restrict__ upos,
                                                        const int* __restrict__ dslot,
                                                        const signed char* __restrict__ kind, double* __restrict__ lu,
                                                        double* __restrict__ dinv, int* __restrict__ err) {
  extern __shared__ double sm[];
  const int i = rows[blockIdx.x];
  const int r0 = rowptr[i], len = rowptr[i + 1] - r0, B2 = bs * bs, t = threadIdx.x;
  double* R = sm;             // the row's blocks
  double* M = R + len * B2;   // diagonal block being inverted
  double* V = M + B2;         // its inverse
  double* T = V + B2;         // scratch block
  double* W = T + B2;         // staged block of an earlier row (Uhat_kj or Dinv_k)
  for (int e = t; e < len * B2; e += FT) R[e] = vals[(size_t)r0 * B2 + e];
  __syncthreads();
  for (int l = lptr[i]; l < lptr[i + 1]; ++l) {
    const int k = lcol[l], sl = lpos[l] - r0;
    const double* Rk = R + sl * B2;
    // A'_ij -= A'_ik Uhat_kj  (slot j != slot k: no lane reads what another writes)
    for (int q = pptr[l]; q < pptr[l + 1]; ++q) {
      const double* Ug = lu + (size_t)ppos[q] * B2;
      for (int e = t; e < B2; e += FT) W[e] = Ug[e];
      __syncthreads();
      double* Rj = R + pslot[q] * B2;
      for (int e = t; e < B2; e += FT) {
        const int r = e / bs, c = e - r * bs;
        double acc = 0.0;
        for (int m = 0; m < bs; ++m) acc += Rk[r * bs + m] * W[m * bs + c];
        Rj[e] -= acc;
      }
      __syncthreads();  // W is restaged by the next product
    }
    // L_ik = A'_ik Dinv_k
    const double* Dk = dinv + (size_t)k * B2;
    for (int e = t; e < B2; e += FT) W[e] = Dk[e];
    __syncthreads();
    for (int e = t; e < B2; e += FT) {
      const int r = e / bs, c = e - r * bs;
      double acc = 0.0;
      for (int m = 0; m < bs; ++m) acc += Rk[r * bs + m] * W[m * bs + c];
      T[e] = acc;
    }
    __syncthreads();
    for (int e = t; e < B2; e += FT) R[sl * B2 + e] = T[e];
    __syncthreads();
  }
  // Dinv_i = (A'_ii)^-1, Gauss-Jordan with partial pivoting
  const int ds = dslot[i];
  for (int e = t; e < B2; e += FT) {
    const int r = e / bs, c = e - r * bs;
    M[e] = R[ds * B2 + e];
    V[e] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int c = 0; c < bs; ++c) {
    // pivot: the first row r >= c of largest |M[r][c]| (NaN rows never win);
    // the diagonal's row when |M[c][c]| is NaN
    const double dc = fabs(M[c * bs + c]);
    double v = -1.0;
    if (t >= c && t < bs) {
      v = fabs(M[t * bs + c]);
      if (v != v) v = -1.0;
    }
    int idx = t;
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(v, off, 64);
      const int oi = __shfl_xor(idx, off, 64);
      if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
    }
    const int p = (dc != dc) ? c : idx;
    const double best = (dc != dc) ? dc : v;
    if (t == 0 && (!(best > 0.0) || !isfinite(best))) err[0] = i + 1;
    if (p != c && t < bs) {
      double a = M[c * bs + t];
      M[c * bs + t] = M[p * bs + t];
      M[p * bs + t] = a;
      a = V[c * bs + t];
      V[c * bs + t] = V[p * bs + t];
      V[p * bs + t] = a;
    }
    __syncthreads();
    const double inv = 1.0 / M[c * bs + c];
    __syncthreads();
    if (t < bs) {
      M[c * bs + t] *= inv;
      V[c * bs + t] *= inv;
    }
    __syncthreads();
    double f[4];  // bs <= 16: at most 4 entries per lane
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = t + u * FT;
      f[u] = (e < B2 && e / bs != c) ? M[(e / bs) * bs + c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = t + u * FT;
      if (e < B2 && e / bs != c) {
        const int j = e - (e / bs) * bs;
        M[e] -= f[u] * M[c * bs + j];
        V[e] -= f[u] * V[c * bs + j];
      }
    }
    __syncthreads();
  }
  for (int e = t; e < B2; e += FT) dinv[(size_t)i * B2 + e] = V[e];
  // earlier blocks and the diagonal as they are; later blocks scaled by Dinv_i
  for (int e = t; e < len * B2; e += FT)
    if (kind[r0 + e / B2] <= 0) lu[(size_t)r0 * B2 + e] = R[e];
  for (int u = uptr[i]; u < uptr[i + 1]; ++u) {
    const double* Rj = R + (upos[u] - r0) * B2;
    for (int e = t; e < B2; e += FT) {
      const int r = e / bs, c = e - r * bs;
      double acc = 0.0;
      for (int m = 0; m < bs; ++m) acc += V[r * bs + m] * Rj[m * bs + c];
      lu[(size_t)upos[u] * B2 + e] = acc;
    }
  }
}

// Forward sweep over one level: y_i = x_i - sum_k L_ik y_k.
template <int BS, typename V>
__global__ __launch_bounds__(AT) void ilu_fwd_kernel(int nrows, const int* __restrict__ rows,
                                                     const int* __restrict__ lptr, const int* __restrict__ lpos,
                                                     const int* __restrict__ lcol, const V* __restrict__ lu,
                                                     const double* __restrict__ x, double* __restrict__ y) {
  const int t = blockIdx.x * AT + threadIdx.x;
  if (t >= nrows * BS) return;
  const int i = rows[t / BS], r = t % BS;
  y[(size_t)i * BS + r] = x[(size_t)i * BS + r] - bsr_row_sum<BS, V>(lptr[i], lptr[i + 1], r, lpos, lcol, lu, y);
}

// Backward sweep over one level: z_i = Dinv_i y_i - sum_j Uhat_ij z_j.
template <int BS, typename V>
__global__ __launch_bounds__(AT) void ilu_bwd_kernel(int nrows, const int* __restrict__ rows,
                                                     const int* __restrict__ uptr, const int* __restrict__ upos,
                                                     const int* __restrict__ ucol, const V* __restrict__ lu,
                                                     const V* __restrict__ dinv, const double* __restrict__ y,
                                                     double* __restrict__ z) {
  const int t = blockIdx.x * AT + threadIdx.x;
  if (t >= nrows * BS) return;
  const int i = rows[t / BS], r = t % BS;
  z[(size_t)i * BS + r] = bsr_row_dot<BS, V>(dinv + ((size_t)i * BS + r) * BS, y + (size_t)i * BS) -
                          bsr_row_sum<BS, V>(uptr[i], uptr[i + 1], r, upos, ucol, lu, z);
}

__global__ void to_float_kernel(size_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

template <int BS, typename V>
void sweeps_v(const Ilu& f, const V* lu, const V* dinv, const double* x, double* y, double* z, hipStream_t s) {
  const IluPlan& P = f.plan;
  for (int l = 0; l < P.nlev_f; ++l) {
    const int n = P.fwd_off[l + 1] - P.fwd_off[l];
    hipLaunchKernelGGL((ilu_fwd_kernel<BS, V>), dim3(ceil_div((long long)n * BS, AT)), dim3(AT), 0, s, n,
                       f.fwd_rows.p + P.fwd_off[l], f.lptr.p, f.lpos.p, f.lcol.p, lu, x, y);
  }
  for (int l = 0; l < P.nlev_b; ++l) {
    const int n = P.bwd_off[l + 1] - P.bwd_off[l];
    hipLaunchKernelGGL((ilu_bwd_kernel<BS, V>), dim3(ceil_div((long long)n * BS, AT)), dim3(AT), 0, s, n,
                       f.bwd_rows.p + P.bwd_off[l], f.uptr.p, f.upos.p, f.ucol.p, lu, dinv, y, z);
  }
  HFX_HIP(hipGetLastError());
}

template <int BS>
void sweeps(const Ilu& f, const double* x, double* y, double* z, hipStream_t s) {
  if (f.low) sweeps_v<BS, float>(f, f.lu32.p, f.dinv32.p, x, y, z, s);
  else sweeps_v<BS, double>(f, f.lu.p, f.dinv.p, x, y, z, s);
}

}  // namespace

void Ilu::setup(int nb_, int bs_, const int* rowptr_host, const int* col_host, int ordering_, hipStream_t s) {
  // the block sizes the sweeps of apply() are instantiated for; refused here,
  // before the analysis and factorisation, rather than at the first apply
  if (!(bs_ >= 1 && bs_ <= 12 && bs_ != 7 && bs_ != 11))
    throw std::runtime_error("ILU(0): block size must be one of 1-6, 8, 9, 10, 12");
  plan = ilu_analyse(nb_, rowptr_host, col_host, ordering_);
  nb = nb_;
  bs = bs_;
  ordering = ordering_;
  const size_t lds = ((size_t)plan.maxlen + 4) * bs * bs * sizeof(double);
  if (lds > 65536) throw std::runtime_error("ILU(0): a block row is too long for the factorisation kernel");
  rowptr.upload(rowptr_host, nb + 1, s);
  fwd_rows.upload(plan.fwd_rows, s);
  bwd_rows.upload(plan.bwd_rows, s);
  lptr.upload(plan.lptr, s);
  lpos.upload(plan.lpos, s);
  lcol.upload(plan.lcol, s);
  uptr.upload(plan.uptr, s);
  upos.upload(plan.upos, s);
  ucol.upload(plan.ucol, s);
  dslot.upload(plan.dslot, s);
  pptr.upload(plan.pptr, s);
  pslot.upload(plan.pslot, s);
  ppos.upload(plan.ppos, s);
  kind.upload(plan.kind, s);
  err.alloc(1);
  lu.alloc((size_t)rowptr_host[nb] * bs * bs);
  dinv.alloc((size_t)nb * bs * bs);
  tmp.alloc((size_t)nb * bs);
  HFX_HIP(hipStreamSynchronize(s));
}

void Ilu::factor(const BsrDev& A, hipStream_t s) {
  if (A.nb != nb || A.bs != bs || (size_t)A.nnzb * bs * bs != lu.n)
    throw std::runtime_error("ILU(0): matrix does not match the analysed pattern");
  HFX_HIP(hipMemsetAsync(err.p, 0, sizeof(int), s));
  const size_t lds = ((size_t)plan.maxlen + 4) * bs * bs * sizeof(double);
  for (int l = 0; l < plan.nlev_f; ++l) {
    const int n = plan.fwd_off[l + 1] - plan.fwd_off[l];
    hipLaunchKernelGGL(ilu_factor_kernel, dim3(n), dim3(FT), lds, s, bs, fwd_rows.p + plan.fwd_off[l], rowptr.p,
                       A.vals.p, lptr.p, lpos.p, lcol.p, pptr.p, pslot.p, ppos.p, uptr.p, upos.p, dslot.p, kind.p,
                       lu.p, dinv.p, err.p);
  }
  HFX_HIP(hipGetLastError());
  int e = 0;
  HFX_HIP(hipMemcpyAsync(&e, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HFX_HIP(hipStreamSynchronize(s));
  if (e) throw std::runtime_error("ILU(0): singular pivot block in block row " + std::to_string(e - 1));
  if (low) {
    lu32.alloc(lu.n);
    dinv32.alloc(dinv.n);
    hipLaunchKernelGGL(to_float_kernel, dim3(std::min<long long>(ceil_div((long long)lu.n, 256), 8192)), dim3(256), 0,
                       s, lu.n, lu.p, lu32.p);
    hipLaunchKernelGGL(to_float_kernel, dim3(std::min<long long>(ceil_div((long long)dinv.n, 256), 8192)), dim3(256),
                       0, s, dinv.n, dinv.p, dinv32.p);
    HFX_HIP(hipGetLastError());
  }
}

void Ilu::apply(const double* x, double* y, hipStream_t s) {
  switch (bs) {
    case 1: sweeps<1>(*this, x, tmp.p, y, s); break;
    case 2: sweeps<2>(*this, x, tmp.p, y, s); break;
    case 3: sweeps<3>(*this, x, tmp.p, y, s); break;
    case 4: sweeps<4>(*this, x, tmp.p, y, s); break;
    case 5: sweeps<5>(*this, x, tmp.p, y, s); break;
    case 6: sweeps<6>(*this, x, tmp.p, y, s); break;
    case 8: sweeps<8>(*this, x, tmp.p, y, s); break;
    case 9: sweeps<9>(*this, x, tmp.p, y, s); break;
    case 10: sweeps<10>(*this, x, tmp.p, y, s); break;
    case 12: sweeps<12>(*this, x, tmp.p, y, s); break;
    default: throw std::runtime_error("ILU(0): unsupported block size for the sweeps");
  }
}

}  // namespace hfx
