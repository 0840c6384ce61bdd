// Fused HDG assembly + static condensation for affine (straight-sided)
// simplices, FP64, one wavefront per element.
//
// With a constant Jacobian J the mass and gradient blocks are reference
// matrices scaled by det J and J^-1 (Operator.cpp:14-84 evaluated once per
// element), so Sqq^-1 [Squ | Sql] needs no per-element factorisation:
//   A_k = sum_a Jinv[k][a] Ahat_a,          Ahat_a = Mhat^-1 Chat_a
//   B_k[:, (g,b)] = -(|f_g| / det J) n_gk G_g[:, b],  G_g = Mhat^-1[:, fn_g] Mfhat
// (HDGBase.cpp:106-157 with Mass.cpp:5-38 on flat faces).  The remaining
// element work is the variable-coefficient integrals at the cubature points
// (HDGDiffusion.cpp:31-145, HDGConvection.cpp:31-104, Source.cpp:5-48), one
// LU of K = Suu - sum_k Suq_k A_k, and the trace-column solves of
// HDGSolver.cpp:331-348:
//   U = -K^-1 (Sul - sum_k Suq_k B_k), U0 = K^-1 F_u, Q_k = -A_k U - B_k,
//   S = Slu U + sum_k Slq_k Q_k + Sll, S0 = -Slu U0 - sum_k Slq_k Q0_k.
//
// Work mapping on the 64-wide wavefront: the element-integral phases spread
// matrix entries over the lanes; the condensation gives each lane one trace
// column c (plus one lane for the F_u column) and keeps that column of U and
// Q in registers, reading the shared factors (LU of K, Ahat, face masses)
// with uniform, broadcast LDS addresses.  Waves never wait on each other:
// the only workgroup barrier is the one after the reference tables are staged.
#include "hdg_element.h"
#include "hfx_internal.h"

namespace hfx {

template <int D, int P>
struct RefConst;
#include "data/refel_const.inc"

namespace {

typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef double dvec2 __attribute__((ext_vector_type(2)));

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Scheduling hint for a fully unrolled block of NM MFMAs fed by ND LDS
// reads: alternate one MFMA with its share of the reads and a few VALU
// instructions, so that the operands of the next MFMAs are fetched while the
// current ones run (one wave per SIMD: nothing else hides the LDS latency).
template <int NM, int ND>
__device__ __forceinline__ void mfma_interleave() {
  constexpr int RD = (ND + NM - 1) / NM;
#pragma unroll
  for (int i = 0; i < NM; ++i) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x100, RD, 0);
    __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);
  }
}

// 1 / x from the hardware estimate and two Newton steps (within an ulp or so
// of the rounded quotient): a shorter dependent chain than the IEEE division
// sequence, on the pivot path of the LU
__device__ inline double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// value of v in lane l (l uniform across the wave)
__device__ inline double readlane_d(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

template <int D, int P, bool TD>
struct ALay {
  using Z = ESz<D, P>;
  static constexpr int NN = Z::NN, NNF = Z::NNF, NIP = Z::NIP, NIPF = Z::NIPF, NF = Z::NF, NL = Z::NL;
  static constexpr int ND = TD ? D * D : 1;      // stored diffusion components per node
  static constexpr int NS = 2 + (TD ? D : 1);    // face weights: tau, b.n, D (or (D n)_k)
  static constexpr int NCOL = NL + 1;            // trace columns + the F_u column
  // shared reference tables (packed by the host in this order)
  static constexpr int T_PHI = 0;
  static constexpr int T_DPHI = T_PHI + NIP * NN;
  static constexpr int T_FPHI = T_DPHI + NIP * NN * D;
  static constexpr int T_W = T_FPHI + NIPF * NNF;
  static constexpr int T_FW = T_W + NIP;
  static constexpr int T_AH = T_FW + NIPF;
  static constexpr int T_G = T_AH + D * NN * NN;
  static constexpr int TAB_ALL = T_G + NF * NN * NNF;
  // all staged in LDS, shared by the waves; the per-wave scratch starts at
  // an even offset (with O_KM and PER even, the LU factors are 16-byte
  // aligned for paired reads)
  static constexpr int TAB = (TAB_ALL + 1) & ~1;
  // per-wave scratch
  static constexpr int NSF = NS;                  // LDS face masses: tau, b.n, D
  static constexpr int O_FM = 0;
  static constexpr int O_SUQ = (O_FM + NF * NSF * NNF * NNF + 1) & ~1;
  static constexpr int O_KM = (O_SUQ + D * NN * NN + 1) & ~1;  // Suu, then K, then its LU factors
  // after the U solve, Suq and the K factors are dead: the face traces n.q of the Q
  // columns take their place
  static constexpr int NQN = NF * NNF * NCOL;
  static constexpr int END_SK = O_KM + NN * NN;
  static constexpr int O_FU = (O_SUQ + NQN > END_SK) ? O_SUQ + NQN : END_SK;
  static constexpr int O_FG = O_FU + NN;          // per face: n (D), |f|
  static constexpr int O_VX = O_FG + NF * (D + 1);
  static constexpr int O_FWT = O_VX + (D + 1) * D;  // face-IP weights (kept for the b.n columns)
  static constexpr int O_RD = O_FWT + NF * NIPF * NS;  // 1 / pivot of each row of U (K = P^T L U)
  // per-cell indices (ints) for the fused trace scatter, gathered with the
  // cell: face-node positions, BSR block of each face pair, face ids, side |
  // Dirichlet flags
  static constexpr int IX_PERM = 0;
  static constexpr int IX_BPOS = IX_PERM + NF * NNF;
  static constexpr int IX_FACE = IX_BPOS + NF * NF;
  static constexpr int IX_FLAG = IX_FACE + NF;
  static constexpr int NIX = IX_FLAG + NF;
  static constexpr int O_PL = O_RD + NN;             // pivot row of each LU step (ints)
  static constexpr int O_IX = O_PL + (NN + 1) / 2;
  static constexpr int O_OV = O_IX + (NIX + 1) / 2;
  // overlay, phase 1
  static constexpr int O_DN = O_OV;
  static constexpr int O_BN = O_DN + NN * ND;
  static constexpr int O_TAU = O_BN + NN * D;
  static constexpr int O_FQ = O_TAU + NF * NNF;
  static constexpr int O_RQ = O_FQ + NIP;
  static constexpr int O_SD = O_RQ + NIP;          // w det D at each cubature point
  static constexpr int O_BJ = O_SD + NIP;
  static constexpr int O_WR = O_BJ + NIP * D;
  static constexpr int O_WF = O_WR + NIP;
  static constexpr int END1 = O_WF + NIP;
  // overlay, phase 2: R columns, overwritten in place by the U columns
  static constexpr int O_RB = O_OV;
  static constexpr int END2 = O_RB + NN * NCOL;
  static constexpr int PER = ((END1 > END2 ? END1 : END2) + 1) & ~1;
  static_assert(NF * NIPF * NS >= NCOL, "the face-IP weights hold one dead row of trace columns");
};

// Wave-wide argmax of a non-negative value over the valid lanes among 0..31;
// ties go to the lowest lane.  The key keeps the top 26 bits of |v| as a
// float and the lane in the low 6 bits; a DPP prefix max inside each 16-lane
// row (row_shr 1, 2, 4, 8) leaves each row's maximum in its last lane, and
// two readlanes combine rows 0 and 1.
__device__ inline int dpp_argmax(double v, bool valid) {
  // branch-free, so that the LU's steps stay one basic block the compiler
  // can overlap
  float fv = (float)v;
  fv = fv >= 0.0f ? fv : 0.0f;
  unsigned key = ((__float_as_uint(fv) & ~63u) | (63u - (unsigned)(threadIdx.x & 63))) + (1u << 6);
  key = valid ? key : 0u;
  unsigned m = key;
  m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x111, 0xF, 0xF, true));
  m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x112, 0xF, 0xF, true));
  m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x114, 0xF, 0xF, true));
  m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x118, 0xF, 0xF, true));
  const unsigned b0 = (unsigned)__builtin_amdgcn_readlane((int)m, 15);
  const unsigned b1 = (unsigned)__builtin_amdgcn_readlane((int)m, 31);
  const unsigned best = b0 > b1 ? b0 : b1;
  return 63 - (int)(best & 63u);
}

// Next-cell gather held in registers, one entry per lane: fetch_idx issues
// the loads that depend on the cell number only (cell nodes, faces, face
// positions, BSR blocks, cubature-point data), fetch_data the loads through
// those indices (coordinates, D, b, tau on both sides, face owners,
// Dirichlet flags).  A wave issues fetch_idx for its next cell once the
// current cell's gather is in LDS and fetch_data late in the current cell,
// so both land while the current cell computes and the top of the next
// iteration only writes them to LDS.
template <int D, int P>
struct GatherPf {
  using Z = ESz<D, P>;
  static constexpr int NN = Z::NN, NNF = Z::NNF, NIP = Z::NIP, NF = Z::NF;
  static_assert((D + 1) * D <= 64 && NN * D <= 64 && NF * NNF <= 64 && NF * NF <= 64 && NIP <= 64,
                "one gathered entry per lane");
  int cn, cd, fc, pm, bp, fx;  // node of entry lane / D, node lane; face, position of entry lane; block; face lane
  double fq, rq, vx, dn, bn, t0, t1;
  int s2, s2x, dr;             // owner cell of face fc and of face fx, Dirichlet flag of fx

  __device__ __forceinline__ void fetch_idx(const ElemArgs& a, int e, int lane, bool fused) {
    const int* cell = a.cells + (size_t)e * NN;
    if (lane < NN * D) cn = cell[lane / D];
    if (lane < NN && a.ndD == 1) cd = cell[lane];
    if (lane < NF * NNF) {
      fc = a.cell2face[(size_t)e * NF + lane / NNF];
      pm = a.face_perm[(size_t)e * NF * NNF + lane];
    }
    if (fused) {
      if (lane < NF * NF) bp = a.bpos[(size_t)e * NF * NF + lane];
      if (lane < NF) fx = a.cell2face[(size_t)e * NF + lane];
    }
    fq = (lane < NIP && a.hasF && a.kind != 0) ? a.f_ip[(size_t)e * NIP + lane] : 0.0;
    rq = (lane < NIP && a.hasR) ? a.r_ip[(size_t)e * NIP + lane] : 0.0;
  }
  __device__ __forceinline__ void fetch_data(const ElemArgs& a, int lane, bool fused) {
    if (lane < (D + 1) * D) vx = a.coords[(size_t)cn * D + lane % D];
    if (lane < NN * D && a.hasB) bn = a.bn[(size_t)cn * D + lane % D];
    if (lane < NN && a.ndD == 1) dn = a.Dn[cd];
    if (lane < NF * NNF) {
      const size_t ti = ((size_t)fc * NNF + pm) * a.tau_sides;
      t0 = a.tau[ti];
      if (a.tau_sides == 2) {
        t1 = a.tau[ti + 1];
        s2 = a.face2cell[2 * (size_t)fc];
      }
    }
    if (fused && lane < NF) {
      s2x = a.face2cell[2 * (size_t)fx];
      dr = a.is_dir ? a.is_dir[fx] : 0;
    }
  }
};

template <int D, int P, bool TD, int W>
__device__ __forceinline__ void affine_load_tables(const double* __restrict__ tabs, double* sm, int* s_pos,
                                                   int* s_fn) {
  using L = ALay<D, P, TD>;
  using RC = RefConst<D, P>;
  for (int t = threadIdx.x; t < L::TAB_ALL; t += 64 * W) sm[t] = tabs[t];
  // s_pos[i]: byte f = position of node i on face f, 0xFF when i is not on f
  static_assert(L::NF <= 4 && L::NNF < 255, "face positions packed in one word");
  for (int t = threadIdx.x; t < L::NN; t += 64 * W) {
    unsigned w = 0u;
    for (int f = 0; f < L::NF; ++f) {
      const int ps = RC::POS[f][t];
      w |= (unsigned)(ps >= 0 ? ps : 0xFF) << (8 * f);
    }
    s_pos[t] = (int)w;
  }
  for (int t = threadIdx.x; t < L::NF * L::NNF; t += 64 * W) s_fn[t] = RC::FN[t / L::NNF][t % L::NNF];
  __syncthreads();
}

// Cells e = e_first + wave, e_first + wave + e_stride, ... below e_end of one
// problem, one wavefront per cell, with the tables already staged: assembles
// and condenses each cell.  Shared by the per-problem kernel and the batched
// kernels.
template <int D, int P, bool TD, int W, bool TSG = false>
__device__ __forceinline__ void affine_cells(const ElemArgs& a, int e_first, int e_end, int e_stride, double* sm,
                                             int* s_pos, int* s_fn) {
  using L = ALay<D, P, TD>;
  constexpr int NN = L::NN, NNF = L::NNF, NIP = L::NIP, NIPF = L::NIPF, NF = L::NF, NL = L::NL, NS = L::NS;
  constexpr int NCOL = L::NCOL, ND = L::ND;
  static_assert(!TD, "tensor diffusion runs on the general element kernel");

  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;

  const double* PHI = sm + L::T_PHI;
  const double* DPHI = sm + L::T_DPHI;
  const double* FPHI = sm + L::T_FPHI;
  const double* WQ = sm + L::T_W;
  const double* FWQ = sm + L::T_FW;
  const double* AH = sm + L::T_AH;  // read with uniform (broadcast) addresses
  const double* GT = sm + L::T_G;
  double* ws = sm + L::TAB + wv * L::PER;
  double* FM = ws + L::O_FM;
  double* SUQ = ws + L::O_SUQ;
  double* KM = ws + L::O_KM;
  double* FU = ws + L::O_FU;
  double* FG = ws + L::O_FG;
  double* VX = ws + L::O_VX;
  double* DN = ws + L::O_DN;
  double* BN = ws + L::O_BN;
  double* TAU = ws + L::O_TAU;
  double* FQ = ws + L::O_FQ;
  double* RQ = ws + L::O_RQ;
  double* SD = ws + L::O_SD;
  double* BJ = ws + L::O_BJ;
  double* WR = ws + L::O_WR;
  double* WF = ws + L::O_WF;
  double* FWT = ws + L::O_FWT;
  double* RB = ws + L::O_RB;
  int* IX = reinterpret_cast<int*>(ws + L::O_IX);
  const bool fused = a.bvals != nullptr;

  // The gather is software-pipelined across a wave's cells (GatherPf) at the
  // orders that run one wave per SIMD anyway; below them the prefetch
  // registers would cost the second wave that hides the gather already.
  GatherPf<D, P> pf;
  constexpr bool PIPE = D == 3 ? P >= 2 : P >= 4;
  if (PIPE && e_first + wv < e_end) {
    pf.fetch_idx(a, e_first + wv, lane, fused);
    pf.fetch_data(a, lane, fused);
  }
  for (int e = e_first + wv; e < e_end; e += e_stride) {
    // ------------------------------------------------------------ gather
    if constexpr (!PIPE) {
      const int* cell = a.cells + (size_t)e * NN;
      for (int t = lane; t < (D + 1) * D; t += 64) VX[t] = a.coords[(size_t)cell[t / D] * D + t % D];
      for (int t = lane; t < NN * ND; t += 64) {
        const int i = t / ND, c = t % ND;
        double v;
        (void)c;
        v = a.ndD == 1 ? a.Dn[cell[i]] : (a.useDiff ? 1.0 : 0.0);
        DN[t] = v;
      }
      for (int t = lane; t < NN * D; t += 64) BN[t] = a.hasB ? a.bn[(size_t)cell[t / D] * D + t % D] : 0.0;
      for (int t = lane; t < NF * NNF; t += 64) {
        const int f = t / NNF, j = t % NNF;
        const int fc = a.cell2face[(size_t)e * NF + f];
        const int side = (a.tau_sides == 2 && a.face2cell[2 * (size_t)fc] != e) ? 1 : 0;
        const int p = a.face_perm[((size_t)e * NF + f) * NNF + j];
        TAU[t] = a.tau[((size_t)fc * NNF + p) * a.tau_sides + side];
        IX[L::IX_PERM + t] = p;
      }
      if (fused) {
        for (int t = lane; t < NF * NF; t += 64) IX[L::IX_BPOS + t] = a.bpos[(size_t)e * NF * NF + t];
        for (int t = lane; t < NF; t += 64) {
          const int fc = a.cell2face[(size_t)e * NF + t];
          IX[L::IX_FACE + t] = fc;
          IX[L::IX_FLAG + t] = (a.face2cell[2 * (size_t)fc] == e ? 0 : 1) | (a.is_dir && a.is_dir[fc] ? 2 : 0);
        }
      }
      for (int t = lane; t < NIP; t += 64) {
        double fq = (a.hasF && a.kind != 0) ? a.f_ip[(size_t)e * NIP + t] : 0.0;
        if (a.u_old && a.ts_idt != 0.0) {  // implicit Euler: + u_old(x_q) / dt
          double uo = 0.0;
          for (int j = 0; j < NN; ++j) uo += PHI[t * NN + j] * a.u_old[(size_t)e * NN + j];
          fq += a.ts_idt * uo;
        }
        FQ[t] = fq;
        RQ[t] = (a.hasR ? a.r_ip[(size_t)e * NIP + t] : 0.0) + a.ts_idt;
      }
    } else {
      if (lane < (D + 1) * D) VX[lane] = pf.vx;
      if (lane < NN) DN[lane] = a.ndD == 1 ? pf.dn : (a.useDiff ? 1.0 : 0.0);
      if (lane < NN * D) BN[lane] = a.hasB ? pf.bn : 0.0;
      if (lane < NF * NNF) {
        TAU[lane] = (a.tau_sides == 2 && pf.s2 != e) ? pf.t1 : pf.t0;
        IX[L::IX_PERM + lane] = pf.pm;
      }
      if (fused) {
        if (lane < NF * NF) IX[L::IX_BPOS + lane] = pf.bp;
        if (lane < NF) {
          IX[L::IX_FACE + lane] = pf.fx;
          IX[L::IX_FLAG + lane] = (pf.s2x == e ? 0 : 1) | (pf.dr ? 2 : 0);
        }
      }
      if (lane < NIP) {
        double fq = pf.fq;
        if (a.u_old && a.ts_idt != 0.0) {  // implicit Euler: + u_old(x_q) / dt
          double uo = 0.0;
          for (int j = 0; j < NN; ++j) uo += PHI[lane * NN + j] * a.u_old[(size_t)e * NN + j];
          fq += a.ts_idt * uo;
        }
        FQ[lane] = fq;
        RQ[lane] = pf.rq + a.ts_idt;
      }
    }
    const int en = e + e_stride;
    const bool more = en < e_end;
    if (PIPE && more) pf.fetch_idx(a, en, lane, fused);
    wave_sync();
    // ------------------------------------------------ element geometry
    // J[r][c] = dx_c/dxi_r (transposed convention of Operator.cpp:31-36)
    double J[D * D], JI[D * D], det;
    for (int r = 0; r < D; ++r)
      for (int c = 0; c < D; ++c) J[r * D + c] = 0.5 * (VX[(r + 1) * D + c] - VX[c]);
    if constexpr (D == 2) {
      det = J[0] * J[3] - J[1] * J[2];
      const double r = 1.0 / det;
      JI[0] = J[3] * r; JI[1] = -J[1] * r; JI[2] = -J[2] * r; JI[3] = J[0] * r;
    } else {
      det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) + J[2] * (J[3] * J[7] - J[4] * J[6]);
      const double r = 1.0 / det;
      JI[0] = (J[4] * J[8] - J[5] * J[7]) * r; JI[1] = (J[2] * J[7] - J[1] * J[8]) * r; JI[2] = (J[1] * J[5] - J[2] * J[4]) * r;
      JI[3] = (J[5] * J[6] - J[3] * J[8]) * r; JI[4] = (J[0] * J[8] - J[2] * J[6]) * r; JI[5] = (J[2] * J[3] - J[0] * J[5]) * r;
      JI[6] = (J[3] * J[7] - J[4] * J[6]) * r; JI[7] = (J[1] * J[6] - J[0] * J[7]) * r; JI[8] = (J[0] * J[4] - J[1] * J[3]) * r;
    }
    if (lane < NF) {
      // face measure and outward unit normal (HDGBase.cpp:34-65)
      const int f = lane;
      const double* x0 = VX + s_fn[f * NNF] * D;
      double t0[D], t1[D], nv[D], meas;
      for (int c = 0; c < D; ++c) t0[c] = 0.5 * (VX[s_fn[f * NNF + 1] * D + c] - x0[c]);
      if constexpr (D == 2) {
        meas = sqrt(t0[0] * t0[0] + t0[1] * t0[1]);
        nv[0] = t0[1];
        nv[1] = -t0[0];
      } else {
        for (int c = 0; c < D; ++c) t1[c] = 0.5 * (VX[s_fn[f * NNF + 2] * D + c] - x0[c]);
        nv[0] = t0[1] * t1[2] - t0[2] * t1[1];
        nv[1] = t0[2] * t1[0] - t0[0] * t1[2];
        nv[2] = t0[0] * t1[1] - t0[1] * t1[0];
        meas = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
      }
      double nn = 0.0;
      for (int c = 0; c < D; ++c) nn += nv[c] * nv[c];
      nn = 1.0 / sqrt(nn);
      int vo = 0;
      for (int i = D; i >= 0; --i)
        if (((unsigned)s_pos[i] >> (8 * f) & 0xFFu) == 0xFFu) vo = i;
      double prod = 0.0;
      for (int c = 0; c < D; ++c) {
        nv[c] *= nn;
        prod += (VX[vo * D + c] - x0[c]) * nv[c];
      }
      const double sg = prod > 0 ? -1.0 : 1.0;
      for (int c = 0; c < D; ++c) FG[f * (D + 1) + c] = sg * nv[c];
      FG[f * (D + 1) + D] = meas;
    }
    // ------------------------------------------------ cubature-point data
    // one cubature point (then one face point) per lane, as straight-line
    // code at clamped indices with the stores masked: the reads pipeline
    static_assert(NIP <= 64 && NF * NIPF <= 64, "one point per lane");
    {
      const int q = lane < NIP ? lane : 0;
      double dq = 0.0, bq[D];
      for (int c = 0; c < D; ++c) bq[c] = 0.0;
#pragma unroll
      for (int i = 0; i < NN; ++i) {
        const double ph = PHI[q * NN + i];
        dq += ph * DN[i];
        for (int c = 0; c < D; ++c) bq[c] += ph * BN[i * D + c];
      }
      const double wd = WQ[q] * det;
      double bj[D];
      for (int r = 0; r < D; ++r) {
        double s = 0.0;
        for (int k = 0; k < D; ++k) s += bq[k] * JI[k * D + r];
        bj[r] = wd * s;
      }
      const double wr = wd * RQ[q], wf = wd * FQ[q];
      if (lane < NIP) {
        SD[q] = wd * dq;
        for (int r = 0; r < D; ++r) BJ[q * D + r] = bj[r];
        WR[q] = wr;
        WF[q] = wf;
      }
    }
    wave_sync();
    {
      const int t = lane < NF * NIPF ? lane : 0;
      const int f = t / NIPF, q = t % NIPF;
      double tq = 0.0, bq[D], dq = 0.0;
      for (int c = 0; c < D; ++c) bq[c] = 0.0;
#pragma unroll
      for (int j = 0; j < NNF; ++j) {
        const double ps = FPHI[q * NNF + j];
        const int node = s_fn[f * NNF + j];
        tq += ps * TAU[f * NNF + j];
        for (int c = 0; c < D; ++c) bq[c] += ps * BN[node * D + c];
        dq += ps * DN[node];
      }
      const double* nf = FG + f * (D + 1);
      const double ds = FWQ[q] * nf[D];
      double bn = 0.0;
      for (int c = 0; c < D; ++c) bn += bq[c] * nf[c];
      if (lane < NF * NIPF) {
        double* w = FWT + t * NS;
        w[0] = ds * tq;
        w[1] = ds * bn;
        w[2] = ds * dq;
      }
    }
    wave_sync();
    // weighted face masses FM[f][s][a][b] = sum_q w_s psi_a psi_b, s = 0: tau, 1: b.n, 2: D
    // on FP64 MFMA: FM[f][s] = (w_s psi)^T psi over the face points (K = NIPF);
    // lane (a, q) supplies psi_a(q) to both operands
    constexpr int NSF = L::NSF;
    {
      constexpr int KSF = (NIPF + 3) / 4;
      static_assert(NNF <= 16, "one MFMA tile per face mass");
      const int li = lane & 15, lk = lane >> 4;
      // psi_a(q) is the same on every face: read once, at clamped (in-range)
      // addresses with the padding zeroed after the load (no exec-masked
      // loads, so the reads pipeline); the faces' products run back to back
      double ps[KSF];
#pragma unroll
      for (int ks = 0; ks < KSF; ++ks) {
        const int q = ks * 4 + lk;
        const double x = FPHI[(q < NIPF ? q : 0) * NNF + (li < NNF ? li : 0)];
        ps[ks] = (q < NIPF && li < NNF) ? x : 0.0;
      }
      // the next face's weights are read before this face's masses are
      // stored (the compiler cannot move reads past the stores)
      double wt[2][KSF][NSF];
      auto wload = [&](int f, double (*w)[NSF]) {
#pragma unroll
        for (int ks = 0; ks < KSF; ++ks) {
          const int q = ks * 4 + lk, qc = q < NIPF ? q : 0;
#pragma unroll
          for (int sf = 0; sf < NSF; ++sf) w[ks][sf] = FWT[(f * NIPF + qc) * NS + sf];  // ps is zero on padding q
        }
      };
      wload(0, wt[0]);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        if (f + 1 < NF) wload(f + 1, wt[(f + 1) & 1]);
        dvec4 acc[NSF];
#pragma unroll
        for (int sf = 0; sf < NSF; ++sf) acc[sf] = dvec4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < KSF; ++ks)
#pragma unroll
          for (int sf = 0; sf < NSF; ++sf)
            acc[sf] = __builtin_amdgcn_mfma_f64_16x16x4f64(ps[ks] * wt[f & 1][ks][sf], ps[ks], acc[sf], 0, 0, 0);
#pragma unroll
        for (int sf = 0; sf < NSF; ++sf)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int aa = 4 * r + lk;
            if (aa < NNF && li < NNF) FM[((f * NSF + sf) * NNF + aa) * NNF + li] = acc[sf][r];
          }
      }
    }
    wave_sync();
    // ----------------------------------------- Suq_k and Suu (entries)
    // On FP64 MFMA over the cubature points: Suq_k = G_k PHI^T and
    // Suu = G_u PHI^T with G_k[i][q] = SD[q] sum_c Jinv[k][c] dphi_c(i, q) and
    // G_u[i][q] = WR[q] phi_i(q) - sum_k BJ[q][k] dphi_k(i, q), K = NIP;
    // then the face terms, and Suq moved to the reference frame
    // SUQ_a = sum_k Jinv[k][a] Suq_k.  The first padding row of G_u carries
    // the source weights WF[q], so that the same product gives
    // F_u[j] = sum_q WF[q] phi_j(q) in row NN of its result.
    static_assert(NN % 16 != 0, "a padding row for F_u");
    {
      constexpr int MT = (NN + 15) / 16, KS = (NIP + 3) / 4, NX = D + 1;
      const int li = lane & 15, lk = lane >> 4;
      double nfe[NF][D];
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int k = 0; k < D; ++k) nfe[f][k] = FG[f * (D + 1) + k];
#pragma unroll 1
      for (int m = 0; m < MT; ++m) {
        const int ia = m * 16 + li;  // A row this lane supplies
        double av[KS][NX];
        // table and per-point reads at clamped indices (no exec-masked loads)
        const int ic = ia < NN ? ia : 0;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const int q = ks * 4 + lk;
          const bool v = q < NIP && ia < NN;
          const int qq = q < NIP ? q : 0;
          double g[D], bj[D];
          for (int c2 = 0; c2 < D; ++c2) g[c2] = DPHI[(qq * NN + ic) * D + c2];
          for (int k = 0; k < D; ++k) bj[k] = BJ[qq * D + k];
          const double sd = SD[qq], wr = WR[qq], ph = PHI[qq * NN + ic], wf = WF[qq];
          double bgs = 0.0;
          for (int k = 0; k < D; ++k) {
            double s2 = 0.0;
            for (int c2 = 0; c2 < D; ++c2) s2 += JI[k * D + c2] * g[c2];
            av[ks][k] = v ? sd * s2 : 0.0;
            bgs += bj[k] * g[k];
          }
          av[ks][D] = v ? wr * ph - bgs : ((ia == NN && q < NIP) ? wf : 0.0);
        }
#pragma unroll 1
        for (int nt = 0; nt < MT; ++nt) {
          const int jb = nt * 16 + li;
          // face positions of the tile's rows and column, read before the
          // MFMAs (the face-mass reads of the epilogue depend on them)
          unsigned pir[4];
          const unsigned pj = (unsigned)s_pos[jb < NN ? jb : 0];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = m * 16 + 4 * r + lk;
            pir[r] = (unsigned)s_pos[i < NN ? i : 0];
          }
          dvec4 acc[NX];
#pragma unroll
          for (int x = 0; x < NX; ++x) acc[x] = dvec4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int ks = 0; ks < KS; ++ks) {
            const int q = ks * 4 + lk;
            const double y = PHI[(q < NIP ? q : 0) * NN + (jb < NN ? jb : 0)];
            const double bv = (q < NIP && jb < NN) ? y : 0.0;
#pragma unroll
            for (int x = 0; x < NX; ++x) acc[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ks][x], bv, acc[x], 0, 0, 0);
          }
          mfma_interleave<KS * NX, KS>();
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = m * 16 + 4 * r + lk, j = jb;
            if (i == NN && j < NN) FU[j] = acc[D][r];
            // face terms without branches: every face's masses are read (at
            // clamped positions) and the entries off the face masked out
            const bool ok = i < NN && j < NN;
            double aq[D], au = acc[D][r];
            for (int k = 0; k < D; ++k) aq[k] = acc[k][r];
            const unsigned pi = pir[r];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
              const int ai = (int)(pi >> (8 * f) & 0xFFu), bj = (int)(pj >> (8 * f) & 0xFFu);
              const bool on = ai != 0xFF && bj != 0xFF;
              const double* F0 = FM + (f * NSF) * NNF * NNF + (on ? ai : 0) * NNF + (on ? bj : 0);
              const double t0 = F0[0], t2 = F0[2 * NNF * NNF];
              au += on ? t0 : 0.0;
              for (int k = 0; k < D; ++k) aq[k] -= nfe[f][k] * (on ? t2 : 0.0);
            }
            for (int r2 = 0; r2 < D; ++r2) {
              double s2 = 0.0;
              for (int k = 0; k < D; ++k) s2 += JI[k * D + r2] * aq[k];
              if (ok) SUQ[(r2 * NN + i) * NN + j] = s2;
            }
            if (ok) KM[i * NN + j] = au;
          }
        }
      }
    }
    wave_sync();
    if constexpr (TSG) {
      // general time transform of the u rows (the TSG kernel variant).  The
      // cell's w, u_old and u_old at the cubature points go to the phase-1
      // overlay (dead after the entries); lane i applies the model's row i
      // [Suu | Suq | Sul] to w and forms (M u_old)_i = det sum_q W_q phi_i(q)
      // u_old(x_q); then, spread over the lanes, every KM entry takes
      // ts_auu KM + ts_mu M with M_ij = det sum_q W_q phi_i(q) phi_j(q), and
      // SUQ is scaled.  Suq_k = sum_a J[a][k] SUQ_a, so
      // sum_k Suq_k w_q,k = sum_a SUQ_a (sum_k J[a][k] w_q,k).
      constexpr int NLOC = NN * (D + 1) + NF * NNF;
      static_assert(NLOC + NN + NIP <= L::PER - L::O_OV, "time-transform staging fits the overlay");
      double* TW = ws + L::O_OV;
      double* TU = TW + NLOC;
      double* UQ = TU + NN;
      for (int t = lane; t < NLOC; t += 64) TW[t] = a.ts_w[(size_t)e * NLOC + t];
      for (int t = lane; t < NN; t += 64) TU[t] = a.u_old[(size_t)e * NN + t];
      wave_sync();
      for (int q = lane; q < NIP; q += 64) {
        double v = 0.0;
        for (int j = 0; j < NN; ++j) v += PHI[q * NN + j] * TU[j];
        UQ[q] = v;
      }
      wave_sync();
      if (lane < NN) {
        const int i = lane;
        double c = 0.0, mu = 0.0;
        for (int q = 0; q < NIP; ++q) mu += WQ[q] * PHI[q * NN + i] * UQ[q];
        for (int j = 0; j < NN; ++j) {
          c += KM[i * NN + j] * TW[j];
          for (int a2 = 0; a2 < D; ++a2) {
            double wt = 0.0;
            for (int k = 0; k < D; ++k) wt += J[a2 * D + k] * TW[NN + j * D + k];
            c += SUQ[(a2 * NN + i) * NN + j] * wt;
          }
        }
        const unsigned pw = (unsigned)s_pos[i];
        for (int f = 0; f < NF; ++f) {
          const int pf = (int)(pw >> (8 * f) & 0xFFu);
          if (pf == 0xFF) continue;
          for (int b = 0; b < NNF; ++b)
            c += (-FM[((f * NSF + 0) * NNF + pf) * NNF + b] + FM[((f * NSF + 1) * NNF + pf) * NNF + b]) *
                 TW[NN * (D + 1) + f * NNF + b];
        }
        FU[i] = a.ts_phi * FU[i] + a.ts_mu * (det * mu) - c;
      }
      wave_sync();  // every row has read KM and SUQ before they change
      for (int t = lane; t < NN * NN; t += 64) {
        const int i = t / NN, j = t % NN;
        double m = 0.0;
        for (int q = 0; q < NIP; ++q) m += WQ[q] * PHI[q * NN + i] * PHI[q * NN + j];
        KM[t] = a.ts_auu * KM[t] + a.ts_mu * (det * m);
      }
      for (int t = lane; t < D * NN * NN; t += 64) SUQ[t] *= a.ts_arest;
      wave_sync();
    }
    // K = Suu - sum_k Suq_k A_k = Suu - sum_a SUQ_a Ahat_a on FP64 MFMA, in
    // place in KM: [SUQ_0 | .. | SUQ_{D-1}] (NN x D NN) times [Ahat_0; ..] (D NN x NN)
    {
      constexpr int MT = (NN + 15) / 16, KD = D * NN, KS = (KD + 3) / 4;
      const int li = lane & 15, lk = lane >> 4;
      // every output tile's MFMA chain in flight at once: with one wave per
      // SIMD, independent accumulators are what hides the MFMA latency
      dvec4 acc[MT][MT];
#pragma unroll
      for (int nt = 0; nt < MT; ++nt)
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[nt][m] = dvec4{0.0, 0.0, 0.0, 0.0};
      // operands loaded at clamped (in-range) addresses and zeroed on the
      // padding lanes afterwards: no exec-masked loads, so they pipeline
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int kk = ks * 4 + lk;
        const bool kv = kk < KD;
        const int a2 = kv ? kk / NN : 0, l = kv ? kk % NN : 0;
        double av[MT], bv[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          const int i = m * 16 + li;
          const double x = SUQ[(a2 * NN + (i < NN ? i : 0)) * NN + l];
          av[m] = (kv && i < NN) ? x : 0.0;
          const double y = AH[(a2 * NN + l) * NN + (i < NN ? i : 0)];
          bv[m] = (kv && i < NN) ? y : 0.0;
        }
#pragma unroll
        for (int nt = 0; nt < MT; ++nt)
#pragma unroll
          for (int m = 0; m < MT; ++m) acc[nt][m] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m], bv[nt], acc[nt][m], 0, 0, 0);
      }
      mfma_interleave<KS * MT * MT, KS * 2 * MT>();
      // Suu entries read (at clamped addresses) before any K entry is written:
      // the compiler cannot tell the entries apart and would otherwise wait
      // on every read-modify-write
      double kv0[MT][MT][4];
#pragma unroll
      for (int nt = 0; nt < MT; ++nt)
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = m * 16 + 4 * r + lk, jb = nt * 16 + li;
            kv0[nt][m][r] = KM[(i < NN ? i : 0) * NN + (jb < NN ? jb : 0)];
          }
#pragma unroll
      for (int nt = 0; nt < MT; ++nt)
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = m * 16 + 4 * r + lk, jb = nt * 16 + li;
            if (i < NN && jb < NN) KM[i * NN + jb] = kv0[nt][m][r] - acc[nt][m][r];
          }
    }
    wave_sync();
    // K = P^T L U by Gaussian elimination with partial pivoting, in place: one
    // row of K per lane in registers, the pivot row fetched with readlane.  At
    // step k the pivot p is the largest |K[i][k]| among the rows not chosen
    // yet (DPP argmax); every other such row takes l = K[i][k] / K[p][k],
    // K[i][j] -= l K[p][j] for j > k and keeps l in K[i][k].  Row p is then
    // row k of U (with the multipliers of the earlier steps to its left) and
    // is stored as column k of KM (KM[j * NN + k]), its reciprocal pivot in
    // RD[k], and the pivot row of step k in PL[k].  The R columns
    // are solved against these factors below: forward and back substitution
    // is backward stable, whereas multiplying by an explicitly formed K^-1
    // leaves an error in U that does not cancel in S = Slu U + Slq Q + Sll
    // (a relative 1e-12 in S for the 2-D p = 4 cells of a 128 x 128 mesh,
    // where cond(K) ~ 3e4).
    int* PL = reinterpret_cast<int*>(ws + L::O_PL);
    {
      double kr[NN];
      const int rl = lane < NN ? lane : 0;
#pragma unroll
      for (int j = 0; j < NN; ++j) kr[j] = KM[rl * NN + j];
      static_assert(NN <= 32, "pivot search covers lanes 0..31");
      unsigned long long used = 0ull;
      int myk = 0;
      double myrd = 0.0;
#pragma unroll
      for (int k = 0; k < NN; ++k) {
        const bool cand = lane < NN && !((used >> lane) & 1ull);
        const int p = __builtin_amdgcn_readfirstlane(dpp_argmax(fabs(kr[k]), cand));
        used |= 1ull << p;
        const bool elim = cand && lane != p;
        const double rp = recip(readlane_d(kr[k], p));
        myk = lane == p ? k : myk;
        myrd = lane == p ? rp : myrd;
        const double l = elim ? kr[k] * rp : 0.0;
#pragma unroll
        for (int j = k + 1; j < NN; ++j) kr[j] = fma(-l, readlane_d(kr[j], p), kr[j]);
        kr[k] = elim ? l : kr[k];
      }
      // stored column by column, KM[j * NN + k] = LU[k][j]: a substitution
      // step reads one contiguous column
      if (lane < NN) {
#pragma unroll
        for (int j = 0; j < NN; ++j) KM[j * NN + myk] = kr[j];
        ws[L::O_RD + myk] = myrd;
        PL[myk] = lane;
      }
    }
    wave_sync();
    // ---------------------------------------------- trace-column phase
    // lane c < NL owns trace column c = (g, bp); lane NL owns the F_u column.
    // Column vectors live in registers (compile-time row index) or in
    // lane-contiguous LDS rows [row][NCOL] when read at face-node positions.
    const int c = lane;
    double* UB = RB;               // U columns replace R columns lane by lane
    double* QN = ws + L::O_SUQ;    // face traces of n.q, over the dead Suq and K factors
    // R[:, c] = Sul[:, c] - sum_k Suq_k B_k[:, c] (or -F_u on lane NL), with
    // B_k[:, (g,b)] = -rho_g n_gk G_g[:, b] and sum_k n_gk Suq_k = sum_a nt_a SUQ_a;
    // Sul[i][(g, b)] = -FMt_g[a][b] + FMb_g[a][b] when node i is node a of
    // face g, else 0 (added to the product in its epilogue)
    // ... + sum_a nt_a(c) SUQ_a G_g(c)[:, bp(c)] for the trace columns on FP64
    // MFMA: X = [SUQ_0 | .. | SUQ_{D-1}] (NN x D NN) times Bt (D NN x NL),
    // Bt[(a, j)][c] = nt_a(c) G_g(c)[j][bp(c)], nt_a = rho_g sum_k J[a][k] n_gk.
    // Operand lanes: A[lane % 16][lane / 16], B[lane / 16][lane % 16]; result
    // row 4 r + lane / 16, column lane % 16 sits in accumulator element r.
    {
      constexpr int MT = (NN + 15) / 16, NT = (NL + 15) / 16, KD = D * NN, KS = (KD + 3) / 4;
      const int li = lane & 15, lk = lane >> 4;
      // all NT x MT tiles accumulated together (independent MFMA chains)
      int ggs[NT], bbs[NT];
      bool cvs[NT];
      dvec4 ntc[NT];  // a vector value, so that the lane-dependent component index needs no memory
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const int cc = nt * 16 + li;
        cvs[nt] = cc < NL;
        ggs[nt] = cvs[nt] ? cc / NNF : 0;
        bbs[nt] = cvs[nt] ? cc % NNF : 0;
        const double rg = FG[ggs[nt] * (D + 1) + D] / det;
        ntc[nt] = dvec4{0.0, 0.0, 0.0, 0.0};
        for (int a2 = 0; a2 < D; ++a2) {
          double s = 0.0;
          for (int k = 0; k < D; ++k) s += J[a2 * D + k] * FG[ggs[nt] * (D + 1) + k];
          ntc[nt][a2] = cvs[nt] ? rg * s : 0.0;
        }
      }
      dvec4 acc[NT][MT];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[nt][m] = dvec4{0.0, 0.0, 0.0, 0.0};
      // operands at clamped addresses, padding zeroed after the load
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int kk = ks * 4 + lk;
        const bool kv = kk < KD;
        const int a2 = kv ? kk / NN : 0, j = kv ? kk % NN : 0;
        double av[MT], bv[NT];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          const int i = m * 16 + li;
          const double x = SUQ[(a2 * NN + (i < NN ? i : 0)) * NN + j];
          av[m] = (kv && i < NN) ? x : 0.0;
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const double ntv = ntc[nt][a2];
          const double y = GT[(ggs[nt] * NN + j) * NNF + bbs[nt]];
          bv[nt] = kv ? ntv * y : 0.0;
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
          for (int m = 0; m < MT; ++m) acc[nt][m] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[m], bv[nt], acc[nt][m], 0, 0, 0);
      }
      // R = Sul + product (F_u column: -F_u), every entry written once; the
      // Sul entries are read at clamped face positions, masked off the face
      // (those reads are interleaved with the MFMAs too)
      static_assert(NL % 16 != 0, "the F_u column lies in the last column tile");
      const double sul_s = TSG ? a.ts_arest : 1.0;  // the time transform scales Sul
      double rv[NT][MT][4];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = m * 16 + 4 * r + lk, ic = i < NN ? i : 0;
            const int pa = (int)((unsigned)s_pos[ic] >> (8 * ggs[nt]) & 0xFFu);
            const bool on = pa != 0xFF;
            const int pac = on ? pa : 0;
            const double t = -FM[((ggs[nt] * NSF + 0) * NNF + pac) * NNF + bbs[nt]] +
                             FM[((ggs[nt] * NSF + 1) * NNF + pac) * NNF + bbs[nt]];
            const double fu = FU[ic];
            rv[nt][m][r] = cvs[nt] ? (on ? sul_s * t : 0.0) : -fu;
          }
      mfma_interleave<KS * MT * NT, KS * (MT + NT) + NT * MT * 4 * 4>();
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = m * 16 + 4 * r + lk, cc = nt * 16 + li;
            if (cc <= NL && i < NN) RB[i * NCOL + cc] = rv[nt][m][r] + acc[nt][m][r];
          }
    }
    wave_sync();
    // U = -K^-1 R: lane c <= NL solves its column against the LU factors,
    // y = L^-1 P r, then x = U^-1 y, with the factors read at uniform
    // (broadcast) LDS addresses; the column is read whole before it is
    // overwritten, so the solve is in place
    if (c <= NL) {
      const double* RDp = ws + L::O_RD;
      double y[NN];
#pragma unroll
      for (int k = 0; k < NN; ++k) y[k] = RB[PL[k] * NCOL + c];
      // column-oriented (right-looking) substitution: once y[k] is final,
      // every later row takes its update at once, so each step issues
      // independent FMAs and the dependent chain is NN long, not NN^2 / 2;
      // the factor column of the next step is read while the current
      // step's FMAs run
      // (factor columns read as aligned pairs: KM is 16-byte aligned)
      double fc[NN], fn[NN];
      auto column = [&](int k, int lo, int hi, double* f) {  // f[i] = LU[i][k], lo <= i < hi
#pragma unroll
        for (int t = (k * NN + lo) & ~1; t < k * NN + hi; t += 2) {
          const dvec2 v = *reinterpret_cast<const dvec2*>(KM + t);
          if (t >= k * NN + lo) f[t - k * NN] = v.x;
          if (t + 1 < k * NN + hi) f[t + 1 - k * NN] = v.y;
        }
      };
      // (the prefetch registers would cost the second wave per SIMD at the
      // low orders: there the FMAs read the factors directly)
      constexpr bool SPF = D == 3 ? P >= 2 : P >= 4;
      if constexpr (SPF) {
      column(0, 1, NN, fc);
#pragma unroll
      for (int k = 0; k < NN - 1; ++k) {
        if (k + 2 < NN) column(k + 1, k + 2, NN, fn);
#pragma unroll
        for (int i = k + 1; i < NN; ++i) y[i] = fma(-fc[i], y[k], y[i]);
#pragma unroll
        for (int i = k + 2; i < NN; ++i) fc[i] = fn[i];
      }
      column(NN - 1, 0, NN - 1, fc);
#pragma unroll
      for (int kk = 0; kk < NN; ++kk) {
        const int k = NN - 1 - kk;
        if (k >= 2) column(k - 1, 0, k - 1, fn);
        y[k] *= RDp[k];
#pragma unroll
        for (int i = 0; i < k; ++i) y[i] = fma(-fc[i], y[k], y[i]);
#pragma unroll
        for (int i = 0; i + 1 < k; ++i) fc[i] = fn[i];
      }
      } else {
#pragma unroll
        for (int k = 0; k < NN - 1; ++k)
#pragma unroll
          for (int i = k + 1; i < NN; ++i) y[i] = fma(-KM[k * NN + i], y[k], y[i]);
#pragma unroll
        for (int kk = 0; kk < NN; ++kk) {
          const int k = NN - 1 - kk;
          y[k] *= RDp[k];
#pragma unroll
          for (int i = 0; i < k; ++i) y[i] = fma(-KM[k * NN + i], y[k], y[i]);
        }
      }
#pragma unroll
      for (int i = 0; i < NN; ++i) UB[i * NCOL + c] = -y[i];
    }
    wave_sync();  // every lane is done with Suq and the K factors before QN overwrites them
    // Q_k = -A_k U - B_k on FP64 MFMA: Y_a = Ahat_a U (NN x NN times NN x NCOL),
    // then per entry q_k = -sum_a Jinv[k][a] Y_a + n_gk rho_g G_g[i][bp];
    // writes U, Q (U0, Q0 for column NL) and keeps n_f . q on the faces (QN)
    {
      constexpr int MT = (NN + 15) / 16, NT = (NCOL + 15) / 16, KS = (NN + 3) / 4;
      const int li = lane & 15, lk = lane >> 4;
      // face normals in registers: the QN stores below would otherwise force
      // the compiler to reload them from LDS after every store
      double nfk[NF][D];
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int k = 0; k < D; ++k) nfk[f][k] = FG[f * (D + 1) + k];
      // the Ahat operands are the same for every column tile: loaded once,
      // at clamped addresses with the padding zeroed after the load
      double av[KS][MT][D];
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int j = ks * 4 + lk, jc = j < NN ? j : 0;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          const int i = m * 16 + li, ic = i < NN ? i : 0;
#pragma unroll
          for (int a2 = 0; a2 < D; ++a2) {
            const double x = AH[(a2 * NN + ic) * NN + jc];
            av[ks][m][a2] = (j < NN && i < NN) ? x : 0.0;
          }
        }
      }
      // column tiles unrolled: the next tile's reads overlap this tile's
      // MFMAs and stores (not in the time-transform variant)
      constexpr int UT = TSG ? 1 : NT;
#pragma unroll UT
      for (int nt = 0; nt < NT; ++nt) {
        const int cc = nt * 16 + li, ccc = cc < NCOL ? cc : NCOL - 1;
        double bv[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const int j = ks * 4 + lk;
          const double x = UB[(j < NN ? j : 0) * NCOL + ccc];
          bv[ks] = (j < NN && cc < NCOL) ? x : 0.0;
        }
        // epilogue inputs, fetched while the MFMAs run: G_g at each row, the
        // U column, each row's face positions
        const bool trc = cc < NL;
        const int gg = trc ? cc / NNF : 0, bb = trc ? cc % NNF : 0;
        double gbv[MT][4], ubv[MT][4];
        unsigned spw[MT][4];
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = m * 16 + 4 * r + lk, ic = i < NN ? i : 0;
            gbv[m][r] = GT[(gg * NN + ic) * NNF + bb];
            ubv[m][r] = UB[ic * NCOL + ccc];
            spw[m][r] = (unsigned)s_pos[ic];
          }
        dvec4 y[MT][D];
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
          for (int a2 = 0; a2 < D; ++a2) y[m][a2] = dvec4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
          for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int a2 = 0; a2 < D; ++a2)
              y[m][a2] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ks][m][a2], bv[ks], y[m][a2], 0, 0, 0);
        mfma_interleave<KS * MT * D, KS>();
        if (cc <= NL) {
          const double rg = FG[gg * (D + 1) + D] / det;
          double ng[D];
#pragma unroll
          for (int k = 0; k < D; ++k) ng[k] = FG[gg * (D + 1) + k];
          double* Qe = trc ? a.Q + (size_t)e * NN * D * NL : a.Q0 + (size_t)e * NN * D;
          double* Ue = trc ? a.U + (size_t)e * NN * NL : a.U0 + (size_t)e * NN;
#pragma unroll
          for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int i = m * 16 + 4 * r + lk;
              if (i >= NN) continue;
              const double gb = trc ? rg * gbv[m][r] : 0.0;
              double qk[D];
              for (int k = 0; k < D; ++k) {
                double s = 0.0;
                for (int a2 = 0; a2 < D; ++a2) s -= JI[k * D + a2] * y[m][a2][r];
                qk[k] = s + ng[k] * gb;
              }
              // streaming stores for U and Q (read again only by the
              // reconstruction); the F_u column's U0 and Q0 are small and
              // go through the cache (a streaming store of a lone 8-byte
              // entry writes a partial line)
              if (trc) {
                __builtin_nontemporal_store(ubv[m][r], &Ue[(size_t)i * NL + cc]);
                for (int k = 0; k < D; ++k) __builtin_nontemporal_store(qk[k], &Qe[(size_t)(i * D + k) * NL + cc]);
              } else {
                Ue[i] = ubv[m][r];
                for (int k = 0; k < D; ++k) Qe[i * D + k] = qk[k];
              }
              // n_f . q on every face of node i; a node off face f writes its
              // value to a dead row instead (the face-IP weights, unused after
              // the face masses), which keeps the stores free of branches
              const unsigned pw = spw[m][r];
#pragma unroll
              for (int f = 0; f < NF; ++f) {
                const int pf = (int)(pw >> (8 * f) & 0xFFu);
                double sq = 0.0;
                for (int k = 0; k < D; ++k) sq += nfk[f][k] * qk[k];
                double* dst = pf != 0xFF ? QN + (f * NNF + pf) * NCOL + cc : FWT + cc;
                *dst = sq;
              }
            }
        }
      }
    }
    wave_sync();  // QN complete before the S rows read it
    if (PIPE && more) pf.fetch_data(a, lane, fused);
    // S rows of face f on FP64 MFMA: [Slu_f | Slq_f] (NNF x 2 NNF) times
    // [U at the face nodes; n.q at the face nodes] (2 NNF x NCOL), i.e.
    // S[(f, aa)][c] = sum_b FMt[aa][b] U[fn_b][c] - FMd[aa][b] QN[(f, b)][c],
    // plus -FMt[aa][bp] + FMb[aa][bp] on the trace columns of face f (Sll);
    // the F_u column gives S0 = -(...).  Explicit HDG modes: S = Sll only, and
    // the F_u column (U0 and Q0 already stored) is replaced by the old
    // solution and the normal old flux at the face nodes.
    const bool expl = a.hdg_explicit != 0;
    if (expl) {
      for (int i = lane; i < NN; i += 64) UB[i * NCOL + NL] = a.sol_old[(size_t)e * NN + i];
      for (int t = lane; t < NF * NNF; t += 64) {
        const int f = t / NNF, node = s_fn[t];
        double qn = 0.0;
        for (int k = 0; k < D; ++k) qn += FG[f * (D + 1) + k] * a.flux_old[((size_t)e * NN + node) * D + k];
        QN[t * NCOL + NL] = qn;
      }
      wave_sync();
    }
    {
      constexpr int NT = (NCOL + 15) / 16, KF = 2 * NNF, KS = (KF + 3) / 4;
      const int li = lane & 15, lk = lane >> 4;
      static_assert(NNF <= 16, "one row tile per face");
      // fused trace scatter: no S; each face block goes to its place in the
      // BSR values in face-node order (off-diagonal blocks have this cell as
      // their only contributor; Dirichlet rows get zeros), the diagonal block
      // to this cell's side of dside, summed by the diagonal pass.
      // Every LDS operand of a face is loaded before its MFMAs, at an
      // in-range (clamped) address with padding lanes zeroed afterwards, so
      // that the loads issue back to back instead of one wait per MFMA.
      const int lic = li < NNF ? li : 0;
      int ccl[NT], gcol[NT], bcol[NT], pcs[NT];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const int cc = nt * 16 + li;
        ccl[nt] = cc < NCOL ? cc : NCOL - 1;
        gcol[nt] = cc < NL ? cc / NNF : 0;
        bcol[nt] = cc < NL ? cc % NNF : 0;
        pcs[nt] = IX[L::IX_PERM + gcol[nt] * NNF + bcol[nt]];  // block column of a trace column
      }
      // faces unrolled: the next face's operand reads overlap this face's
      // MFMAs and stores (not in the time-transform variant, whose registers
      // that would spill)
      constexpr int UF = TSG ? 1 : NF;
#pragma unroll UF
      for (int f = 0; f < NF; ++f) {
        const double* fmt = FM + (f * NSF + 0) * NNF * NNF;
        const double* fmb = FM + (f * NSF + 1) * NNF * NNF;
        // operands at computed offsets (no selected pointers: those become
        // branches with a wait each); Slu_f = FMt, Slq_f = -FMd
        double av[KS], bv[KS][NT];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const int kk = ks * 4 + lk;
          const bool kv = kk < KF, isu = kk < NNF;
          const int b = kv ? (isu ? kk : kk - NNF) : 0;
          const double x = FM[((f * NSF + (isu ? 0 : 2)) * NNF + lic) * NNF + b];
          av[ks] = (kv && li < NNF) ? (isu ? x : -x) : 0.0;
          const int node = s_fn[f * NNF + b];
          const int roff = isu ? L::O_RB + node * NCOL : L::O_SUQ + (f * NNF + b) * NCOL;
#pragma unroll
          for (int nt = 0; nt < NT; ++nt) {
            const double y = ws[roff + ccl[nt]];
            bv[ks][nt] = (kv && nt * 16 + li < NCOL) ? y : 0.0;
          }
        }
        // epilogue inputs, loaded while the MFMAs run: the face's flags and
        // node permutation, each column's block, the Sll = -FMt + FMb terms
        // (the index words are read in both modes; only the fused scatter
        // uses them)
        const int Ff = IX[L::IX_FACE + f];
        const int flg = IX[L::IX_FLAG + f];
        int prm[4], bpo[NT];
        double sll[NT][4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int aa = 4 * r + lk;
          prm[r] = IX[L::IX_PERM + f * NNF + (aa < NNF ? aa : 0)];
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          bpo[nt] = IX[L::IX_BPOS + f * NF + gcol[nt]];
          // (only a tile holding some of face f's own columns needs them)
          const bool own = nt * 16 <= f * NNF + NNF - 1 && f * NNF <= nt * 16 + 15;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int aa = 4 * r + lk, ac = aa < NNF ? aa : 0;
            sll[nt][r] = own ? -fmt[ac * NNF + bcol[nt]] + fmb[ac * NNF + bcol[nt]] : 0.0;
          }
        }
        // the NT column tiles of face f accumulated together (independent
        // MFMA chains sharing the A operand)
        dvec4 accs[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) accs[nt] = dvec4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
          for (int nt = 0; nt < NT; ++nt)
            accs[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ks], bv[ks][nt], accs[nt], 0, 0, 0);
        mfma_interleave<KS * NT, KS * (NT + 1)>();
        const bool dirf = (flg & 2) != 0;
        const int sidef = flg & 1;
        // one streaming store per trace-column entry, its address and value
        // chosen by selects: to its block of the trace matrix (fused; zero on
        // a Dirichlet face's off-diagonal blocks) or to S; the F_u column to
        // S0 (negated)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const int cc = nt * 16 + li;
          const dvec4 acc = accs[nt];
          const bool trc = cc < NL;
          const int gg = gcol[nt];
          double* blk = (gg == f) ? a.dside + (size_t)(2 * Ff + sidef) * NNF * NNF + pcs[nt]
                                  : a.bvals + (size_t)bpo[nt] * NNF * NNF + pcs[nt];
          double* srow = a.S + (size_t)e * NL * NL + cc;
          double* s0 = a.S0 + (size_t)e * NL;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int aa = 4 * r + lk;
            double sv = (expl && trc) ? 0.0 : acc[r];
            sv += (trc && f == gg) ? sll[nt][r] : 0.0;
            double* dst = fused ? blk + (size_t)prm[r] * NNF : srow + (size_t)(f * NNF + aa) * NL;
            const double val = (fused && dirf && gg != f) ? 0.0 : sv;
            if (trc && aa < NNF) __builtin_nontemporal_store(val, dst);
            if (cc == NL && aa < NNF) s0[f * NNF + aa] = -sv;  // S0: small, through the cache
          }
        }
      }
    }
    wave_sync();
  }
}

// Orders whose registers fit two waves per SIMD (the compiler is asked to
// keep them there: the second wave hides the LDS and MFMA latencies; the
// accumulators then live in ordinary VGPRs)
template <int D, int P, bool TSG>
constexpr bool two_waves() {
  return D == 2 ? P <= (TSG ? 2 : 3) : P == 1;
}

template <int D, int P, bool TD, int W, bool TSG>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(two_waves<D, P, TSG>() ? 2 : 1, 8))) void
hdg_affine_kernel(ElemArgs a) {
  using L = ALay<D, P, TD>;
  extern __shared__ __align__(16) double sm[];
  __shared__ int s_pos[L::NN];
  __shared__ int s_fn[L::NF * L::NNF];
  affine_load_tables<D, P, TD, W>(a.tabs, sm, s_pos, s_fn);
  affine_cells<D, P, TD, W, TSG>(a, a.el_offset + blockIdx.x * W, a.el_offset + a.nEl, gridDim.x * W, sm, s_pos,
                                 s_fn);
}

// Persistent kernel over the cells of many 2-D problems of one order P:
// workgroup b stages the order's reference tables once, then takes the items
// b, b + gridDim.x, ... of a cost-sorted list; an item is a range of cells of
// one problem.  Each order has its own kernel so that its register and LDS
// budget is that of the order alone.
template <int P, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(two_waves<2, P, false>() ? 2 : 1, 8))) void
hdg_affine_batch_kernel(const ElemArgs* __restrict__ probs,
                                                                  const BatchItem* __restrict__ items, int nitems) {
  using L = ALay<2, P, false>;
  extern __shared__ __align__(16) double sm[];
  __shared__ int s_pos[L::NN];
  __shared__ int s_fn[L::NF * L::NNF];
  if (blockIdx.x >= nitems) return;
  // every problem of one order shares the reference tables: stage them once
  affine_load_tables<2, P, false, W>(probs[items[blockIdx.x].prob].tabs, sm, s_pos, s_fn);
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const BatchItem w = items[it];
    const ElemArgs a = probs[w.prob];
    affine_cells<2, P, false, W>(a, w.e0, w.e1, W, sm, s_pos, s_fn);
  }
}

template <int D, int P, bool TD, int W, bool TSG = false>
void launch_aff(const ElemArgs& a, int grid_cap, hipStream_t s) {
  using L = ALay<D, P, TD>;
  const size_t lds = (size_t)(L::TAB + W * L::PER) * sizeof(double);
  auto kern = hdg_affine_kernel<D, P, TD, W, TSG>;
  static PerDeviceLaunch once;
  once.lds_attr((const void*)kern, lds);
  // persistent grid by default: as many workgroups as are resident at once
  // (the LDS allows one per CU for 3-D p=3), each looping over its cells, so
  // that the reference tables are staged once per workgroup instead of once
  // per W cells (C3: 5.28 -> 4.98 ms)
  const int resident = once.resident_wgs((const void*)kern, 64 * W, lds);
  long long grid = ((long long)a.nEl + W - 1) / W;
  const long long cap = grid_cap > 0 ? grid_cap : resident;
  if (grid > cap) grid = cap;
  if (grid <= 0) return;
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * W), lds, s, a);
  HFX_HIP(hipGetLastError());
}

}  // namespace

template <int D, int P>
static size_t aff_lds(bool td, int w) {
  return td ? (size_t)(ALay<D, P, true>::TAB + w * ALay<D, P, true>::PER) * 8
            : (size_t)(ALay<D, P, false>::TAB + w * ALay<D, P, false>::PER) * 8;
}

size_t affine_lds_bytes(int dim, int order, bool tensor_d, int waves) {
#define HFX_CASE(D, P) \
  if (dim == D && order == P) return aff_lds<D, P>(tensor_d, waves);
  HFX_CASE(2, 1) HFX_CASE(2, 2) HFX_CASE(2, 3) HFX_CASE(2, 4) HFX_CASE(2, 5)
  HFX_CASE(3, 1) HFX_CASE(3, 2) HFX_CASE(3, 3)
#undef HFX_CASE
  throw std::runtime_error("affine element kernel: (dim, order) not supported");
}

// waves per workgroup that fit the LDS (with 4 KiB to spare) for (D, P);
// the host picks the same number
template <int D, int P>
constexpr int aff_wave_fit() {
  using L = ALay<D, P, false>;
  for (int w = 4; w > 1; --w)
    if ((L::TAB + w * L::PER) * 8 + 4096 <= 160 * 1024) return w;
  return 1;
}

void launch_affine(int dim, int order, bool tensor_d, int waves, const ElemArgs& a, int grid_cap, hipStream_t s) {
  if (tensor_d) throw std::runtime_error("affine element kernel: tensor diffusion uses the general kernel");
  if (a.ts_gen) {
    // the time-transform variant exists at the LDS-fitting wave count only
#define HFX_CASE(D, P)                                                                     \
    if (dim == D && order == P) {                                                          \
      if (waves != aff_wave_fit<D, P>())                                                   \
        throw std::runtime_error("affine element kernel: time transform needs the default waves per workgroup"); \
      launch_aff<D, P, false, aff_wave_fit<D, P>(), true>(a, grid_cap, s);                 \
      return;                                                                              \
    }
    HFX_CASE(2, 1) HFX_CASE(2, 2) HFX_CASE(2, 3) HFX_CASE(2, 4) HFX_CASE(2, 5)
    HFX_CASE(3, 1) HFX_CASE(3, 2) HFX_CASE(3, 3)
#undef HFX_CASE
    throw std::runtime_error("affine element kernel: (dim, order) not supported");
  }
#define HFX_CASE(D, P)                                                   \
  if (dim == D && order == P) {                                          \
    switch (waves) {                                                     \
      case 1: launch_aff<D, P, false, 1>(a, grid_cap, s); return;       \
      case 2: launch_aff<D, P, false, 2>(a, grid_cap, s); return;       \
      case 3: launch_aff<D, P, false, 3>(a, grid_cap, s); return;       \
      case 4: launch_aff<D, P, false, 4>(a, grid_cap, s); return;       \
      default: throw std::runtime_error("affine element kernel: waves per workgroup must be 1 to 4"); \
    }                                                                    \
  }
  HFX_CASE(2, 1) HFX_CASE(2, 2) HFX_CASE(2, 3) HFX_CASE(2, 4) HFX_CASE(2, 5)
  HFX_CASE(3, 1) HFX_CASE(3, 2) HFX_CASE(3, 3)
#undef HFX_CASE
  throw std::runtime_error("affine element kernel: (dim, order) not supported");
}

}  // namespace hfx

namespace hfx {

size_t affine_batch_lds_bytes(int order, int waves) {
  switch (order) {
    case 1: return aff_lds<2, 1>(false, waves);
    case 2: return aff_lds<2, 2>(false, waves);
    case 3: return aff_lds<2, 3>(false, waves);
    case 4: return aff_lds<2, 4>(false, waves);
    case 5: return aff_lds<2, 5>(false, waves);
    default: throw std::runtime_error("batched element kernel: order must be 1..5");
  }
}

void launch_affine_batch(int order, int waves, const ElemArgs* probs, const BatchItem* items, int nitems, int grid,
                         hipStream_t s, bool attribute_only) {
  if (nitems <= 0) return;
  const size_t lds = affine_batch_lds_bytes(order, waves);
  if (lds > 160 * 1024) throw std::runtime_error("batched element kernel: LDS exceeds 160 KB for this wave count");
  if (grid <= 0) throw std::runtime_error("batched element kernel: empty grid");
  auto go = [&](auto kern) {
    // the LDS limit is raised once per device, outside any stream capture
    // (prepare_affine_batch)
    static PerDeviceLaunch once;
    once.lds_attr((const void*)kern, lds);
    if (attribute_only) return;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * waves), lds, s, probs, items, nitems);
    HFX_HIP(hipGetLastError());
  };
#define HFX_BW(P)                                                                     \
  case P:                                                                             \
    switch (waves) {                                                                  \
      case 1: go(hdg_affine_batch_kernel<P, 1>); break;                              \
      case 2: go(hdg_affine_batch_kernel<P, 2>); break;                              \
      case 4: go(hdg_affine_batch_kernel<P, 4>); break;                              \
      default: throw std::runtime_error("batched element kernel: waves must be 1, 2 or 4"); \
    }                                                                                 \
    break;
  switch (order) {
    HFX_BW(1) HFX_BW(2) HFX_BW(3) HFX_BW(4) HFX_BW(5)
    default: throw std::runtime_error("batched element kernel: order must be 1..5");
  }
#undef HFX_BW
}

void prepare_affine_batch(int order, int waves) {
  const BatchItem* none = nullptr;
  launch_affine_batch(order, waves, nullptr, none, 1, 1, nullptr, true);
}

}  // namespace hfx
