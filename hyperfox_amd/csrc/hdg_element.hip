// Fused per-element HDG assembly + static condensation (FP64).
//
// One workgroup owns one element at a time (grid-stride over elements).  The
// element's operator blocks are built at the cubature points and condensed
// onto the face traces without leaving LDS; only the condensed blocks
// S, S0, U, Q, U0, Q0 are written to HBM (element-major, coalesced).
//
// Math (reference: HDGBase.cpp:67-158, HDGDiffusion.cpp:31-145,
// HDGConvection.cpp:31-104, Source.cpp:5-48, HDGSolver.cpp:331-348), with
// Sqq = I_d (x) M, so Sqq^-1 [Squ | Sql] is d independent mass solves:
//   A_k = M^-1 Cg_k,  B_k = M^-1 Sql_k,  K = Suu - sum_k Suq_k A_k
//   U = -K^-1 (Sul - sum_k (Suq_k M^-1) Sql_k),  U0 = K^-1 F_u
//   Q_k = -A_k U - B_k,  Q0_k = -A_k U0
//   S = Slu U + sum_k Slq_k Q_k + Sll,  S0 = -Slu U0 - sum_k Slq_k Q0_k
// Face blocks are kept as per-face weighted face masses
//   FM[f][s][a][b] = sum_q w_s(f,q) psi_a psi_b,
//   s = 0: tau, 1: b.n, 2..2+d-1: n_k, 2+d..2+2d-1: (D n)_k   (all times dS).
#include "dense_gj.h"
#include "hdg_element.h"
#include "hfx_internal.h"

namespace hfx {

namespace {

template <int D, int P, int NT>
struct Layout {
  using Z = ESz<D, P>;
  static constexpr int NN = Z::NN, NNF = Z::NNF, NIP = Z::NIP, NIPF = Z::NIPF, NF = Z::NF, NL = Z::NL;
  static constexpr int NS = 2 + 2 * D;
  static constexpr int NC1 = NN + D * NN + NN;  // [M | Cg_0..Cg_{d-1} | I]
  static constexpr int NC2 = NN + NL + 1;       // [K | -R | Fu]
  // persistent regions
  static constexpr int O_TAB = 0;
  static constexpr int O_FM = O_TAB + Z::TAB;
  static constexpr int O_AUG1 = O_FM + NF * NS * NNF * NNF;
  static constexpr int O_AUG2 = O_AUG1 + NN * NC1;
  static constexpr int O_R5 = O_AUG2 + NN * NC2;  // SUQ[D][NN][NN] SUU[NN][NN] FU[NN]
  static constexpr int O_SUQ = O_R5;
  static constexpr int O_SUU = O_SUQ + D * NN * NN;
  static constexpr int O_FU = O_SUU + NN * NN;
  static constexpr int O_R4 = O_FU + NN;
  // R4 phase 1: gathered inputs and IP data
  static constexpr int O_X = O_R4;
  static constexpr int O_DN = O_X + NN * D;
  static constexpr int O_BN = O_DN + NN * D * D;
  static constexpr int O_TAU = O_BN + NN * D;
  static constexpr int O_FQ = O_TAU + NF * NNF;
  static constexpr int O_RQ = O_FQ + NIP;
  static constexpr int O_DV = O_RQ + NIP;
  static constexpr int O_JI = O_DV + NIP;
  static constexpr int O_DQ = O_JI + NIP * D * D;
  static constexpr int O_BQ = O_DQ + NIP * D * D;
  static constexpr int O_GR = O_BQ + NIP * D;
  static constexpr int O_FWT = O_GR + NIP * NN * D;
  static constexpr int END1 = O_FWT + NF * NIPF * NS;
  // R4 phase 2: Z_k = Suq_k M^-1
  static constexpr int O_Z = O_R4;
  static constexpr int END2 = O_Z + D * NN * NN;
  // R4 phase 3: Q_k (LDS copy) and Q0_k
  static constexpr int O_QL = O_R4;
  static constexpr int O_Q0L = O_QL + D * NN * NL;
  static constexpr int END3 = O_Q0L + D * NN;
  static constexpr int TOTAL = END1 > END2 ? (END1 > END3 ? END1 : END3) : (END2 > END3 ? END2 : END3);
};

template <int D>
__device__ inline double det_small(const double* J) {
  if (D == 2) return J[0] * J[3] - J[1] * J[2];
  return J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) + J[2] * (J[3] * J[7] - J[4] * J[6]);
}

template <int D>
__device__ inline void inv_small(const double* J, double dt, double* Ji) {
  const double r = 1.0 / dt;
  if (D == 2) {
    Ji[0] = J[3] * r; Ji[1] = -J[1] * r; Ji[2] = -J[2] * r; Ji[3] = J[0] * r;
  } else {
    Ji[0] = (J[4] * J[8] - J[5] * J[7]) * r; Ji[1] = (J[2] * J[7] - J[1] * J[8]) * r; Ji[2] = (J[1] * J[5] - J[2] * J[4]) * r;
    Ji[3] = (J[5] * J[6] - J[3] * J[8]) * r; Ji[4] = (J[0] * J[8] - J[2] * J[6]) * r; Ji[5] = (J[2] * J[3] - J[0] * J[5]) * r;
    Ji[6] = (J[3] * J[7] - J[4] * J[6]) * r; Ji[7] = (J[1] * J[6] - J[0] * J[7]) * r; Ji[8] = (J[0] * J[4] - J[1] * J[3]) * r;
  }
}

// Orders whose working set does not fit the 160 KB of LDS keep it in a
// device workspace, one slab per workgroup (the same layout, read and written
// through the caches); the workgroup barriers order those accesses as they
// order the LDS ones.
template <int D, int P>
constexpr bool element_in_workspace() {
  return (size_t)Layout<D, P, 64>::TOTAL * sizeof(double) > 160 * 1024;
}

template <int D, int P, int NT>
__global__ __launch_bounds__(NT) void hdg_element_kernel(ElemArgs a) {
  using L = Layout<D, P, NT>;
  using Z = ESz<D, P>;
  constexpr int NN = L::NN, NNF = L::NNF, NIP = L::NIP, NIPF = L::NIPF, NF = L::NF, NL = L::NL, NS = L::NS;
  constexpr int NC1 = L::NC1, NC2 = L::NC2;
  extern __shared__ double lds_ws[];
  double* sm = lds_ws;
  if constexpr (element_in_workspace<D, P>()) sm = a.gws + (size_t)blockIdx.x * L::TOTAL;
  __shared__ int fn[NF * NNF];
  __shared__ int pos[NF * NN];
  __shared__ int vopp[NF];
  __shared__ int perm1[NN], perm2[NN];
  const int tid = threadIdx.x;

  for (int t = tid; t < Z::TAB; t += NT) sm[L::O_TAB + t] = a.tabs[t];
  for (int t = tid; t < NF * NN; t += NT) pos[t] = -1;
  __syncthreads();
  for (int t = tid; t < NF * NNF; t += NT) {
    fn[t] = a.faceNodes[t];
    pos[(t / NNF) * NN + a.faceNodes[t]] = t % NNF;
  }
  __syncthreads();
  if (tid < NF) {
    int v = -1;
    for (int i = 0; i < NN && v < 0; ++i)
      if (pos[tid * NN + i] < 0) v = i;
    vopp[tid] = v;
  }
  __syncthreads();

  const double* PHI = sm + L::O_TAB + Z::T_PHI;
  const double* DPHI = sm + L::O_TAB + Z::T_DPHI;
  const double* FPHI = sm + L::O_TAB + Z::T_FPHI;
  const double* FDPHI = sm + L::O_TAB + Z::T_FDPHI;
  const double* W = sm + L::O_TAB + Z::T_W;
  const double* FW = sm + L::O_TAB + Z::T_FW;
  double* FM = sm + L::O_FM;
  double* AUG1 = sm + L::O_AUG1;
  double* AUG2 = sm + L::O_AUG2;
  double* SUQ = sm + L::O_SUQ;
  double* SUU = sm + L::O_SUU;
  double* FU = sm + L::O_FU;
  double* X = sm + L::O_X;
  double* DN = sm + L::O_DN;
  double* BN = sm + L::O_BN;
  double* TAU = sm + L::O_TAU;
  double* FQ = sm + L::O_FQ;
  double* RQ = sm + L::O_RQ;
  double* DV = sm + L::O_DV;
  double* JI = sm + L::O_JI;
  double* DQ = sm + L::O_DQ;
  double* BQ = sm + L::O_BQ;
  double* GR = sm + L::O_GR;
  double* FWT = sm + L::O_FWT;
  double* ZB = sm + L::O_Z;
  double* QL = sm + L::O_QL;
  double* Q0L = sm + L::O_Q0L;

  const int e_end = a.el_offset + a.nEl;
  for (int e = a.el_offset + blockIdx.x; e < e_end; e += gridDim.x) {
    // ---------------------------------------------------------- gather
    for (int t = tid; t < NN * D; t += NT) {
      int i = t / D, k = t % D;
      X[t] = a.coords[(size_t)a.cells[(size_t)e * NN + i] * D + k];
    }
    for (int t = tid; t < NN * D * D; t += NT) {
      int i = t / (D * D), r = (t / D) % D, c = t % D;
      double v = (r == c) ? 1.0 : 0.0;
      if (!a.useDiff) v = 0.0;
      else if (a.ndD == 1) v = (r == c) ? a.Dn[a.cells[(size_t)e * NN + i]] : 0.0;
      else if (a.ndD == D * D) v = a.Dn[(size_t)a.cells[(size_t)e * NN + i] * D * D + c * D + r];
      DN[t] = v;  // DN[i][r][c] = D_rc at node i
    }
    for (int t = tid; t < NN * D; t += NT)
      BN[t] = a.hasB ? a.bn[(size_t)a.cells[(size_t)e * NN + t / D] * D + t % D] : 0.0;
    for (int t = tid; t < NF * NNF; t += NT) {
      int f = t / NNF, j = t % NNF;
      int fc = a.cell2face[(size_t)e * NF + f];
      int side = (a.tau_sides == 2 && a.face2cell[2 * (size_t)fc] != e) ? 1 : 0;
      int p = a.face_perm[((size_t)e * NF + f) * NNF + j];
      TAU[t] = a.tau[((size_t)fc * NNF + p) * a.tau_sides + side];
    }
    for (int t = tid; t < NIP; t += NT) {
      double fq = (a.hasF && a.kind != 0) ? a.f_ip[(size_t)e * NIP + t] : 0.0;
      if (a.u_old && a.ts_idt != 0.0) {  // implicit Euler: + u_old(x_q) / dt
        double uo = 0.0;
        for (int j = 0; j < NN; ++j) uo += PHI[t * NN + j] * a.u_old[(size_t)e * NN + j];
        fq += a.ts_idt * uo;
      }
      FQ[t] = fq;
      RQ[t] = (a.hasR ? a.r_ip[(size_t)e * NIP + t] : 0.0) + a.ts_idt;
    }
    __syncthreads();
    // ---------------------------------------------- IP geometry (element)
    for (int q = tid; q < NIP; q += NT) {
      double J[D * D];
      for (int t = 0; t < D * D; ++t) J[t] = 0.0;
      for (int i = 0; i < NN; ++i)
        for (int r = 0; r < D; ++r)
          for (int c = 0; c < D; ++c) J[r * D + c] += DPHI[(q * NN + i) * D + r] * X[i * D + c];
      double dt = det_small<D>(J);
      DV[q] = W[q] * dt;
      inv_small<D>(J, dt, JI + q * D * D);
      for (int t = 0; t < D * D; ++t) {
        double s = 0.0;
        for (int i = 0; i < NN; ++i) s += PHI[q * NN + i] * DN[i * D * D + t];
        DQ[q * D * D + t] = s;
      }
      for (int k = 0; k < D; ++k) {
        double s = 0.0;
        for (int i = 0; i < NN; ++i) s += PHI[q * NN + i] * BN[i * D + k];
        BQ[q * D + k] = s;
      }
    }
    // ------------------------------------------------- IP geometry (faces)
    for (int t = tid; t < NF * NIPF; t += NT) {
      int f = t / NIPF, q = t % NIPF;
      double Jf[(D - 1) * D];
      for (int u = 0; u < (D - 1) * D; ++u) Jf[u] = 0.0;
      for (int j = 0; j < NNF; ++j) {
        const double* xj = X + fn[f * NNF + j] * D;
        for (int r = 0; r < D - 1; ++r)
          for (int c = 0; c < D; ++c) Jf[r * D + c] += FDPHI[(q * NNF + j) * (D - 1) + r] * xj[c];
      }
      double nv[D];
      double meas;
      if (D == 2) {
        meas = sqrt(Jf[0] * Jf[0] + Jf[1] * Jf[1]);
        nv[0] = Jf[1];
        nv[1] = -Jf[0];
      } else {
        double g00 = 0, g01 = 0, g11 = 0;
        for (int c = 0; c < D; ++c) {
          g00 += Jf[c] * Jf[c];
          g01 += Jf[c] * Jf[D + c];
          g11 += Jf[D + c] * Jf[D + c];
        }
        meas = sqrt(g00 * g11 - g01 * g01);
        nv[0] = Jf[1] * Jf[D + 2] - Jf[2] * Jf[D + 1];
        nv[1] = Jf[2] * Jf[D + 0] - Jf[0] * Jf[D + 2];
        nv[2] = Jf[0] * Jf[D + 1] - Jf[1] * Jf[D + 0];
      }
      double nn = 0.0;
      for (int c = 0; c < D; ++c) nn += nv[c] * nv[c];
      nn = 1.0 / sqrt(nn);
      double prod = 0.0;
      const int v0 = fn[f * NNF], vn = vopp[f];
      for (int c = 0; c < D; ++c) {
        nv[c] *= nn;
        prod += (X[vn * D + c] - X[v0 * D + c]) * nv[c];
      }
      if (prod > 0)
        for (int c = 0; c < D; ++c) nv[c] = -nv[c];
      double tq = 0.0, bq[D], dq[D * D];
      for (int c = 0; c < D; ++c) bq[c] = 0.0;
      for (int u = 0; u < D * D; ++u) dq[u] = 0.0;
      for (int j = 0; j < NNF; ++j) {
        double ps = FPHI[q * NNF + j];
        int node = fn[f * NNF + j];
        tq += ps * TAU[f * NNF + j];
        for (int c = 0; c < D; ++c) bq[c] += ps * BN[node * D + c];
        for (int u = 0; u < D * D; ++u) dq[u] += ps * DN[node * D * D + u];
      }
      const double dvf = FW[q] * meas;
      double bn = 0.0;
      for (int c = 0; c < D; ++c) bn += bq[c] * nv[c];
      double* w = FWT + (f * NIPF + q) * NS;
      w[0] = dvf * tq;
      w[1] = dvf * bn;
      for (int k = 0; k < D; ++k) {
        w[2 + k] = dvf * nv[k];
        double s = 0.0;
        for (int c = 0; c < D; ++c) s += dq[k * D + c] * nv[c];
        w[2 + D + k] = dvf * s;
      }
    }
    __syncthreads();
    // physical gradients GR[q][i][k] = (J^-1 grad_xi phi_i)_k
    for (int t = tid; t < NIP * NN; t += NT) {
      int q = t / NN, i = t % NN;
      for (int k = 0; k < D; ++k) {
        double s = 0.0;
        for (int r = 0; r < D; ++r) s += JI[q * D * D + k * D + r] * DPHI[(q * NN + i) * D + r];
        GR[t * D + k] = s;
      }
    }
    // weighted face masses
    for (int t = tid; t < NF * NS * NNF * NNF; t += NT) {
      int b = t % NNF, aa = (t / NNF) % NNF, s = (t / (NNF * NNF)) % NS, f = t / (NS * NNF * NNF);
      double acc = 0.0;
      for (int q = 0; q < NIPF; ++q) acc += FWT[(f * NIPF + q) * NS + s] * FPHI[q * NNF + aa] * FPHI[q * NNF + b];
      FM[t] = acc;
    }
    __syncthreads();
    // -------------------------------------------------- volume blocks
    for (int t = tid; t < NN * NN; t += NT) {
      const int i = t / NN, j = t % NN;
      double m = 0.0, suu = 0.0, cg[D], sq[D];
      for (int k = 0; k < D; ++k) { cg[k] = 0.0; sq[k] = 0.0; }
      for (int q = 0; q < NIP; ++q) {
        const double w = DV[q];
        const double pi = PHI[q * NN + i], pj = w * PHI[q * NN + j];
        const double* g = GR + (q * NN + i) * D;
        m += pi * pj;
        double bg = 0.0;
        for (int k = 0; k < D; ++k) {
          cg[k] += g[k] * pj;
          double s = 0.0;
          for (int c = 0; c < D; ++c) s += DQ[q * D * D + k * D + c] * g[c];
          sq[k] += s * pj;
          bg += BQ[q * D + k] * g[k];
        }
        suu += (RQ[q] * pi - bg) * pj;
      }
      for (int f = 0; f < NF; ++f) {
        const int ai = pos[f * NN + i], bj = pos[f * NN + j];
        if (ai >= 0 && bj >= 0) {
          const double* F0 = FM + (f * NS) * NNF * NNF + ai * NNF + bj;
          suu += F0[0];
          for (int k = 0; k < D; ++k) sq[k] -= F0[(2 + D + k) * NNF * NNF];
        }
      }
      AUG1[i * NC1 + j] = m;
      for (int k = 0; k < D; ++k) {
        AUG1[i * NC1 + NN + k * NN + j] = cg[k];
        SUQ[(k * NN + i) * NN + j] = sq[k];
      }
      AUG1[i * NC1 + NN + D * NN + j] = (i == j) ? 1.0 : 0.0;
      SUU[i * NN + j] = suu;
    }
    for (int i = tid; i < NN; i += NT) {
      double s = 0.0;
      for (int q = 0; q < NIP; ++q) s += DV[q] * FQ[q] * PHI[q * NN + i];
      FU[i] = s;
    }
    __syncthreads();
    if (a.ts_gen) {
      // general time transform of the u rows: F_u first (it applies the
      // model's rows to w), then the rows; M is the mass block of AUG1
      const double* w = a.ts_w + (size_t)e * Z::NLOC;
      const double* uo = a.u_old + (size_t)e * NN;
      for (int i = tid; i < NN; i += NT) {
        double c = 0.0, mu = 0.0;
        for (int j = 0; j < NN; ++j) {
          c += SUU[i * NN + j] * w[j];
          mu += AUG1[i * NC1 + j] * uo[j];
          for (int k = 0; k < D; ++k) c += SUQ[(k * NN + i) * NN + j] * w[NN + j * D + k];
        }
        for (int f = 0; f < NF; ++f) {
          const int ai = pos[f * NN + i];
          if (ai < 0) continue;
          for (int b = 0; b < NNF; ++b)
            c += (-FM[((f * NS + 0) * NNF + ai) * NNF + b] + FM[((f * NS + 1) * NNF + ai) * NNF + b]) *
                 w[NN * (D + 1) + f * NNF + b];
        }
        FU[i] = a.ts_phi * FU[i] + a.ts_mu * mu - c;
      }
      __syncthreads();
      for (int t = tid; t < NN * NN; t += NT) SUU[t] = a.ts_auu * SUU[t] + a.ts_mu * AUG1[(t / NN) * NC1 + t % NN];
      for (int t = tid; t < D * NN * NN; t += NT) SUQ[t] *= a.ts_arest;
      __syncthreads();
    }

    if (a.Kloc) {
      // assembly-only path: write the local matrix in the reference layout
      constexpr int NLOC = Z::NLOC;
      constexpr int OQ = NN, OL = NN * (D + 1);
      double* K = a.Kloc + (size_t)(e - a.el_offset) * NLOC * NLOC;
      for (int t = tid; t < NLOC * NLOC; t += NT) {
        const int r = t / NLOC, c = t % NLOC;
        double v = 0.0;
        if (r < OQ) {
          const int i = r;
          if (c < OQ) v = SUU[i * NN + c];
          else if (c < OL) v = SUQ[(((c - OQ) % D) * NN + i) * NN + (c - OQ) / D];
          else {
            const int f = (c - OL) / NNF, b = (c - OL) % NNF, ai = pos[f * NN + i];
            if (ai >= 0)
              v = a.ts_arest * (-FM[((f * NS + 0) * NNF + ai) * NNF + b] + FM[((f * NS + 1) * NNF + ai) * NNF + b]);
          }
        } else if (r < OL) {
          const int i = (r - OQ) / D, k = (r - OQ) % D;
          if (c < OQ) v = AUG1[i * NC1 + NN + k * NN + c];
          else if (c < OL) { if ((c - OQ) % D == k) v = AUG1[i * NC1 + (c - OQ) / D]; }
          else {
            const int f = (c - OL) / NNF, b = (c - OL) % NNF, ai = pos[f * NN + i];
            if (ai >= 0) v = -FM[((f * NS + 2 + k) * NNF + ai) * NNF + b];
          }
        } else {
          const int f = (r - OL) / NNF, aa = (r - OL) % NNF;
          if (c < OQ) {
            const int bj = pos[f * NN + c];
            if (bj >= 0) v = FM[((f * NS + 0) * NNF + aa) * NNF + bj];
          } else if (c < OL) {
            const int j = (c - OQ) / D, k = (c - OQ) % D, bj = pos[f * NN + j];
            if (bj >= 0) v = -FM[((f * NS + 2 + D + k) * NNF + aa) * NNF + bj];
          } else {
            const int g = (c - OL) / NNF, b = (c - OL) % NNF;
            if (g == f) v = -FM[((f * NS + 0) * NNF + aa) * NNF + b] + FM[((f * NS + 1) * NNF + aa) * NNF + b];
          }
        }
        // time schemes: the reference's u rows are ts_dt x (these rows)
        K[t] = (a.ts_dt > 0.0 && r < OQ) ? a.ts_dt * v : v;
      }
      double* F = a.Floc + (size_t)(e - a.el_offset) * NLOC;
      for (int t = tid; t < NLOC; t += NT) F[t] = t < NN ? (a.ts_dt > 0.0 ? a.ts_dt * FU[t] : FU[t]) : 0.0;
      __syncthreads();
      continue;
    }

    // ---------------------------------- [M | Cg | I] -> [I | A | M^-1]
    gauss_jordan<NN, NC1, NT>(AUG1, perm1);
    // K = Suu - sum_k Suq_k A_k ;  Z_k = Suq_k M^-1
    for (int t = tid; t < NN * NN; t += NT) {
      const int i = t / NN, j = t % NN;
      double s = SUU[t];
      for (int k = 0; k < D; ++k)
        for (int l = 0; l < NN; ++l) s -= SUQ[(k * NN + i) * NN + l] * AUG1[perm1[l] * NC1 + NN + k * NN + j];
      AUG2[i * NC2 + j] = s;
    }
    for (int t = tid; t < D * NN * NN; t += NT) {
      const int k = t / (NN * NN), i = (t / NN) % NN, j = t % NN;
      double s = 0.0;
      for (int l = 0; l < NN; ++l) s += SUQ[(k * NN + i) * NN + l] * AUG1[perm1[l] * NC1 + NN + D * NN + j];
      ZB[t] = s;
    }
    __syncthreads();
    // -R = -(Sul - sum_k Z_k Sql_k) ; Fu
    for (int t = tid; t < NN * NL; t += NT) {
      const int i = t / NL, c = t % NL, f = c / NNF, b = c % NNF;
      double r = 0.0;
      const int ai = pos[f * NN + i];
      if (ai >= 0) r = a.ts_arest * (-FM[((f * NS + 0) * NNF + ai) * NNF + b] + FM[((f * NS + 1) * NNF + ai) * NNF + b]);
      for (int k = 0; k < D; ++k)
        for (int aa = 0; aa < NNF; ++aa)
          r += ZB[(k * NN + i) * NN + fn[f * NNF + aa]] * FM[((f * NS + 2 + k) * NNF + aa) * NNF + b];
      AUG2[i * NC2 + NN + c] = -r;
    }
    for (int i = tid; i < NN; i += NT) AUG2[i * NC2 + NN + NL] = FU[i];
    __syncthreads();
    // ---------------------------------------- [K | -R | Fu] -> [I | U | U0]
    gauss_jordan<NN, NC2, NT>(AUG2, perm2);
    // Q_k = -A_k U - M^-1 Sql_k ;  Q0_k = -A_k U0 ; write U, U0, Q, Q0
    double* Ue = a.U + (size_t)e * NN * NL;
    double* Qe = a.Q + (size_t)e * NN * D * NL;
    for (int t = tid; t < NN * NL; t += NT) Ue[t] = AUG2[perm2[t / NL] * NC2 + NN + t % NL];
    for (int t = tid; t < NN; t += NT) a.U0[(size_t)e * NN + t] = AUG2[perm2[t] * NC2 + NN + NL];
    for (int t = tid; t < D * NN * NL; t += NT) {
      const int k = t / (NN * NL), i = (t / NL) % NN, c = t % NL, f = c / NNF, b = c % NNF;
      double s = 0.0;
      const int pi = perm1[i];
      for (int j = 0; j < NN; ++j) s -= AUG1[pi * NC1 + NN + k * NN + j] * AUG2[perm2[j] * NC2 + NN + c];
      for (int aa = 0; aa < NNF; ++aa)
        s += AUG1[pi * NC1 + NN + D * NN + fn[f * NNF + aa]] * FM[((f * NS + 2 + k) * NNF + aa) * NNF + b];
      QL[t] = s;
      Qe[(size_t)(i * D + k) * NL + c] = s;
    }
    for (int t = tid; t < D * NN; t += NT) {
      const int k = t / NN, i = t % NN;
      double s = 0.0;
      for (int j = 0; j < NN; ++j) s -= AUG1[perm1[i] * NC1 + NN + k * NN + j] * AUG2[perm2[j] * NC2 + NN + NL];
      Q0L[t] = s;
      a.Q0[(size_t)e * NN * D + i * D + k] = s;
    }
    __syncthreads();
    // S = Slu U + sum_k Slq_k Q_k + Sll ; S0 = -Slu U0 - sum_k Slq_k Q0_k
    // (explicit HDG modes: S = Sll, S0 = -Slu sol_old - Slq flux_old)
    double* Se = a.S + (size_t)e * NL * NL;
    for (int t = tid; t < NL * NL; t += NT) {
      const int r = t / NL, c = t % NL, f = r / NNF, aa = r % NNF;
      double s = 0.0;
      for (int b = 0; b < NNF && !a.hdg_explicit; ++b) {
        const int j = fn[f * NNF + b];
        s += FM[((f * NS + 0) * NNF + aa) * NNF + b] * AUG2[perm2[j] * NC2 + NN + c];
        for (int k = 0; k < D; ++k) s -= FM[((f * NS + 2 + D + k) * NNF + aa) * NNF + b] * QL[(k * NN + j) * NL + c];
      }
      if (c / NNF == f) {
        const int b = c % NNF;
        s += -FM[((f * NS + 0) * NNF + aa) * NNF + b] + FM[((f * NS + 1) * NNF + aa) * NNF + b];
      }
      Se[t] = s;
    }
    for (int r = tid; r < NL; r += NT) {
      const int f = r / NNF, aa = r % NNF;
      double s = 0.0;
      for (int b = 0; b < NNF; ++b) {
        const int j = fn[f * NNF + b];
        const double uj = a.hdg_explicit ? a.sol_old[(size_t)e * NN + j] : AUG2[perm2[j] * NC2 + NN + NL];
        s -= FM[((f * NS + 0) * NNF + aa) * NNF + b] * uj;
        for (int k = 0; k < D; ++k) {
          const double qkj = a.hdg_explicit ? a.flux_old[((size_t)e * NN + j) * D + k] : Q0L[k * NN + j];
          s += FM[((f * NS + 2 + D + k) * NNF + aa) * NNF + b] * qkj;
        }
      }
      a.S0[(size_t)e * NL + r] = s;
    }
    __syncthreads();
  }
}

template <int D, int P>
__global__ void ip_coords_kernel(int nEl, const double* __restrict__ tabs, const double* __restrict__ coords,
                                 const int* __restrict__ cells, double* __restrict__ out) {
  using Z = ESz<D, P>;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nEl * Z::NIP) return;
  const int e = (int)(t / Z::NIP), q = (int)(t % Z::NIP);
  double x[D];
  for (int k = 0; k < D; ++k) x[k] = 0.0;
  for (int i = 0; i < Z::NN; ++i) {
    const double ph = tabs[Z::T_PHI + q * Z::NN + i];
    const int node = cells[(size_t)e * Z::NN + i];
    for (int k = 0; k < D; ++k) x[k] += ph * coords[(size_t)node * D + k];
  }
  for (int k = 0; k < D; ++k) out[(size_t)t * D + k] = x[k];
}

template <int D, int P, int NT>
void launch_t(const ElemArgs& a, int grid_cap, hipStream_t s) {
  using L = Layout<D, P, NT>;
  constexpr bool gws = element_in_workspace<D, P>();
  const size_t lds = gws ? 0 : (size_t)L::TOTAL * sizeof(double);
  auto kern = hdg_element_kernel<D, P, NT>;
  static PerDeviceLaunch once;
  if (!gws) once.lds_attr((const void*)kern, lds);
  int grid = a.nEl;
  if (grid_cap > 0 && grid > grid_cap) grid = grid_cap;
  if (gws) {
    if (!a.gws || a.gws_slots <= 0) throw std::runtime_error("element kernel: this order needs its device workspace");
    if (grid > a.gws_slots) grid = a.gws_slots;
  }
  if (grid <= 0) return;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), lds, s, a);
  HFX_HIP(hipGetLastError());
}

template <int D, int P>
void launch_dp(const ElemArgs& a, int nt, int grid_cap, hipStream_t s) {
  switch (nt) {
    case 64: launch_t<D, P, 64>(a, grid_cap, s); break;
    case 128: launch_t<D, P, 128>(a, grid_cap, s); break;
    case 256: launch_t<D, P, 256>(a, grid_cap, s); break;
    default: throw std::runtime_error("element kernel: threads per element must be 64, 128 or 256");
  }
}

}  // namespace

size_t element_lds_bytes(int dim, int order) {
#define HFX_CASE(D, P)                                                                             \
  if (dim == D && order == P)                                                                      \
    return element_in_workspace<D, P>() ? 0 : (size_t)Layout<D, P, 64>::TOTAL * sizeof(double);
  HFX_CASE(2, 1) HFX_CASE(2, 2) HFX_CASE(2, 3) HFX_CASE(2, 4) HFX_CASE(2, 5)
  HFX_CASE(2, 6) HFX_CASE(2, 7) HFX_CASE(2, 8) HFX_CASE(2, 9) HFX_CASE(2, 10)
  HFX_CASE(3, 1) HFX_CASE(3, 2) HFX_CASE(3, 3) HFX_CASE(3, 4) HFX_CASE(3, 5)
#undef HFX_CASE
  throw std::runtime_error("element kernel: (dim, order) not supported on the GPU path");
}

size_t element_workspace_doubles(int dim, int order) {
#define HFX_CASE(D, P) \
  if (dim == D && order == P) return element_in_workspace<D, P>() ? (size_t)Layout<D, P, 64>::TOTAL : 0;
  HFX_CASE(2, 1) HFX_CASE(2, 2) HFX_CASE(2, 3) HFX_CASE(2, 4) HFX_CASE(2, 5)
  HFX_CASE(2, 6) HFX_CASE(2, 7) HFX_CASE(2, 8) HFX_CASE(2, 9) HFX_CASE(2, 10)
  HFX_CASE(3, 1) HFX_CASE(3, 2) HFX_CASE(3, 3) HFX_CASE(3, 4) HFX_CASE(3, 5)
#undef HFX_CASE
  throw std::runtime_error("element kernel: (dim, order) not supported on the GPU path");
}

void launch_element(int dim, int order, const ElemArgs& a, int nt, int grid_cap, hipStream_t s) {
#define HFX_CASE(D, P) \
  if (dim == D && order == P) { launch_dp<D, P>(a, nt, grid_cap, s); return; }
  HFX_CASE(2, 1) HFX_CASE(2, 2) HFX_CASE(2, 3) HFX_CASE(2, 4) HFX_CASE(2, 5)
  HFX_CASE(2, 6) HFX_CASE(2, 7) HFX_CASE(2, 8) HFX_CASE(2, 9) HFX_CASE(2, 10)
  HFX_CASE(3, 1) HFX_CASE(3, 2) HFX_CASE(3, 3) HFX_CASE(3, 4) HFX_CASE(3, 5)
#undef HFX_CASE
  throw std::runtime_error("element kernel: (dim, order) not supported on the GPU path");
}

void launch_ip_coords(int dim, int order, int nEl, const double* tabs, const double* coords, const int* cells,
                      double* out, hipStream_t s) {
#define HFX_CASE(D, P)                                                                                   \
  if (dim == D && order == P) {                                                                          \
    long long n = (long long)nEl * ESz<D, P>::NIP;                                                       \
    if (n > 0) hipLaunchKernelGGL((ip_coords_kernel<D, P>), dim3(ceil_div(n, 256)), dim3(256), 0, s, nEl, \
                                  tabs, coords, cells, out);                                             \
    HFX_HIP(hipGetLastError());                                                                          \
    return;                                                                                              \
  }
  HFX_CASE(2, 1) HFX_CASE(2, 2) HFX_CASE(2, 3) HFX_CASE(2, 4) HFX_CASE(2, 5)
  HFX_CASE(2, 6) HFX_CASE(2, 7) HFX_CASE(2, 8) HFX_CASE(2, 9) HFX_CASE(2, 10)
  HFX_CASE(3, 1) HFX_CASE(3, 2) HFX_CASE(3, 3) HFX_CASE(3, 4) HFX_CASE(3, 5)
#undef HFX_CASE
  throw std::runtime_error("ip coords: (dim, order) not supported on the GPU path");
}

}  // namespace hfx
