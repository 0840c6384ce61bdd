// Burgers cell kernel (FP64): HDGBurgersModel's local system with m = d
// unknowns per node, condensed onto the face traces without leaving LDS.
//
// Local system (HDGBurgersModel.cpp:86-132): HDGBase with an m x m tau per
// face node (HDGBase.cpp:18-32, 67-158), HDGDiffusion with a scalar
// coefficient per node (HDGDiffusion.cpp:31-145), HDGUNabU
// (HDGUNabU.cpp:86-190) and one Source per component (Source.cpp:5-48), in
// the order u (i m + a), q ((i d + k) m + a), lambda ((f nNf + j) m + a).
//
// Every q block acts per component: Sqq = I (x) M, Squ_k, Sql_k, Suq_k and
// Slq_k are the scalar blocks Kronecker-multiplied by I_m.  Condensation
// (HDGSolver.cpp:331-348) therefore eliminates q with one scalar mass solve,
//   A_k = M^-1 Squ_k,  B_k = M^-1 Sql_k,
// and only the coupled u system of m nN rows is dense:
//   K = Suu - (sum_k Suq_k A_k) (x) I,  R = -(Sul - (sum_k Suq_k B_k) (x) I)
//   U = K^-1 R,  U0 = K^-1 F_u,  Q_k = -(A_k (x) I) U - B_k (x) I,
//   Q0_k = -(A_k (x) I) U0,
//   S = (Slu - P (x) I) U + Sll - Pb (x) I,  S0 = F_l - (Slu - P (x) I) U0,
// P = sum_k Slq_k A_k, Pb = sum_k Slq_k B_k.  HDGUNabU's right-hand side
// op [u/2; lambda/2] is evaluated in closed form at the IPs:
//   F_u(i, a) = -int u_a (phi_i div u + u . grad phi_i)
//               + sum over the faces of i of int psi (lambda.n)(lambda_a + u_a)/2,
//   F_l(f, j, a) = int_f psi_j (lambda.n)(lambda_a + u_a)/2.
#include "burgers.h"

#include <stdexcept>

#include "dense_gj.h"
#include "hdg_element.h"
#include "hfx_internal.h"

namespace hfx {

namespace {

constexpr int BT = 256;  // threads per cell

template <int D>
__device__ inline double bdet(const double* J) {
  if (D == 2) return J[0] * J[3] - J[1] * J[2];
  return J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) + J[2] * (J[3] * J[7] - J[4] * J[6]);
}

template <int D>
__device__ inline void binv(const double* J, double dt, double* Ji) {
  const double r = 1.0 / dt;
  if (D == 2) {
    Ji[0] = J[3] * r; Ji[1] = -J[1] * r; Ji[2] = -J[2] * r; Ji[3] = J[0] * r;
  } else {
    Ji[0] = (J[4] * J[8] - J[5] * J[7]) * r; Ji[1] = (J[2] * J[7] - J[1] * J[8]) * r; Ji[2] = (J[1] * J[5] - J[2] * J[4]) * r;
    Ji[3] = (J[5] * J[6] - J[3] * J[8]) * r; Ji[4] = (J[0] * J[8] - J[2] * J[6]) * r; Ji[5] = (J[2] * J[3] - J[0] * J[5]) * r;
    Ji[6] = (J[3] * J[7] - J[4] * J[6]) * r; Ji[7] = (J[1] * J[6] - J[0] * J[7]) * r; Ji[8] = (J[0] * J[4] - J[1] * J[3]) * r;
  }
}

// MDL 0: HDGBurgersModel (one diffusion coefficient for every component,
// HDGUNabU); MDL 1: HDGnGammaModel (D on n, G on Gamma, four HDGConvection
// velocities, no HDGUNabU); MDL 2: a scalar model with m = dim DOFs per node
// (HDGLaplaceModel / HDGDiffusionSource, HDGOperator::multiplyDOFs,
// HDGOperator.cpp:17-28): the MDL 0 system without HDGUNabU, i.e. one
// diffusion operator repeated on every component, coupled only by the
// matrix tau.
template <int D, int P, int MDL>
struct BLay {
  using Z = ESz<D, P>;
  static constexpr int M = D;
  static constexpr int NDC = MDL == 1 ? M : 1;  // diffusion coefficient sets
  static constexpr int NV = MDL == 1 ? M * M : 0;  // convection velocities
  static constexpr int NN = Z::NN, NNF = Z::NNF, NIP = Z::NIP, NIPF = Z::NIPF, NF = Z::NF, NL = Z::NL;
  static constexpr int MU = M * NN, ML = M * NL, MM = M * M, NFQ = NF * NIPF;
  static constexpr int NC1 = NN + D * NN + D * NL;  // [M | Squ_k | Sql_k]
  static constexpr int NC2 = MU + ML + 1;           // [K | R | F_u]
  static constexpr int O_X = Z::TAB;
  static constexpr int O_DN = O_X + NN * D;     // [c][i]
  static constexpr int O_VN = O_DN + NDC * NN;  // [i][t][k] nodal velocities
  static constexpr int O_TAU = O_VN + NN * NV * D;  // [f][j][t]
  static constexpr int O_SOL = O_TAU + NL * MM;  // [i][a]
  static constexpr int O_TR = O_SOL + MU;        // [f][j][a]
  static constexpr int O_FIP = O_TR + ML;        // [q][a]
  static constexpr int O_DV = O_FIP + NIP * M;
  static constexpr int O_GR = O_DV + NIP;        // [q][i][k] physical grad phi_i
  static constexpr int O_UQ = O_GR + NIP * NN * D;  // [q][a]
  static constexpr int O_DIV = O_UQ + NIP * M;
  static constexpr int O_DQ = O_DIV + NIP;      // [c][q]
  static constexpr int O_VQ = O_DQ + NDC * NIP;  // [q][t][k]
  static constexpr int O_FDS = O_VQ + NIP * NV * D;  // [f][q] face measure
  static constexpr int O_FN = O_FDS + NFQ;       // [f][q][k] outward normal
  static constexpr int O_LAM = O_FN + NFQ * D;   // [f][q][a] trace
  static constexpr int O_USF = O_LAM + NFQ * M;  // [f][q][a] cell solution
  static constexpr int O_TQ = O_USF + NFQ * M;   // [f][q][t] tau
  static constexpr int O_DFQ = O_TQ + NFQ * MM;  // [c][f q]
  static constexpr int O_VFN = O_DFQ + NDC * NFQ;  // [f q][t] v_t . n
  static constexpr int O_FL = O_VFN + NFQ * NV;  // F_l [ML]
  static constexpr int O_AUG1 = O_FL + ML;
  static constexpr int O_SUQ = O_AUG1 + NN * NC1;          // [c][k][i][j]
  static constexpr int O_SLQ = O_SUQ + NDC * D * NN * NN;  // [c][k][f nNf + a][j]
  static constexpr int O_KS = O_SLQ + NDC * D * NL * NN;   // [c] sum_k Suq_k A_k
  static constexpr int O_RS = O_KS + NDC * NN * NN;        // [c] sum_k Suq_k B_k
  static constexpr int O_P = O_RS + NDC * NN * NL;         // [c] sum_k Slq_k A_k
  static constexpr int O_PB = O_P + NDC * NL * NN;         // [c] sum_k Slq_k B_k
  static constexpr int O_SLU = O_PB + NDC * NL * NL;       // Slu - P (x) I  [ML][MU]
  static constexpr int O_AUG2 = O_SLU + ML * MU;
  static constexpr int TOTAL = O_AUG2 + MU * NC2;
  static_assert(MU <= 64, "Burgers kernel: at most 64 coupled u rows");
};

// Waves per SIMD the register allocation is held to (each workgroup puts
// one wave on each SIMD).  Without a target the allocator takes ~170 VGPRs,
// two waves per SIMD, and small cells wait on each other's barriers.  When
// three or more workgroups fit a CU's 160 KB of LDS the target is four
// (128 VGPRs, a few spills in the set-up phases): measured on gfx950 for
// 32,768 2-D p=3 cells, Burgers 4.9 -> 3.6 ms, nGamma 3.4 -> 3.1 ms (a target
// of three ran nGamma at 3.8 ms).  Larger cells keep their LDS-bound count.
template <int D, int P, int MDL>
constexpr int burgers_waves() {
  constexpr long bytes = (long)BLay<D, P, MDL>::TOTAL * 8 + 1024;
  constexpr long w = (160L * 1024) / bytes;
  return w >= 3 ? 4 : (w < 1 ? 1 : (int)w);
}

// Cells whose working set does not fit the LDS (3-D p = 3: 279 KB) keep it
// in a device workspace, one slab per workgroup in the same layout; the
// workgroup barriers order those accesses as they order the LDS ones.
template <int D, int P, int MDL>
constexpr bool burgers_in_workspace() {
  return (size_t)BLay<D, P, MDL>::TOTAL * sizeof(double) > 150 * 1024;
}

template <int D, int P, int MDL>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(burgers_waves<D, P, MDL>())))
void burgers_cell_kernel(BurgersArgs a) {
  using L = BLay<D, P, MDL>;
  constexpr int NDC = L::NDC, NV = L::NV;
  using Z = ESz<D, P>;
  constexpr int M = L::M, NN = L::NN, NNF = L::NNF, NIP = L::NIP, NIPF = L::NIPF, NF = L::NF, NL = L::NL;
  constexpr int MU = L::MU, ML = L::ML, MM = L::MM, NFQ = L::NFQ, NC1 = L::NC1, NC2 = L::NC2;
  extern __shared__ double lds_ws[];
  double* sm = lds_ws;
  if constexpr (burgers_in_workspace<D, P, MDL>()) sm = a.gws + (size_t)blockIdx.x * L::TOTAL;
  __shared__ int fn[NF * NNF];
  __shared__ int posf[NF * NN];
  __shared__ int vopp[NF];
  __shared__ int perm1[NN], perm2[MU];
  const int tid = threadIdx.x;

  for (int t = tid; t < Z::TAB; t += BT) sm[t] = a.tabs[t];
  for (int t = tid; t < NF * NN; t += BT) posf[t] = -1;
  __syncthreads();
  for (int t = tid; t < NF * NNF; t += BT) {
    fn[t] = a.faceNodes[t];
    posf[(t / NNF) * NN + a.faceNodes[t]] = t % NNF;
  }
  __syncthreads();
  if (tid < NF) {
    int v = -1;
    for (int i = 0; i < NN && v < 0; ++i)
      if (posf[tid * NN + i] < 0) v = i;
    vopp[tid] = v;
  }
  __syncthreads();

  const double* PHI = sm + Z::T_PHI;
  const double* DPHI = sm + Z::T_DPHI;
  const double* FPHI = sm + Z::T_FPHI;
  const double* FDPHI = sm + Z::T_FDPHI;
  const double* W = sm + Z::T_W;
  const double* FW = sm + Z::T_FW;
  double* X = sm + L::O_X;
  double* DN = sm + L::O_DN;
  double* TAU = sm + L::O_TAU;
  double* SOL = sm + L::O_SOL;
  double* TR = sm + L::O_TR;
  double* FIP = sm + L::O_FIP;
  double* DV = sm + L::O_DV;
  double* GR = sm + L::O_GR;
  double* UQ = sm + L::O_UQ;
  double* DIV = sm + L::O_DIV;
  double* DQ = sm + L::O_DQ;
  double* VN = sm + L::O_VN;
  double* VQ = sm + L::O_VQ;
  double* VFN = sm + L::O_VFN;
  double* FDS = sm + L::O_FDS;
  double* FN = sm + L::O_FN;
  double* LAM = sm + L::O_LAM;
  double* USF = sm + L::O_USF;
  double* TQ = sm + L::O_TQ;
  double* DFQ = sm + L::O_DFQ;
  double* FL = sm + L::O_FL;
  double* AUG1 = sm + L::O_AUG1;
  double* SUQ = sm + L::O_SUQ;
  double* SLQ = sm + L::O_SLQ;
  double* KS = sm + L::O_KS;
  double* RS = sm + L::O_RS;
  double* PP = sm + L::O_P;
  double* PB = sm + L::O_PB;
  double* SLU = sm + L::O_SLU;
  double* AUG2 = sm + L::O_AUG2;
  // coefficient set c of component a: cs_of(a); set c present: has_c(c)
  auto cs_of = [](int comp) { return MDL == 1 ? comp : 0; };
  auto has_c = [&](int c) { return c == 0 ? a.Dn != nullptr : a.G != nullptr; };

  for (int e = blockIdx.x; e < a.nEl; e += gridDim.x) {
    __syncthreads();  // the previous cell's outputs are written
    // ------------------------------------------------------------ gather
    for (int t = tid; t < NN * D; t += BT) X[t] = a.coords[(size_t)a.cells[(size_t)e * NN + t / D] * D + t % D];
    for (int t = tid; t < NDC * NN; t += BT) {
      const int c = t / NN, node = a.cells[(size_t)e * NN + t % NN];
      const double* src = c == 0 ? a.Dn : a.G;
      DN[t] = src ? src[node] : 0.0;
    }
    for (int t = tid; t < NL * MM; t += BT) {
      const int fj = t / MM, c = t % MM, f = fj / NNF;
      const int F = a.cell2face[(size_t)e * NF + f];
      const int side = (a.tau_sides == 2 && a.face2cell[2 * (size_t)F] != e) ? 1 : 0;
      const int p = a.face_perm[(size_t)e * NL + fj];
      TAU[t] = a.tau[(((size_t)F * NNF + p) * a.tau_sides + side) * MM + c];
    }
    for (int t = tid; t < MU; t += BT) SOL[t] = a.sol[(size_t)e * MU + t];
    for (int t = tid; t < ML; t += BT) {
      const int fj = t / M, f = fj / NNF;
      const int F = a.cell2face[(size_t)e * NF + f];
      TR[t] = a.trace[((size_t)F * NNF + a.face_perm[(size_t)e * NL + fj]) * M + t % M];
    }
    for (int t = tid; t < NIP * M; t += BT) {
      double fq = a.f_ip ? a.f_ip[(size_t)e * NIP * M + t] : 0.0;
      if (a.u_old && a.ts_idt != 0.0) {  // implicit Euler: + u_old(x_q) / dt
        const int q = t / M, c = t % M;
        double uo = 0.0;
        for (int j = 0; j < NN; ++j) uo += PHI[q * NN + j] * a.u_old[((size_t)e * NN + j) * M + c];
        fq += a.ts_idt * uo;
      }
      FIP[t] = fq;
    }
    __syncthreads();
    if constexpr (MDL == 1) {
      // HDGnGammaModel.cpp:116-124: v_00 = 0, v_01 = b, v_10 = (c^2 - (G/n)^2) b,
      // v_11 = 2 (G/n) b at the nodes, from the BufferSolution (n, Gamma)
      for (int t = tid; t < NN * D; t += BT) {
        const int i = t / D, k = t % D;
        const double bk = a.bfield[(size_t)a.cells[(size_t)e * NN + i] * D + k];
        const double r = SOL[i * M + 1] / SOL[i * M];
        VN[(i * NV + 0) * D + k] = 0.0;
        VN[(i * NV + 1) * D + k] = bk;
        VN[(i * NV + 2) * D + k] = (a.sound_speed * a.sound_speed - r * r) * bk;
        VN[(i * NV + 3) * D + k] = r * 2.0 * bk;
      }
      __syncthreads();
    }
    // ------------------------------------------------------- IP geometry
    for (int q = tid; q < NIP; q += BT) {
      double J[D * D];
      for (int t = 0; t < D * D; ++t) J[t] = 0.0;
      for (int i = 0; i < NN; ++i)
        for (int r = 0; r < D; ++r)
          for (int c = 0; c < D; ++c) J[r * D + c] += DPHI[(q * NN + i) * D + r] * X[i * D + c];
      const double dt = bdet<D>(J);
      double Ji[D * D];
      binv<D>(J, dt, Ji);
      DV[q] = W[q] * dt;
      double uq[D], dv = 0.0, dq[NDC];
      for (int c = 0; c < NDC; ++c) dq[c] = 0.0;
      for (int c = 0; c < M; ++c) uq[c] = 0.0;
      for (int i = 0; i < NN; ++i) {
        const double ph = PHI[q * NN + i];
        for (int k = 0; k < D; ++k) {
          double s = 0.0;
          for (int r = 0; r < D; ++r) s += Ji[k * D + r] * DPHI[(q * NN + i) * D + r];
          GR[(q * NN + i) * D + k] = s;
          dv += s * SOL[i * M + k];
        }
        for (int c = 0; c < M; ++c) uq[c] += ph * SOL[i * M + c];
        for (int c = 0; c < NDC; ++c) dq[c] += ph * DN[c * NN + i];
      }
      for (int c = 0; c < M; ++c) UQ[q * M + c] = uq[c];
      DIV[q] = dv;
      for (int c = 0; c < NDC; ++c) DQ[c * NIP + q] = dq[c];
      for (int tk = 0; tk < NV * D; ++tk) {
        double v = 0.0;
        for (int i = 0; i < NN; ++i) v += PHI[q * NN + i] * VN[i * NV * D + tk];
        VQ[q * NV * D + tk] = v;
      }
    }
    for (int t = tid; t < NFQ; t += BT) {
      const int f = t / NIPF, q = t % NIPF;
      double Jf[(D - 1) * D];
      for (int u = 0; u < (D - 1) * D; ++u) Jf[u] = 0.0;
      for (int j = 0; j < NNF; ++j) {
        const double* xj = X + fn[f * NNF + j] * D;
        for (int r = 0; r < D - 1; ++r)
          for (int c = 0; c < D; ++c) Jf[r * D + c] += FDPHI[(q * NNF + j) * (D - 1) + r] * xj[c];
      }
      double nv[D], meas;
      if constexpr (D == 2) {
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
      double nn = 0.0, prod = 0.0;
      for (int c = 0; c < D; ++c) nn += nv[c] * nv[c];
      nn = sqrt(nn);
      const int v0 = fn[f * NNF], vn = vopp[f];
      for (int c = 0; c < D; ++c) {
        nv[c] /= nn;
        prod += (X[vn * D + c] - X[v0 * D + c]) * nv[c];
      }
      if (prod > 0.0)
        for (int c = 0; c < D; ++c) nv[c] = -nv[c];
      FDS[t] = FW[q] * meas;
      for (int c = 0; c < D; ++c) FN[t * D + c] = nv[c];
      double lam[D], us[D], tq[MM], dfq[NDC];
      for (int c = 0; c < NDC; ++c) dfq[c] = 0.0;
      for (int c = 0; c < M; ++c) lam[c] = us[c] = 0.0;
      for (int c = 0; c < MM; ++c) tq[c] = 0.0;
      for (int j = 0; j < NNF; ++j) {
        const double ps = FPHI[q * NNF + j];
        const int nj = fn[f * NNF + j];
        for (int c = 0; c < M; ++c) {
          lam[c] += ps * TR[(f * NNF + j) * M + c];
          us[c] += ps * SOL[nj * M + c];
        }
        for (int c = 0; c < MM; ++c) tq[c] += ps * TAU[(f * NNF + j) * MM + c];
        for (int c = 0; c < NDC; ++c) dfq[c] += ps * DN[c * NN + nj];
      }
      for (int c = 0; c < M; ++c) {
        LAM[t * M + c] = lam[c];
        USF[t * M + c] = us[c];
      }
      for (int c = 0; c < MM; ++c) TQ[t * MM + c] = tq[c];
      for (int c = 0; c < NDC; ++c) DFQ[c * NFQ + t] = dfq[c];
      for (int tv = 0; tv < NV; ++tv) {
        double vn = 0.0;
        for (int k = 0; k < D; ++k) {
          double v = 0.0;
          for (int j = 0; j < NNF; ++j) v += FPHI[q * NNF + j] * VN[(fn[f * NNF + j] * NV + tv) * D + k];
          vn += v * nv[k];
        }
        VFN[t * (NV > 0 ? NV : 1) + tv] = vn;
      }
    }
    __syncthreads();
    // ------------------------------------------------ scalar q blocks
    for (int t = tid; t < NN * NC1; t += BT) {
      const int i = t / NC1, c = t % NC1;
      double v = 0.0;
      if (c < NN) {
        for (int q = 0; q < NIP; ++q) v += DV[q] * PHI[q * NN + i] * PHI[q * NN + c];
      } else if (c < NN + D * NN) {
        const int k = (c - NN) / NN, j = (c - NN) % NN;
        for (int q = 0; q < NIP; ++q) v += DV[q] * GR[(q * NN + i) * D + k] * PHI[q * NN + j];
      } else {
        const int cc = c - NN - D * NN, k = cc / NL, fc = cc % NL, f = fc / NNF, cn = fc % NNF;
        const int ai = posf[f * NN + i];
        if (ai >= 0)
          for (int q = 0; q < NIPF; ++q) {
            const int o = f * NIPF + q;
            v -= FDS[o] * FN[o * D + k] * FPHI[q * NNF + ai] * FPHI[q * NNF + cn];
          }
      }
      AUG1[t] = v;
    }
    for (int t = tid; t < NDC * D * NN * NN; t += BT) {  // Suq_k per coefficient set
      const int c = t / (D * NN * NN), k = (t / (NN * NN)) % D, i = (t / NN) % NN, j = t % NN;
      const double* dqc = DQ + c * NIP;
      const double* dfc = DFQ + c * NFQ;
      double v = 0.0;
      if (has_c(c)) {
        for (int q = 0; q < NIP; ++q) v += DV[q] * dqc[q] * GR[(q * NN + i) * D + k] * PHI[q * NN + j];
        for (int f = 0; f < NF; ++f) {
          const int ai = posf[f * NN + i], cj = posf[f * NN + j];
          if (ai < 0 || cj < 0) continue;
          for (int q = 0; q < NIPF; ++q) {
            const int o = f * NIPF + q;
            v -= FDS[o] * dfc[o] * FN[o * D + k] * FPHI[q * NNF + ai] * FPHI[q * NNF + cj];
          }
        }
      }
      SUQ[t] = v;
    }
    for (int t = tid; t < NDC * D * NL * NN; t += BT) {  // Slq_k per coefficient set
      const int c = t / (D * NL * NN), k = (t / (NL * NN)) % D, fa = (t / NN) % NL, j = t % NN;
      const int f = fa / NNF, ai = fa % NNF;
      const int cj = posf[f * NN + j];
      const double* dfc = DFQ + c * NFQ;
      double v = 0.0;
      if (has_c(c) && cj >= 0)
        for (int q = 0; q < NIPF; ++q) {
          const int o = f * NIPF + q;
          v -= FDS[o] * dfc[o] * FN[o * D + k] * FPHI[q * NNF + ai] * FPHI[q * NNF + cj];
        }
      SLQ[t] = v;
    }
    __syncthreads();
    gauss_jordan<NN, NC1, BT>(AUG1, perm1);
    // A_k[i][j] = AUG1[perm1[i] NC1 + NN + k NN + j], B_k[i][fc] = AUG1[perm1[i] NC1 + NN + D NN + k NL + fc]
    for (int t = tid; t < NDC * (NN * NN + NN * NL + NL * NN + NL * NL); t += BT) {
      constexpr int PER = NN * NN + NN * NL + NL * NN + NL * NL;
      const int c = t / PER;
      int u = t % PER;
      const double* suq = SUQ + c * D * NN * NN;
      const double* slq = SLQ + c * D * NL * NN;
      const double* Lrow;
      int ncol, col;
      double* dst;
      bool isA;
      if (u < NN * NN) {
        Lrow = suq + (u / NN) * NN; ncol = NN * NN; col = u % NN; isA = true; dst = KS + c * NN * NN + u;
      } else if ((u -= NN * NN) < NN * NL) {
        Lrow = suq + (u / NL) * NN; ncol = NN * NN; col = u % NL; isA = false; dst = RS + c * NN * NL + u;
      } else if ((u -= NN * NL) < NL * NN) {
        Lrow = slq + (u / NN) * NN; ncol = NL * NN; col = u % NN; isA = true; dst = PP + c * NL * NN + u;
      } else {
        u -= NL * NN;
        Lrow = slq + (u / NL) * NN; ncol = NL * NN; col = u % NL; isA = false; dst = PB + c * NL * NL + u;
      }
      double v = 0.0;
      for (int k = 0; k < D; ++k)
        for (int l = 0; l < NN; ++l) {
          const double* xr = AUG1 + perm1[l] * NC1 + NN + (isA ? k * NN : D * NN + k * NL);
          v += Lrow[k * ncol + l] * xr[col];
        }
      *dst = v;
    }
    __syncthreads();
    // ----------------------------------- coupled blocks [K | R | F_u], Slu, F_l
    for (int t = tid; t < MU * NC2; t += BT) {
      const int r = t / NC2, c = t % NC2, i = r / M, na = r % M;
      double v = 0.0;
      if (c < MU) {
        const int j = c / M, mb = c % M;
        for (int f = 0; f < NF; ++f) {  // HDGBase tau
          const int ai = posf[f * NN + i], cj = posf[f * NN + j];
          if (ai < 0 || cj < 0) continue;
          for (int q = 0; q < NIPF; ++q) {
            const int o = f * NIPF + q;
            v += FDS[o] * TQ[o * MM + mb * M + na] * FPHI[q * NNF + ai] * FPHI[q * NNF + cj];
          }
        }
        if constexpr (MDL == 0) {
          for (int q = 0; q < NIP; ++q) {  // HDGUNabU bulk
            const double* gi = GR + (q * NN + i) * D;
            const double* gj = GR + (q * NN + j) * D;
            const double pi = PHI[q * NN + i], pj = PHI[q * NN + j];
            double w = UQ[q * M + na] * (gj[mb] * pi + gi[mb] * pj);
            if (na == mb) {
              double ug = 0.0;
              for (int k = 0; k < D; ++k) ug += UQ[q * M + k] * gi[k];
              w += (DIV[q] * pi + ug) * pj;
            }
            v -= DV[q] * w;
          }
        } else if constexpr (MDL == 1) {
          for (int q = 0; q < NIP; ++q) {  // HDGConvection of v_(na mb): Suu -= C^T
            const double* gi = GR + (q * NN + i) * D;
            const double* vq = VQ + (q * NV + na * M + mb) * D;
            double vg = 0.0;
            for (int k = 0; k < D; ++k) vg += vq[k] * gi[k];
            v -= DV[q] * vg * PHI[q * NN + j];
          }
        }
        if (na == mb) {
          v -= KS[cs_of(na) * NN * NN + i * NN + j];
          if (a.ts_idt != 0.0) {  // implicit Euler: + M / dt
            double mij = 0.0;
            for (int q = 0; q < NIP; ++q) mij += DV[q] * PHI[q * NN + i] * PHI[q * NN + j];
            v += a.ts_idt * mij;
          }
        }
      } else if (c < MU + ML) {
        const int l = c - MU, f = l / (NNF * M), cn = (l / M) % NNF, mb = l % M;
        const int ai = posf[f * NN + i];
        if (ai >= 0)
          for (int q = 0; q < NIPF; ++q) {
            const int o = f * NIPF + q;
            double w = -TQ[o * MM + mb * M + na];
            if constexpr (MDL == 0) {  // HDGUNabU face rows summed into the u rows
              double ln = 0.0;
              for (int k = 0; k < D; ++k) ln += LAM[o * M + k] * FN[o * D + k];
              w += USF[o * M + na] * FN[o * D + mb];
              if (na == mb) w += ln;
            } else if constexpr (MDL == 1) {  // HDGConvection: Sul += (v.n) face mass
              w += VFN[o * (NV > 0 ? NV : 1) + na * M + mb];
            }
            v += FDS[o] * w * FPHI[q * NNF + ai] * FPHI[q * NNF + cn];
          }
        if (na == mb) v -= RS[cs_of(na) * NN * NL + i * NL + f * NNF + cn];
        v = -v;
      } else {
        for (int q = 0; q < NIP; ++q) {
          const double* gi = GR + (q * NN + i) * D;
          double ug = 0.0;
          for (int k = 0; k < D; ++k) ug += UQ[q * M + k] * gi[k];
          double w = PHI[q * NN + i] * FIP[q * M + na];
          if constexpr (MDL == 0) w -= UQ[q * M + na] * (DIV[q] * PHI[q * NN + i] + ug);
          v += DV[q] * w;
        }
        for (int f = 0; f < NF && MDL == 0; ++f) {
          const int ai = posf[f * NN + i];
          if (ai < 0) continue;
          for (int q = 0; q < NIPF; ++q) {
            const int o = f * NIPF + q;
            double ln = 0.0;
            for (int k = 0; k < D; ++k) ln += LAM[o * M + k] * FN[o * D + k];
            v += FDS[o] * FPHI[q * NNF + ai] * 0.5 * ln * (LAM[o * M + na] + USF[o * M + na]);
          }
        }
      }
      AUG2[t] = v;
    }
    for (int t = tid; t < ML * MU + ML; t += BT) {
      if (t < ML * MU) {
        const int rl = t / MU, c = t % MU, f = rl / (NNF * M), ai = (rl / M) % NNF, na = rl % M;
        const int j = c / M, mb = c % M, cj = posf[f * NN + j];
        double v = 0.0;
        if (cj >= 0)
          for (int q = 0; q < NIPF; ++q) {
            const int o = f * NIPF + q;
            v += FDS[o] * TQ[o * MM + mb * M + na] * FPHI[q * NNF + ai] * FPHI[q * NNF + cj];
          }
        if (na == mb) v -= PP[cs_of(na) * NL * NN + (f * NNF + ai) * NN + j];
        SLU[t] = v;
      } else {
        const int rl = t - ML * MU, f = rl / (NNF * M), ai = (rl / M) % NNF, na = rl % M;
        double v = 0.0;
        for (int q = 0; q < NIPF && MDL == 0; ++q) {
          const int o = f * NIPF + q;
          double ln = 0.0;
          for (int k = 0; k < D; ++k) ln += LAM[o * M + k] * FN[o * D + k];
          v += FDS[o] * FPHI[q * NNF + ai] * 0.5 * ln * (LAM[o * M + na] + USF[o * M + na]);
        }
        FL[rl] = v;
      }
    }
    __syncthreads();
    gauss_jordan<MU, NC2, BT>(AUG2, perm2);
    // U[r][l] = AUG2[perm2[r] NC2 + MU + l], U0[r] = AUG2[perm2[r] NC2 + MU + ML]
    double* Ue = a.U + (size_t)e * MU * ML;
    double* Qe = a.Q + (size_t)e * D * MU * ML;
    for (int t = tid; t < MU * (ML + 1); t += BT) {
      const int r = t / (ML + 1), l = t % (ML + 1);
      const double v = AUG2[perm2[r] * NC2 + MU + l];
      if (l < ML) Ue[(size_t)r * ML + l] = v;
      else a.U0[(size_t)e * MU + r] = v;
    }
    for (int t = tid; t < D * MU * (ML + 1); t += BT) {
      const int rq = t / (ML + 1), l = t % (ML + 1);
      const int i = rq / (D * M), k = (rq / M) % D, na = rq % M;
      const double* Ak = AUG1 + NN + k * NN;
      double v = 0.0;
      for (int j = 0; j < NN; ++j) v -= Ak[perm1[i] * NC1 + j] * AUG2[perm2[j * M + na] * NC2 + MU + l];
      if (l < ML) {
        const int fc = l / M, mb = l % M;
        if (na == mb) v -= AUG1[perm1[i] * NC1 + NN + D * NN + k * NL + fc];
        Qe[(size_t)rq * ML + l] = v;
      } else {
        a.Q0[(size_t)e * D * MU + rq] = v;
      }
    }
    for (int t = tid; t < ML * (ML + 1); t += BT) {
      const int rl = t / (ML + 1), l = t % (ML + 1);
      const double* sr = SLU + rl * MU;
      if (l < ML) {
        double v = 0.0;
        for (int r = 0; r < MU; ++r) v += sr[r] * AUG2[perm2[r] * NC2 + MU + l];
        const int f = rl / (NNF * M), ai = (rl / M) % NNF, na = rl % M;
        const int g = l / (NNF * M), cn = (l / M) % NNF, mb = l % M;
        if (f == g)
          for (int q = 0; q < NIPF; ++q) {  // Sll: HDGBase tau and HDGUNabU / HDGConvection face terms
            const int o = f * NIPF + q;
            double w = -TQ[o * MM + mb * M + na];
            if constexpr (MDL == 0) {
              double ln = 0.0;
              for (int k = 0; k < D; ++k) ln += LAM[o * M + k] * FN[o * D + k];
              w += USF[o * M + na] * FN[o * D + mb];
              if (na == mb) w += ln;
            } else if constexpr (MDL == 1) {
              w += VFN[o * (NV > 0 ? NV : 1) + na * M + mb];
            }
            v += FDS[o] * w * FPHI[q * NNF + ai] * FPHI[q * NNF + cn];
          }
        if (na == mb) v -= PB[cs_of(na) * NL * NL + (f * NNF + ai) * NL + g * NNF + cn];
        a.S[((size_t)e * ML + rl) * ML + l] = v;
      } else {
        double v = FL[rl];
        for (int r = 0; r < MU; ++r) v -= sr[r] * AUG2[perm2[r] * NC2 + MU + ML];
        a.S0[(size_t)e * ML + rl] = v;
      }
    }
  }
}

template <int D, int P, int MDL>
void launch_dp(const BurgersArgs& a, hipStream_t s) {
  constexpr bool gws = burgers_in_workspace<D, P, MDL>();
  const size_t lds = gws ? 0 : (size_t)BLay<D, P, MDL>::TOTAL * sizeof(double);
  static PerDeviceLaunch once;
  if (!gws) once.lds_attr((const void*)burgers_cell_kernel<D, P, MDL>, lds);
  int grid = a.nEl < 65536 ? a.nEl : 65536;
  if (gws) {
    if (!a.gws || a.gws_slots <= 0) throw std::runtime_error("Burgers kernel: this order needs its device workspace");
    if (grid > a.gws_slots) grid = a.gws_slots;
  }
  hipLaunchKernelGGL((burgers_cell_kernel<D, P, MDL>), dim3(grid), dim3(BT), lds, s, a);
  HFX_HIP(hipGetLastError());
}

__global__ void expand_perm_kernel(long long nblk, int nNf, int m, const int8_t* __restrict__ perm,
                                   int8_t* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = nblk * nNf * m;
  if (t >= n) return;
  const long long blk = t / ((long long)nNf * m);
  const int j = (int)((t / m) % nNf), c = (int)(t % m);
  out[t] = (int8_t)(perm[blk * nNf + j] * m + c);
}

constexpr int RT = 256;

__global__ __launch_bounds__(RT) void residual_partial_kernel(long long n, const double* __restrict__ cur,
                                                              const double* __restrict__ prev,
                                                              double* __restrict__ partial) {
  __shared__ double red[2][RT];
  double d2 = 0.0, p2 = 0.0;
  for (long long i = (long long)blockIdx.x * RT + threadIdx.x; i < n; i += (long long)gridDim.x * RT) {
    const double d = cur[i] - prev[i];
    d2 += d * d;
    p2 += prev[i] * prev[i];
  }
  red[0][threadIdx.x] = d2;
  red[1][threadIdx.x] = p2;
  __syncthreads();
  for (int w = RT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = red[0][0];
    partial[2 * blockIdx.x + 1] = red[1][0];
  }
}

__global__ __launch_bounds__(64) void residual_final_kernel(int g, const double* __restrict__ partial,
                                                            double* __restrict__ out2) {
  if (threadIdx.x == 0) {
    double d2 = 0.0, p2 = 0.0;
    for (int b = 0; b < g; ++b) {
      d2 += partial[2 * b];
      p2 += partial[2 * b + 1];
    }
    out2[0] = d2;
    out2[1] = p2;
  }
}

__global__ void damp_kernel(long long n, double damp, double* __restrict__ cur, double* __restrict__ prev) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double v = (1.0 - damp) * cur[t] + damp * prev[t];
  cur[t] = v;
  prev[t] = v;
}

}  // namespace

bool burgers_supported(int dim, int order) {
  return (dim == 2 && order >= 1 && order <= 5) || (dim == 3 && order >= 1 && order <= 3);
}

size_t burgers_workspace_doubles(int dim, int order, int model) {
  if (model == 1) return 0;
  if (dim == 3 && order == 3)
    return model == 2 ? (burgers_in_workspace<3, 3, 2>() ? BLay<3, 3, 2>::TOTAL : 0)
                      : (burgers_in_workspace<3, 3, 0>() ? BLay<3, 3, 0>::TOTAL : 0);
  return 0;
}

void launch_burgers(int dim, int order, const BurgersArgs& a, hipStream_t s) {
  if (a.nEl <= 0) return;
  if (a.model == 1) {
    if (dim != 2) throw std::runtime_error("nGamma kernel: 2-D only");
    switch (order) {
      case 1: launch_dp<2, 1, 1>(a, s); return;
      case 2: launch_dp<2, 2, 1>(a, s); return;
      case 3: launch_dp<2, 3, 1>(a, s); return;
      case 4: launch_dp<2, 4, 1>(a, s); return;
      case 5: launch_dp<2, 5, 1>(a, s); return;
    }
  } else if (a.model == 2) {
    if (dim == 2) {
      switch (order) {
        case 1: launch_dp<2, 1, 2>(a, s); return;
        case 2: launch_dp<2, 2, 2>(a, s); return;
        case 3: launch_dp<2, 3, 2>(a, s); return;
        case 4: launch_dp<2, 4, 2>(a, s); return;
        case 5: launch_dp<2, 5, 2>(a, s); return;
      }
    } else if (dim == 3) {
      switch (order) {
        case 1: launch_dp<3, 1, 2>(a, s); return;
        case 2: launch_dp<3, 2, 2>(a, s); return;
        case 3: launch_dp<3, 3, 2>(a, s); return;
      }
    }
  } else if (dim == 2) {
    switch (order) {
      case 1: launch_dp<2, 1, 0>(a, s); return;
      case 2: launch_dp<2, 2, 0>(a, s); return;
      case 3: launch_dp<2, 3, 0>(a, s); return;
      case 4: launch_dp<2, 4, 0>(a, s); return;
      case 5: launch_dp<2, 5, 0>(a, s); return;
    }
  } else if (dim == 3) {
    switch (order) {
      case 1: launch_dp<3, 1, 0>(a, s); return;
      case 2: launch_dp<3, 2, 0>(a, s); return;
      case 3: launch_dp<3, 3, 0>(a, s); return;
    }
  }
  throw std::runtime_error("Burgers kernel: unsupported dimension/order");
}

void launch_expand_perm(long long nblk, int nNf, int m, const int8_t* perm, int8_t* out, hipStream_t s) {
  const long long n = nblk * nNf * m;
  if (n == 0) return;
  hipLaunchKernelGGL(expand_perm_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, nblk, nNf, m, perm, out);
  HFX_HIP(hipGetLastError());
}

// partial must hold 2 x 256 doubles
void burgers_residual(long long n, const double* cur, const double* prev, double* partial, double* out2,
                      hipStream_t s) {
  long long g = (n + RT * 8 - 1) / (RT * 8);
  if (g < 1) g = 1;
  if (g > 256) g = 256;
  hipLaunchKernelGGL(residual_partial_kernel, dim3((int)g), dim3(RT), 0, s, n, cur, prev, partial);
  HFX_HIP(hipGetLastError());
  hipLaunchKernelGGL(residual_final_kernel, dim3(1), dim3(64), 0, s, (int)g, partial, out2);
  HFX_HIP(hipGetLastError());
}

void burgers_damp(long long n, double damp, double* cur, double* prev, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(damp_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, damp, cur, prev);
  HFX_HIP(hipGetLastError());
}

}  // namespace hfx
