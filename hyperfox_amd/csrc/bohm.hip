// HDGBohmModel (src/model/HDGBohmModel.cpp:97-268) on the GPU: the
// Bohm-Chodura boundary rows of the 2-D (n, Gamma) problem, rebuilt for every
// Bohm face before each assembly from the current Solution, Flux and Trace.
//
// Per face, with the face-local fields HDGSolver::applyBoundaryConditions
// gives a boundary model (HDGSolver.cpp:404-476): at each face cubature point
// q the model's rows are
//   n rows      tau_0.(u - lambda) - (D n).q_n
//   Gamma rows  (1 - T)(tau_1.(u - lambda) - (G n).q_Gamma)
//               + T (u_Gamma b.n - u_n c_s |b.n|)
// with T = T(u_Gamma b.n, u_n c_s b.n) the transfer function; the matrix is
// their linearisation with T frozen (the "original system") plus, on the
// Gamma rows' u columns, dT/du (Dirichlet - Neumann), and the RHS is that
// derivative part applied to the current state.  Every entry
// B(m 2 + a, col(n)) is a sum over q of dV phi_m phi_n c_q(a, col), so the
// kernel forms the twelve coefficients c_q per cubature point in LDS and then
// one thread per entry sums over q.  One workgroup per face; the rows then go
// through the HDG-type boundary pass (HDGSolver.cpp:503-518) like any
// caller-supplied B and r.
#include "burgers.h"
#include "hfx_internal.h"

namespace hfx {

namespace {

constexpr int BT_BOHM = 64;
constexpr int MAX_NNF = 6;   // 2-D p <= 5
constexpr int MAX_NIPF = 8;

// the transfer functions of tests/regression/HDG/TestHDGBohm.cpp:37-130
__device__ inline double num_heaviside(double x) { return 0.5 * (1.0 + tanh(x / 0.001)); }
__device__ inline double num_dheaviside(double x) {
  const double t = tanh(x / 0.001);
  return (0.5 / 0.001) * (1.0 - t * t);
}

__device__ void transfer(int kind, double x, double y, double* T, double* d0, double* d1) {
  const double tol = 1e-5;
  *d0 = 0.0;
  *d1 = 0.0;
  if (kind == HFX_BOHM_NEUMANN) {
    *T = 0.0;
  } else if (kind == HFX_BOHM_DIRICHLET) {
    *T = 1.0;
  } else if (kind == HFX_BOHM_HEAVISIDE) {
    double r = 0.0;
    if (fabs(y) > tol) {
      if ((x > 0) != (y > 0)) r = 1.0;
      else if (fabs(x) - fabs(y) < 0.0) r = 1.0;
    }
    *T = r;
  } else {
    const double Hx = num_heaviside(x), Hy = num_heaviside(y), Hmx = num_heaviside(-x), Hmy = num_heaviside(-y);
    const double dHx = num_dheaviside(x), dHy = num_dheaviside(y), dHmx = num_dheaviside(-x), dHmy = num_dheaviside(-y);
    const double Hyx = num_heaviside(y - x), Hxy = num_heaviside(x - y);
    const double dHyx = num_dheaviside(y - x), dHxy = num_dheaviside(x - y);
    const double Htol = num_heaviside(tol - fabs(y)), dHtol = num_dheaviside(tol - fabs(y));
    const double base = Hx * Hmy + Hmx * Hy + Hx * Hy * Hyx + Hmx * Hmy * Hxy;
    *T = base * (1.0 - Htol);
    *d0 = (dHx * Hmy - dHmx * Hy + dHx * Hy * Hyx - Hx * Hy * dHyx - dHmx * Hmy * Hxy + Hmx * Hmy * dHxy) * (1.0 - Htol);
    *d1 = (-Hx * dHmy + Hmx * dHy + Hx * dHy * Hyx + Hx * Hy * dHyx - Hmx * dHmy * Hxy - Hmx * Hmy * dHxy) *
              (1.0 - Htol) + base * dHtol;
  }
}

// one workgroup per Bohm face i: B[i] (2 nNf x 8 nNf, row-major) and r[i] (2 nNf)
__global__ __launch_bounds__(BT_BOHM) void bohm_kernel(BohmArgs a) {
  const int i = blockIdx.x;
  const int t = threadIdx.x;
  const int nNf = a.nNf, nipf = a.nipf, nN = a.nN, nF = a.nFpc;
  __shared__ double xs[MAX_NNF][2], nr[MAX_NNF][2], bb[MAX_NNF][2], tr[MAX_NNF][2], so[MAX_NNF][2];
  __shared__ double fl[MAX_NNF][4], ta[MAX_NNF][4], dd[MAX_NNF], gg[MAX_NNF];
  __shared__ double cu[MAX_NIPF][2][2], cl[MAX_NIPF][2][2], cq[MAX_NIPF][2][2], dv[MAX_NIPF], rq[MAX_NIPF][2];
  __shared__ int cell_s, lf_s;
  const int F = a.faces[i];
  if (t == 0) {
    const int e = a.face2cell[2 * F];
    int lf = 0;
    for (int f = 0; f < nF; ++f)
      if (a.cell2face[(size_t)e * nF + f] == F) lf = f;
    cell_s = e;
    lf_s = lf;
  }
  __syncthreads();
  const int e = cell_s, lf = lf_s;
  // face-local fields in face-node order: face node j is the cell's local
  // node faceNodes[lf][invperm[e][lf][j]]
  if (t < nNf) {
    const int j = t;
    const int pos = a.invperm[((size_t)e * nF + lf) * nNf + j];
    const int cn = a.faceNodes[lf * nNf + pos];
    const int node = a.cells[(size_t)e * nN + cn];
    for (int c = 0; c < 2; ++c) {
      xs[j][c] = a.coords[(size_t)node * 2 + c];
      nr[j][c] = a.normals[((size_t)i * nNf + j) * 2 + c];
      bb[j][c] = a.bfield[(size_t)node * 2 + c];
      tr[j][c] = a.trace[((size_t)F * nNf + j) * 2 + c];
      so[j][c] = a.sol[((size_t)e * nN + cn) * 2 + c];
    }
    for (int c = 0; c < 4; ++c) {
      fl[j][c] = a.flux[((size_t)e * nN + cn) * 4 + c];                                  // ((i d + k) m + a)
      ta[j][c] = a.tau[(((size_t)F * nNf + j) * a.tau_sides + 0) * 4 + c];               // side 0
    }
    dd[j] = a.D[node];
    gg[j] = a.G[node];
  }
  __syncthreads();
  if (t < nipf) {
    const int q = t;
    const double* ph = a.fphi + (size_t)q * nNf;
    const double* dph = a.fdphi + (size_t)q * nNf;
    double J[2] = {0.0, 0.0}, n[2] = {0.0, 0.0}, b[2] = {0.0, 0.0}, lam[2] = {0.0, 0.0}, u[2] = {0.0, 0.0};
    double Fq[4] = {0.0, 0.0, 0.0, 0.0}, tq[4] = {0.0, 0.0, 0.0, 0.0}, Dq = 0.0, Gq = 0.0;
    for (int j = 0; j < nNf; ++j) {
      const double s = ph[j];
      for (int c = 0; c < 2; ++c) {
        J[c] += dph[j] * xs[j][c];
        n[c] += s * nr[j][c];
        b[c] += s * bb[j][c];
        lam[c] += s * tr[j][c];
        u[c] += s * so[j][c];
      }
      for (int c = 0; c < 4; ++c) {
        Fq[c] += s * fl[j][c];
        tq[c] += s * ta[j][c];
      }
      Dq += s * dd[j];
      Gq += s * gg[j];
    }
    const double cs = a.sound_speed;
    const double bn = b[0] * n[0] + b[1] * n[1];
    dv[q] = a.fw[q] * sqrt(J[0] * J[0] + J[1] * J[1]);
    double T, d0, d1;
    transfer(a.kind, u[1] * bn, u[0] * cs * bn, &T, &d0, &d1);
    // tau column-major: tau(r, c) = tq[c 2 + r]; flux q_{a,k} = Fq[k 2 + a];
    // scalar D and G: D n = D_s n, G n = G_s n
    const double Dn[2] = {Dq * n[0], Dq * n[1]}, Gn[2] = {Gq * n[0], Gq * n[1]};
    const double dirv = u[1] * bn - u[0] * cs * fabs(bn);
    const double neuv = -(Gn[0] * Fq[1] + Gn[1] * Fq[3]) + tq[3] * (u[1] - lam[1]) + tq[1] * (u[0] - lam[0]);
    const double lm0 = cs * bn * d1 * (dirv - neuv), lm1 = bn * d0 * (dirv - neuv);
    for (int c = 0; c < 2; ++c) {
      cu[q][0][c] = tq[c * 2 + 0];
      cl[q][0][c] = -tq[c * 2 + 0];
      cl[q][1][c] = -tq[c * 2 + 1] * (1.0 - T);
      cq[q][0][c] = -Dn[c];
      cq[q][1][c] = -(1.0 - T) * Gn[c];
    }
    cu[q][1][0] = tq[1] * (1.0 - T) - T * cs * fabs(bn) + lm0;
    cu[q][1][1] = tq[3] * (1.0 - T) + T * bn + lm1;
    rq[q][0] = 0.0;
    rq[q][1] = lm0 * u[0] + lm1 * u[1];      // (B - original) U0 at this point
  }
  __syncthreads();
  const int bs = 2 * nNf, nc = 8 * nNf;
  double* Bf = a.B + (size_t)i * bs * nc;
  for (int k = t; k < bs * nc; k += BT_BOHM) {
    const int row = k / nc, col = k % nc;
    const int m = row / 2, ar = row % 2;
    int nd, kind, comp;
    if (col < 2 * nNf) {
      nd = col / 2; comp = col % 2; kind = 0;                       // u (n 2 + b)
    } else if (col < 6 * nNf) {
      const int x = col - 2 * nNf;                                  // q ((n 2 + k) 2 + b)
      nd = x / 4; comp = (x / 2) % 2; kind = (x % 2 == ar) ? 1 : 3;
    } else {
      const int x = col - 6 * nNf;                                  // lambda (n 2 + b)
      nd = x / 2; comp = x % 2; kind = 2;
    }
    double s = 0.0;
    if (kind != 3)
      for (int q = 0; q < nipf; ++q) {
        const double c = kind == 0 ? cu[q][ar][comp] : (kind == 1 ? cq[q][ar][comp] : cl[q][ar][comp]);
        s += dv[q] * a.fphi[(size_t)q * nNf + m] * a.fphi[(size_t)q * nNf + nd] * c;
      }
    Bf[k] = s;
  }
  for (int row = t; row < bs; row += BT_BOHM) {
    const int m = row / 2, ar = row % 2;
    double s = 0.0;
    for (int q = 0; q < nipf; ++q) s += dv[q] * a.fphi[(size_t)q * nNf + m] * rq[q][ar];
    a.r[(size_t)i * bs + row] = s;
  }
}

}  // namespace

void launch_bohm(const BohmArgs& a, int nfaces, hipStream_t s) {
  if (nfaces <= 0) return;
  if (a.nNf > MAX_NNF || a.nipf > MAX_NIPF) throw std::runtime_error("Bohm model: face order above 2-D p = 5");
  hipLaunchKernelGGL(bohm_kernel, dim3(nfaces), dim3(BT_BOHM), 0, s, a);
  HFX_HIP(hipGetLastError());
}

}  // namespace hfx
