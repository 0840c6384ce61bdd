// Global trace system: BSR pattern, deterministic gather-scatter of the
// condensed element blocks, and element reconstruction.  See trace_system.h.
#include <algorithm>
#include <stdexcept>

#include "trace_system.h"

namespace hfx {

BsrPattern build_bsr_pattern(int nCells, int nFaces, int nFpc, int bs, const int* cell2face, const int* face2cell) {
  BsrPattern P;
  P.nb = nFaces;
  P.bs = bs;
  P.rowptr.assign(nFaces + 1, 0);
  P.face_cells.assign((size_t)4 * nFaces, -1);
  auto local_face = [&](int e, int F) {
    for (int l = 0; l < nFpc; ++l)
      if (cell2face[(size_t)e * nFpc + l] == F) return l;
    return -1;
  };
  std::vector<int> cols;
  std::vector<std::vector<int>> rows(nFaces);
  for (int F = 0; F < nFaces; ++F) {
    cols.clear();
    for (int s = 0; s < 2; ++s) {
      int e = face2cell[2 * (size_t)F + s];
      if (e < 0) continue;
      int lf = local_face(e, F);
      if (lf < 0) throw std::runtime_error("bsr: face2cell and cell2face disagree");
      P.face_cells[4 * (size_t)F + 2 * s] = e;
      P.face_cells[4 * (size_t)F + 2 * s + 1] = lf;
      for (int l = 0; l < nFpc; ++l) cols.push_back(cell2face[(size_t)e * nFpc + l]);
    }
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    rows[F] = cols;
    P.rowptr[F + 1] = P.rowptr[F] + (int)cols.size();
  }
  const int nnzb = P.rowptr[nFaces];
  P.col.resize(nnzb);
  P.rowf.resize(nnzb);
  P.contrib.assign((size_t)4 * nnzb, -1);
  P.diag.assign(nFaces, -1);
  for (int F = 0; F < nFaces; ++F) {
    int k = P.rowptr[F];
    const int e0 = P.face_cells[4 * (size_t)F], a0 = P.face_cells[4 * (size_t)F + 1];
    const int e1 = P.face_cells[4 * (size_t)F + 2], a1 = P.face_cells[4 * (size_t)F + 3];
    for (int G : rows[F]) {
      P.col[k] = G;
      P.rowf[k] = F;
      int* c = &P.contrib[4 * (size_t)k];
      if (G == F) {
        P.diag[F] = k;
        c[0] = e0; c[1] = a0 * 16 + a0; c[2] = e1; c[3] = e1 >= 0 ? a1 : 0;
      } else {
        int b = e0 >= 0 ? local_face(e0, G) : -1;
        if (b >= 0) { c[0] = e0; c[1] = a0 * 16 + b; }
        else {
          b = local_face(e1, G);
          if (b < 0) throw std::runtime_error("bsr: coupling without a common cell");
          c[0] = e1; c[1] = a1 * 16 + b;
        }
        c[2] = -1; c[3] = 0;
      }
      ++k;
    }
  }
  return P;
}

namespace {

template <int BS>
__global__ void scatter_kernel(int nnzb, int nFpc, const int* __restrict__ col, const int* __restrict__ rowf,
                               const int* __restrict__ contrib, const double* __restrict__ S,
                               const int8_t* __restrict__ invperm, const uint8_t* __restrict__ is_dir,
                               double* __restrict__ vals) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nnzb * BS * BS) return;
  const int k = (int)(t / (BS * BS));
  const int r = (int)((t / BS) % BS), c = (int)(t % BS);
  const int F = rowf[k];
  const int NL = nFpc * BS;
  double v;
  if (is_dir && is_dir[F]) {
    v = (col[k] == F && r == c) ? 1.0 : 0.0;
  } else {
    const int e0 = contrib[4 * k], ab = contrib[4 * k + 1], e1 = contrib[4 * k + 2], a1 = contrib[4 * k + 3];
    const int a0 = ab >> 4, b0 = ab & 15;
    const int lr = a0 * BS + invperm[((size_t)e0 * nFpc + a0) * BS + r];
    const int lc = b0 * BS + invperm[((size_t)e0 * nFpc + b0) * BS + c];
    v = S[((size_t)e0 * NL + lr) * NL + lc];
    if (e1 >= 0) {
      const int lr1 = a1 * BS + invperm[((size_t)e1 * nFpc + a1) * BS + r];
      const int lc1 = a1 * BS + invperm[((size_t)e1 * nFpc + a1) * BS + c];
      v += S[((size_t)e1 * NL + lr1) * NL + lc1];
    }
  }
  vals[t] = v;
}

template <int BS>
__global__ void scatter_diag_kernel(int nb, const int* __restrict__ diag, const int* __restrict__ face_cells,
                                    const double* __restrict__ dside, const uint8_t* __restrict__ is_dir,
                                    double* __restrict__ vals) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nb * BS * BS) return;
  const int F = (int)(t / (BS * BS)), rc = (int)(t % (BS * BS));
  double v;
  if (is_dir && is_dir[F]) {
    v = (rc / BS == rc % BS) ? 1.0 : 0.0;
  } else {
    v = dside[(size_t)2 * F * BS * BS + rc];
    if (face_cells[4 * (size_t)F + 2] >= 0) v += dside[(size_t)(2 * F + 1) * BS * BS + rc];
  }
  vals[(size_t)diag[F] * BS * BS + rc] = v;
}

__global__ void rhs_kernel(int nb, int bs, int nFpc, const int* __restrict__ face_cells, const double* __restrict__ S0,
                           const int8_t* __restrict__ invperm, const uint8_t* __restrict__ is_dir,
                           const double* __restrict__ g, double* __restrict__ rhs) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nb * bs) return;
  const int F = (int)(t / bs), r = (int)(t % bs);
  const int NL = nFpc * bs;
  double v = 0.0;
  if (is_dir && is_dir[F]) {
    v = g ? g[t] : 0.0;
  } else {
    for (int s = 0; s < 2; ++s) {
      const int e = face_cells[4 * F + 2 * s], a = face_cells[4 * F + 2 * s + 1];
      if (e < 0) continue;
      v += S0[(size_t)e * NL + a * bs + invperm[((size_t)e * nFpc + a) * bs + r]];
    }
  }
  rhs[t] = v;
}

// IntegratedDirichletModel rows (IntegratedDirichletModel.cpp:35-56): the
// diagonal block of boundary face F = bfaces[ib] becomes the face mass
// M_ab = sum_q w_q sqrt(det(J_f J_f^T)) psi_a psi_b of the face's nodes,
// taken from its cell (face position pp[a] of cell-local face node a).  One
// thread per block entry.
__global__ void integrated_dirichlet_kernel(int nbf, const int* __restrict__ bfaces, int dim, int nN, int nNf,
                                            int nipf, int nFpc, const int* __restrict__ face_cells,
                                            const int* __restrict__ fnodes, const int8_t* __restrict__ face_perm,
                                            const int* __restrict__ cells, const double* __restrict__ coords,
                                            const double* __restrict__ fphi, const double* __restrict__ fdphi,
                                            const double* __restrict__ fw, const int* __restrict__ diag, int m,
                                            double* __restrict__ vals) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nbf * nNf * nNf) return;
  const int ib = (int)(t / (nNf * nNf)), a = (int)((t / nNf) % nNf), b = (int)(t % nNf);
  const int F = bfaces[ib];
  const int e = face_cells[4 * (size_t)F], lf = face_cells[4 * (size_t)F + 1];
  const int* cn = cells + (size_t)e * nN;
  const int fd = dim - 1;
  double mass = 0.0;
  for (int q = 0; q < nipf; ++q) {
    double J[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // fd x dim
    for (int j = 0; j < nNf; ++j) {
      const double* x = coords + (size_t)cn[fnodes[lf * nNf + j]] * dim;
      for (int r = 0; r < fd; ++r) {
        const double dp = fdphi[(q * nNf + j) * fd + r];
        for (int c = 0; c < dim; ++c) J[r * dim + c] += dp * x[c];
      }
    }
    double meas;
    if (dim == 2) {
      meas = sqrt(J[0] * J[0] + J[1] * J[1]);
    } else {
      const double g00 = J[0] * J[0] + J[1] * J[1] + J[2] * J[2];
      const double g01 = J[0] * J[3] + J[1] * J[4] + J[2] * J[5];
      const double g11 = J[3] * J[3] + J[4] * J[4] + J[5] * J[5];
      meas = sqrt(g00 * g11 - g01 * g01);
    }
    mass += fphi[q * nNf + a] * fphi[q * nNf + b] * (fw[q] * meas);
  }
  const int8_t* pp = face_perm + ((size_t)e * nFpc + lf) * nNf;
  const int bs = nNf * m;
  for (int c = 0; c < m; ++c) vals[(size_t)diag[F] * bs * bs + (pp[a] * m + c) * bs + pp[b] * m + c] = mass;
}

// rhs of those rows: M g (g in face-node order; null: zero)
__global__ void integrated_dirichlet_rhs_kernel(int nbf, const int* __restrict__ bfaces, int nNf,
                                                const int* __restrict__ diag, const double* __restrict__ vals,
                                                const double* __restrict__ g, double* __restrict__ rhs) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nbf * nNf) return;
  const int F = bfaces[t / nNf], r = (int)(t % nNf);
  double v = 0.0;
  if (g) {
    const double* row = vals + (size_t)diag[F] * nNf * nNf + r * nNf;
    for (int c = 0; c < nNf; ++c) v += row[c] * g[(size_t)F * nNf + c];
  }
  rhs[(size_t)F * nNf + r] = v;
}

// HDG-type boundary model rows (HDGSolver.cpp:503-518) of boundary face
// F = hfaces[ib], whose rows the scatter left as the cell's own, with m
// unknowns per node (block size bs = nNf m, face order p m + a): row pa over
// every cell-local trace column c becomes (Set) or gains (Add)
//   sum_(pj, b) B_u[pa][pj m + b] U[i_pj m + b][c]
//   + sum_(pj, k, b) B_q[pa][(pj d + k) m + b] Q[(i_pj d + k) m + b][c]
// (+ B_lambda[pa][pc] on the face's own block), i_pj the cell node at face
// position pj (the U and Q rows of the face's nodes, HDGSolver.cpp:508-511).
// One thread per (face, row, column).
__global__ void hdg_boundary_kernel(int nhf, const int* __restrict__ hfaces, int dim, int nN, int nNf, int m,
                                    int nFpc, const int* __restrict__ face_cells, const int* __restrict__ fnodes,
                                    const int8_t* __restrict__ face_perm, const int8_t* __restrict__ invperm,
                                    const int* __restrict__ bpos, const double* __restrict__ U,
                                    const double* __restrict__ Q, const double* __restrict__ B, int mset,
                                    double* __restrict__ vals) {
  const int bs = nNf * m, NL = nFpc * bs, nc = bs * (2 + dim);
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nhf * bs * NL) return;
  const int ib = (int)(t / ((long long)bs * NL)), pa = (int)((t / NL) % bs), c = (int)(t % NL);
  const int F = hfaces[ib];
  const int e = face_cells[4 * (size_t)F], lf = face_cells[4 * (size_t)F + 1];
  const int8_t* ipf = invperm + ((size_t)e * nFpc + lf) * nNf;
  const double* Bi = B + ((size_t)ib * bs + pa) * nc;
  const int mu = nN * m, mq = mu * dim;
  double v = 0.0;
  for (int pj = 0; pj < nNf; ++pj) {
    const int i = fnodes[lf * nNf + ipf[pj]];
    for (int b = 0; b < m; ++b) {
      v += Bi[pj * m + b] * U[((size_t)e * mu + i * m + b) * NL + c];
      for (int k = 0; k < dim; ++k)
        v += Bi[bs + (pj * dim + k) * m + b] * Q[((size_t)e * mq + (i * dim + k) * m + b) * NL + c];
    }
  }
  const int G = c / bs, bn = (c % bs) / m, bc = c % m;
  const int pc = face_perm[((size_t)e * nFpc + G) * nNf + bn] * m + bc;
  if (G == lf) v += Bi[bs * (1 + dim) + pc];
  double* dst = vals + (size_t)bpos[((size_t)e * nFpc + lf) * nFpc + G] * bs * bs + pa * bs + pc;
  *dst = mset ? v : *dst + v;
}

// rhs of those rows: r - B_u U0_F - B_q Q0_F, set or added
__global__ void hdg_boundary_rhs_kernel(int nhf, const int* __restrict__ hfaces, int dim, int nN, int nNf, int m,
                                        int nFpc, const int* __restrict__ face_cells, const int* __restrict__ fnodes,
                                        const int8_t* __restrict__ invperm, const double* __restrict__ U0,
                                        const double* __restrict__ Q0, const double* __restrict__ B,
                                        const double* __restrict__ r, int rset, double* __restrict__ rhs) {
  const int bs = nNf * m, nc = bs * (2 + dim);
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nhf * bs) return;
  const int ib = (int)(t / bs), pa = (int)(t % bs);
  const int F = hfaces[ib];
  const int e = face_cells[4 * (size_t)F], lf = face_cells[4 * (size_t)F + 1];
  const int8_t* ipf = invperm + ((size_t)e * nFpc + lf) * nNf;
  const double* Bi = B + ((size_t)ib * bs + pa) * nc;
  const int mu = nN * m, mq = mu * dim;
  double v = r[(size_t)ib * bs + pa];
  for (int pj = 0; pj < nNf; ++pj) {
    const int i = fnodes[lf * nNf + ipf[pj]];
    for (int b = 0; b < m; ++b) {
      v -= Bi[pj * m + b] * U0[(size_t)e * mu + i * m + b];
      for (int k = 0; k < dim; ++k) v -= Bi[bs + (pj * dim + k) * m + b] * Q0[(size_t)e * mq + (i * dim + k) * m + b];
    }
  }
  double* dst = rhs + (size_t)F * bs + pa;
  *dst = rset ? v : *dst + v;
}

// u = U lambda_e + U0, q = Q lambda_e + Q0 for one cell per wavefront: the
// cell's lambda is gathered into LDS, then up to 64 rows of its U and Q blocks
// at a time are staged in LDS with coalesced loads (rows padded to NL + 1
// doubles against bank conflicts) and lane r sums row r, from U0 / Q0 and over
// the columns in order.
__global__ __launch_bounds__(64) void reconstruct_kernel(int nEl, int nN, int nQ, int nFpc, int bs,
                                                         const double* __restrict__ U, const double* __restrict__ U0,
                                                         const double* __restrict__ Q, const double* __restrict__ Q0,
                                                         const int* __restrict__ cell2face,
                                                         const int8_t* __restrict__ face_perm,
                                                         const double* __restrict__ lambda, double* __restrict__ u,
                                                         double* __restrict__ q) {
  extern __shared__ double rsm[];
  const int NL = nFpc * bs, rows = nN + nQ, ld = NL + 1;
  double* L = rsm;
  double* T = rsm + NL;
  const int lane = threadIdx.x;
  for (int e = blockIdx.x; e < nEl; e += gridDim.x) {
    __syncthreads();  // the previous cell is done with L and T
    for (int t = lane; t < NL; t += 64) {
      const int f = t / bs, j = t % bs;
      L[t] = lambda[(size_t)cell2face[(size_t)e * nFpc + f] * bs + face_perm[((size_t)e * nFpc + f) * bs + j]];
    }
    const double* Ue = U + (size_t)e * nN * NL;
    const double* Qe = Q + (size_t)e * nQ * NL;
    for (int r0 = 0; r0 < rows; r0 += 64) {
      const int nr = min(64, rows - r0);
      if (r0 > 0) __syncthreads();  // the rows of the previous pass are consumed
      for (int k = lane; k < nr * NL; k += 64) {
        const int rr = r0 + k / NL, c = k % NL;
        T[(k / NL) * ld + c] = rr < nN ? Ue[(size_t)rr * NL + c] : Qe[(size_t)(rr - nN) * NL + c];
      }
      __syncthreads();
      if (lane < nr) {
        const int r = r0 + lane;
        double acc = r < nN ? U0[(size_t)e * nN + r] : Q0[(size_t)e * nQ + (r - nN)];
        const double* row = T + lane * ld;
        for (int c = 0; c < NL; ++c) acc += row[c] * L[c];
        if (r < nN) u[(size_t)e * nN + r] = acc;
        else q[(size_t)e * nQ + (r - nN)] = acc;
      }
    }
  }
}

}  // namespace

void launch_scatter(const BsrDev& A, int nFpc, const double* S, const double* S0, const int8_t* invperm,
                    const uint8_t* is_dir, const double* g, double* rhs, hipStream_t s) {
  const long long n = (long long)A.nnzb * A.bs * A.bs;
  const int nt = 256;
#define HFX_BS(B)                                                                                               \
  case B:                                                                                                       \
    hipLaunchKernelGGL((scatter_kernel<B>), dim3(ceil_div(n, nt)), dim3(nt), 0, s, A.nnzb, nFpc, A.col.p,      \
                       A.rowf.p, A.contrib.p, S, invperm, is_dir, A.vals.p);                                    \
    break;
  if (n > 0) {
    switch (A.bs) {
      HFX_BS(2) HFX_BS(3) HFX_BS(4) HFX_BS(5) HFX_BS(6) HFX_BS(7) HFX_BS(8) HFX_BS(9) HFX_BS(10) HFX_BS(11) HFX_BS(12) HFX_BS(15) HFX_BS(18) HFX_BS(21) HFX_BS(30)
      default: throw std::runtime_error("scatter: unsupported block size");
    }
    HFX_HIP(hipGetLastError());
  }
#undef HFX_BS
  const long long nr = A.nrows();
  if (nr > 0) {
    hipLaunchKernelGGL(rhs_kernel, dim3(ceil_div(nr, nt)), dim3(nt), 0, s, A.nb, A.bs, nFpc, A.face_cells.p, S0,
                       invperm, is_dir, g, rhs);
    HFX_HIP(hipGetLastError());
  }
}

void launch_scatter_diag(const BsrDev& A, int nFpc, const double* dside, const double* S0, const int8_t* invperm,
                         const uint8_t* is_dir, const double* g, double* rhs, hipStream_t s) {
  const int nt = 256;
  const long long n = (long long)A.nb * A.bs * A.bs;
#define HFX_BS(B)                                                                                                 \
  case B:                                                                                                         \
    hipLaunchKernelGGL((scatter_diag_kernel<B>), dim3(ceil_div(n, nt)), dim3(nt), 0, s, A.nb, A.diag.p,           \
                       A.face_cells.p, dside, is_dir, A.vals.p);                                                  \
    break;
  if (n > 0) {
    switch (A.bs) {
      HFX_BS(2) HFX_BS(3) HFX_BS(4) HFX_BS(5) HFX_BS(6) HFX_BS(10)
      default: throw std::runtime_error("scatter: unsupported block size");
    }
    HFX_HIP(hipGetLastError());
  }
#undef HFX_BS
  const long long nr = A.nrows();
  if (nr > 0) {
    hipLaunchKernelGGL(rhs_kernel, dim3(ceil_div(nr, nt)), dim3(nt), 0, s, A.nb, A.bs, nFpc, A.face_cells.p, S0,
                       invperm, is_dir, g, rhs);
    HFX_HIP(hipGetLastError());
  }
}

void launch_integrated_dirichlet(const BsrDev& A, int nbf, const int* bfaces, int dim, int nN, int nipf, int nFpc,
                                 const int* fnodes, const int8_t* face_perm, const int* cells, const double* coords,
                                 const double* fphi, const double* fdphi, const double* fw, const double* g,
                                 double* rhs, hipStream_t s, int m) {
  if (nbf <= 0) return;
  const int nt = 256, bs = A.bs, nNf = bs / m;
  if (nNf * m != bs) throw std::runtime_error("integrated Dirichlet: block size is not a multiple of m");
  const long long n = (long long)nbf * nNf * nNf;
  hipLaunchKernelGGL(integrated_dirichlet_kernel, dim3(ceil_div(n, nt)), dim3(nt), 0, s, nbf, bfaces, dim, nN, nNf,
                     nipf, nFpc, A.face_cells.p, fnodes, face_perm, cells, coords, fphi, fdphi, fw, A.diag.p, m,
                     A.vals.p);
  HFX_HIP(hipGetLastError());
  hipLaunchKernelGGL(integrated_dirichlet_rhs_kernel, dim3(ceil_div((long long)nbf * bs, nt)), dim3(nt), 0, s, nbf,
                     bfaces, bs, A.diag.p, A.vals.p, g, rhs);
  HFX_HIP(hipGetLastError());
}

void launch_hdg_boundary(const BsrDev& A, int nhf, const int* hfaces, int dim, int nN, int nFpc, const int* fnodes,
                         const int8_t* face_perm, const int8_t* invperm, const int* bpos, const double* U,
                         const double* Q, const double* U0, const double* Q0, const double* B, const double* r,
                         int matrix_set, int rhs_set, double* rhs, hipStream_t s, int m) {
  if (nhf <= 0) return;
  const int nt = 256, bs = A.bs, nNf = bs / m;
  if (nNf * m != bs) throw std::runtime_error("HDG-type boundary: block size is not a multiple of m");
  const long long n = (long long)nhf * bs * nFpc * bs;
  hipLaunchKernelGGL(hdg_boundary_kernel, dim3(ceil_div(n, nt)), dim3(nt), 0, s, nhf, hfaces, dim, nN, nNf, m, nFpc,
                     A.face_cells.p, fnodes, face_perm, invperm, bpos, U, Q, B, matrix_set, A.vals.p);
  HFX_HIP(hipGetLastError());
  hipLaunchKernelGGL(hdg_boundary_rhs_kernel, dim3(ceil_div((long long)nhf * bs, nt)), dim3(nt), 0, s, nhf, hfaces,
                     dim, nN, nNf, m, nFpc, A.face_cells.p, fnodes, invperm, U0, Q0, B, r, rhs_set, rhs);
  HFX_HIP(hipGetLastError());
}

std::vector<int> cell_block_positions(const BsrPattern& P, int nCells, int nFpc, const int* cell2face) {
  std::vector<int> out((size_t)nCells * nFpc * nFpc);
  for (int e = 0; e < nCells; ++e)
    for (int f = 0; f < nFpc; ++f) {
      const int F = cell2face[(size_t)e * nFpc + f];
      const int* c0 = P.col.data() + P.rowptr[F];
      const int* c1 = P.col.data() + P.rowptr[F + 1];
      for (int g = 0; g < nFpc; ++g) {
        const int G = cell2face[(size_t)e * nFpc + g];
        const int* it = std::lower_bound(c0, c1, G);
        if (it == c1 || *it != G) throw std::runtime_error("bsr: a cell's face pair is missing from the pattern");
        out[((size_t)e * nFpc + f) * nFpc + g] = (int)(it - P.col.data());
      }
    }
  return out;
}

void launch_reconstruct(int nEl, int nN, int nQ, int nFpc, int bs, const double* U, const double* U0, const double* Q,
                        const double* Q0, const int* cell2face, const int8_t* face_perm, const double* lambda, double* u,
                        double* q, hipStream_t s) {
  if (nEl <= 0) return;
  const int NL = nFpc * bs;
  const size_t lds = ((size_t)NL + 64 * ((size_t)NL + 1)) * sizeof(double);
  if (lds > 64 * 1024) throw std::runtime_error("reconstruct: trace block too long");
  const int grid = std::min(nEl, 65536);
  hipLaunchKernelGGL(reconstruct_kernel, dim3(grid), dim3(64), lds, s, nEl, nN, nQ, nFpc, bs, U, U0, Q, Q0, cell2face,
                     face_perm, lambda, u, q);
  HFX_HIP(hipGetLastError());
}

}  // namespace hfx
