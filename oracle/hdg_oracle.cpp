// FP64 CPU oracle for the HDG element-by-element path (test infrastructure
// only; see hdg_oracle.h).  Every routine cites the reference code it restates.
#include "hdg_oracle.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

#include "../hyperfox_amd/csrc/data/cubature_data.inc"

typedef std::vector<double> vec;

// ---------------------------------------------------------------- 1-D rules
// Legendre P_n(x) and P_n'(x) by the three-term recurrence.
void leg(int n, long double x, long double* p, long double* dp) {
  long double a = 1.0L, b = x;
  if (n == 0) { *p = 1; *dp = 0; return; }
  for (int k = 2; k <= n; ++k) { long double c = ((2 * k - 1) * x * b - (k - 1) * a) / k; a = b; b = c; }
  *p = b;
  *dp = n * (a - x * b) / (1.0L - x * x);
}

// Gauss-Legendre n points (Cubature.cpp:106-280 tabulates the same rules).
void gauss(int n, vec& x, vec& w) {
  x.assign(n, 0); w.assign(n, 0);
  for (int k = 0; k < n; ++k) {
    long double z = cosl(3.14159265358979323846264338327950288L * (k + 0.75L) / (n + 0.5L));
    for (int it = 0; it < 200; ++it) {
      long double p, dp; leg(n, z, &p, &dp);
      long double dz = p / dp; z -= dz;
      if (fabsl(dz) < 1e-19L) break;
    }
    long double p, dp; leg(n, z, &p, &dp);
    x[k] = (double)z; w[k] = (double)(2.0L / ((1.0L - z * z) * dp * dp));
  }
}

// GLL interior points of order p by bisection + Newton on P_p' (sign changes
// of P_p' bracket its roots between the Gauss points of P_p).
vec gll_interior(int p) {
  vec r;
  if (p < 2) return r;
  vec gx, gw; gauss(p, gx, gw);
  std::sort(gx.begin(), gx.end());
  for (int k = 0; k + 1 < p; ++k) {
    long double lo = gx[k], hi = gx[k + 1];
    long double plo, dlo, phi_, dhi;
    leg(p, lo, &plo, &dlo); leg(p, hi, &phi_, &dhi);
    for (int it = 0; it < 200; ++it) {
      long double mid = 0.5L * (lo + hi), pm, dm;
      leg(p, mid, &pm, &dm);
      if ((dm > 0) == (dlo > 0)) { lo = mid; dlo = dm; } else { hi = mid; }
      if (hi - lo < 1e-20L) break;
    }
    r.push_back((double)(0.5L * (lo + hi)));
  }
  return r;
}

// ------------------------------------------------------------ node database
// The reference builds simplex nodes as: vertices, then the nodes of every
// sub-entity (edges, then 3-D faces) mapped affinely from the lower-dimensional
// node set, then interior nodes of the warped Lobatto grid; later duplicates
// are dropped (ReferenceElement.cpp:1045-1183).
const double VT2[3][2] = {{-1, -1}, {1, -1}, {-1, 1}};
const double VT3[4][3] = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
const int ED2[3][2] = {{0, 1}, {1, 2}, {2, 0}};
const int ED3[6][2] = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 1}, {3, 2}};
const int FC3[4][3] = {{3, 1, 0}, {2, 1, 3}, {2, 3, 0}, {0, 1, 2}};

double vtx(int dim, int v, int k) { return dim == 2 ? VT2[v][k] : VT3[v][k]; }

vec nodes_of(int dim, int p);

// point of the q-simplex (coords xq) carried to the entity with vertex list ent
void carry(int dim, int q, const int* ent, const double* xq, double* out) {
  for (int k = 0; k < dim; ++k) {
    double v = vtx(dim, ent[0], k);
    for (int l = 0; l < q; ++l) v += 0.5 * (xq[l] + 1.0) * (vtx(dim, ent[l + 1], k) - vtx(dim, ent[0], k));
    out[k] = v;
  }
}

vec nodes_of(int dim, int p) {
  if (dim == 1) {
    if (p == 0) return {0.0};
    vec x = {-1.0, 1.0};
    vec in = gll_interior(p);
    x.insert(x.end(), in.begin(), in.end());
    return x;
  }
  std::vector<vec> all;
  for (int v = 0; v <= dim; ++v) { vec x(dim); for (int k = 0; k < dim; ++k) x[k] = vtx(dim, v, k); all.push_back(x); }
  vec e1 = nodes_of(1, p);
  int nedge = dim == 2 ? 3 : 6;
  for (int e = 0; e < nedge; ++e)
    for (size_t l = 0; l < e1.size(); ++l) {
      vec x(dim);
      carry(dim, 1, dim == 2 ? ED2[e] : ED3[e], &e1[l], x.data());
      all.push_back(x);
    }
  if (dim == 3) {
    vec f2 = nodes_of(2, p);
    for (int f = 0; f < 4; ++f)
      for (size_t l = 0; l < f2.size() / 2; ++l) {
        vec x(3);
        carry(3, 2, FC3[f], &f2[2 * l], x.data());
        all.push_back(x);
      }
  }
  if (p > 1) {
    vec in = gll_interior(p);
    vec L(p + 1); L[0] = -1; L[p] = 1;
    for (int k = 1; k < p; ++k) L[k] = in[k - 1];
    auto coef = [&](const int* c) {
      int s = 0; for (int l = 0; l < dim; ++l) s += c[l];
      double v = 2.0 + dim * L[c[0]];
      for (int l = 1; l < dim; ++l) v -= L[c[l]];
      v -= L[p - s];
      return v / (dim + 1.0) - 1.0;
    };
    if (dim == 2) {
      for (int a = 1; a < p; ++a)
        for (int b = 1; b < p; ++b)
          if (a + b < p) {
            int c1[2] = {a, b}, c2[2] = {b, a};
            all.push_back({coef(c1), coef(c2)});
          }
    } else {
      for (int a = 1; a < p; ++a)
        for (int b = 1; b < p; ++b)
          for (int c = 1; c < p; ++c)
            if (a + b + c < p) {
              int c1[3] = {a, b, c}, c2[3] = {b, a, c}, c3[3] = {c, b, a};
              all.push_back({coef(c1), coef(c2), coef(c3)});
            }
    }
  }
  vec out;
  for (size_t i = 0; i < all.size(); ++i) {
    bool dup = false;
    for (size_t j = 0; j < i && !dup; ++j) {
      bool same = true;
      for (int k = 0; k < dim; ++k) same = same && std::fabs(all[i][k] - all[j][k]) <= 1e-12;
      dup = same;
    }
    if (!dup) out.insert(out.end(), all[i].begin(), all[i].end());
  }
  return out;
}

// ----------------------------------------------------------------- cubature
int cubature(int dim, int order, vec* pts, vec* w) {
  if (dim == 1) {
    int n = std::max(1, (order + 2) / 2);
    vec x, ww; gauss(n, x, ww);
    if (pts) *pts = x;
    if (w) *w = ww;
    return n;
  }
  const int *o2n, *list, *off; const double* data; int nl, maxo;
  if (dim == 2) { o2n = wv_tri_order_to_npts; list = wv_tri_npts_list; off = wv_tri_npts_offset; data = wv_tri_data; nl = 19; maxo = 20; }
  else { o2n = wv_tet_order_to_npts; list = wv_tet_npts_list; off = wv_tet_npts_offset; data = wv_tet_data; nl = 9; maxo = 10; }
  if (order < 0 || order > maxo) throw std::runtime_error("cubature order out of range");
  int n = o2n[order], slot = -1;
  for (int i = 0; i < nl; ++i) if (list[i] == n) slot = i;
  if (pts) pts->assign(n * dim, 0);
  if (w) w->assign(n, 0);
  for (int i = 0; i < n; ++i) {
    const double* r = data + (size_t)(off[slot] + i) * (dim + 1);
    if (pts) for (int k = 0; k < dim; ++k) (*pts)[i * dim + k] = r[k];
    if (w) (*w)[i] = r[dim];
  }
  return n;
}

// ------------------------------------------------ Lagrange basis (monomial)
// Lagrange functions of the node set through a monomial basis inverted in
// long double; independent of the modal basis the reference uses
// (ReferenceElement.cpp:269-480), which spans the same P_p.
// Lagrange basis on the given nodes through the monomials: phi_j = sum_m
// mono_m W[m][j], W the inverse Vandermonde.  The monomial Vandermonde of the
// 2-D p = 10 nodes has a condition number near 1e9, so the inversion and the
// sums run in IEEE quadruple precision (__float128, 113-bit mantissa): the
// tables then carry full double accuracy at every order (long double left
// 1e-10 in the p = 10 derivatives).
typedef __float128 qf;
static inline qf qabs(qf v) { return v < 0 ? -v : v; }

struct Basis {
  int dim, p, n;
  std::vector<int> ex;             // n x dim exponents
  std::vector<qf> W;               // n x n, phi_j = sum_m mono_m W[m][j]
  void init(int d, int pp, const vec& nodes) {
    dim = d; p = pp; ex.clear();
    for (int a = 0; a <= p; ++a)
      for (int b = 0; b <= (d > 1 ? p - a : 0); ++b)
        for (int c = 0; c <= (d > 2 ? p - a - b : 0); ++c) { ex.push_back(a); if (d > 1) ex.push_back(b); if (d > 2) ex.push_back(c); }
    n = (int)ex.size() / d;
    std::vector<qf> V(n * n), I(n * n, 0);
    for (int i = 0; i < n; ++i) for (int m = 0; m < n; ++m) V[i * n + m] = mono(m, &nodes[i * d]);
    for (int i = 0; i < n; ++i) I[i * n + i] = 1;
    for (int k = 0; k < n; ++k) {   // Gauss-Jordan, partial pivoting
      int pv = k;
      for (int i = k + 1; i < n; ++i) if (qabs(V[i * n + k]) > qabs(V[pv * n + k])) pv = i;
      for (int j = 0; j < n; ++j) { std::swap(V[k * n + j], V[pv * n + j]); std::swap(I[k * n + j], I[pv * n + j]); }
      qf dk = 1 / V[k * n + k];
      for (int j = 0; j < n; ++j) { V[k * n + j] *= dk; I[k * n + j] *= dk; }
      for (int i = 0; i < n; ++i) if (i != k) {
        qf f = V[i * n + k];
        if (f != 0) for (int j = 0; j < n; ++j) { V[i * n + j] -= f * V[k * n + j]; I[i * n + j] -= f * I[k * n + j]; }
      }
    }
    W = I;
  }
  qf mono(int m, const double* x) const {
    qf v = 1;
    for (int k = 0; k < dim; ++k) for (int e = 0; e < ex[m * dim + k]; ++e) v *= (qf)x[k];
    return v;
  }
  void eval(const double* x, double* phi, double* dphi) const {
    std::vector<qf> mv(n), mg(n * dim);
    for (int m = 0; m < n; ++m) {
      mv[m] = mono(m, x);
      for (int k = 0; k < dim; ++k) {
        int e = ex[m * dim + k];
        qf g = e;
        if (e == 0) g = 0;
        for (int l = 0; l < dim; ++l) {
          int el = ex[m * dim + l] - (l == k ? 1 : 0);
          for (int t = 0; t < el; ++t) g *= (qf)x[l];
        }
        mg[m * dim + k] = g;
      }
    }
    for (int j = 0; j < n; ++j) {
      qf s = 0;
      for (int m = 0; m < n; ++m) s += mv[m] * W[m * n + j];
      phi[j] = (double)s;
      if (dphi) for (int k = 0; k < dim; ++k) {
        qf g = 0;
        for (int m = 0; m < n; ++m) g += mg[m * dim + k] * W[m * n + j];
        dphi[j * dim + k] = (double)g;
      }
    }
  }
};

struct Ref {
  int d, p, nN, nF, nNf, nip, nipf;
  vec nodes, ipx, ipw, phi, dphi;      // element
  vec fnodes, fipx, fipw, fphi, fdphi; // face element
  std::vector<std::vector<int>> fmap;  // nF x nNf
  int lenU() const { return nN; }
  int lenQ() const { return nN * d; }
  int lenL() const { return nF * nNf; }
  int n() const { return nN * (d + 1) + nF * nNf; }
};

std::map<std::pair<int, int>, std::shared_ptr<Ref>> g_refs;

std::mutex g_ref_mu;

const Ref& ref(int d, int p) {
  auto key = std::make_pair(d, p);
  std::lock_guard<std::mutex> lk(g_ref_mu);
  {
    if (!g_refs.count(key)) {
      if (d < 2 || d > 3 || p < 1 || p > (d == 2 ? 10 : 5))
        throw std::runtime_error("oracle supports 2-D orders 1-10 and 3-D orders 1-5");
      auto r = std::make_shared<Ref>();
      r->d = d; r->p = p; r->nF = d + 1;
      r->nodes = nodes_of(d, p);
      r->nN = (int)r->nodes.size() / d;
      Basis B; B.init(d, p, r->nodes);
      if (B.n != r->nN) throw std::runtime_error("oracle: node count mismatch");
      r->nip = cubature(d, 2 * p, &r->ipx, &r->ipw);
      r->phi.resize(r->nip * r->nN); r->dphi.resize(r->nip * r->nN * d);
      for (int q = 0; q < r->nip; ++q) B.eval(&r->ipx[q * d], &r->phi[q * r->nN], &r->dphi[q * r->nN * d]);
      int fd = d - 1;
      r->fnodes = nodes_of(fd, p);
      r->nNf = (int)r->fnodes.size() / fd;
      Basis Bf; Bf.init(fd, p, r->fnodes);
      r->nipf = cubature(fd, 2 * p, &r->fipx, &r->fipw);
      r->fphi.resize(r->nipf * r->nNf); r->fdphi.resize(r->nipf * r->nNf * fd);
      for (int q = 0; q < r->nipf; ++q) Bf.eval(&r->fipx[q * fd], &r->fphi[q * r->nNf], &r->fdphi[q * r->nNf * fd]);
      // face-node maps: locate the carried face nodes (ReferenceElement.cpp:136-202)
      r->fmap.assign(r->nF, {});
      for (int f = 0; f < r->nF; ++f) {
        const int* ent = d == 2 ? ED2[f] : FC3[f];
        for (int l = 0; l < r->nNf; ++l) {
          double x[3];
          carry(d, fd, ent, &r->fnodes[l * fd], x);
          int idx = -1;
          for (int k = 0; k < r->nN && idx < 0; ++k) {
            bool same = true;
            for (int c = 0; c < d; ++c) same = same && std::fabs(r->nodes[k * d + c] - x[c]) <= 1e-12;
            if (same) idx = k;
          }
          if (idx < 0) throw std::runtime_error("oracle: face node not found");
          r->fmap[f].push_back(idx);
        }
      }
      g_refs[key] = r;
    }
  }
  return *g_refs.at(key);
}

// ------------------------------------------------------------- small dense
double det(int n, const double* J) {
  if (n == 1) return J[0];
  if (n == 2) return J[0] * J[3] - J[1] * J[2];
  return J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) + J[2] * (J[3] * J[7] - J[4] * J[6]);
}
void inv(int n, const double* J, double* Ji) {
  double dt = det(n, J);
  if (n == 2) { Ji[0] = J[3] / dt; Ji[1] = -J[1] / dt; Ji[2] = -J[2] / dt; Ji[3] = J[0] / dt; return; }
  Ji[0] = (J[4] * J[8] - J[5] * J[7]) / dt; Ji[1] = (J[2] * J[7] - J[1] * J[8]) / dt; Ji[2] = (J[1] * J[5] - J[2] * J[4]) / dt;
  Ji[3] = (J[5] * J[6] - J[3] * J[8]) / dt; Ji[4] = (J[0] * J[8] - J[2] * J[6]) / dt; Ji[5] = (J[2] * J[3] - J[0] * J[5]) / dt;
  Ji[6] = (J[3] * J[7] - J[4] * J[6]) / dt; Ji[7] = (J[1] * J[6] - J[0] * J[7]) / dt; Ji[8] = (J[0] * J[4] - J[1] * J[3]) / dt;
}

// Householder QR solve of A X = B (A n x n row-major, B n x m row-major,
// overwritten by X).  Restates the role of Eigen::HouseholderQR at
// HDGSolver.cpp:217-218, 331-345.
void qr_solve(int n, vec A, double* B, int m) {
  vec tau(n);
  for (int k = 0; k < n; ++k) {
    double nrm = 0;
    for (int i = k; i < n; ++i) nrm += A[i * n + k] * A[i * n + k];
    nrm = std::sqrt(nrm);
    if (nrm == 0) throw std::runtime_error("oracle: singular local block");
    double alpha = A[k * n + k] > 0 ? -nrm : nrm;
    // v = x - alpha e1, stored in column k below and at the diagonal
    double v0 = A[k * n + k] - alpha;
    double vn = v0 * v0;
    for (int i = k + 1; i < n; ++i) vn += A[i * n + k] * A[i * n + k];
    A[k * n + k] = v0;
    tau[k] = vn > 0 ? 2.0 / vn : 0.0;
    // apply H = I - tau v v^T to the trailing columns of A and to B
    for (int j = k + 1; j < n; ++j) {
      double s = 0;
      for (int i = k; i < n; ++i) s += A[i * n + k] * A[i * n + j];
      s *= tau[k];
      for (int i = k; i < n; ++i) A[i * n + j] -= s * A[i * n + k];
    }
    for (int j = 0; j < m; ++j) {
      double s = 0;
      for (int i = k; i < n; ++i) s += A[i * n + k] * B[i * m + j];
      s *= tau[k];
      for (int i = k; i < n; ++i) B[i * m + j] -= s * A[i * n + k];
    }
    A[k * n + k] = alpha;  // R diagonal
  }
  for (int j = 0; j < m; ++j)
    for (int i = n - 1; i >= 0; --i) {
      double s = B[i * m + j];
      for (int c = i + 1; c < n; ++c) s -= A[i * n + c] * B[c * m + j];
      B[i * m + j] = s / A[i * n + i];
    }
}

// ------------------------------------------------------- element geometry
struct Geo {
  vec dV;        // nip + nF*nipf
  vec Jinv;      // nip * d*d
  vec nrm;       // nF*nipf*d
};

// Jacobians, inverse Jacobians, measures and outward normals at the element
// and face IPs (Operator.cpp:14-84; HDGModel.cpp:53-85; HDGBase.cpp:34-65).
void geometry(const Ref& R, const double* x, Geo& g) {
  const int d = R.d, nN = R.nN;
  g.dV.assign(R.nip + R.nF * R.nipf, 0);
  g.Jinv.assign(R.nip * d * d, 0);
  g.nrm.assign(R.nF * R.nipf * d, 0);
  for (int q = 0; q < R.nip; ++q) {
    double J[9] = {0};
    for (int i = 0; i < nN; ++i)
      for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) J[a * d + b] += R.dphi[(q * nN + i) * d + a] * x[i * d + b];
    double dt = det(d, J);
    g.dV[q] = R.ipw[q] * dt;
    inv(d, J, &g.Jinv[q * d * d]);
  }
  const int fd = d - 1;
  for (int f = 0; f < R.nF; ++f) {
    int v0 = R.fmap[f][0], vn = -1;
    for (int k = 0; k < nN && vn < 0; ++k)
      if (std::find(R.fmap[f].begin(), R.fmap[f].end(), k) == R.fmap[f].end()) vn = k;
    for (int q = 0; q < R.nipf; ++q) {
      double J[6] = {0};  // fd x d
      for (int i = 0; i < R.nNf; ++i) {
        const double* xi = x + R.fmap[f][i] * d;
        for (int a = 0; a < fd; ++a)
          for (int b = 0; b < d; ++b) J[a * d + b] += R.fdphi[(q * R.nNf + i) * fd + a] * xi[b];
      }
      double JJt[9];  // fd x fd with fd <= 2; det() is sized for 3 x 3
      for (int a = 0; a < fd; ++a)
        for (int c = 0; c < fd; ++c) {
          double s = 0;
          for (int b = 0; b < d; ++b) s += J[a * d + b] * J[c * d + b];
          JJt[a * fd + c] = s;
        }
      double meas = std::sqrt(det(fd, JJt));
      double nv[3];
      if (d == 2) { nv[0] = J[1]; nv[1] = -J[0]; }
      else {
        nv[0] = J[1] * J[5] - J[2] * J[4];
        nv[1] = J[2] * J[3] - J[0] * J[5];
        nv[2] = J[0] * J[4] - J[1] * J[3];
      }
      double nn = 0;
      for (int b = 0; b < d; ++b) nn += nv[b] * nv[b];
      nn = std::sqrt(nn);
      double prod = 0;
      for (int b = 0; b < d; ++b) { nv[b] /= nn; prod += (x[vn * d + b] - x[v0 * d + b]) * nv[b]; }
      if (prod > 0) for (int b = 0; b < d; ++b) nv[b] = -nv[b];
      int o = f * R.nipf + q;
      g.dV[R.nip + o] = R.fipw[q] * meas;
      for (int b = 0; b < d; ++b) g.nrm[o * d + b] = nv[b];
    }
  }
}

// -------------------------------------------------------- local assembly
// Sum of the HDG operators of the model (HDGBase.cpp:67-158,
// HDGDiffusion.cpp:31-145, HDGConvection.cpp:31-104, Convection.cpp:5-49,
// Mass.cpp:5-38, Source.cpp:5-48, Reaction.cpp:5-36) for one element.
void local_system(const Ref& R, orc_model mdl, const double* x, const double* D, const double* b, const double* tau,
                  const double* f_ip, const double* r_ip, double* K, double* F) {
  const int d = R.d, nN = R.nN, nNf = R.nNf, nF = R.nF, n = R.n();
  const int oq = nN, ol = nN * (d + 1);
  // the Reaction operator exists in HDGConvectionDiffusionReactionSource only
  // (HDGConvectionDiffusionReactionSource.cpp:90-95)
  if (r_ip && mdl.kind != 2) throw std::runtime_error("oracle: a reaction needs the CDRS model (kind 2)");
  Geo g; geometry(R, x, g);
  std::fill(K, K + (size_t)n * n, 0.0);
  std::fill(F, F + n, 0.0);
  auto at = [&](int r, int c) -> double& { return K[(size_t)r * n + c]; };
  // diffusion tensor per node (column-major d x d); identity if absent
  bool useDiff = (mdl.kind != 2) || (D != nullptr);
  std::vector<double> Dn(nN * d * d, 0.0);
  for (int i = 0; i < nN; ++i)
    for (int a = 0; a < d; ++a)
      for (int c = 0; c < d; ++c) {
        double v = (a == c) ? 1.0 : 0.0;
        if (D) v = (mdl.ndD == 1) ? (a == c ? D[i] : 0.0) : D[i * d * d + c * d + a];
        Dn[(i * d + a) * d + c] = v;  // Dn[i][a][c] = D_ac
      }
  // --- HDGBase face terms
  for (int f = 0; f < nF; ++f)
    for (int q = 0; q < R.nipf; ++q) {
      int o = f * R.nipf + q;
      double dv = g.dV[R.nip + o];
      const double* ps = &R.fphi[q * nNf];
      double t = 0;
      for (int a = 0; a < nNf; ++a) t += tau[f * nNf + a] * ps[a];
      for (int a = 0; a < nNf; ++a)
        for (int c = 0; c < nNf; ++c) {
          double val = dv * t * ps[a] * ps[c];
          int ia = R.fmap[f][a], ic = R.fmap[f][c];
          at(ol + f * nNf + a, ol + f * nNf + c) -= val;
          at(ol + f * nNf + a, ic) += val;
          at(ia, ic) += val;
          at(ia, ol + f * nNf + c) -= val;
          for (int k = 0; k < d; ++k) at(oq + ia * d + k, ol + f * nNf + c) -= dv * g.nrm[o * d + k] * ps[a] * ps[c];
        }
    }
  // --- HDGBase volume terms
  for (int q = 0; q < R.nip; ++q) {
    double dv = g.dV[q];
    const double* ph = &R.phi[q * nN];
    const double* Ji = &g.Jinv[q * d * d];
    for (int i = 0; i < nN; ++i) {
      double gr[3];
      for (int k = 0; k < d; ++k) {
        double s = 0;
        for (int a = 0; a < d; ++a) s += Ji[k * d + a] * R.dphi[(q * nN + i) * d + a];
        gr[k] = s;
      }
      for (int k = 0; k < d; ++k)
        for (int j = 0; j < nN; ++j) {
          at(oq + i * d + k, j) += dv * gr[k] * ph[j];
          at(oq + i * d + k, oq + j * d + k) += dv * ph[i] * ph[j];
        }
    }
  }
  // --- HDGDiffusion
  if (useDiff) {
    for (int f = 0; f < nF; ++f)
      for (int q = 0; q < R.nipf; ++q) {
        int o = f * R.nipf + q;
        double dv = g.dV[R.nip + o];
        const double* ps = &R.fphi[q * nNf];
        double Dq[9] = {0};
        for (int a = 0; a < nNf; ++a)
          for (int t = 0; t < d * d; ++t) Dq[t] += ps[a] * Dn[R.fmap[f][a] * d * d + t];
        double v[3];
        for (int k = 0; k < d; ++k) {
          double s = 0;
          for (int c = 0; c < d; ++c) s += Dq[k * d + c] * g.nrm[o * d + c];
          v[k] = s * dv;
        }
        for (int a = 0; a < nNf; ++a)
          for (int c = 0; c < nNf; ++c)
            for (int k = 0; k < d; ++k) {
              double val = v[k] * ps[a] * ps[c];
              at(ol + f * nNf + a, oq + R.fmap[f][c] * d + k) -= val;
              at(R.fmap[f][a], oq + R.fmap[f][c] * d + k) -= val;
            }
      }
    for (int q = 0; q < R.nip; ++q) {
      double dv = g.dV[q];
      const double* ph = &R.phi[q * nN];
      const double* Ji = &g.Jinv[q * d * d];
      double Dq[9] = {0};
      for (int i = 0; i < nN; ++i)
        for (int t = 0; t < d * d; ++t) Dq[t] += ph[i] * Dn[i * d * d + t];
      for (int i = 0; i < nN; ++i) {
        double gr[3], dg[3];
        for (int k = 0; k < d; ++k) {
          double s = 0;
          for (int a = 0; a < d; ++a) s += Ji[k * d + a] * R.dphi[(q * nN + i) * d + a];
          gr[k] = s;
        }
        for (int k = 0; k < d; ++k) {
          double s = 0;
          for (int c = 0; c < d; ++c) s += Dq[k * d + c] * gr[c];
          dg[k] = s * dv;
        }
        for (int j = 0; j < nN; ++j)
          for (int k = 0; k < d; ++k) at(i, oq + j * d + k) += dg[k] * ph[j];
      }
    }
  }
  // --- HDGConvection
  if (b != nullptr) {
    for (int q = 0; q < R.nip; ++q) {
      double dv = g.dV[q];
      const double* ph = &R.phi[q * nN];
      const double* Ji = &g.Jinv[q * d * d];
      double bq[3] = {0};
      for (int i = 0; i < nN; ++i) for (int k = 0; k < d; ++k) bq[k] += ph[i] * b[i * d + k];
      for (int l = 0; l < nN; ++l) {
        double s = 0;
        for (int k = 0; k < d; ++k) {
          double gr = 0;
          for (int a = 0; a < d; ++a) gr += Ji[k * d + a] * R.dphi[(q * nN + l) * d + a];
          s += bq[k] * gr;
        }
        // Convection C(k,l) += dV (b.grad phi_l) phi_k, and Suu -= C^T
        for (int kk = 0; kk < nN; ++kk) at(l, kk) -= dv * s * ph[kk];
      }
    }
    for (int f = 0; f < nF; ++f)
      for (int q = 0; q < R.nipf; ++q) {
        int o = f * R.nipf + q;
        const double* ps = &R.fphi[q * nNf];
        double bq[3] = {0};
        for (int a = 0; a < nNf; ++a) for (int k = 0; k < d; ++k) bq[k] += ps[a] * b[R.fmap[f][a] * d + k];
        double bn = 0;
        for (int k = 0; k < d; ++k) bn += bq[k] * g.nrm[o * d + k];
        double m = g.dV[R.nip + o] * bn;
        for (int a = 0; a < nNf; ++a)
          for (int c = 0; c < nNf; ++c) {
            double val = m * ps[a] * ps[c];
            at(R.fmap[f][a], ol + f * nNf + c) += val;
            at(ol + f * nNf + a, ol + f * nNf + c) += val;
          }
      }
  }
  // --- Reaction and Source
  for (int q = 0; q < R.nip; ++q) {
    double dv = g.dV[q];
    const double* ph = &R.phi[q * nN];
    if (r_ip)
      for (int i = 0; i < nN; ++i)
        for (int j = 0; j < nN; ++j) at(i, j) += dv * r_ip[q] * ph[i] * ph[j];
    if (f_ip && mdl.kind != 0)
      for (int i = 0; i < nN; ++i) F[i] += dv * f_ip[q] * ph[i];
  }
}

// Static condensation (HDGSolver.cpp:331-348) of a local system with u, q
// and lambda blocks of lu, lq and ll rows.  Row-major outputs.
void condense_sizes(int lu, int lq, int ll, const double* K, const double* F, double* S, double* S0, double* U,
                    double* Q, double* U0, double* Q0) {
  const int n = lu + lq + ll;
  const int ou = 0, oq = lu, ol = lu + lq;
  auto blk = [&](int r0, int nr, int c0, int nc) {
    vec o((size_t)nr * nc);
    for (int i = 0; i < nr; ++i) for (int j = 0; j < nc; ++j) o[(size_t)i * nc + j] = K[(size_t)(r0 + i) * n + c0 + j];
    return o;
  };
  vec Sqq = blk(oq, lq, oq, lq);
  vec AB((size_t)lq * (lu + ll));  // [Sqq^-1 Squ | Sqq^-1 Sql]
  for (int i = 0; i < lq; ++i) {
    for (int j = 0; j < lu; ++j) AB[(size_t)i * (lu + ll) + j] = K[(size_t)(oq + i) * n + ou + j];
    for (int j = 0; j < ll; ++j) AB[(size_t)i * (lu + ll) + lu + j] = K[(size_t)(oq + i) * n + ol + j];
  }
  qr_solve(lq, Sqq, AB.data(), lu + ll);
  vec Kuu = blk(ou, lu, ou, lu);
  vec Suq = blk(ou, lu, oq, lq);
  vec R2((size_t)lu * (ll + 1));  // [-(Sul - Suq B) | Fu]
  for (int i = 0; i < lu; ++i) {
    for (int j = 0; j < lu; ++j) {
      double s = 0;
      for (int k = 0; k < lq; ++k) s += Suq[(size_t)i * lq + k] * AB[(size_t)k * (lu + ll) + j];
      Kuu[(size_t)i * lu + j] -= s;
    }
    for (int j = 0; j < ll; ++j) {
      double s = 0;
      for (int k = 0; k < lq; ++k) s += Suq[(size_t)i * lq + k] * AB[(size_t)k * (lu + ll) + lu + j];
      R2[(size_t)i * (ll + 1) + j] = -(K[(size_t)(ou + i) * n + ol + j] - s);
    }
    R2[(size_t)i * (ll + 1) + ll] = F[ou + i];
  }
  qr_solve(lu, Kuu, R2.data(), ll + 1);
  for (int i = 0; i < lu; ++i) {
    for (int j = 0; j < ll; ++j) U[(size_t)i * ll + j] = R2[(size_t)i * (ll + 1) + j];
    U0[i] = R2[(size_t)i * (ll + 1) + ll];
  }
  for (int i = 0; i < lq; ++i) {
    for (int j = 0; j < ll; ++j) {
      double s = 0;
      for (int k = 0; k < lu; ++k) s += AB[(size_t)i * (lu + ll) + k] * U[(size_t)k * ll + j];
      Q[(size_t)i * ll + j] = -s - AB[(size_t)i * (lu + ll) + lu + j];
    }
    double s = 0;
    for (int k = 0; k < lu; ++k) s += AB[(size_t)i * (lu + ll) + k] * U0[k];
    Q0[i] = -s;
  }
  for (int i = 0; i < ll; ++i) {
    for (int j = 0; j < ll; ++j) {
      double s = K[(size_t)(ol + i) * n + ol + j];
      for (int k = 0; k < lu; ++k) s += K[(size_t)(ol + i) * n + ou + k] * U[(size_t)k * ll + j];
      for (int k = 0; k < lq; ++k) s += K[(size_t)(ol + i) * n + oq + k] * Q[(size_t)k * ll + j];
      S[(size_t)i * ll + j] = s;
    }
    double s = F[ol + i];
    for (int k = 0; k < lu; ++k) s -= K[(size_t)(ol + i) * n + ou + k] * U0[k];
    for (int k = 0; k < lq; ++k) s -= K[(size_t)(ol + i) * n + oq + k] * Q0[k];
    S0[i] = s;
  }
}

void condense(const Ref& R, const double* K, const double* F, double* S, double* S0, double* U, double* Q, double* U0,
              double* Q0) {
  condense_sizes(R.lenU(), R.lenQ(), R.lenL(), K, F, S, S0, U, Q, U0, Q0);
}

// Explicit HDG modes (HDGSolver.cpp:346-353): S = Sll and
// S0 = F_lambda - Slu sol - Slq flux with the element's old solution and flux.
void explicit_blocks(const Ref& R, const double* K, const double* F, const double* sol, const double* flux,
                     double* S, double* S0) {
  const int n = R.n(), lu = R.lenU(), lq = R.lenQ(), ll = R.lenL();
  const int ou = 0, oq = lu, ol = lu + lq;
  for (int i = 0; i < ll; ++i) {
    for (int j = 0; j < ll; ++j) S[(size_t)i * ll + j] = K[(size_t)(ol + i) * n + ol + j];
    double v = F[ol + i];
    for (int k = 0; k < lu; ++k) v -= K[(size_t)(ol + i) * n + ou + k] * sol[k];
    for (int k = 0; k < lq; ++k) v -= K[(size_t)(ol + i) * n + oq + k] * flux[k];
    S0[i] = v;
  }
}

void check_dim(int dim, int p) {
  if (dim < 2 || dim > 3 || p < 1 || p > (dim == 2 ? 10 : 5))
    throw std::runtime_error("oracle supports 2-D orders 1-10 and 3-D orders 1-5");
}

}  // namespace

// =================================================================== C ABI

#define ORC_TRY try {
#define ORC_END                        \
  }                                    \
  catch (const std::exception& e) {    \
    g_err = e.what();                  \
    return -1;                         \
  }                                    \
  return 0;

extern "C" const char* orc_last_error(void) { return g_err.c_str(); }

extern "C" int orc_sizes(int dim, int p, int* out) {
  ORC_TRY
  check_dim(dim, p);
  const Ref& R = ref(dim, p);
  out[0] = R.nN; out[1] = R.nF; out[2] = R.nNf; out[3] = R.nip; out[4] = R.nipf; out[5] = R.n(); out[6] = R.lenL();
  ORC_END
}

extern "C" int orc_cubature(int dim, int order, double* pts, double* w) {
  try {
    vec P, W;
    int n = cubature(dim, order, &P, &W);
    if (pts) std::copy(P.begin(), P.end(), pts);
    if (w) std::copy(W.begin(), W.end(), w);
    return n;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

extern "C" int orc_refel(int dim, int p, double* nodes, int* face_nodes, double* ipx, double* ipw, double* phi,
                         double* dphi, double* fipx, double* fipw, double* fphi) {
  ORC_TRY
  check_dim(dim, p);
  const Ref& R = ref(dim, p);
  if (nodes) std::copy(R.nodes.begin(), R.nodes.end(), nodes);
  if (face_nodes) for (int f = 0; f < R.nF; ++f) for (int j = 0; j < R.nNf; ++j) face_nodes[f * R.nNf + j] = R.fmap[f][j];
  if (ipx) std::copy(R.ipx.begin(), R.ipx.end(), ipx);
  if (ipw) std::copy(R.ipw.begin(), R.ipw.end(), ipw);
  if (phi) std::copy(R.phi.begin(), R.phi.end(), phi);
  if (dphi) std::copy(R.dphi.begin(), R.dphi.end(), dphi);
  if (fipx) std::copy(R.fipx.begin(), R.fipx.end(), fipx);
  if (fipw) std::copy(R.fipw.begin(), R.fipw.end(), fipw);
  if (fphi) std::copy(R.fphi.begin(), R.fphi.end(), fphi);
  ORC_END
}

namespace {

// Euler::apply with isExplicit = false (Euler.cpp:18-37), called from
// HDGModel::compute (HDGModel.cpp:39-47) on the u rows of the local system:
// u rows (every column) and F_u scaled by dt, then the element mass M added
// to the u-u block and M u_old to F_u.  M is the u-u mass of the TimeScheme's
// Mass operator (Mass.cpp:5-38), which is the first component block of
// Sqq = I_d (x) M (HDGBase.cpp:139-157) on the same cubature.
std::vector<double> time_mass(const Ref& R, const double* K) {
  const int n = R.n(), nN = R.nN, d = R.d;
  std::vector<double> M((size_t)nN * nN);
  for (int i = 0; i < nN; ++i)
    for (int j = 0; j < nN; ++j) M[(size_t)i * nN + j] = K[(size_t)(nN + i * d) * n + nN + j * d];
  return M;
}

void apply_euler(const Ref& R, double dt, const double* u_old, double* K, double* F) {
  const int n = R.n(), nN = R.nN;
  const std::vector<double> M = time_mass(R, K);
  for (int r = 0; r < nN; ++r) {
    for (int c = 0; c < n; ++c) K[(size_t)r * n + c] *= dt;
    F[r] *= dt;
  }
  for (int i = 0; i < nN; ++i) {
    double mu = 0.0;
    for (int j = 0; j < nN; ++j) {
      K[(size_t)i * n + j] += M[(size_t)i * nN + j];
      mu += M[(size_t)i * nN + j] * u_old[j];
    }
    F[i] += mu;
  }
}

// Euler::apply with isExplicit = true (Euler.cpp:30-35): u rows and F_u
// scaled by dt, then F_u += (M - Suu) u_old with the scaled Suu, and Suu = M.
void apply_euler_explicit(const Ref& R, double dt, const double* u_old, double* K, double* F) {
  const int n = R.n(), nN = R.nN;
  const std::vector<double> M = time_mass(R, K);
  for (int r = 0; r < nN; ++r) {
    for (int c = 0; c < n; ++c) K[(size_t)r * n + c] *= dt;
    F[r] *= dt;
  }
  for (int i = 0; i < nN; ++i) {
    double acc = 0.0;
    for (int j = 0; j < nN; ++j) acc += (M[(size_t)i * nN + j] - K[(size_t)i * n + j]) * u_old[j];
    F[i] += acc;
    for (int j = 0; j < nN; ++j) K[(size_t)i * n + j] = M[(size_t)i * nN + j];
  }
}

// RungeKutta::apply (RungeKutta.cpp:113-159) on the u rows [Suu | Suq | Sul]
// of a local system, for the auxiliary fields {Flux, Trace}: a = row `stage`
// of the Butcher table without its first column (nst entries), ut = the
// cell-local [OldSolution | OldFlux | OldTrace] and kst[j * n ..] the
// cell-local [RKStage_j | RKStage_Flux_j | RKStage_Trace_j], in the local
// system's column order.  Steps, as the reference takes them:
//   uj = dt sum_{i<stage} a_i k_i + ut;  K_u *= dt;  F_u *= dt;
//   F_u -= K_u uj;  K_u *= a_stage;  Suu += M;  F_u += K_u ut.
void apply_rk(const Ref& R, double dt, int nst, const double* a, int stage, const double* ut, const double* kst,
              double* K, double* F) {
  const int n = R.n(), nN = R.nN;
  const std::vector<double> M = time_mass(R, K);
  std::vector<double> uj(n, 0.0);
  for (int i = 0; i < stage; ++i)
    for (int c = 0; c < n; ++c) uj[c] += a[i] * kst[(size_t)i * n + c];
  for (int c = 0; c < n; ++c) uj[c] = uj[c] * dt + ut[c];
  for (int r = 0; r < nN; ++r) {
    for (int c = 0; c < n; ++c) K[(size_t)r * n + c] *= dt;
    F[r] *= dt;
  }
  for (int r = 0; r < nN; ++r) {
    double acc = 0.0;
    for (int c = 0; c < n; ++c) acc += K[(size_t)r * n + c] * uj[c];
    F[r] -= acc;
  }
  for (int r = 0; r < nN; ++r)
    for (int c = 0; c < n; ++c) K[(size_t)r * n + c] *= a[stage];
  for (int i = 0; i < nN; ++i)
    for (int j = 0; j < nN; ++j) K[(size_t)i * n + j] += M[(size_t)i * nN + j];
  for (int r = 0; r < nN; ++r) {
    double acc = 0.0;
    for (int c = 0; c < n; ++c) acc += K[(size_t)r * n + c] * ut[c];
    F[r] += acc;
  }
}

// Mass operator on the face element (Mass.cpp:5-38) with the face's own
// nodes xf (nNf x d, face-node order), as IntegratedDirichletModel computes
// it (IntegratedDirichletModel.cpp:35-44): Jacobians of the face map
// (Operator.cpp:14-36), measure sqrt(det(J J^T)) times the face weights,
// M_ab = sum_q psi_a psi_b dV_q.
void face_mass(const Ref& R, const double* xf, double* M) {
  const int d = R.d, fd = d - 1, nNf = R.nNf;
  std::fill(M, M + (size_t)nNf * nNf, 0.0);
  for (int q = 0; q < R.nipf; ++q) {
    double J[6] = {0};
    for (int i = 0; i < nNf; ++i)
      for (int a = 0; a < fd; ++a)
        for (int b = 0; b < d; ++b) J[a * d + b] += R.fdphi[(q * nNf + i) * fd + a] * xf[i * d + b];
    double JJt[9];  // fd x fd with fd <= 2; det() is sized for 3 x 3
    for (int a = 0; a < fd; ++a)
      for (int c = 0; c < fd; ++c) {
        double s = 0;
        for (int b = 0; b < d; ++b) s += J[a * d + b] * J[c * d + b];
        JJt[a * fd + c] = s;
      }
    const double dv = R.fipw[q] * std::sqrt(det(fd, JJt));
    for (int a = 0; a < nNf; ++a)
      for (int b = 0; b < nNf; ++b) M[a * nNf + b] += R.fphi[q * nNf + a] * R.fphi[q * nNf + b] * dv;
  }
}

void check_rk(double dt, int nst, int stage) {
  if (dt == 0.0)
    throw std::runtime_error("RungeKutta: the time step needs to be set before applying and it should not be 0");
  if (nst < 1 || stage < 0 || stage >= nst) throw std::runtime_error("RungeKutta: stage out of range");
}


// HDGBase (HDGBase.cpp:18-32, 67-158) with m unknowns per node and an m x m
// tau per face node (tau_ab at (f nNf + j) m^2 + b m + a), added into K
// (order u (i m + a), q ((i d + k) m + a), lambda ((f nNf + j) m + a)).
void base_m(const Ref& R, const Geo& g, const double* tau, int m, double* K) {
  const int d = R.d, nN = R.nN, nNf = R.nNf, nF = R.nF;
  const int lu = nN * m, lq = nN * d * m, ll = nF * nNf * m, n = lu + lq + ll;
  const int oq = lu, ol = lu + lq;
  auto at = [&](int r, int c) -> double& { return K[(size_t)r * n + c]; };
  auto L = [&](int f, int j, int a) { return ol + (f * nNf + j) * m + a; };
  for (int f = 0; f < nF; ++f)
    for (int q = 0; q < R.nipf; ++q) {
      const int o = f * R.nipf + q;
      const double dv = g.dV[R.nip + o];
      const double* ps = &R.fphi[q * nNf];
      double T[9] = {0};
      for (int t = 0; t < m * m; ++t)
        for (int j = 0; j < nNf; ++j) T[t] += tau[(f * nNf + j) * m * m + t] * ps[j];
      for (int a = 0; a < nNf; ++a)
        for (int c = 0; c < nNf; ++c) {
          const int ia = R.fmap[f][a], ic = R.fmap[f][c];
          const double w = dv * ps[a] * ps[c];
          for (int na = 0; na < m; ++na)
            for (int mb = 0; mb < m; ++mb) {
              const double val = T[mb * m + na] * w;
              at(L(f, a, na), L(f, c, mb)) -= val;
              at(L(f, a, na), ic * m + mb) += val;
              at(ia * m + na, ic * m + mb) += val;
              at(ia * m + na, L(f, c, mb)) -= val;
            }
          for (int na = 0; na < m; ++na)
            for (int k = 0; k < d; ++k) at(oq + (ia * d + k) * m + na, L(f, c, na)) -= w * g.nrm[o * d + k];
        }
    }
  for (int q = 0; q < R.nip; ++q) {
    const double dv = g.dV[q];
    const double* ph = &R.phi[q * nN];
    const double* Ji = &g.Jinv[q * d * d];
    for (int i = 0; i < nN; ++i) {
      double gr[3];
      for (int k = 0; k < d; ++k) {
        double s = 0;
        for (int a = 0; a < d; ++a) s += Ji[k * d + a] * R.dphi[(q * nN + i) * d + a];
        gr[k] = s;
      }
      for (int k = 0; k < d; ++k)
        for (int a = 0; a < m; ++a)
          for (int j = 0; j < nN; ++j) {
            at(oq + (i * d + k) * m + a, j * m + a) += dv * gr[k] * ph[j];
            at(oq + (i * d + k) * m + a, oq + (j * d + k) * m + a) += dv * ph[i] * ph[j];
          }
    }
  }
}

// Scalar HDGConvection (HDGConvection.cpp:31-104 + Convection.cpp:5-49, face
// Mass) with the nodal velocity v (nN x d), added into the scalar local
// matrix K1 (n = R.n(), the m = 1 order): Suu -= C^T, Sul and Sll += (v.n)
// face mass.
void convection_op(const Ref& R, const Geo& g, const double* v, double* K1) {
  const int d = R.d, nN = R.nN, nNf = R.nNf, nF = R.nF, n = R.n(), ol = nN * (d + 1);
  auto at = [&](int r, int c) -> double& { return K1[(size_t)r * n + c]; };
  for (int q = 0; q < R.nip; ++q) {
    const double dv = g.dV[q];
    const double* ph = &R.phi[q * nN];
    const double* Ji = &g.Jinv[q * d * d];
    double vq[3] = {0};
    for (int i = 0; i < nN; ++i) for (int k = 0; k < d; ++k) vq[k] += ph[i] * v[i * d + k];
    for (int l = 0; l < nN; ++l) {
      double s = 0;
      for (int k = 0; k < d; ++k) {
        double gr = 0;
        for (int a = 0; a < d; ++a) gr += Ji[k * d + a] * R.dphi[(q * nN + l) * d + a];
        s += vq[k] * gr;
      }
      for (int kk = 0; kk < nN; ++kk) at(l, kk) -= dv * s * ph[kk];
    }
  }
  for (int f = 0; f < nF; ++f)
    for (int q = 0; q < R.nipf; ++q) {
      const int o = f * R.nipf + q;
      const double* ps = &R.fphi[q * nNf];
      double vq[3] = {0};
      for (int a = 0; a < nNf; ++a) for (int k = 0; k < d; ++k) vq[k] += ps[a] * v[R.fmap[f][a] * d + k];
      double vn = 0;
      for (int k = 0; k < d; ++k) vn += vq[k] * g.nrm[o * d + k];
      const double w = g.dV[R.nip + o] * vn;
      for (int a = 0; a < nNf; ++a)
        for (int c = 0; c < nNf; ++c) {
          const double val = w * ps[a] * ps[c];
          at(R.fmap[f][a], ol + f * nNf + c) += val;
          at(ol + f * nNf + a, ol + f * nNf + c) += val;
        }
    }
}

// Scalar HDGDiffusion (HDGDiffusion.cpp:31-145) with the coefficient Dn I at
// each node, added into the scalar local matrix K1.
void diffusion_op(const Ref& R, const Geo& g, const double* Dn, double* K1) {
  const int d = R.d, nN = R.nN, nNf = R.nNf, nF = R.nF, n = R.n(), oq = nN, ol = nN * (d + 1);
  auto at = [&](int r, int c) -> double& { return K1[(size_t)r * n + c]; };
  for (int f = 0; f < nF; ++f)
    for (int q = 0; q < R.nipf; ++q) {
      const int o = f * R.nipf + q;
      const double dv = g.dV[R.nip + o];
      const double* ps = &R.fphi[q * nNf];
      double Dq = 0;
      for (int a = 0; a < nNf; ++a) Dq += ps[a] * Dn[R.fmap[f][a]];
      for (int a = 0; a < nNf; ++a)
        for (int c = 0; c < nNf; ++c)
          for (int k = 0; k < d; ++k) {
            const double val = Dq * g.nrm[o * d + k] * dv * ps[a] * ps[c];
            at(ol + f * nNf + a, oq + R.fmap[f][c] * d + k) -= val;
            at(R.fmap[f][a], oq + R.fmap[f][c] * d + k) -= val;
          }
    }
  for (int q = 0; q < R.nip; ++q) {
    const double dv = g.dV[q];
    const double* ph = &R.phi[q * nN];
    const double* Ji = &g.Jinv[q * d * d];
    double Dq = 0;
    for (int i = 0; i < nN; ++i) Dq += ph[i] * Dn[i];
    for (int i = 0; i < nN; ++i) {
      double gr[3];
      for (int k = 0; k < d; ++k) {
        double s = 0;
        for (int a = 0; a < d; ++a) s += Ji[k * d + a] * R.dphi[(q * nN + i) * d + a];
        gr[k] = s;
      }
      for (int j = 0; j < nN; ++j)
        for (int k = 0; k < d; ++k) at(i, oq + j * d + k) += Dq * gr[k] * dv * ph[j];
    }
  }
}

// HDGnGammaModel::computeLocalMatrix / computeLocalRHS
// (HDGnGammaModel.cpp:105-170), 2-D, m = 2 (n, Gamma): HDGBase with the
// 2 x 2 tau, then four scalar HDGConvection operators with the nodal
// velocities v_00 = 0, v_01 = b, v_10 = (c^2 - (Gamma/n)^2) b,
// v_11 = 2 (Gamma/n) b from the BufferSolution, HDGDiffusion with D on
// (n, n) and G on (Gamma, Gamma) (each only when given), interleaved into the
// component blocks (scalar index i -> 2 i + a); the RHS is one Source per
// component.
void ngamma_local(const Ref& R, const double* x, const double* D, const double* G, const double* b, const double* sol,
                  const double* tau, double cs, const double* f_ip, double* K, double* F) {
  const int d = R.d, m = 2, nN = R.nN, n1 = R.n(), n = n1 * m;
  if (d != 2) throw std::runtime_error("nGamma: 2-D only (b has 2 components)");
  Geo g; geometry(R, x, g);
  std::fill(K, K + (size_t)n * n, 0.0);
  std::fill(F, F + n, 0.0);
  base_m(R, g, tau, m, K);
  std::vector<double> v(4 * nN * d, 0.0);
  for (int i = 0; i < nN; ++i) {
    const double r = sol[i * 2 + 1] / sol[i * 2];
    for (int k = 0; k < d; ++k) {
      v[(1 * nN + i) * d + k] = b[i * 2 + k];
      v[(2 * nN + i) * d + k] = (cs * cs - r * r) * b[i * 2 + k];
      v[(3 * nN + i) * d + k] = r * 2.0 * b[i * 2 + k];
    }
  }
  std::vector<double> C((size_t)4 * n1 * n1, 0.0), Df((size_t)2 * n1 * n1, 0.0);
  for (int t = 0; t < 4; ++t) convection_op(R, g, &v[(size_t)t * nN * d], &C[(size_t)t * n1 * n1]);
  if (D) diffusion_op(R, g, D, &Df[0]);
  if (G) diffusion_op(R, g, G, &Df[(size_t)n1 * n1]);
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n1; ++j) {
      const size_t ij = (size_t)i * n1 + j, nn = (size_t)n1 * n1;
      K[(size_t)(2 * i) * n + 2 * j] += Df[ij] + C[ij];
      K[(size_t)(2 * i) * n + 2 * j + 1] += C[nn + ij];
      K[(size_t)(2 * i + 1) * n + 2 * j] += C[2 * nn + ij];
      K[(size_t)(2 * i + 1) * n + 2 * j + 1] += Df[nn + ij] + C[3 * nn + ij];
    }
  if (f_ip)
    for (int q = 0; q < R.nip; ++q)
      for (int i = 0; i < nN; ++i)
        for (int a = 0; a < m; ++a) F[i * m + a] += g.dV[q] * f_ip[q * m + a] * R.phi[q * nN + i];
}

// HDGBurgersModel::computeLocalMatrix / computeLocalRHS
// (HDGBurgersModel.cpp:86-132) with m = d unknowns per node: HDGBase with an
// m x m tau per face node (HDGBase.cpp:18-32, 67-158), HDGDiffusion with a
// scalar coefficient per node (HDGDiffusion.cpp:31-145, Dn null: no
// diffusion operator, HDGBurgersModel.cpp:99-104), HDGUNabU
// (HDGUNabU.cpp:86-190) and one Source per component (Source.cpp:5-48).
// Local order: u (i m + a), q ((i d + k) m + a), lambda ((f nNf + j) m + a).
// tau: (f nNf + j) m^2 + (b m + a) for tau_ab (column-major per node);
// sol: i m + a (the "BufferSolution"); trace: lambda in the cell's local
// face-node order; f_ip: q m + a (null: no source).
void burgers_local(const Ref& R, const double* x, const double* Dn, const double* tau, const double* sol,
                   const double* trace, const double* f_ip, double* K, double* F) {
  const int d = R.d, m = d, nN = R.nN, nNf = R.nNf, nF = R.nF;
  const int lu = nN * m, lq = nN * d * m, ll = nF * nNf * m, n = lu + lq + ll;
  const int oq = lu, ol = lu + lq;
  Geo g; geometry(R, x, g);
  std::fill(K, K + (size_t)n * n, 0.0);
  std::fill(F, F + n, 0.0);
  auto at = [&](int r, int c) -> double& { return K[(size_t)r * n + c]; };
  auto grad = [&](int q, int i, double* gr) {
    const double* Ji = &g.Jinv[q * d * d];
    for (int k = 0; k < d; ++k) {
      double s = 0;
      for (int a = 0; a < d; ++a) s += Ji[k * d + a] * R.dphi[(q * nN + i) * d + a];
      gr[k] = s;
    }
  };
  auto L = [&](int f, int j, int a) { return ol + (f * nNf + j) * m + a; };
  base_m(R, g, tau, m, K);
  // --- HDGUNabU, in its own matrix for the right-hand side
  std::vector<double> op((size_t)n * n, 0.0);
  auto opa = [&](int r, int c) -> double& { return op[(size_t)r * n + c]; };
  for (int f = 0; f < nF; ++f)
    for (int q = 0; q < R.nipf; ++q) {
      const int o = f * R.nipf + q;
      const double dv = g.dV[R.nip + o];
      const double* ps = &R.fphi[q * nNf];
      const double* nv = &g.nrm[o * d];
      double lam[3] = {0}, us[3] = {0};
      for (int j = 0; j < nNf; ++j)
        for (int a = 0; a < m; ++a) {
          lam[a] += ps[j] * trace[(f * nNf + j) * m + a];
          us[a] += ps[j] * sol[R.fmap[f][j] * m + a];
        }
      double ln = 0;
      for (int k = 0; k < d; ++k) ln += lam[k] * nv[k];
      for (int i = 0; i < nNf; ++i)
        for (int a = 0; a < m; ++a)
          for (int j = 0; j < nNf; ++j) {
            opa(L(f, i, a), L(f, j, a)) += ln * dv * ps[j] * ps[i];
            for (int b = 0; b < m; ++b) opa(L(f, i, a), L(f, j, b)) += us[a] * ps[j] * ps[i] * nv[b] * dv;
          }
    }
  for (int f = 0; f < nF; ++f)
    for (int i = 0; i < nNf; ++i)
      for (int a = 0; a < m; ++a)
        for (int c = 0; c < ll; ++c) opa(R.fmap[f][i] * m + a, ol + c) += opa(L(f, i, a), ol + c);
  for (int q = 0; q < R.nip; ++q) {
    const double dv = g.dV[q];
    const double* ph = &R.phi[q * nN];
    std::vector<double> gm((size_t)nN * d);
    for (int i = 0; i < nN; ++i) grad(q, i, &gm[(size_t)i * d]);
    double gs[9] = {0}, uq[3] = {0};  // gs[k][a] = d_k u_a
    for (int i = 0; i < nN; ++i)
      for (int a = 0; a < m; ++a) {
        uq[a] += ph[i] * sol[i * m + a];
        for (int k = 0; k < d; ++k) gs[k * m + a] += gm[(size_t)i * d + k] * sol[i * m + a];
      }
    double div = 0;
    for (int k = 0; k < d; ++k) div += gs[k * m + k];
    for (int i = 0; i < nN; ++i) {
      double ug = 0;
      for (int k = 0; k < d; ++k) ug += uq[k] * gm[(size_t)i * d + k];
      for (int a = 0; a < m; ++a)
        for (int j = 0; j < nN; ++j) {
          opa(i * m + a, j * m + a) -= (div * ph[i] + ug) * dv * ph[j];
          for (int b = 0; b < m; ++b)
            opa(i * m + a, j * m + b) -= uq[a] * (gm[(size_t)j * d + b] * ph[i] + gm[(size_t)i * d + b] * ph[j]) * dv;
        }
    }
  }
  // rhs = op [sol/2 ; trace/2] on the u and lambda rows (HDGUNabU.cpp:180-189)
  for (int r = 0; r < n; ++r) {
    if (r >= oq && r < ol) continue;
    double s = 0;
    for (int c = 0; c < lu; ++c) s += opa(r, c) * (sol[c] / 2.0);
    for (int c = 0; c < ll; ++c) s += opa(r, ol + c) * (trace[c] / 2.0);
    F[r] = s;
  }
  for (size_t t = 0; t < (size_t)n * n; ++t) K[t] += op[t];
  // --- HDGDiffusion with D I at each node
  if (Dn) {
    for (int f = 0; f < nF; ++f)
      for (int q = 0; q < R.nipf; ++q) {
        const int o = f * R.nipf + q;
        const double dv = g.dV[R.nip + o];
        const double* ps = &R.fphi[q * nNf];
        double Dq = 0;
        for (int a = 0; a < nNf; ++a) Dq += ps[a] * Dn[R.fmap[f][a]];
        for (int a = 0; a < nNf; ++a)
          for (int c = 0; c < nNf; ++c)
            for (int k = 0; k < d; ++k) {
              const double val = Dq * g.nrm[o * d + k] * dv * ps[a] * ps[c];
              for (int na = 0; na < m; ++na) {
                at(L(f, a, na), oq + (R.fmap[f][c] * d + k) * m + na) -= val;
                at(R.fmap[f][a] * m + na, oq + (R.fmap[f][c] * d + k) * m + na) -= val;
              }
            }
      }
    for (int q = 0; q < R.nip; ++q) {
      const double dv = g.dV[q];
      const double* ph = &R.phi[q * nN];
      double Dq = 0;
      for (int i = 0; i < nN; ++i) Dq += ph[i] * Dn[i];
      for (int i = 0; i < nN; ++i) {
        double gr[3];
        grad(q, i, gr);
        for (int a = 0; a < m; ++a)
          for (int j = 0; j < nN; ++j)
            for (int k = 0; k < d; ++k) at(i * m + a, oq + (j * d + k) * m + a) += Dq * gr[k] * dv * ph[j];
      }
    }
  }
  // --- Source_a on the u rows of component a (HDGBurgersModel.cpp:118-128)
  if (f_ip)
    for (int q = 0; q < R.nip; ++q)
      for (int i = 0; i < nN; ++i)
        for (int a = 0; a < m; ++a) F[i * m + a] += g.dV[q] * f_ip[q * m + a] * R.phi[q * nN + i];
}
}  // namespace

extern "C" int orc_burgers_local(int dim, int p, const double* x, const double* Dn, const double* tau,
                                 const double* sol, const double* trace, const double* f_ip, double* K, double* F) {
  ORC_TRY
  check_dim(dim, p);
  burgers_local(ref(dim, p), x, Dn, tau, sol, trace, f_ip, K, F);
  ORC_END
}

extern "C" int orc_ngamma_local(int p, const double* x, const double* D, const double* G, const double* b,
                                const double* sol, const double* tau, double cs, const double* f_ip, double* K,
                                double* F) {
  ORC_TRY
  check_dim(2, p);
  ngamma_local(ref(2, p), x, D, G, b, sol, tau, cs, f_ip, K, F);
  ORC_END
}

extern "C" int orc_condense_sizes(int lu, int lq, int ll, const double* K, const double* F, double* S, double* S0,
                                  double* U, double* Q, double* U0, double* Q0) {
  ORC_TRY
  if (lu <= 0 || lq <= 0 || ll <= 0) throw std::runtime_error("condense: empty blocks");
  condense_sizes(lu, lq, ll, K, F, S, S0, U, Q, U0, Q0);
  ORC_END
}

extern "C" int orc_apply_euler_explicit(int dim, int p, double dt, const double* u_old, double* K, double* F) {
  ORC_TRY
  check_dim(dim, p);
  if (!(dt > 0.0)) throw std::runtime_error("Euler: the time step needs to be set and it should not be 0");
  apply_euler_explicit(ref(dim, p), dt, u_old, K, F);
  ORC_END
}

extern "C" int orc_face_mass(int dim, int p, const double* xf, double* M) {
  ORC_TRY
  check_dim(dim, p);
  face_mass(ref(dim, p), xf, M);
  ORC_END
}

extern "C" int orc_apply_rk(int dim, int p, double dt, int nst, const double* a, int stage, const double* ut,
                            const double* kst, double* K, double* F) {
  ORC_TRY
  check_dim(dim, p);
  check_rk(dt, nst, stage);
  apply_rk(ref(dim, p), dt, nst, a, stage, ut, kst, K, F);
  ORC_END
}

extern "C" int orc_apply_euler(int dim, int p, double dt, const double* u_old, double* K, double* F) {
  ORC_TRY
  check_dim(dim, p);
  if (!(dt > 0.0)) throw std::runtime_error("Euler: the time step needs to be set and it should not be 0");
  apply_euler(ref(dim, p), dt, u_old, K, F);
  ORC_END
}

extern "C" int orc_local_system(int dim, int p, orc_model model, const double* x, const double* D, const double* b,
                                const double* tau, const double* f_ip, const double* r_ip, double* K, double* F) {
  ORC_TRY
  check_dim(dim, p);
  local_system(ref(dim, p), model, x, D, b, tau, f_ip, r_ip, K, F);
  ORC_END
}

extern "C" int orc_explicit_blocks(int dim, int p, const double* K, const double* F, const double* sol,
                                   const double* flux, double* S, double* S0) {
  ORC_TRY
  check_dim(dim, p);
  explicit_blocks(ref(dim, p), K, F, sol, flux, S, S0);
  ORC_END
}

extern "C" int orc_ip_coords(int dim, int p, const double* x, double* out) {
  ORC_TRY
  check_dim(dim, p);
  const Ref& R = ref(dim, p);
  for (int q = 0; q < R.nip; ++q)
    for (int k = 0; k < dim; ++k) {
      double s = 0;
      for (int i = 0; i < R.nN; ++i) s += R.phi[q * R.nN + i] * x[i * dim + k];
      out[q * dim + k] = s;
    }
  ORC_END
}

extern "C" int orc_condense(int dim, int p, const double* K, const double* F, double* S, double* S0, double* U,
                            double* Q, double* U0, double* Q0) {
  ORC_TRY
  check_dim(dim, p);
  condense(ref(dim, p), K, F, S, S0, U, Q, U0, Q0);
  ORC_END
}

struct orc_system {
  int dim, p, nEl, nFaces, ndofs;
  const Ref* R;
  orc_model model;
  std::vector<double> coords, D, b, tau, f_ip, r_ip, dir;
  bool hasD = false, hasB = false, hasF = false, hasR = false;
  int tau_sides = 1;
  std::vector<int> cells, faces, c2f, f2c;
  std::vector<int> perm;              // nEl*nF*nNf : face2CellMap
  std::vector<char> is_bnd;           // nFaces
  std::vector<char> bc_kind;          // per face: 0 DirichletModel, 1 IntegratedDirichletModel, 2 HDG type
  // HDG-type boundary model per face: index into hb/hr, local matrix
  // [B_u | B_q | B_lambda] (nNf x nNf (2 + d)) and rhs (nNf), assembly types
  std::vector<int> hdg_idx;
  std::vector<double> hb, hr;
  int hdg_mset = 1, hdg_rset = 1;
  std::vector<double> S, S0, U, Q, U0, Q0;
  double ts_dt = 0.0;          // time step (0: no time scheme)
  int ts_kind = 0;             // 1 implicit Euler, 2 explicit Euler, 3 Runge-Kutta stage
  std::vector<double> u_old;   // nEl * nN "Solution" field (Euler)
  int rk_nst = 0, rk_stage = 0;
  std::vector<double> rk_a;    // Butcher row of the stage (nst)
  std::vector<double> rk_old;  // nEl * n cell-local [OldSolution | OldFlux | OldTrace]
  std::vector<double> rk_k;    // nEl * nst * n cell-local stage vectors
  int hdg_type = 0;            // HDGSolverType: 0 IMPLICIT, 1 WEXPLICIT, 2 SEXPLICIT
  std::vector<double> sol, flux;  // nEl * nN and nEl * nN * dim for the explicit modes
  std::vector<long long> rowptr;
  std::vector<int> cols;
  std::vector<double> vals, rhs;
};

namespace {

// Per-element gather, local system, condensation and Dirichlet rows
// (HDGSolver.cpp:176-529), then the ADD scatter into the CSR rows
// (HDGSolver.cpp:531-675).
void system_assemble(orc_system* s, int nEl, int nthreads) {
  const Ref& R = *s->R;
  const int d = s->dim, nN = R.nN, nNf = R.nNf, nF = R.nF, n = R.n();
  const int lu = R.lenU(), lq = R.lenQ(), ll = R.lenL();
#pragma omp parallel num_threads(nthreads)
  {
    std::vector<double> x(nN * d), Dl, bl, tl(nF * nNf), K((size_t)n * n), F(n);
#pragma omp for schedule(dynamic, 64)
    for (int e = 0; e < nEl; ++e) {
      const int* c = &s->cells[(size_t)e * nN];
      for (int i = 0; i < nN; ++i) for (int k = 0; k < d; ++k) x[i * d + k] = s->coords[(size_t)c[i] * d + k];
      if (s->hasD) { Dl.resize(nN * s->model.ndD); for (int i = 0; i < nN; ++i) for (int t = 0; t < s->model.ndD; ++t) Dl[i * s->model.ndD + t] = s->D[(size_t)c[i] * s->model.ndD + t]; }
      if (s->hasB) { bl.resize(nN * d); for (int i = 0; i < nN; ++i) for (int k = 0; k < d; ++k) bl[i * d + k] = s->b[(size_t)c[i] * d + k]; }
      for (int f = 0; f < nF; ++f) {
        int fc = s->c2f[(size_t)e * nF + f];
        int side = (s->tau_sides == 2 && s->f2c[2 * fc] != e) ? 1 : 0;
        for (int j = 0; j < nNf; ++j) {
          int pos = s->perm[((size_t)e * nF + f) * nNf + j];
          tl[f * nNf + j] = s->tau[((size_t)fc * nNf + pos) * s->tau_sides + side];
        }
      }
      local_system(R, s->model, x.data(), s->hasD ? Dl.data() : nullptr, s->hasB ? bl.data() : nullptr, tl.data(),
                   s->hasF ? &s->f_ip[(size_t)e * R.nip] : nullptr, s->hasR ? &s->r_ip[(size_t)e * R.nip] : nullptr,
                   K.data(), F.data());
      if (s->ts_kind == 1) apply_euler(R, s->ts_dt, &s->u_old[(size_t)e * nN], K.data(), F.data());
      else if (s->ts_kind == 2) apply_euler_explicit(R, s->ts_dt, &s->u_old[(size_t)e * nN], K.data(), F.data());
      else if (s->ts_kind == 3)
        apply_rk(R, s->ts_dt, s->rk_nst, s->rk_a.data(), s->rk_stage, &s->rk_old[(size_t)e * n],
                 &s->rk_k[(size_t)e * s->rk_nst * n], K.data(), F.data());
      double* Se = &s->S[(size_t)e * ll * ll];
      double* S0e = &s->S0[(size_t)e * ll];
      condense(R, K.data(), F.data(), Se, S0e, &s->U[(size_t)e * lu * ll], &s->Q[(size_t)e * lq * ll],
               &s->U0[(size_t)e * lu], &s->Q0[(size_t)e * lq]);
      if (s->hdg_type != 0)
        explicit_blocks(R, K.data(), F.data(), &s->sol[(size_t)e * lu], &s->flux[(size_t)e * lq], Se, S0e);
      // Dirichlet (DirichletModel: Set/Set, identity matrix, rhs = g in the
      // face's node order, which is this cell's order on a boundary face;
      // IntegratedDirichletModel: Set/Set, the face mass M and rhs M g,
      // HDGSolver.cpp:489-502)
      for (int f = 0; f < nF; ++f) {
        int fc = s->c2f[(size_t)e * nF + f];
        if (!s->is_bnd[fc]) continue;
        if (!s->bc_kind.empty() && s->bc_kind[fc] == 2) {
          // HDG-type boundary model (HDGSolver.cpp:489-518): Set zeroes the
          // face's rows of S / S0; then S_f += B_u U_F + B_q Q_F,
          // S_ff += B_lambda, S0_f += r - B_u U0_F - B_q Q0_F, with U_F, Q_F the
          // rows of U, Q at the face's nodes (face order = this cell's order)
          const int ib = s->hdg_idx[fc];
          const int nc = nNf * (2 + d);
          const double* B = &s->hb[(size_t)ib * nNf * nc];
          const double* r = &s->hr[(size_t)ib * nNf];
          const double* Ue = &s->U[(size_t)e * lu * ll];
          const double* Qe = &s->Q[(size_t)e * lq * ll];
          const double* U0e = &s->U0[(size_t)e * lu];
          const double* Q0e = &s->Q0[(size_t)e * lq];
          for (int a = 0; a < nNf; ++a) {
            double* row = Se + (size_t)(f * nNf + a) * ll;
            if (s->hdg_mset)
              for (int j = 0; j < ll; ++j) row[j] = 0.0;
            if (s->hdg_rset) S0e[f * nNf + a] = 0.0;
            double rr = r[a];
            for (int jn = 0; jn < nNf; ++jn) {
              const int i = R.fmap[f][jn];
              const double bu = B[a * nc + jn];
              for (int c = 0; c < ll; ++c) row[c] += bu * Ue[(size_t)i * ll + c];
              rr -= bu * U0e[i];
              for (int k = 0; k < d; ++k) {
                const double bq = B[a * nc + nNf + jn * d + k];
                for (int c = 0; c < ll; ++c) row[c] += bq * Qe[(size_t)(i * d + k) * ll + c];
                rr -= bq * Q0e[i * d + k];
              }
            }
            for (int b = 0; b < nNf; ++b) row[f * nNf + b] += B[a * nc + nNf * (1 + d) + b];
            S0e[f * nNf + a] += rr;
          }
          continue;
        }
        const bool integ = !s->bc_kind.empty() && s->bc_kind[fc] == 1;
        std::vector<double> xf, M;
        if (integ) {
          xf.resize((size_t)nNf * d);
          M.resize((size_t)nNf * nNf);
          for (int j = 0; j < nNf; ++j)
            for (int k = 0; k < d; ++k) xf[j * d + k] = s->coords[(size_t)s->faces[(size_t)fc * nNf + j] * d + k];
          face_mass(R, xf.data(), M.data());
        }
        for (int a = 0; a < nNf; ++a) {
          for (int j = 0; j < ll; ++j) Se[(size_t)(f * nNf + a) * ll + j] = 0.0;
          S0e[f * nNf + a] = 0.0;
          if (!integ) {
            Se[(size_t)(f * nNf + a) * ll + f * nNf + a] += 1.0;
            S0e[f * nNf + a] += s->dir.empty() ? 0.0 : s->dir[(size_t)fc * nNf + a];
          } else {
            double r = 0.0;
            for (int b = 0; b < nNf; ++b) {
              Se[(size_t)(f * nNf + a) * ll + f * nNf + b] += M[a * nNf + b];
              r += M[a * nNf + b] * (s->dir.empty() ? 0.0 : s->dir[(size_t)fc * nNf + b]);
            }
            S0e[f * nNf + a] += r;
          }
        }
      }
    }
  }
  // scatter (serial over elements: ADD_VALUES semantics, fixed order)
  std::fill(s->vals.begin(), s->vals.end(), 0.0);
  std::fill(s->rhs.begin(), s->rhs.end(), 0.0);
  std::vector<int> rows(ll);
  for (int e = 0; e < nEl; ++e) {
    for (int f = 0; f < nF; ++f) {
      int fc = s->c2f[(size_t)e * nF + f];
      for (int j = 0; j < nNf; ++j) rows[f * nNf + j] = fc * nNf + s->perm[((size_t)e * nF + f) * nNf + j];
    }
    const double* Se = &s->S[(size_t)e * ll * ll];
    for (int i = 0; i < ll; ++i) {
      int r = rows[i];
      long long b0 = s->rowptr[r], b1 = s->rowptr[r + 1];
      for (int j = 0; j < ll; ++j) {
        int cc = rows[j];
        long long lo = b0, hi = b1;
        while (lo < hi) { long long mid = (lo + hi) / 2; if (s->cols[mid] < cc) lo = mid + 1; else hi = mid; }
        s->vals[lo] += Se[(size_t)i * ll + j];
      }
      s->rhs[r] += s->S0[(size_t)e * ll + i];
    }
  }
}

// vector loops below this length run on one thread: a parallel region costs
// more than the loop on the small systems of the known-answer tests
constexpr int PAR_MIN = 16384;

void spmv(const orc_system* s, const double* x, double* y) {
#pragma omp parallel for schedule(static) if (s->ndofs >= PAR_MIN)
  for (int r = 0; r < s->ndofs; ++r) {
    double acc = 0;
    for (long long k = s->rowptr[r]; k < s->rowptr[r + 1]; ++k) acc += s->vals[k] * x[s->cols[k]];
    y[r] = acc;
  }
}

double dot(int n, const double* a, const double* b) {
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static) if (n >= PAR_MIN)
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// GMRES(restart) with left preconditioning by P^-1 = diag(A)^-1 and
// classical Gram-Schmidt, x0 = 0; convergence on the preconditioned residual
// relative to ||P^-1 b|| (PETSc KSPGMRES/PCJACOBI defaults used by
// PetscInterface.cpp:60-86, 222-247).
template <typename Apply>
int gmres_left_jacobi(int n, Apply A, const std::vector<double>& diag, const double* b, double rtol, int maxits,
                      int restart, double* x, bool* conv) {
  std::vector<double> dinv(n);
  for (int i = 0; i < n; ++i) dinv[i] = diag[i] != 0.0 ? 1.0 / diag[i] : 1.0;
  std::fill(x, x + n, 0.0);
  std::vector<double> V((size_t)(restart + 1) * n), H((size_t)(restart + 1) * restart), cs(restart), sn(restart),
      gg(restart + 1), w(n), t(n);
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
  for (int i = 0; i < n; ++i) w[i] = dinv[i] * b[i];
  double r0 = std::sqrt(dot(n, w.data(), w.data()));
  *conv = false;
  if (r0 == 0.0) { *conv = true; return 0; }
  double target = rtol * r0;
  int it = 0;
  while (it < maxits) {
    A(x, t.data());
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
    for (int i = 0; i < n; ++i) w[i] = dinv[i] * (b[i] - t[i]);
    double beta = std::sqrt(dot(n, w.data(), w.data()));
    if (beta <= target) { *conv = true; break; }
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
    for (int i = 0; i < n; ++i) V[i] = w[i] / beta;
    std::fill(gg.begin(), gg.end(), 0.0);
    gg[0] = beta;
    int k = 0;
    for (; k < restart && it < maxits; ++k, ++it) {
      A(&V[(size_t)k * n], t.data());
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
      for (int i = 0; i < n; ++i) w[i] = dinv[i] * t[i];
      for (int j = 0; j <= k; ++j) H[(size_t)j * restart + k] = dot(n, &V[(size_t)j * n], w.data());
      // classical Gram-Schmidt: w -= sum_j h_j v_j, the j sum in order for each i
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
      for (int i = 0; i < n; ++i) {
        double wi = w[i];
        for (int j = 0; j <= k; ++j) wi -= H[(size_t)j * restart + k] * V[(size_t)j * n + i];
        w[i] = wi;
      }
      double hn = std::sqrt(dot(n, w.data(), w.data()));
      H[(size_t)(k + 1) * restart + k] = hn;
      if (hn != 0.0) {
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
        for (int i = 0; i < n; ++i) V[(size_t)(k + 1) * n + i] = w[i] / hn;
      }
      for (int j = 0; j < k; ++j) {
        double a = H[(size_t)j * restart + k], c = H[(size_t)(j + 1) * restart + k];
        H[(size_t)j * restart + k] = cs[j] * a + sn[j] * c;
        H[(size_t)(j + 1) * restart + k] = -sn[j] * a + cs[j] * c;
      }
      double a = H[(size_t)k * restart + k], c = H[(size_t)(k + 1) * restart + k];
      double rr = std::hypot(a, c);
      cs[k] = rr != 0 ? a / rr : 1.0;
      sn[k] = rr != 0 ? c / rr : 0.0;
      H[(size_t)k * restart + k] = rr;
      H[(size_t)(k + 1) * restart + k] = 0.0;
      gg[k + 1] = -sn[k] * gg[k];
      gg[k] = cs[k] * gg[k];
      if (std::fabs(gg[k + 1]) <= target || hn == 0.0) { ++k; ++it; *conv = true; break; }
    }
    // x += V y, with H y = g
    std::vector<double> y(k);
    for (int i = k - 1; i >= 0; --i) {
      double sacc = gg[i];
      for (int j = i + 1; j < k; ++j) sacc -= H[(size_t)i * restart + j] * y[j];
      y[i] = sacc / H[(size_t)i * restart + i];
    }
#pragma omp parallel for schedule(static) if (n >= PAR_MIN)
    for (int i = 0; i < n; ++i) {
      double xi = x[i];
      for (int j = 0; j < k; ++j) xi += y[j] * V[(size_t)j * n + i];
      x[i] = xi;
    }
    if (*conv) break;
  }
  return it;
}

}  // namespace

extern "C" orc_system* orc_system_build(int dim, int p, orc_model model, int nNodes, const double* coords, int nEl,
                                        const int* cells, int nFaces, const int* faces, const int* cell2face,
                                        const int* face2cell, const double* D, const double* b, const double* tau,
                                        int tau_sides, const double* f_ip, const double* r_ip,
                                        const double* dirichlet, int nthreads) {
  try {
    check_dim(dim, p);
    auto s = std::unique_ptr<orc_system>(new orc_system());
    const Ref& R = ref(dim, p);
    s->dim = dim; s->p = p; s->R = &R; s->model = model; s->nEl = nEl; s->nFaces = nFaces;
    const int nN = R.nN, nNf = R.nNf, nF = R.nF, ll = R.lenL();
    s->coords.assign(coords, coords + (size_t)nNodes * dim);
    s->cells.assign(cells, cells + (size_t)nEl * nN);
    s->faces.assign(faces, faces + (size_t)nFaces * nNf);
    s->c2f.assign(cell2face, cell2face + (size_t)nEl * nF);
    s->f2c.assign(face2cell, face2cell + (size_t)nFaces * 2);
    if (D) { s->hasD = true; s->D.assign(D, D + (size_t)nNodes * model.ndD); }
    if (b) { s->hasB = true; s->b.assign(b, b + (size_t)nNodes * dim); }
    s->tau_sides = tau_sides;
    s->tau.assign(tau, tau + (size_t)nFaces * nNf * tau_sides);
    if (f_ip) { s->hasF = true; s->f_ip.assign(f_ip, f_ip + (size_t)nEl * R.nip); }
    if (r_ip) { s->hasR = true; s->r_ip.assign(r_ip, r_ip + (size_t)nEl * R.nip); }
    if (dirichlet) s->dir.assign(dirichlet, dirichlet + (size_t)nFaces * nNf);
    s->is_bnd.assign(nFaces, 0);
    for (int f = 0; f < nFaces; ++f) s->is_bnd[f] = (s->f2c[2 * f + 1] == -1);
    // face2CellMap: position of each cell-local face node in the face list
    s->perm.resize((size_t)nEl * nF * nNf);
    for (int e = 0; e < nEl; ++e)
      for (int f = 0; f < nF; ++f) {
        int fc = s->c2f[(size_t)e * nF + f];
        for (int j = 0; j < nNf; ++j) {
          int node = s->cells[(size_t)e * nN + R.fmap[f][j]];
          int pos = -1;
          for (int k = 0; k < nNf; ++k) if (s->faces[(size_t)fc * nNf + k] == node) pos = k;
          if (pos < 0) throw std::runtime_error("oracle: couldn't find cell node in face");
          s->perm[((size_t)e * nF + f) * nNf + j] = pos;
        }
      }
    // sparsity (HDGSolver.cpp:117-164): row block of face F couples with
    // every face of every cell containing F
    s->ndofs = nFaces * nNf;
    std::vector<std::vector<int>> fcols(nFaces);
    for (int e = 0; e < nEl; ++e)
      for (int f = 0; f < nF; ++f) {
        int F = s->c2f[(size_t)e * nF + f];
        for (int g = 0; g < nF; ++g) fcols[F].push_back(s->c2f[(size_t)e * nF + g]);
      }
    s->rowptr.assign(s->ndofs + 1, 0);
    for (int F = 0; F < nFaces; ++F) {
      auto& v = fcols[F];
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
      for (int j = 0; j < nNf; ++j) s->rowptr[(size_t)F * nNf + j + 1] = (long long)v.size() * nNf;
    }
    for (int r = 0; r < s->ndofs; ++r) s->rowptr[r + 1] += s->rowptr[r];
    s->cols.resize(s->rowptr[s->ndofs]);
    for (int F = 0; F < nFaces; ++F)
      for (int j = 0; j < nNf; ++j) {
        long long o = s->rowptr[(size_t)F * nNf + j];
        for (int G : fcols[F]) for (int k = 0; k < nNf; ++k) s->cols[o++] = G * nNf + k;
      }
    s->vals.assign(s->cols.size(), 0.0);
    s->rhs.assign(s->ndofs, 0.0);
    s->S.assign((size_t)nEl * ll * ll, 0); s->S0.assign((size_t)nEl * ll, 0);
    s->U.assign((size_t)nEl * R.lenU() * ll, 0); s->Q.assign((size_t)nEl * R.lenQ() * ll, 0);
    s->U0.assign((size_t)nEl * R.lenU(), 0); s->Q0.assign((size_t)nEl * R.lenQ(), 0);
    system_assemble(s.get(), nEl, nthreads > 0 ? nthreads : omp_get_max_threads());
    return s.release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

extern "C" void orc_system_free(orc_system* s) { delete s; }

extern "C" int orc_system_set_hdg_type(orc_system* s, int type, const double* sol, const double* flux, int nthreads) {
  ORC_TRY
  if (type < 0 || type > 2) throw std::runtime_error("HDGSolver: unknown solver type");
  s->hdg_type = type;
  if (type != 0) {
    s->sol.assign(sol, sol + (size_t)s->nEl * s->R->lenU());
    s->flux.assign(flux, flux + (size_t)s->nEl * s->R->lenQ());
  }
  system_assemble(s, s->nEl, nthreads > 0 ? nthreads : omp_get_max_threads());
  ORC_END
}

extern "C" int orc_system_set_euler(orc_system* s, double dt, const double* u_old, int nthreads) {
  ORC_TRY
  if (dt < 0.0) throw std::runtime_error("Euler: negative time step");
  s->ts_dt = dt;
  s->ts_kind = dt > 0.0 ? 1 : 0;
  if (dt > 0.0) s->u_old.assign(u_old, u_old + (size_t)s->nEl * s->R->nN);
  system_assemble(s, s->nEl, nthreads > 0 ? nthreads : omp_get_max_threads());
  ORC_END
}

extern "C" int orc_system_set_time_scheme(orc_system* s, int kind, double dt, int nst, const double* a, int stage,
                                          const double* u_old, const double* old_local, const double* stage_local,
                                          int nthreads) {
  ORC_TRY
  const int n = s->R->n();
  if (kind < 0 || kind > 3) throw std::runtime_error("time scheme: unknown kind");
  if (kind != 0 && !(dt > 0.0)) throw std::runtime_error("time scheme: the time step should be > 0");
  s->ts_kind = kind;
  s->ts_dt = kind ? dt : 0.0;
  if (kind == 1 || kind == 2) s->u_old.assign(u_old, u_old + (size_t)s->nEl * s->R->nN);
  if (kind == 3) {
    check_rk(dt, nst, stage);
    s->rk_nst = nst;
    s->rk_stage = stage;
    s->rk_a.assign(a, a + nst);
    s->rk_old.assign(old_local, old_local + (size_t)s->nEl * n);
    s->rk_k.assign((size_t)s->nEl * nst * n, 0.0);
    for (int j = 0; j < nst; ++j)
      for (int e = 0; e < s->nEl; ++e)
        std::copy(stage_local + ((size_t)j * s->nEl + e) * n, stage_local + ((size_t)j * s->nEl + e + 1) * n,
                  &s->rk_k[((size_t)e * nst + j) * n]);
  }
  system_assemble(s, s->nEl, nthreads > 0 ? nthreads : omp_get_max_threads());
  ORC_END
}

extern "C" int orc_system_set_boundary_model(orc_system* s, int kind, const int* faces, int nfaces, int nthreads) {
  ORC_TRY
  if (kind < 0 || kind > 1) throw std::runtime_error("boundary model: unknown kind");
  if (s->bc_kind.empty()) s->bc_kind.assign(s->nFaces, 0);
  if (!faces) {
    for (int F = 0; F < s->nFaces; ++F)
      if (s->is_bnd[F]) s->bc_kind[F] = (char)kind;
  } else {
    for (int i = 0; i < nfaces; ++i) {
      if (faces[i] < 0 || faces[i] >= s->nFaces || !s->is_bnd[faces[i]])
        throw std::runtime_error("boundary model: not a boundary face");
      s->bc_kind[faces[i]] = (char)kind;
    }
  }
  system_assemble(s, s->nEl, nthreads > 0 ? nthreads : omp_get_max_threads());
  ORC_END
}

extern "C" int orc_system_set_boundary_hdg(orc_system* s, const int* faces, int nfaces, const double* B,
                                           const double* r, int matrix_set, int rhs_set, int nthreads) {
  ORC_TRY
  const int nNf = s->R->nNf, nc = nNf * (2 + s->dim);
  if (s->bc_kind.empty()) s->bc_kind.assign(s->nFaces, 0);
  s->hdg_idx.assign(s->nFaces, -1);
  for (int F = 0; F < s->nFaces; ++F)
    if (s->bc_kind[F] == 2) s->bc_kind[F] = 0;
  for (int i = 0; i < nfaces; ++i) {
    if (faces[i] < 0 || faces[i] >= s->nFaces || !s->is_bnd[faces[i]])
      throw std::runtime_error("boundary model: not a boundary face");
    s->bc_kind[faces[i]] = 2;
    s->hdg_idx[faces[i]] = i;
  }
  s->hb.assign(B, B + (size_t)nfaces * nNf * nc);
  s->hr.assign(r, r + (size_t)nfaces * nNf);
  s->hdg_mset = matrix_set != 0;
  s->hdg_rset = rhs_set != 0;
  system_assemble(s, s->nEl, nthreads > 0 ? nthreads : omp_get_max_threads());
  ORC_END
}

extern "C" int orc_system_local_vectors(const orc_system* s, const double* u, const double* q, const double* lambda,
                                        double* out) {
  ORC_TRY
  const Ref& R = *s->R;
  const int n = R.n(), nF = R.nF, nNf = R.nNf, lu = R.lenU(), lq = R.lenQ();
  for (int e = 0; e < s->nEl; ++e) {
    double* o = out + (size_t)e * n;
    for (int i = 0; i < lu; ++i) o[i] = u[(size_t)e * lu + i];
    for (int i = 0; i < lq; ++i) o[lu + i] = q[(size_t)e * lq + i];
    for (int f = 0; f < nF; ++f) {
      const int fc = s->c2f[(size_t)e * nF + f];
      for (int j = 0; j < nNf; ++j)
        o[lu + lq + f * nNf + j] = lambda[(size_t)fc * nNf + s->perm[((size_t)e * nF + f) * nNf + j]];
    }
  }
  ORC_END
}

extern "C" int orc_system_sizes(const orc_system* s, long long* out) {
  out[0] = s->ndofs;
  out[1] = (long long)s->cols.size();
  return 0;
}

extern "C" int orc_system_csr(const orc_system* s, long long* rowptr, int* cols, double* vals, double* rhs) {
  if (rowptr) std::copy(s->rowptr.begin(), s->rowptr.end(), rowptr);
  if (cols) std::copy(s->cols.begin(), s->cols.end(), cols);
  if (vals) std::copy(s->vals.begin(), s->vals.end(), vals);
  if (rhs) std::copy(s->rhs.begin(), s->rhs.end(), rhs);
  return 0;
}

extern "C" int orc_system_condensed(const orc_system* s, double* S, double* S0, double* U, double* Q, double* U0,
                                    double* Q0) {
  if (S) std::copy(s->S.begin(), s->S.end(), S);
  if (S0) std::copy(s->S0.begin(), s->S0.end(), S0);
  if (U) std::copy(s->U.begin(), s->U.end(), U);
  if (Q) std::copy(s->Q.begin(), s->Q.end(), Q);
  if (U0) std::copy(s->U0.begin(), s->U0.end(), U0);
  if (Q0) std::copy(s->Q0.begin(), s->Q0.end(), Q0);
  return 0;
}

extern "C" int orc_system_solve(orc_system* s, int method, double rtol, int maxits, int restart, double* lambda,
                                int* info, double* relres) {
  ORC_TRY
  const int n = s->ndofs;
  if (method == 1) {
    if (n > 8000) throw std::runtime_error("dense LU limited to 8000 unknowns");
    std::vector<double> A((size_t)n * n, 0.0), bb(s->rhs);
    for (int r = 0; r < n; ++r)
      for (long long k = s->rowptr[r]; k < s->rowptr[r + 1]; ++k) A[(size_t)r * n + s->cols[k]] = s->vals[k];
    for (int k = 0; k < n; ++k) {
      int pv = k;
      for (int i = k + 1; i < n; ++i) if (std::fabs(A[(size_t)i * n + k]) > std::fabs(A[(size_t)pv * n + k])) pv = i;
      if (A[(size_t)pv * n + k] == 0.0) throw std::runtime_error("singular trace system");
      if (pv != k) { for (int j = 0; j < n; ++j) std::swap(A[(size_t)k * n + j], A[(size_t)pv * n + j]); std::swap(bb[k], bb[pv]); }
#pragma omp parallel for schedule(static) if (n - k > 256)
      for (int i = k + 1; i < n; ++i) {
        double f = A[(size_t)i * n + k] / A[(size_t)k * n + k];
        if (f == 0.0) continue;
        for (int j = k; j < n; ++j) A[(size_t)i * n + j] -= f * A[(size_t)k * n + j];
        bb[i] -= f * bb[k];
      }
    }
    for (int i = n - 1; i >= 0; --i) {
      double acc = bb[i];
      for (int j = i + 1; j < n; ++j) acc -= A[(size_t)i * n + j] * lambda[j];
      lambda[i] = acc / A[(size_t)i * n + i];
    }
    info[0] = 1; info[1] = 1;
  } else {
    std::vector<double> diag(n, 0.0);
    for (int r = 0; r < n; ++r)
      for (long long k = s->rowptr[r]; k < s->rowptr[r + 1]; ++k) if (s->cols[k] == r) diag[r] = s->vals[k];
    bool conv = false;
    info[0] = gmres_left_jacobi(n, [&](const double* x, double* y) { spmv(s, x, y); }, diag, s->rhs.data(), rtol,
                                maxits, restart, lambda, &conv);
    info[1] = conv ? 1 : 0;
  }
  std::vector<double> t(n);
  spmv(s, lambda, t.data());
  double rn = 0, bn = 0;
  for (int i = 0; i < n; ++i) { rn += (s->rhs[i] - t[i]) * (s->rhs[i] - t[i]); bn += s->rhs[i] * s->rhs[i]; }
  relres[0] = bn > 0 ? std::sqrt(rn / bn) : std::sqrt(rn);
  ORC_END
}

extern "C" int orc_system_reconstruct(const orc_system* s, const double* lambda, double* u, double* q) {
  ORC_TRY
  const Ref& R = *s->R;
  const int nF = R.nF, nNf = R.nNf, lu = R.lenU(), lq = R.lenQ(), ll = R.lenL();
#pragma omp parallel for schedule(static)
  for (int e = 0; e < s->nEl; ++e) {
    std::vector<double> le(ll);
    for (int f = 0; f < nF; ++f) {
      int fc = s->c2f[(size_t)e * nF + f];
      for (int j = 0; j < nNf; ++j) le[f * nNf + j] = lambda[(size_t)fc * nNf + s->perm[((size_t)e * nF + f) * nNf + j]];
    }
    for (int i = 0; i < lu; ++i) {
      double acc = s->U0[(size_t)e * lu + i];
      for (int j = 0; j < ll; ++j) acc += s->U[((size_t)e * lu + i) * ll + j] * le[j];
      u[(size_t)e * lu + i] = acc;
    }
    for (int i = 0; i < lq; ++i) {
      double acc = s->Q0[(size_t)e * lq + i];
      for (int j = 0; j < ll; ++j) acc += s->Q[((size_t)e * lq + i) * ll + j] * le[j];
      q[(size_t)e * lq + i] = acc;
    }
  }
  ORC_END
}

extern "C" double orc_system_time_assembly(orc_system* s, int nEl_sample, int nthreads) {
  try {
    int ne = std::min(std::max(nEl_sample, 1), s->nEl);
    auto t0 = std::chrono::steady_clock::now();
    system_assemble(s, ne, nthreads > 0 ? nthreads : omp_get_max_threads());
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1.0;
  }
}

extern "C" int orc_dense_gmres(int n, const double* A, const double* b, double rtol, int maxits, int restart,
                               double* x, int* iters) {
  ORC_TRY
  std::vector<double> diag(n);
  for (int i = 0; i < n; ++i) diag[i] = A[(size_t)i * n + i];
  bool conv = false;
  auto ap = [&](const double* v, double* y) {
    for (int i = 0; i < n; ++i) { double s = 0; for (int j = 0; j < n; ++j) s += A[(size_t)i * n + j] * v[j]; y[i] = s; }
  };
  *iters = gmres_left_jacobi(n, ap, diag, b, rtol, maxits, restart, x, &conv);
  ORC_END
}
