// Reference simplex element: see refel.h.
#include "refel.h"

#include <cmath>
#include <stdexcept>
#include <string>

namespace hfx {

namespace {
#include "data/cubature_data.inc"

// Legendre P_n and its first two derivatives at x.
void legendre(int n, double x, double* p, double* dp) {
  double p0 = 1.0, p1 = x;
  if (n == 0) { *p = 1.0; *dp = 0.0; return; }
  for (int k = 2; k <= n; ++k) {
    double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = pk;
  }
  *p = p1;
  *dp = n * (x * p1 - p0) / (x * x - 1.0);  // valid away from |x| = 1
}

// Gauss-Legendre rule with n points by Newton on P_n.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.resize(n);
  w.resize(n);
  for (int k = 0; k < n; ++k) {
    double z = std::cos(M_PI * (k + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p, dp;
      legendre(n, z, &p, &dp);
      double dz = p / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    double p, dp;
    legendre(n, z, &p, &dp);
    x[k] = z;
    w[k] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

// First-order forward-mode dual number with N tangent directions.
template <int N>
struct Dual {
  double v;
  double d[N];
  Dual(double a = 0.0) : v(a) { for (int i = 0; i < N; ++i) d[i] = 0.0; }
};
template <int N> Dual<N> operator+(Dual<N> a, const Dual<N>& b) { a.v += b.v; for (int i = 0; i < N; ++i) a.d[i] += b.d[i]; return a; }
template <int N> Dual<N> operator-(Dual<N> a, const Dual<N>& b) { a.v -= b.v; for (int i = 0; i < N; ++i) a.d[i] -= b.d[i]; return a; }
template <int N> Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v * b.v);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N> Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.v / b.v);
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v);
  return r;
}
template <int N> Dual<N> operator*(double s, Dual<N> a) { a.v *= s; for (int i = 0; i < N; ++i) a.d[i] *= s; return a; }
template <int N> Dual<N> operator+(double s, Dual<N> a) { a.v += s; return a; }
template <int N> Dual<N> operator+(Dual<N> a, double s) { a.v += s; return a; }

// Jacobi polynomial P_n^{(alpha,0)}(x) by the three-term recurrence.
template <typename T>
T jacobi(int n, double alpha, const T& x) {
  T p0(1.0);
  if (n == 0) return p0;
  T p1 = 0.5 * (alpha + 2.0) * x + (0.5 * alpha);  // ((a+2)x + a)/2
  for (int k = 2; k <= n; ++k) {
    double a = alpha;
    double c = 2.0 * k + a;  // 2k + alpha + beta, beta = 0
    double a1 = 2.0 * k * (k + a) * (c - 2.0);
    double a2 = (c - 1.0) * a * a;
    double a3 = (c - 1.0) * c * (c - 2.0);
    double a4 = 2.0 * (k + a - 1.0) * (k - 1.0) * c;
    T pk = (1.0 / a1) * ((a2 + a3 * x) * p1 - a4 * p0);
    p0 = p1;
    p1 = pk;
  }
  return p1;
}

template <typename T>
T ipow(const T& x, int k) {
  T r(1.0);
  for (int i = 0; i < k; ++i) r = r * x;
  return r;
}

// Proriol-Koornwinder-Dubiner mode (unnormalised) with multi-index m at the
// reference point xi, in collapsed coordinates.  Any basis of P_p gives the
// same Lagrange functions; this one keeps the Vandermonde well conditioned.
template <typename T>
T pkd_mode(int dim, const int* m, const T* xi) {
  if (dim == 1) return jacobi(m[0], 0.0, xi[0]);
  if (dim == 2) {
    const T& r = xi[0];
    const T& s = xi[1];
    T a = (std::fabs(1.0 - s.v) > 1e-14) ? (2.0 * ((1.0 + r) / (1.0 + (-1.0) * s)) + (-1.0)) : T(-1.0);
    return jacobi(m[0], 0.0, a) * jacobi(m[1], 2.0 * m[0] + 1.0, s) * ipow(1.0 + (-1.0) * s, m[0]);
  }
  const T& r = xi[0];
  const T& s = xi[1];
  const T& t = xi[2];
  T st = s + t;
  T a = (std::fabs(st.v) > 1e-14) ? ((-2.0) * ((1.0 + r) / st) + (-1.0)) : T(-1.0);
  T b = (std::fabs(1.0 - t.v) > 1e-14) ? (2.0 * ((1.0 + s) / (1.0 + (-1.0) * t)) + (-1.0)) : T(-1.0);
  return jacobi(m[0], 0.0, a) * jacobi(m[1], 2.0 * m[0] + 1.0, b) * ipow(1.0 + (-1.0) * b, m[0]) *
         jacobi(m[2], 2.0 * (m[0] + m[1]) + 2.0, t) * ipow(1.0 + (-1.0) * t, m[0] + m[1]);
}

// Dense LU inverse with partial pivoting (small matrices only).
std::vector<double> invert(std::vector<double> a, int n) {
  std::vector<double> inv(n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(a[i * n + k]) > std::fabs(a[piv * n + k])) piv = i;
    if (a[piv * n + k] == 0.0) throw std::runtime_error("RefEl: singular Vandermonde");
    if (piv != k)
      for (int j = 0; j < n; ++j) { std::swap(a[k * n + j], a[piv * n + j]); std::swap(inv[k * n + j], inv[piv * n + j]); }
    double d = 1.0 / a[k * n + k];
    for (int j = 0; j < n; ++j) { a[k * n + j] *= d; inv[k * n + j] *= d; }
    for (int i = 0; i < n; ++i) {
      if (i == k) continue;
      double f = a[i * n + k];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) { a[i * n + j] -= f * a[k * n + j]; inv[i * n + j] -= f * inv[k * n + j]; }
    }
  }
  return inv;
}

// Principal vertices and sub-entity lists of the reference simplices
// (ReferenceElement.cpp:1011-1043).  Entities of q+1 vertices are q-dimensional.
const std::vector<std::vector<double>>& principal(int dim) {
  static const std::vector<std::vector<double>> p1 = {{-1.0}, {1.0}};
  static const std::vector<std::vector<double>> p2 = {{-1.0, -1.0}, {1.0, -1.0}, {-1.0, 1.0}};
  static const std::vector<std::vector<double>> p3 = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
  return dim == 1 ? p1 : (dim == 2 ? p2 : p3);
}
const std::vector<std::vector<int>>& entities(int dim) {
  static const std::vector<std::vector<int>> e1 = {{0}, {1}};
  static const std::vector<std::vector<int>> e2 = {{0, 1}, {1, 2}, {2, 0}};
  static const std::vector<std::vector<int>> e3 = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 1}, {3, 2},
                                                   {3, 1, 0}, {2, 1, 3}, {2, 3, 0}, {0, 1, 2}};
  return dim == 1 ? e1 : (dim == 2 ? e2 : e3);
}

// Affine map of the q-simplex reference point x (q coords) onto the entity
// `ent` of the dim-simplex: x -> T (x - v0_q) + P[ent[0]], T mapping the
// q-simplex vertices onto the entity's vertices.
std::vector<double> map_to_entity(int dim, int q, const std::vector<int>& ent, const double* x) {
  const auto& P = principal(dim);
  const auto& Pq = principal(q);
  // Pq vertices l+1 - v0 are 2 e_l, so inv(inter) = I/2.
  std::vector<double> out(dim);
  for (int d = 0; d < dim; ++d) {
    double v = P[ent[0]][d];
    for (int l = 0; l < q; ++l) v += 0.5 * (P[ent[l + 1]][d] - P[ent[0]][d]) * (x[l] - Pq[0][l]);
    out[d] = v;
  }
  return out;
}

bool same_node(const double* a, const double* b, int dim) {
  for (int d = 0; d < dim; ++d)
    if (std::fabs(a[d] - b[d]) > 1e-12) return false;
  return true;
}

// All multi-indices of length `len` over {lo..hi}, first index slowest
// (the order of generateCombinationsWithRepetition, ReferenceElement.cpp:580-611).
void combinations(int len, int lo, int hi, std::vector<std::vector<int>>& out) {
  std::vector<int> c(len, lo);
  if (hi < lo) return;
  while (true) {
    out.push_back(c);
    int k = len - 1;
    while (k >= 0 && c[k] == hi) { c[k] = lo; --k; }
    if (k < 0) break;
    ++c[k];
  }
}

}  // namespace

int simplex_num_nodes(int dim, int p) {
  if (dim == 0) return 1;
  int n = 1;
  for (int k = 1; k <= dim; ++k) n = n * (p + k) / k;  // C(p+dim, dim)
  return n;
}

std::vector<double> gll_nodes_ref_order(int p) {
  if (p == 0) return {0.0};
  std::vector<double> x = {-1.0, 1.0};
  // interior nodes: roots of P_p', Newton from Chebyshev-Lobatto guesses
  for (int k = p - 1; k >= 1; --k) {
    double z = std::cos(M_PI * k / p);
    for (int it = 0; it < 100; ++it) {
      double P, dP;
      legendre(p, z, &P, &dP);
      double d2P = (2.0 * z * dP - p * (p + 1.0) * P) / (1.0 - z * z);
      double dz = dP / d2P;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    x.push_back(z);
  }
  // cos(pi k/p) for k = p-1..1 runs ascending; symmetrise exactly
  int ni = p - 1;
  for (int i = 0; i < ni / 2; ++i) {
    double a = 0.5 * (x[2 + ni - 1 - i] - x[2 + i]);
    x[2 + i] = -a;
    x[2 + ni - 1 - i] = a;
  }
  if (ni % 2 == 1) x[2 + ni / 2] = 0.0;
  return x;
}

std::vector<double> simplex_nodes(int dim, int p) {
  if (dim == 1) return gll_nodes_ref_order(p);
  if (p == 0) return std::vector<double>(dim, 0.0);
  std::vector<std::vector<double>> nodes = principal(dim);
  for (int q = 1; q < dim; ++q) {
    std::vector<double> sub = simplex_nodes(q, p);
    int nsub = (int)sub.size() / q;
    for (const auto& ent : entities(dim)) {
      if ((int)ent.size() != q + 1) continue;
      for (int l = 0; l < nsub; ++l) nodes.push_back(map_to_entity(dim, q, ent, &sub[l * q]));
    }
  }
  if (p > 1) {
    // interior nodes of the warped Lobatto grid (ReferenceElement.cpp:1097-1144)
    std::vector<double> g = gll_nodes_ref_order(p);
    std::vector<double> L(p + 1);  // ascending GLL points
    L[0] = -1.0;
    for (int k = 1; k < p; ++k) L[k] = g[k + 1];
    L[p] = 1.0;
    std::vector<std::vector<int>> all, red;
    combinations(dim, 1, p - 1, all);
    for (auto& c : all) {
      int s = 0;
      for (int v : c) s += v;
      if (s < p) red.push_back(c);
    }
    auto coef = [&](const std::vector<int>& c) {
      int s = 0;
      for (int v : c) s += v;
      double v = 2.0 + dim * L[c[0]];
      for (int l = 1; l < dim; ++l) v -= L[c[l]];
      v -= L[p - s];
      return v / (dim + 1.0) - 1.0;
    };
    for (auto& c : red) {
      std::vector<double> x(dim);
      if (dim == 2) {
        x[0] = coef(c);
        x[1] = coef({c[1], c[0]});
      } else {
        x[0] = coef(c);
        x[1] = coef({c[1], c[0], c[2]});
        x[2] = coef({c[2], c[1], c[0]});
      }
      nodes.push_back(x);
    }
  }
  // drop later duplicates (ReferenceElement.cpp:1161-1183)
  std::vector<double> out;
  for (size_t i = 0; i < nodes.size(); ++i) {
    bool dup = false;
    for (size_t j = 0; j < i && !dup; ++j) dup = same_node(nodes[i].data(), nodes[j].data(), dim);
    if (!dup) out.insert(out.end(), nodes[i].begin(), nodes[i].end());
  }
  return out;
}

Cubature make_cubature(int dim, int order) {
  Cubature c;
  c.dim = dim;
  c.order = order;
  if (order < 0) throw std::runtime_error("Cubature: negative order");
  if (dim == 1) {
    int n = (order + 2) / 2;  // ceil((order+1)/2)
    if (n < 1) n = 1;
    std::vector<double> x, w;
    gauss_legendre(n, x, w);
    c.nip = n;
    c.pts = x;
    c.w = w;
    return c;
  }
  const int* o2n;
  const int* list;
  const int* off;
  const double* data;
  int nlist, maxo;
  if (dim == 2) {
    o2n = wv_tri_order_to_npts; list = wv_tri_npts_list; off = wv_tri_npts_offset; data = wv_tri_data;
    nlist = sizeof(wv_tri_npts_list) / sizeof(int); maxo = sizeof(wv_tri_order_to_npts) / sizeof(int) - 1;
  } else if (dim == 3) {
    o2n = wv_tet_order_to_npts; list = wv_tet_npts_list; off = wv_tet_npts_offset; data = wv_tet_data;
    nlist = sizeof(wv_tet_npts_list) / sizeof(int); maxo = sizeof(wv_tet_order_to_npts) / sizeof(int) - 1;
  } else {
    throw std::runtime_error("Cubature: unsupported dimension " + std::to_string(dim));
  }
  if (order > maxo)
    throw std::runtime_error("Cubature: order " + std::to_string(order) + " not available in dimension " + std::to_string(dim));
  int n = o2n[order];
  int slot = -1;
  for (int i = 0; i < nlist; ++i)
    if (list[i] == n) slot = i;
  c.nip = n;
  c.pts.resize(n * dim);
  c.w.resize(n);
  for (int i = 0; i < n; ++i) {
    const double* r = data + (size_t)(off[slot] + i) * (dim + 1);
    for (int d = 0; d < dim; ++d) c.pts[i * dim + d] = r[d];
    c.w[i] = r[dim];
  }
  return c;
}

RefEl::RefEl(int dim_, int order_) : dim(dim_), order(order_) {
  if (dim < 1 || dim > 3) throw std::runtime_error("RefEl: dimension must be 1, 2 or 3");
  if (order < 1 || order > (dim == 3 ? 5 : 10))
    throw std::runtime_error("RefEl: order " + std::to_string(order) + " not supported in dimension " + std::to_string(dim));
  nodes = simplex_nodes(dim, order);
  nN = (int)nodes.size() / dim;
  if (nN != simplex_num_nodes(dim, order)) throw std::runtime_error("RefEl: node count mismatch");
  nF = dim + 1;
  cub = make_cubature(dim, 2 * order);
  if (dim > 1) face.reset(new RefEl(dim - 1, order));
  // face-node maps (ReferenceElement.cpp:136-202)
  faceNodes.assign(nF, {});
  if (dim == 1) {
    faceNodes[0] = {0};
    faceNodes[1] = {1};
  } else {
    int fi = 0;
    for (const auto& ent : entities(dim)) {
      if ((int)ent.size() != dim) continue;
      for (int l = 0; l < face->nN; ++l) {
        std::vector<double> x = map_to_entity(dim, dim - 1, ent, &face->nodes[l * (dim - 1)]);
        int idx = -1;
        for (int k = 0; k < nN && idx < 0; ++k)
          if (same_node(&nodes[k * dim], x.data(), dim)) idx = k;
        if (idx < 0) throw std::runtime_error("RefEl: face node not found");
        faceNodes[fi].push_back(idx);
      }
      ++fi;
    }
  }
  // modal basis and inverse Vandermonde
  std::vector<std::vector<int>> all;
  combinations(dim, 0, order, all);
  for (auto& m : all) {
    int s = 0;
    for (int v : m) s += v;
    if (s <= order) modes_.insert(modes_.end(), m.begin(), m.end());
  }
  std::vector<double> V(nN * nN);
  for (int n = 0; n < nN; ++n)
    for (int m = 0; m < nN; ++m) {
      Dual<1> xi[3];
      for (int d = 0; d < dim; ++d) xi[d] = Dual<1>(nodes[n * dim + d]);
      V[n * nN + m] = pkd_mode(dim, &modes_[m * dim], xi).v;
    }
  vinv_ = invert(V, nN);
  int nip = cub.nip;
  phi.resize(nip * nN);
  dphi.resize(nip * nN * dim);
  for (int q = 0; q < nip; ++q) eval(&cub.pts[q * dim], &phi[q * nN], &dphi[q * nN * dim]);
}

void RefEl::eval(const double* xi, double* phi_out, double* dphi_out) const {
  std::vector<double> md(nN), mg(nN * dim);
  for (int m = 0; m < nN; ++m) {
    Dual<3> x[3];
    for (int d = 0; d < dim; ++d) {
      x[d] = Dual<3>(xi[d]);
      x[d].d[d] = 1.0;
    }
    Dual<3> v = pkd_mode(dim, &modes_[m * dim], x);
    md[m] = v.v;
    for (int d = 0; d < dim; ++d) mg[m * dim + d] = v.d[d];
  }
  // phi_n(x) = sum_m mode_m(x) Vinv[m][n]
  for (int n = 0; n < nN; ++n) {
    double s = 0.0;
    for (int m = 0; m < nN; ++m) s += md[m] * vinv_[m * nN + n];
    phi_out[n] = s;
    if (dphi_out)
      for (int d = 0; d < dim; ++d) {
        double g = 0.0;
        for (int m = 0; m < nN; ++m) g += mg[m * dim + d] * vinv_[m * nN + n];
        dphi_out[n * dim + d] = g;
      }
  }
}

}  // namespace hfx
