// Simplex mesh topology and synthetic structured meshes: see mesh.h.
#include "mesh.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>

namespace hfx {

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double splitmix_uniform(uint64_t& s) { return (splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

namespace {

struct VecHash {
  size_t operator()(const std::vector<int>& v) const {
    size_t h = 1469598103934665603ull;
    for (int x : v) h = (h ^ (size_t)(unsigned)x) * 1099511628211ull;
    return h;
  }
};

}  // namespace

void build_topology(const RefEl& re, Mesh& m) {
  const int d = re.dim;
  m.dim = d;
  m.order = re.order;
  m.nN = re.nN;
  m.nNf = re.nNf();
  m.nFpc = re.nF;
  const int nF = re.nF, nNf = m.nNf;
  std::unordered_map<std::vector<int>, int, VecHash> key2face;
  key2face.reserve((size_t)m.nCells * nF);
  m.cell2face.assign((size_t)m.nCells * nF, -1);
  m.face2cell.clear();
  m.faces.clear();
  std::vector<int> key(d);
  // faces numbered by first appearance over cells in order, so the first
  // cell of every face is its lowest-numbered one
  for (int e = 0; e < m.nCells; ++e) {
    const int* c = &m.cells[(size_t)e * m.nN];
    for (int lf = 0; lf < nF; ++lf) {
      for (int k = 0; k < d; ++k) key[k] = c[re.faceNodes[lf][k]];  // face vertices come first
      std::sort(key.begin(), key.end());
      auto it = key2face.find(key);
      if (it == key2face.end()) {
        int f = (int)key2face.size();
        key2face.emplace(key, f);
        m.face2cell.push_back(e);
        m.face2cell.push_back(-1);
        for (int j = 0; j < nNf; ++j) m.faces.push_back(c[re.faceNodes[lf][j]]);
        m.cell2face[(size_t)e * nF + lf] = f;
      } else {
        int f = it->second;
        if (m.face2cell[2 * f + 1] != -1)
          throw std::runtime_error("mesh: face shared by more than two cells (cell " + std::to_string(e) + ")");
        m.face2cell[2 * f + 1] = e;
        m.cell2face[(size_t)e * nF + lf] = f;
      }
    }
  }
  m.nFaces = (int)key2face.size();
  m.boundary.clear();
  for (int f = 0; f < m.nFaces; ++f)
    if (m.face2cell[2 * f + 1] == -1) m.boundary.push_back(f);
  m.face_perm.assign((size_t)m.nCells * nF * nNf, 0);
  for (int e = 0; e < m.nCells; ++e) {
    const int* c = &m.cells[(size_t)e * m.nN];
    for (int lf = 0; lf < nF; ++lf) {
      int f = m.cell2face[(size_t)e * nF + lf];
      const int* fn = &m.faces[(size_t)f * nNf];
      for (int j = 0; j < nNf; ++j) {
        int node = c[re.faceNodes[lf][j]];
        int pos = -1;
        for (int k = 0; k < nNf; ++k)
          if (fn[k] == node) pos = k;
        if (pos < 0) throw std::runtime_error("mesh: cell node not found in face (cell " + std::to_string(e) + ")");
        m.face_perm[((size_t)e * nF + lf) * nNf + j] = (int8_t)pos;
      }
    }
  }
}

Mesh mesh_from_cells(const RefEl& re, int nNodes, const double* coords, int nCells, const int* cells, bool topology) {
  Mesh m;
  m.dim = re.dim;
  m.nN = re.nN;
  m.nNodes = nNodes;
  m.nCells = nCells;
  m.coords.assign(coords, coords + (size_t)nNodes * re.dim);
  m.cells.assign(cells, cells + (size_t)nCells * re.nN);
  for (int v : m.cells)
    if (v < 0 || v >= nNodes) throw std::runtime_error("mesh: cell node index out of range");
  if (topology) build_topology(re, m);
  return m;
}

Mesh structured_mesh(const RefEl& re, int n, double jit, uint64_t seed, bool topology) {
  const int d = re.dim;
  if (d != 2 && d != 3) throw std::runtime_error("structured_mesh: dimension must be 2 or 3");
  if (n < 1) throw std::runtime_error("structured_mesh: need n >= 1");
  const double h = 1.0 / n;
  const int nv1 = n + 1;
  const int nVerts = d == 2 ? nv1 * nv1 : nv1 * nv1 * nv1;
  std::vector<double> vx((size_t)nVerts * d);
  uint64_t st = seed;
  for (int v = 0; v < nVerts; ++v) {
    int ijk[3] = {v % nv1, (v / nv1) % nv1, d == 3 ? v / (nv1 * nv1) : 0};
    bool interior = true;
    for (int k = 0; k < d; ++k)
      if (ijk[k] == 0 || ijk[k] == n) interior = false;
    for (int k = 0; k < d; ++k) {
      double x = ijk[k] * h;
      double r = splitmix_uniform(st);  // one draw per vertex coordinate, always
      if (interior) x += (2.0 * r - 1.0) * jit * h;
      vx[(size_t)v * d + k] = x;
    }
  }
  // vertex simplices
  std::vector<std::array<int, 4>> simp;
  if (d == 2) {
    simp.reserve((size_t)2 * n * n);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        int v00 = j * nv1 + i, v10 = v00 + 1, v01 = v00 + nv1, v11 = v01 + 1;
        simp.push_back({v00, v10, v11, -1});
        simp.push_back({v00, v11, v01, -1});
      }
  } else {
    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    simp.reserve((size_t)6 * n * n * n);
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          for (auto& p : perms) {
            int pos[3] = {i, j, k};
            std::array<int, 4> t;
            t[0] = (pos[2] * nv1 + pos[1]) * nv1 + pos[0];
            for (int s = 0; s < 3; ++s) {
              pos[p[s]] += 1;
              t[s + 1] = (pos[2] * nv1 + pos[1]) * nv1 + pos[0];
            }
            simp.push_back(t);
          }
        }
  }
  // orient every simplex positively (det of the edge matrix > 0)
  for (auto& t : simp) {
    double e[3][3];
    for (int a = 0; a < d; ++a)
      for (int b = 0; b < d; ++b) e[a][b] = vx[(size_t)t[a + 1] * d + b] - vx[(size_t)t[0] * d + b];
    double det = d == 2 ? e[0][0] * e[1][1] - e[0][1] * e[1][0]
                        : e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) - e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
                              e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
    if (det < 0) std::swap(t[1], t[2]);
  }
  // high-order nodes: vertices keep their ids; other reference nodes are
  // keyed by the (vertex id, barycentric weight) pairs of their support
  Mesh m;
  m.dim = d;
  m.nN = re.nN;
  m.nCells = (int)simp.size();
  m.cells.resize((size_t)m.nCells * re.nN);
  std::vector<double> coords(vx);
  std::map<std::vector<long long>, int> key2node;
  std::vector<std::vector<double>> bary(re.nN, std::vector<double>(d + 1));
  for (int a = 0; a < re.nN; ++a) {
    double s = 0.0;
    for (int k = 0; k < d; ++k) {
      bary[a][k + 1] = 0.5 * (re.nodes[a * d + k] + 1.0);
      s += bary[a][k + 1];
    }
    bary[a][0] = 1.0 - s;
  }
  std::vector<std::pair<long long, long long>> sup;
  for (int e = 0; e < m.nCells; ++e) {
    const auto& t = simp[e];
    int* c = &m.cells[(size_t)e * re.nN];
    for (int a = 0; a < re.nN; ++a) {
      if (a <= d) {
        c[a] = t[a];
        continue;
      }
      sup.clear();
      for (int v = 0; v <= d; ++v)
        if (bary[a][v] > 1e-12) sup.push_back({t[v], std::llround(bary[a][v] * 1099511627776.0)});
      std::sort(sup.begin(), sup.end());
      std::vector<long long> key;
      for (auto& pr : sup) {
        key.push_back(pr.first);
        key.push_back(pr.second);
      }
      auto it = key2node.find(key);
      if (it != key2node.end()) {
        c[a] = it->second;
        continue;
      }
      int id = (int)(coords.size() / d);
      key2node.emplace(key, id);
      // x = X0 + sum_k lambda_{k+1} (X_{k+1} - X0)
      for (int b = 0; b < d; ++b) {
        double x = vx[(size_t)t[0] * d + b];
        for (int k = 0; k < d; ++k) x += bary[a][k + 1] * (vx[(size_t)t[k + 1] * d + b] - vx[(size_t)t[0] * d + b]);
        coords.push_back(x);
      }
      c[a] = id;
    }
  }
  m.coords.swap(coords);
  m.nNodes = (int)(m.coords.size() / d);
  if (topology) build_topology(re, m);
  return m;
}

}  // namespace hfx
