// C ABI of libhfx_hip.so (declared in include/hfx_hip.h).
#include "../../include/hfx_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "hdg_element.h"
#include "hfx_internal.h"
#include "krylov.h"
#include "mesh.h"
#include "refel.h"
#include "topology.h"
#include "trace_system.h"
#include "capi_internal.h"

namespace {
thread_local std::string g_err;
}

namespace hfx {
// Shared with the linear-algebra entry points (capi_la.cpp).
void set_last_error(const std::string& msg) { g_err = msg; }
}  // namespace hfx

#define HFX_TRY try {
#define HFX_END                       \
  }                                   \
  catch (const std::exception& e) {   \
    g_err = e.what();                 \
    return -1;                        \
  }                                   \
  return 0;

namespace hfx {
// Reference tables of the general element kernels (ESz's packed layout);
// shared with the Burgers entry points (capi_burgers.cpp).
void pack_tables(const hfx::RefEl& re, std::vector<double>& t) {
  const int d = re.dim;
  const hfx::RefEl& fe = *re.face;
  t.clear();
  t.insert(t.end(), re.phi.begin(), re.phi.end());
  t.insert(t.end(), re.dphi.begin(), re.dphi.end());
  t.insert(t.end(), fe.phi.begin(), fe.phi.end());
  t.insert(t.end(), fe.dphi.begin(), fe.dphi.end());
  t.insert(t.end(), re.cub.w.begin(), re.cub.w.end());
  t.insert(t.end(), fe.cub.w.begin(), fe.cub.w.end());
  (void)d;
}
}  // namespace hfx

namespace {
using hfx::pack_tables;

// Reference tables of the affine element kernel (order of ALay in
// hdg_affine.hip): PHI, DPHI, FPHI, W, FW, Ahat_a = Mhat^-1 Chat_a,
// G_g = Mhat^-1[:, fn_g] Mfhat.  Built in long double.
void pack_affine_tables(const hfx::RefEl& re, std::vector<double>& t) {
  const int d = re.dim, nN = re.nN, nip = re.nip();
  const hfx::RefEl& fe = *re.face;
  const int nNf = fe.nN, nipf = fe.nip(), nF = re.nF;
  typedef long double ld;
  std::vector<ld> M(nN * nN, 0), Mi(nN * nN, 0);
  for (int q = 0; q < nip; ++q)
    for (int i = 0; i < nN; ++i)
      for (int j = 0; j < nN; ++j) M[i * nN + j] += (ld)re.cub.w[q] * re.phi[q * nN + i] * re.phi[q * nN + j];
  for (int i = 0; i < nN; ++i) Mi[i * nN + i] = 1;
  for (int k = 0; k < nN; ++k) {  // SPD: Gauss-Jordan without pivoting
    ld inv = 1.0L / M[k * nN + k];
    for (int j = 0; j < nN; ++j) { M[k * nN + j] *= inv; Mi[k * nN + j] *= inv; }
    for (int r = 0; r < nN; ++r) {
      if (r == k) continue;
      ld f = M[r * nN + k];
      for (int j = 0; j < nN; ++j) { M[r * nN + j] -= f * M[k * nN + j]; Mi[r * nN + j] -= f * Mi[k * nN + j]; }
    }
  }
  t.clear();
  t.insert(t.end(), re.phi.begin(), re.phi.end());
  t.insert(t.end(), re.dphi.begin(), re.dphi.end());
  t.insert(t.end(), fe.phi.begin(), fe.phi.end());
  t.insert(t.end(), re.cub.w.begin(), re.cub.w.end());
  t.insert(t.end(), fe.cub.w.begin(), fe.cub.w.end());
  for (int a = 0; a < d; ++a) {
    std::vector<ld> C(nN * nN, 0);
    for (int q = 0; q < nip; ++q)
      for (int i = 0; i < nN; ++i)
        for (int j = 0; j < nN; ++j)
          C[i * nN + j] += (ld)re.cub.w[q] * re.dphi[(q * nN + i) * d + a] * re.phi[q * nN + j];
    for (int i = 0; i < nN; ++i)
      for (int j = 0; j < nN; ++j) {
        ld s = 0;
        for (int l = 0; l < nN; ++l) s += Mi[i * nN + l] * C[l * nN + j];
        t.push_back((double)s);
      }
  }
  std::vector<ld> Mf(nNf * nNf, 0);
  for (int q = 0; q < nipf; ++q)
    for (int i = 0; i < nNf; ++i)
      for (int j = 0; j < nNf; ++j) Mf[i * nNf + j] += (ld)fe.cub.w[q] * fe.phi[q * nNf + i] * fe.phi[q * nNf + j];
  for (int g = 0; g < nF; ++g)
    for (int j = 0; j < nN; ++j)
      for (int b = 0; b < nNf; ++b) {
        ld s = 0;
        for (int a2 = 0; a2 < nNf; ++a2) s += Mi[j * nN + re.faceNodes[g][a2]] * Mf[a2 * nNf + b];
        t.push_back((double)s);
      }
}

// True when every cell's nodes are the affine image of its vertices (so
// the Jacobian is constant per cell and faces are flat).
bool mesh_is_affine(const hfx::RefEl& re, const hfx::Mesh& m) {
  const int d = re.dim;
  for (int e = 0; e < m.nCells; ++e) {
    const int* c = &m.cells[(size_t)e * re.nN];
    double h = 0.0;
    for (int v = 1; v <= d; ++v)
      for (int k = 0; k < d; ++k) h = std::max(h, std::fabs(m.coords[(size_t)c[v] * d + k] - m.coords[(size_t)c[0] * d + k]));
    for (int i = d + 1; i < re.nN; ++i)
      for (int k = 0; k < d; ++k) {
        double x = m.coords[(size_t)c[0] * d + k];
        for (int a = 0; a < d; ++a)
          x += 0.5 * (re.nodes[i * d + a] + 1.0) * (m.coords[(size_t)c[a + 1] * d + k] - m.coords[(size_t)c[0] * d + k]);
        if (std::fabs(x - m.coords[(size_t)c[i] * d + k]) > 1e-10 * h) return false;
      }
  }
  return true;
}

void require(bool c, const char* msg) {
  if (!c) throw std::runtime_error(msg);
}

}  // namespace

extern "C" const char* hfx_last_error(void) { return g_err.c_str(); }

extern "C" int hfx_device_count(int* n) {
  HFX_TRY
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  HFX_END
}

extern "C" int hfx_refel_sizes(int dim, int order, int* out) {
  HFX_TRY
  hfx::RefEl re(dim, order);
  require(dim >= 2, "hfx_refel_sizes: dimension must be 2 or 3");
  out[0] = re.nN;
  out[1] = re.nF;
  out[2] = re.nNf();
  out[3] = re.nip();
  out[4] = re.face->nip();
  out[5] = re.nN * (dim + 1) + re.nF * re.nNf();
  out[6] = re.nF * re.nNf();
  HFX_END
}

extern "C" int hfx_refel_tables(int dim, int order, double* nodes, int* face_nodes, double* ipx, double* ipw,
                                double* phi, double* dphi, double* fipx, double* fipw, double* fphi) {
  HFX_TRY
  require(dim >= 2, "hfx_refel_tables: dimension must be 2 or 3");
  hfx::RefEl re(dim, order);
  if (nodes) std::copy(re.nodes.begin(), re.nodes.end(), nodes);
  if (face_nodes)
    for (int f = 0; f < re.nF; ++f)
      for (int j = 0; j < re.nNf(); ++j) face_nodes[f * re.nNf() + j] = re.faceNodes[f][j];
  if (ipx) std::copy(re.cub.pts.begin(), re.cub.pts.end(), ipx);
  if (ipw) std::copy(re.cub.w.begin(), re.cub.w.end(), ipw);
  if (phi) std::copy(re.phi.begin(), re.phi.end(), phi);
  if (dphi) std::copy(re.dphi.begin(), re.dphi.end(), dphi);
  if (fipx) std::copy(re.face->cub.pts.begin(), re.face->cub.pts.end(), fipx);
  if (fipw) std::copy(re.face->cub.w.begin(), re.face->cub.w.end(), fipw);
  if (fphi) std::copy(re.face->phi.begin(), re.face->phi.end(), fphi);
  HFX_END
}

extern "C" int hfx_cubature(int dim, int order, double* pts, double* w) {
  try {
    hfx::Cubature c = hfx::make_cubature(dim, order);
    if (pts) std::copy(c.pts.begin(), c.pts.end(), pts);
    if (w) std::copy(c.w.begin(), c.w.end(), w);
    return c.nip;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

extern "C" int hfx_mesh_structured(int dim, int order, int n, double jitter, unsigned long long seed, hfx_mesh** out) {
  HFX_TRY
  auto m = std::unique_ptr<hfx_mesh>(new hfx_mesh());
  m->re.reset(new hfx::RefEl(dim, order));
  m->m = hfx::structured_mesh(*m->re, n, jitter, seed);
  *out = m.release();
  HFX_END
}

extern "C" int hfx_mesh_from_cells(int dim, int order, int nNodes, const double* coords, int nCells, const int* cells,
                                   hfx_mesh** out) {
  HFX_TRY
  auto m = std::unique_ptr<hfx_mesh>(new hfx_mesh());
  m->re.reset(new hfx::RefEl(dim, order));
  m->m = hfx::mesh_from_cells(*m->re, nNodes, coords, nCells, cells);
  *out = m.release();
  HFX_END
}

extern "C" int hfx_mesh_structured_gpu(int device, int dim, int order, int n, double jitter, unsigned long long seed,
                                       hfx_mesh** out) {
  HFX_TRY
  auto m = std::unique_ptr<hfx_mesh>(new hfx_mesh());
  m->re.reset(new hfx::RefEl(dim, order));
  m->m = hfx::structured_mesh(*m->re, n, jitter, seed, false);
  hfx::build_topology_gpu(*m->re, m->m, device);
  *out = m.release();
  HFX_END
}

extern "C" int hfx_mesh_from_cells_gpu(int device, int dim, int order, int nNodes, const double* coords, int nCells,
                                       const int* cells, hfx_mesh** out) {
  HFX_TRY
  auto m = std::unique_ptr<hfx_mesh>(new hfx_mesh());
  m->re.reset(new hfx::RefEl(dim, order));
  m->m = hfx::mesh_from_cells(*m->re, nNodes, coords, nCells, cells, false);
  hfx::build_topology_gpu(*m->re, m->m, device);
  *out = m.release();
  HFX_END
}

extern "C" int hfx_mesh_sizes(const hfx_mesh* m, long long* out) {
  HFX_TRY
  require(m, "hfx_mesh_sizes: null mesh");
  out[0] = m->m.nNodes;
  out[1] = m->m.nCells;
  out[2] = m->m.nFaces;
  out[3] = (long long)m->m.boundary.size();
  out[4] = (long long)m->m.nFaces * m->m.nNf;
  HFX_END
}

extern "C" int hfx_mesh_get(const hfx_mesh* mm, double* coords, int* cells, int* faces, int* cell2face,
                            int* face2cell, int* boundary, signed char* face_perm) {
  HFX_TRY
  require(mm, "hfx_mesh_get: null mesh");
  const hfx::Mesh& m = mm->m;
  if (coords) std::copy(m.coords.begin(), m.coords.end(), coords);
  if (cells) std::copy(m.cells.begin(), m.cells.end(), cells);
  if (faces) std::copy(m.faces.begin(), m.faces.end(), faces);
  if (cell2face) std::copy(m.cell2face.begin(), m.cell2face.end(), cell2face);
  if (face2cell) std::copy(m.face2cell.begin(), m.face2cell.end(), face2cell);
  if (boundary) std::copy(m.boundary.begin(), m.boundary.end(), boundary);
  if (face_perm) std::memcpy(face_perm, m.face_perm.data(), m.face_perm.size());
  HFX_END
}

extern "C" void hfx_mesh_free(hfx_mesh* m) { delete m; }

extern "C" int hfx_hdg_create(int device, int dim, int order, hfx_hdg** out) {
  HFX_TRY
  int n = 0;
  HFX_HIP(hipGetDeviceCount(&n));
  require(device >= 0 && device < n, "hfx_hdg_create: no such GPU");
  auto h = std::unique_ptr<hfx_hdg>(new hfx_hdg());
  h->device = device;
  h->dim = dim;
  h->order = order;
  h->re.reset(new hfx::RefEl(dim, order));
  (void)hfx::element_lds_bytes(dim, order);  // throws for unsupported (dim, order)
  h->set_device();
  HFX_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  std::vector<double> t;
  pack_tables(*h->re, t);
  h->tabs.upload(t, h->stream);
  pack_affine_tables(*h->re, t);
  h->tabs_aff.upload(t, h->stream);
  std::vector<int> fnodes;
  for (auto& f : h->re->faceNodes) fnodes.insert(fnodes.end(), f.begin(), f.end());
  h->faceNodes.upload(fnodes, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  *out = h.release();
  HFX_END
}

extern "C" void hfx_hdg_destroy(hfx_hdg* h) {
  if (!h) return;
  try {
    h->set_device();
    if (h->stream) (void)hipStreamSynchronize(h->stream);
  } catch (...) {
  }
  hipStream_t s = h->stream;
  delete h;
  if (s) (void)hipStreamDestroy(s);
}

extern "C" int hfx_hdg_set_mesh(hfx_hdg* h, const hfx_mesh* mm) {
  HFX_TRY
  require(h && mm, "hfx_hdg_set_mesh: null handle");
  require(mm->m.dim == h->dim && mm->re->order == h->order, "hfx_hdg_set_mesh: mesh dimension/order mismatch");
  h->set_device();
  static std::atomic<unsigned long long> serial{0};
  h->mesh_serial = ++serial;
  h->lpc_state.reset();
  h->ilu.reset();
  h->rk_started = false;  // Runge-Kutta fields are sized for the mesh
  h->rk_stage = 0;
  h->bc_kind.clear();     // boundary models are per face of the mesh
  h->n_bc_int = 0;
  h->n_bc_hdg = 0;
  h->mesh = mm->m;
  h->mesh.order = h->order;
  h->affine = mesh_is_affine(*h->re, h->mesh);
  const hfx::Mesh& m = h->mesh;
  hipStream_t s = h->stream;
  h->coords.upload(m.coords, s);
  h->cells.upload(m.cells, s);
  h->c2f.upload(m.cell2face, s);
  h->f2c.upload(m.face2cell, s);
  h->perm.upload(m.face_perm, s);
  std::vector<int8_t> inv(m.face_perm.size());
  for (size_t blk = 0; blk < m.face_perm.size() / m.nNf; ++blk)
    for (int j = 0; j < m.nNf; ++j) inv[blk * m.nNf + m.face_perm[blk * m.nNf + j]] = (int8_t)j;
  h->invperm.upload(inv, s);
  h->pat = hfx::build_bsr_pattern(m.nCells, m.nFaces, m.nFpc, m.nNf, m.cell2face.data(), m.face2cell.data());
  h->A.nb = h->pat.nb;
  h->A.bs = h->pat.bs;
  h->A.nnzb = (int)h->pat.col.size();
  h->A.rowptr.upload(h->pat.rowptr, s);
  h->A.col.upload(h->pat.col, s);
  h->A.rowf.upload(h->pat.rowf, s);
  h->A.diag.upload(h->pat.diag, s);
  h->A.contrib.upload(h->pat.contrib, s);
  h->A.face_cells.upload(h->pat.face_cells, s);
  h->bpos.upload(hfx::cell_block_positions(h->pat, m.nCells, m.nFpc, m.cell2face.data()), s);
  h->dside.alloc((size_t)2 * h->pat.nb * h->pat.bs * h->pat.bs);
  h->last_fused = false;
  h->A.vals.alloc((size_t)h->A.nnzb * m.nNf * m.nNf);
  std::vector<uint8_t> dir(m.nFaces, 0);
  for (int f : m.boundary) dir[f] = 1;
  h->is_dir.upload(dir, s);
  const int nN = m.nN, nL = m.nFpc * m.nNf, nQ = nN * h->dim;
  h->S.alloc((size_t)m.nCells * nL * nL);
  h->S0.alloc((size_t)m.nCells * nL);
  h->U.alloc((size_t)m.nCells * nN * nL);
  h->Q.alloc((size_t)m.nCells * nQ * nL);
  h->U0.alloc((size_t)m.nCells * nN);
  h->Q0.alloc((size_t)m.nCells * nQ);
  h->rhs.alloc((size_t)m.nFaces * m.nNf);
  h->lambda.alloc((size_t)m.nFaces * m.nNf);
  h->u.alloc((size_t)m.nCells * nN);
  HFX_HIP(hipMemsetAsync(h->u.p, 0, h->u.n * sizeof(double), s));  // "Solution" starts at 0
  h->q.alloc((size_t)m.nCells * nQ);
  HFX_HIP(hipMemsetAsync(h->q.p, 0, h->q.n * sizeof(double), s));  // "Flux" starts at 0
  h->g.alloc((size_t)m.nFaces * m.nNf);
  HFX_HIP(hipMemsetAsync(h->g.p, 0, h->g.n * sizeof(double), s));
  HFX_HIP(hipStreamSynchronize(s));
  h->has_mesh = true;
  h->has_fields = h->assembled = h->solved = false;
  ++h->args_version;
  HFX_END
}

extern "C" int hfx_hdg_set_model(hfx_hdg* h, int kind) {
  HFX_TRY
  require(h, "hfx_hdg_set_model: null handle");
  require(kind >= 0 && kind <= 2, "hfx_hdg_set_model: unknown model kind");
  // Reaction is an operator of HDGConvectionDiffusionReactionSource only
  // (HDGConvectionDiffusionReactionSource.cpp:90-95); the other models have none
  require(kind == HFX_MODEL_CDR_SOURCE || !h->hasR,
          "hfx_hdg_set_model: a reaction is set; only HDGConvectionDiffusionReactionSource has a Reaction operator");
  h->kind = kind;
  h->assembled = h->solved = false;
  ++h->args_version;
  HFX_END
}

extern "C" int hfx_hdg_set_fields(hfx_hdg* h, const double* tau, int tau_sides, const double* D, int ndD,
                                  const double* b, const double* dirichlet) {
  HFX_TRY
  require(h && h->has_mesh, "hfx_hdg_set_fields: set the mesh first");
  require(tau != nullptr, "hfx_hdg_set_fields: the field map must have a Tau field");
  require(tau_sides == 1 || tau_sides == 2, "hfx_hdg_set_fields: the Tau field must be single or double valued");
  require(D == nullptr || ndD == 1 || ndD == h->dim * h->dim,
          "hfx_hdg_set_fields: DiffusionTensor must be scalar or dim x dim per node");
  h->set_device();
  const hfx::Mesh& m = h->mesh;
  hipStream_t s = h->stream;
  h->tau_sides = tau_sides;
  h->tau.upload(tau, (size_t)m.nFaces * m.nNf * tau_sides, s);
  h->hasD = D != nullptr;
  h->ndD = h->hasD ? ndD : 0;
  if (D) h->Dn.upload(D, (size_t)m.nNodes * ndD, s);
  h->hasB = b != nullptr;
  if (b) h->bn.upload(b, (size_t)m.nNodes * h->dim, s);
  if (dirichlet) h->g.upload(dirichlet, (size_t)m.nFaces * m.nNf, s);
  else HFX_HIP(hipMemsetAsync(h->g.p, 0, h->g.n * sizeof(double), s));
  HFX_HIP(hipStreamSynchronize(s));
  h->has_fields = true;
  h->assembled = h->solved = false;
  ++h->args_version;
  HFX_END
}

extern "C" int hfx_hdg_ip_coords(hfx_hdg* h, double* out) {
  HFX_TRY
  require(h && h->has_mesh, "hfx_hdg_ip_coords: set the mesh first");
  h->set_device();
  const size_t n = (size_t)h->mesh.nCells * h->re->nip() * h->dim;
  hfx::DevBuf<double> d;
  d.alloc(n);
  hfx::launch_ip_coords(h->dim, h->order, h->mesh.nCells, h->tabs.p, h->coords.p, h->cells.p, d.p, h->stream);
  d.download(out, n, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  HFX_END
}

extern "C" int hfx_hdg_set_source(hfx_hdg* h, const double* f_ip, const double* r_ip) {
  HFX_TRY
  require(h && h->has_mesh, "hfx_hdg_set_source: set the mesh first");
  require(r_ip == nullptr || h->kind == HFX_MODEL_CDR_SOURCE,
          "hfx_hdg_set_source: only HDGConvectionDiffusionReactionSource has a Reaction operator "
          "(set that model kind first)");
  h->set_device();
  const size_t n = (size_t)h->mesh.nCells * h->re->nip();
  h->hasF = f_ip != nullptr;
  h->hasR = r_ip != nullptr;
  if (f_ip) h->f_ip.upload(f_ip, n, h->stream);
  if (r_ip) h->r_ip.upload(r_ip, n, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  h->assembled = h->solved = false;
  ++h->args_version;
  HFX_END
}

extern "C" int hfx_hdg_set_launch(hfx_hdg* h, int threads, int grid_cap) {
  HFX_TRY
  require(h, "hfx_hdg_set_launch: null handle");
  require(threads == 0 || threads == 64 || threads == 128 || threads == 256,
          "hfx_hdg_set_launch: threads must be 0, 64, 128 or 256");
  require(grid_cap >= 0, "hfx_hdg_set_launch: grid cap must be >= 0");
  h->nt = threads;
  h->grid_cap = grid_cap;
  HFX_END
}

extern "C" int hfx_hdg_assemble(hfx_hdg* h) {
  HFX_TRY
  require(h && h->has_mesh && h->has_fields, "hfx_hdg_assemble: the solver must be initialized and allocated before assembling");
  h->set_device();
  h->run_element();
  h->run_scatter();
  HFX_HIP(hipStreamSynchronize(h->stream));
  h->assembled = true;
  h->solved = false;
  HFX_END
}

extern "C" int hfx_hdg_local_matrices(hfx_hdg* h, int first, int count, double* K, double* F) {
  HFX_TRY
  require(h && h->has_mesh && h->has_fields, "hfx_hdg_local_matrices: mesh and fields must be set");
  require(first >= 0 && count >= 0 && first + count <= h->mesh.nCells, "hfx_hdg_local_matrices: cell range out of bounds");
  if (count == 0) return 0;
  h->set_device();
  const int n = h->re->nN * (h->dim + 1) + h->re->nF * h->re->nNf();
  hfx::DevBuf<double> dK, dF;
  dK.alloc((size_t)count * n * n);
  dF.alloc((size_t)count * n);
  h->prepare_time_scheme();
  hfx::ElemArgs a = h->elem_args();
  a.el_offset = first;
  a.nEl = count;
  a.Kloc = dK.p;
  a.Floc = dF.p;
  hfx::launch_element(h->dim, h->order, a, h->threads(), h->grid_cap, h->stream);
  dK.download(K, (size_t)count * n * n, h->stream);
  dF.download(F, (size_t)count * n, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  HFX_END
}

extern "C" int hfx_hdg_get_condensed(hfx_hdg* h, double* S, double* S0, double* U, double* Q, double* U0, double* Q0) {
  HFX_TRY
  require(h && h->assembled, "hfx_hdg_get_condensed: assemble first");
  h->set_device();
  // the fused scatter does not store S: run the element kernel once more in
  // its stored-S form (same device code up to the stores)
  if (S && h->last_fused) h->run_element(true);
  if (S) h->S.download(S, h->S.n, h->stream);
  if (S0) h->S0.download(S0, h->S0.n, h->stream);
  if (U) h->U.download(U, h->U.n, h->stream);
  if (Q) h->Q.download(Q, h->Q.n, h->stream);
  if (U0) h->U0.download(U0, h->U0.n, h->stream);
  if (Q0) h->Q0.download(Q0, h->Q0.n, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  HFX_END
}

extern "C" int hfx_hdg_trace_sizes(hfx_hdg* h, long long* out) {
  HFX_TRY
  require(h && h->has_mesh, "hfx_hdg_trace_sizes: set the mesh first");
  out[0] = h->A.nrows();
  out[1] = (long long)h->A.nnzb * h->A.bs * h->A.bs;
  out[2] = h->A.nnzb;
  HFX_END
}

extern "C" int hfx_hdg_get_trace_csr(hfx_hdg* h, long long* rowptr, int* cols, double* vals, double* rhs) {
  HFX_TRY
  require(h && h->assembled, "hfx_hdg_get_trace_csr: assemble first");
  h->set_device();
  const int bs = h->A.bs;
  std::vector<double> bv((size_t)h->A.nnzb * bs * bs);
  h->A.vals.download(bv.data(), bv.size(), h->stream);
  if (rhs) h->rhs.download(rhs, h->rhs.n, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  const auto& P = h->pat;
  long long o = 0;
  for (int F = 0; F < P.nb; ++F)
    for (int r = 0; r < bs; ++r) {
      if (rowptr) rowptr[(long long)F * bs + r] = o;
      for (int k = P.rowptr[F]; k < P.rowptr[F + 1]; ++k)
        for (int c = 0; c < bs; ++c) {
          if (cols) cols[o] = P.col[k] * bs + c;
          if (vals) vals[o] = bv[((size_t)k * bs + r) * bs + c];
          ++o;
        }
    }
  if (rowptr) rowptr[(long long)P.nb * bs] = o;
  HFX_END
}

namespace {

// Set up preconditioner `pc` on the assembled trace matrix; returns its apply
// (empty for none).
std::function<void(const double*, double*)> make_pc(hfx_hdg* h, int pc) {
  const long long n = h->A.nrows();
  hipStream_t s = h->stream;
  std::function<void(const double*, double*)> M;
  const auto t0 = std::chrono::steady_clock::now();
  if (pc == HFX_PC_JACOBI) {
    h->dinv.alloc(n);
    hfx::bsr_inv_point_diag(h->A, h->dinv.p, s);
    M = [h, n, s](const double* in, double* out) { hfx::apply_point_diag(n, h->dinv.p, in, out, s); };
  } else if (pc == HFX_PC_BJACOBI) {
    h->dinv.alloc((size_t)h->A.nb * h->A.bs * h->A.bs);
    hfx::bsr_invert_diag(h->A, h->dinv.p, s);
    M = [h, s](const double* in, double* out) { hfx::apply_block_diag(h->A.nb, h->A.bs, h->dinv.p, in, out, s); };
  } else if (pc == HFX_PC_LEARNED) {
    hfx::lpc_prepare(h);
    const hfx::LpcInfer* st = h->lpc_state.get();
    M = [st, s](const double* in, double* out) { hfx::bsr_spmv_f32vals(st->g, st->mvals.p, in, out, s); };
  } else if (pc == HFX_PC_ILU0 || pc == HFX_PC_ILU0_NATURAL || pc == HFX_PC_ILU0_F32) {
    const int ord = pc == HFX_PC_ILU0_NATURAL ? hfx::ILU_NATURAL : hfx::ILU_MULTICOLOUR;
    if (!h->ilu || h->ilu->ordering != ord) {
      h->ilu.reset(new hfx::Ilu());
      h->ilu->setup(h->A.nb, h->A.bs, h->pat.rowptr.data(), h->pat.col.data(), ord, s);
    }
    h->ilu->low = pc == HFX_PC_ILU0_F32;
    h->ilu->factor(h->A, s);
    hfx::Ilu* f = h->ilu.get();
    M = [f, s](const double* in, double* out) { f->apply(in, out, s); };
  } else if (pc != HFX_PC_NONE) {
    throw std::runtime_error("hfx_hdg_solve: unknown preconditioner");
  }
  if (pc != HFX_PC_LEARNED) {  // the learned set-up times itself
    HFX_HIP(hipStreamSynchronize(s));
    h->last_pc_setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return M;
}

void do_solve(hfx_hdg* h, int pc, double rtol, int maxits, int restart, hfx::KrylovInfo& info) {
  const long long n = h->A.nrows();
  hipStream_t s = h->stream;
  if (h->hdg_type == HFX_HDG_SEXPLICIT) {
    // face-local direct solves (HDGSolver.cpp:709-729): the trace matrix is
    // block diagonal, lambda_F = S_FF^-1 S0_F
    h->dinv.alloc((size_t)h->A.nb * h->A.bs * h->A.bs);
    hfx::bsr_invert_diag(h->A, h->dinv.p, s);
    hfx::apply_block_diag(h->A.nb, h->A.bs, h->dinv.p, h->rhs.p, h->lambda.p, s);
    hfx::Reducer red;
    red.init(n);
    hfx::DevBuf<double> r;
    r.alloc(n);
    hfx::bsr_spmv(h->A, h->lambda.p, r.p, h->rhs.p, s);
    const double bn = red.norm(n, h->rhs.p, s), rn = red.norm(n, r.p, s);
    info = hfx::KrylovInfo();
    info.iters = 1;
    info.true_relres = bn > 0.0 ? rn / bn : rn;
    info.reason = hfx::CONVERGED_RTOL;
    return;
  }
  std::function<void(const double*, double*)> M = make_pc(h, pc);
  HFX_HIP(hipMemsetAsync(h->lambda.p, 0, n * sizeof(double), s));
  hfx::KrylovOpts o;
  o.rtol = rtol;
  o.maxits = maxits;
  o.restart = restart;
  o.graph = h->solver_graphs;
  if (h->krylov == HFX_KRYLOV_BICGSTAB) hfx::bicgstab(h->A, h->rhs.p, h->lambda.p, M, o, info, s);
  else hfx::fgmres(h->A, h->rhs.p, h->lambda.p, M, o, info, s);
}

}  // namespace

extern "C" int hfx_hdg_set_time_scheme(hfx_hdg* h, int scheme, double dt) {
  HFX_TRY
  require(h != nullptr, "hfx_hdg_set_time_scheme: null handle");
  require(scheme == HFX_TS_NONE || scheme == HFX_TS_BEULER || scheme == HFX_TS_FEULER || scheme == HFX_TS_RK,
          "hfx_hdg_set_time_scheme: unknown scheme");
  require(scheme == HFX_TS_NONE || (dt > 0.0 && std::isfinite(dt)),
          "hfx_hdg_set_time_scheme: the time step needs to be set and it should not be 0");
  h->ts_scheme = scheme;
  h->ts_dt = scheme == HFX_TS_NONE ? 0.0 : dt;
  h->assembled = h->solved = false;
  ++h->args_version;
  HFX_END
}

namespace {

// Butcher tables of RungeKutta::setUpDB (RungeKutta.cpp:197-285) in RKType
// order: row i < n-1 is [c_i | a_i0 ..], the last row [0 | b_0 ..].
std::vector<double> rk_db(int type, int* n) {
  const double g = 1.0 - std::sqrt(2.0) / 2.0;
  switch (type) {
    case HFX_RK_FEULER: *n = 2; return {0, 0, 0, 1};
    case HFX_RK_EMIDPOINT: *n = 3; return {0, 0, 0, 0.5, 0.5, 0, 0, 0, 1};
    case HFX_RK_HEUN: *n = 3; return {0, 0, 0, 1, 1, 0, 0, 0.5, 0.5};
    case HFX_RK_KUTTA3: *n = 4; return {0, 0, 0, 0, 0.5, 0.5, 0, 0, 1, -1, 2, 0, 0, 1.0 / 6, 2.0 / 3, 1.0 / 6};
    case HFX_RK_HEUN3: *n = 4; return {0, 0, 0, 0, 1.0 / 3, 1.0 / 3, 0, 0, 2.0 / 3, 0, 2.0 / 3, 0, 0, 0.25, 0, 0.75};
    case HFX_RK_SSPRK3: *n = 4; return {0, 0, 0, 0, 1, 1, 0, 0, 0.5, 0.25, 0.25, 0, 0, 1.0 / 6, 1.0 / 6, 2.0 / 3};
    case HFX_RK_RK4:
      *n = 5;
      return {0, 0, 0, 0, 0, 0.5, 0.5, 0, 0, 0, 0.5, 0, 0.5, 0, 0, 1, 0, 0, 1, 0, 0, 1.0 / 6, 1.0 / 3, 1.0 / 3, 1.0 / 6};
    case HFX_RK_BEULER: *n = 2; return {1, 1, 0, 1};
    case HFX_RK_IMIDPOINT: *n = 2; return {0.5, 0.5, 0, 1};
    case HFX_RK_CRANKNICOLSON: *n = 3; return {0, 0, 0, 1, 0.5, 0.5, 0, 0.5, 0.5};
    case HFX_RK_KS2: *n = 3; return {0.5, 0.5, 0, 1.5, -0.5, 2, 0, -0.5, 1.5};
    case HFX_RK_QZ2: *n = 3; return {0.25, 0.25, 0, 0.75, 0.5, 0.25, 0, 0.5, 0.5};
    case HFX_RK_ALX2: *n = 3; return {g, g, 0, 1, 1 - g, g, 0, 1 - g, g};
    case HFX_RK_RK43:
      *n = 5;
      return {0.5, 0.5, 0, 0, 0, 2.0 / 3, 1.0 / 6, 0.5, 0, 0, 0.5, -0.5, 0.5, 0.5, 0, 1, 1.5, -1.5, 0.5, 0.5, 0, 1.5,
              -1.5, 0.5, 0.5};
    default: throw std::runtime_error("RungeKutta: unknown Butcher table");
  }
}

hfx::RkCoef rk_row(const hfx_hdg* h, int r) {
  hfx::RkCoef c{};
  for (int j = 0; j < h->rk_stages(); ++j) c.c[j] = h->rk_entry(r, 1 + j);
  return c;
}

}  // namespace

extern "C" int hfx_rk_table(int type, double* table, int* n) {
  HFX_TRY
  require(n != nullptr, "hfx_rk_table: null size");
  const std::vector<double> t = rk_db(type, n);
  if (table) std::copy(t.begin(), t.end(), table);
  HFX_END
}

extern "C" int hfx_hdg_set_butcher(hfx_hdg* h, int type, const double* table, int n) {
  HFX_TRY
  require(h != nullptr, "hfx_hdg_set_butcher: null handle");
  std::vector<double> t;
  int nn = 0;
  if (type == HFX_RK_CUSTOM) {
    require(table != nullptr && n >= 2, "hfx_hdg_set_butcher: a custom table needs at least 2 x 2 values");
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j)
        require(table[(size_t)i * n + j] == 0.0,
                "RungeKutta: the upper triangular part of the Butcher table should be null (no fully implicit "
                "implementation as of yet)");
    t.assign(table, table + (size_t)n * n);
    nn = n;
  } else {
    t = rk_db(type, &nn);
  }
  require(nn - 1 <= hfx::RK_MAX_STAGES, "hfx_hdg_set_butcher: at most 8 stages");
  for (double v : t) require(std::isfinite(v), "hfx_hdg_set_butcher: table entries must be finite");
  h->rk_table = t;
  h->rk_n = nn;
  h->rk_stage = 0;
  h->rk_started = false;
  h->assembled = h->solved = false;
  ++h->args_version;
  HFX_END
}

extern "C" int hfx_hdg_rk_start_step(hfx_hdg* h) {
  HFX_TRY
  require(h && h->has_mesh, "hfx_hdg_rk_start_step: set the mesh first");
  require(h->rk_n >= 2, "hfx_hdg_rk_start_step: set a Butcher table first");
  h->set_device();
  const int nst = h->rk_stages();
  h->rk_old_u.alloc(h->u.n);
  h->rk_old_q.alloc(h->q.n);
  h->rk_old_l.alloc(h->lambda.n);
  h->rk_ku.alloc(h->u.n * nst);
  h->rk_kq.alloc(h->q.n * nst);
  h->rk_kl.alloc(h->lambda.n * nst);
  HFX_HIP(hipMemcpyAsync(h->rk_old_u.p, h->u.p, h->u.n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  HFX_HIP(hipMemcpyAsync(h->rk_old_q.p, h->q.p, h->q.n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  HFX_HIP(hipMemcpyAsync(h->rk_old_l.p, h->lambda.p, h->lambda.n * sizeof(double), hipMemcpyDeviceToDevice,
                         h->stream));
  HFX_HIP(hipMemsetAsync(h->rk_ku.p, 0, h->rk_ku.n * sizeof(double), h->stream));
  HFX_HIP(hipMemsetAsync(h->rk_kq.p, 0, h->rk_kq.n * sizeof(double), h->stream));
  HFX_HIP(hipMemsetAsync(h->rk_kl.p, 0, h->rk_kl.n * sizeof(double), h->stream));
  HFX_HIP(hipStreamSynchronize(h->stream));
  h->rk_stage = 0;
  h->rk_started = true;
  HFX_END
}

extern "C" int hfx_hdg_rk_compute_stage(hfx_hdg* h) {
  HFX_TRY
  require(h && h->rk_started, "hfx_hdg_rk_compute_stage: start the time step first");
  require(h->rk_stage < h->rk_stages(),
          "RungeKutta: cannot compute more stages than the method allows, think about computing the solution");
  h->set_device();
  const hfx::RkCoef a = rk_row(h, h->rk_stage);
  hfx::launch_rk_stage(h->u.n, h->u.p, h->rk_old_u.p, h->rk_ku.p, h->rk_stage, a, h->ts_dt, h->stream);
  hfx::launch_rk_stage(h->q.n, h->q.p, h->rk_old_q.p, h->rk_kq.p, h->rk_stage, a, h->ts_dt, h->stream);
  hfx::launch_rk_stage(h->lambda.n, h->lambda.p, h->rk_old_l.p, h->rk_kl.p, h->rk_stage, a, h->ts_dt, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  ++h->rk_stage;
  HFX_END
}

extern "C" int hfx_hdg_rk_compute_solution(hfx_hdg* h) {
  HFX_TRY
  require(h && h->rk_started, "hfx_hdg_rk_compute_solution: start the time step first");
  require(h->rk_stage == h->rk_stages(), "RungeKutta: all stages must be computed before computing the solution");
  h->set_device();
  const int nst = h->rk_stages();
  const hfx::RkCoef b = rk_row(h, nst);
  hfx::launch_rk_solution(h->u.n, h->u.p, h->rk_old_u.p, h->rk_ku.p, nst, b, h->ts_dt, h->stream);
  hfx::launch_rk_solution(h->q.n, h->q.p, h->rk_old_q.p, h->rk_kq.p, nst, b, h->ts_dt, h->stream);
  hfx::launch_rk_solution(h->lambda.n, h->lambda.p, h->rk_old_l.p, h->rk_kl.p, nst, b, h->ts_dt, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  h->rk_stage = 0;
  h->rk_started = false;
  HFX_END
}

extern "C" int hfx_hdg_rk_stages(hfx_hdg* h, int* out) {
  HFX_TRY
  require(h && out, "hfx_hdg_rk_stages: null argument");
  out[0] = h->rk_n ? h->rk_stages() : 0;
  out[1] = h->rk_stage;
  HFX_END
}

extern "C" int hfx_hdg_set_boundary_model(hfx_hdg* h, int kind, const int* faces, int n) {
  HFX_TRY
  require(h && h->has_mesh, "hfx_hdg_set_boundary_model: set the mesh first");
  require(kind == HFX_BC_DIRICHLET || kind == HFX_BC_INTEGRATED_DIRICHLET, "hfx_hdg_set_boundary_model: unknown kind");
  const hfx::Mesh& m = h->mesh;
  std::vector<char> bnd(m.nFaces, 0);
  for (int f : m.boundary) bnd[f] = 1;
  if (h->bc_kind.size() != (size_t)m.nFaces) h->bc_kind.assign(m.nFaces, HFX_BC_DIRICHLET);
  if (!faces) {
    for (int F = 0; F < m.nFaces; ++F)
      if (bnd[F] && h->bc_kind[F] != 2) h->bc_kind[F] = (char)kind;
  } else {
    require(n >= 0, "hfx_hdg_set_boundary_model: negative face count");
    for (int i = 0; i < n; ++i)
      require(faces[i] >= 0 && faces[i] < m.nFaces && bnd[faces[i]],
              "hfx_hdg_set_boundary_model: not a boundary face");
    for (int i = 0; i < n; ++i)
      require(h->bc_kind[faces[i]] != 2,
              "hfx_hdg_set_boundary_model: the face has an HDG-type boundary model (clear it with "
              "hfx_hdg_set_boundary_hdg, n = 0)");
    for (int i = 0; i < n; ++i) h->bc_kind[faces[i]] = (char)kind;
  }
  std::vector<int> lst;
  for (int F = 0; F < m.nFaces; ++F)
    if (h->bc_kind[F] == HFX_BC_INTEGRATED_DIRICHLET) lst.push_back(F);
  h->set_device();
  h->bc_int_faces.upload(lst, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  h->n_bc_int = (int)lst.size();
  h->assembled = h->solved = false;
  ++h->args_version;
  HFX_END
}

extern "C" int hfx_hdg_set_boundary_hdg(hfx_hdg* h, const int* faces, int n, const double* B, const double* r,
                                        int matrix_set, int rhs_set) {
  HFX_TRY
  require(h && h->has_mesh, "hfx_hdg_set_boundary_hdg: set the mesh first");
  require(n >= 0 && (n == 0 || (faces && B && r)), "hfx_hdg_set_boundary_hdg: null argument");
  const hfx::Mesh& m = h->mesh;
  const int nNf = m.nNf, nc = nNf * (2 + h->dim);
  std::vector<char> bnd(m.nFaces, 0);
  for (int f : m.boundary) bnd[f] = 1;
  std::vector<char> seen(m.nFaces, 0);
  for (int i = 0; i < n; ++i) {
    require(faces[i] >= 0 && faces[i] < m.nFaces && bnd[faces[i]], "hfx_hdg_set_boundary_hdg: not a boundary face");
    require(!seen[faces[i]], "hfx_hdg_set_boundary_hdg: a face is listed twice");
    seen[faces[i]] = 1;
  }
  for (size_t i = 0; i < (size_t)n * nNf * nc; ++i)
    require(std::isfinite(B[i]), "hfx_hdg_set_boundary_hdg: local matrices must be finite");
  if (h->bc_kind.size() != (size_t)m.nFaces) h->bc_kind.assign(m.nFaces, HFX_BC_DIRICHLET);
  for (int F = 0; F < m.nFaces; ++F)
    if (h->bc_kind[F] == 2) h->bc_kind[F] = HFX_BC_DIRICHLET;  // the new set replaces the old one
  for (int i = 0; i < n; ++i) h->bc_kind[faces[i]] = 2;
  h->set_device();
  h->bc_hdg_faces.upload(std::vector<int>(faces, faces + n), h->stream);
  h->bc_hdg_B.upload(std::vector<double>(B, B + (size_t)n * nNf * nc), h->stream);
  h->bc_hdg_r.upload(std::vector<double>(r, r + (size_t)n * nNf), h->stream);
  h->n_bc_hdg = n;
  h->bc_hdg_mset = matrix_set != 0;
  h->bc_hdg_rset = rhs_set != 0;
  // the integrated-Dirichlet list no longer holds these faces
  std::vector<int> lst;
  for (int F = 0; F < m.nFaces; ++F)
    if (h->bc_kind[F] == HFX_BC_INTEGRATED_DIRICHLET) lst.push_back(F);
  h->bc_int_faces.upload(lst, h->stream);
  h->n_bc_int = (int)lst.size();
  h->upload_boundary_flags();
  h->assembled = h->solved = false;
  ++h->args_version;
  HFX_END
}

extern "C" int hfx_hdg_set_trace(hfx_hdg* h, const double* lambda) {
  HFX_TRY
  require(h && h->has_mesh, "hfx_hdg_set_trace: set the mesh first");
  require(lambda != nullptr, "hfx_hdg_set_trace: null field");
  h->set_device();
  h->lambda.upload(lambda, h->lambda.n, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  HFX_END
}

extern "C" int hfx_hdg_set_solution(hfx_hdg* h, const double* u) {
  HFX_TRY
  require(h && h->has_mesh, "hfx_hdg_set_solution: set the mesh first");
  require(u != nullptr, "hfx_hdg_set_solution: null field");
  h->set_device();
  h->u.upload(u, h->u.n, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  HFX_END
}

extern "C" int hfx_hdg_set_solver_type(hfx_hdg* h, int type) {
  HFX_TRY
  require(h != nullptr, "hfx_hdg_set_solver_type: null handle");
  require(type == HFX_HDG_IMPLICIT || type == HFX_HDG_WEXPLICIT || type == HFX_HDG_SEXPLICIT,
          "hfx_hdg_set_solver_type: unknown solver type");
  h->hdg_type = type;
  h->assembled = h->solved = false;
  ++h->args_version;
  HFX_END
}

extern "C" int hfx_hdg_set_flux(hfx_hdg* h, const double* q) {
  HFX_TRY
  require(h && h->has_mesh, "hfx_hdg_set_flux: set the mesh first");
  require(q != nullptr, "hfx_hdg_set_flux: null field");
  h->set_device();
  h->q.upload(q, h->q.n, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  HFX_END
}

extern "C" int hfx_hdg_set_fused_scatter(hfx_hdg* h, int on) {
  HFX_TRY
  require(h != nullptr, "hfx_hdg_set_fused_scatter: null handle");
  h->fuse_scatter = on != 0;
  ++h->args_version;
  HFX_END
}

extern "C" int hfx_hdg_set_solver_graphs(hfx_hdg* h, int on) {
  HFX_TRY
  require(h != nullptr, "hfx_hdg_set_solver_graphs: null handle");
  h->solver_graphs = on != 0;
  HFX_END
}

extern "C" int hfx_hdg_set_krylov(hfx_hdg* h, int method) {
  HFX_TRY
  require(h != nullptr, "hfx_hdg_set_krylov: null handle");
  require(method == HFX_KRYLOV_FGMRES || method == HFX_KRYLOV_BICGSTAB, "hfx_hdg_set_krylov: unknown method");
  h->krylov = method;
  HFX_END
}

extern "C" int hfx_hdg_solve(hfx_hdg* h, int pc, double rtol, int maxits, int restart, int* info, double* relres) {
  HFX_TRY
  require(h && h->assembled, "hfx_hdg_solve: system must be assembled before solving");
  require(rtol > 0 && maxits > 0, "hfx_hdg_solve: rtol and maxits must be positive");
  h->set_device();
  hfx::KrylovInfo ki;
  do_solve(h, pc, rtol, maxits, restart, ki);
  const hfx::Mesh& m = h->mesh;
  hfx::launch_reconstruct(m.nCells, m.nN, m.nN * h->dim, m.nFpc, m.nNf, h->U.p, h->U0.p, h->Q.p, h->Q0.p, h->c2f.p,
                          h->perm.p, h->lambda.p, h->u.p, h->q.p, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  if (info) {
    info[0] = ki.iters;
    info[1] = ki.reason;
  }
  if (relres) relres[0] = ki.true_relres;
  h->last_pc = pc;
  h->last_rtol = rtol;
  h->last_maxits = maxits;
  h->last_restart = restart;
  h->solved = true;
  HFX_END
}

extern "C" int hfx_hdg_pc_apply(hfx_hdg* h, int pc, const double* x, double* y) {
  HFX_TRY
  require(h && h->assembled, "hfx_hdg_pc_apply: system must be assembled first");
  require(x && y, "hfx_hdg_pc_apply: null vector");
  h->set_device();
  const long long n = h->A.nrows();
  hipStream_t s = h->stream;
  std::function<void(const double*, double*)> M = make_pc(h, pc);
  hfx::DevBuf<double> dx, dy;
  dx.upload(x, n, s);
  dy.alloc(n);
  if (M) M(dx.p, dy.p);
  else HFX_HIP(hipMemcpyAsync(dy.p, dx.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  dy.download(y, n, s);
  HFX_HIP(hipStreamSynchronize(s));
  HFX_END
}

extern "C" int hfx_hdg_ilu_factors(hfx_hdg* h, int ordering, double* lu, double* dinv, int* info) {
  HFX_TRY
  require(h && h->assembled, "hfx_hdg_ilu_factors: system must be assembled first");
  require(ordering == hfx::ILU_MULTICOLOUR || ordering == hfx::ILU_NATURAL, "hfx_hdg_ilu_factors: unknown ordering");
  h->set_device();
  make_pc(h, ordering == hfx::ILU_MULTICOLOUR ? HFX_PC_ILU0 : HFX_PC_ILU0_NATURAL);
  const hfx::Ilu& f = *h->ilu;
  if (lu) f.lu.download(lu, f.lu.n, h->stream);
  if (dinv) f.dinv.download(dinv, f.dinv.n, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  if (info) {
    info[0] = f.plan.ncolours;
    info[1] = f.plan.nlev_f;
    info[2] = f.plan.nlev_b;
  }
  HFX_END
}

extern "C" int hfx_ilu_analyse(int nb, const int* rowptr, const int* col, int ordering, int* rank, int* levf,
                               int* levb, int* colour, long long* info) {
  HFX_TRY
  require(nb > 0 && rowptr && col, "hfx_ilu_analyse: empty pattern");
  const hfx::IluPlan P = hfx::ilu_analyse(nb, rowptr, col, ordering);
  for (int i = 0; i < nb; ++i) {
    if (rank) rank[i] = P.rank[i];
    if (levf) levf[i] = P.levf[i];
    if (levb) levb[i] = P.levb[i];
    if (colour) colour[i] = P.colour[i];
  }
  if (info) {
    info[0] = P.ncolours;
    info[1] = P.nlev_f;
    info[2] = P.nlev_b;
    info[3] = (long long)P.pslot.size();
  }
  HFX_END
}

extern "C" int hfx_hdg_get_solution(hfx_hdg* h, double* u, double* q, double* lambda) {
  HFX_TRY
  require(h && h->solved, "hfx_hdg_get_solution: solve first");
  h->set_device();
  if (u) h->u.download(u, h->u.n, h->stream);
  if (q) h->q.download(q, h->q.n, h->stream);
  if (lambda) h->lambda.download(lambda, h->lambda.n, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  HFX_END
}

extern "C" int hfx_hdg_time(hfx_hdg* h, int what, int warmup, int reps, double* ms) {
  HFX_TRY
  require(h && h->has_mesh && h->has_fields, "hfx_hdg_time: mesh and fields must be set");
  require(what >= 0 && what <= 2 && warmup >= 0 && reps >= 1, "hfx_hdg_time: bad arguments");
  require(what != 2 || h->solved, "hfx_hdg_time: solve once before timing the solve");
  h->set_device();
  hipEvent_t e0, e1;
  HFX_HIP(hipEventCreate(&e0));
  HFX_HIP(hipEventCreate(&e1));
  for (int i = 0; i < warmup + reps; ++i) {
    HFX_HIP(hipEventRecord(e0, h->stream));
    if (what == 0) h->run_element();
    else if (what == 1) { h->run_element(); h->run_scatter(); }
    else { hfx::KrylovInfo ki; do_solve(h, h->last_pc, h->last_rtol, h->last_maxits, h->last_restart, ki); }
    HFX_HIP(hipEventRecord(e1, h->stream));
    HFX_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    HFX_HIP(hipEventElapsedTime(&t, e0, e1));
    if (i >= warmup) ms[i - warmup] = t;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  h->assembled = true;
  HFX_END
}

extern "C" int hfx_element_lds_bytes(int dim, int order, long long* bytes) {
  HFX_TRY
  *bytes = (long long)hfx::element_lds_bytes(dim, order);
  HFX_END
}

extern "C" int hfx_element_workspace_bytes(int dim, int order, long long* bytes) {
  HFX_TRY
  *bytes = (long long)(hfx::element_workspace_doubles(dim, order) * sizeof(double));
  HFX_END
}

extern "C" int hfx_hdg_set_kernel(hfx_hdg* h, int which) {
  HFX_TRY
  require(h, "hfx_hdg_set_kernel: null handle");
  require(which >= 0 && which <= 2, "hfx_hdg_set_kernel: 0 (auto), 1 (general) or 2 (affine)");
  h->kernel_choice = which;
  HFX_END
}

extern "C" int hfx_hdg_kernel_info(hfx_hdg* h, int* out) {
  HFX_TRY
  require(h && h->has_mesh, "hfx_hdg_kernel_info: set the mesh first");
  const bool aff = h->use_affine();
  out[0] = aff ? 2 : 1;
  out[1] = aff ? 64 * h->affine_waves() : h->threads();
  out[2] = h->affine ? 1 : 0;
  long long lds = aff ? (long long)hfx::affine_lds_bytes(h->dim, h->order, false, h->affine_waves())
                      : (long long)hfx::element_lds_bytes(h->dim, h->order);
  out[3] = (int)lds;
  HFX_END
}
