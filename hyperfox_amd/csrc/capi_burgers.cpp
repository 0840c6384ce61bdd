// C ABI of the Burgers path: HDGBurgersModel (m = dim unknowns per node) on
// HDGSolver's IMPLICIT assembly, the trace solve and NonLinearWrapper's
// loop, resident on one GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hfx_hip.h"
#include "burgers.h"
#include "capi_internal.h"

namespace hfx {
void set_last_error(const std::string& msg);
}

#define HFX_TRY try {
#define HFX_END                         \
  }                                     \
  catch (const std::exception& e) {     \
    hfx::set_last_error(e.what());      \
    return -1;                          \
  }                                     \
  return 0;

struct hfx_burgers {
  int device = 0, dim = 0, order = 0, m = 0;
  hipStream_t stream = nullptr;
  std::unique_ptr<hfx::RefEl> re;
  hfx::Mesh mesh;
  bool has_mesh = false, has_fields = false, assembled = false;
  int tau_sides = 1, bc_kind = HFX_BC_INTEGRATED_DIRICHLET, krylov = HFX_KRYLOV_FGMRES;
  int model = 0;            // 0 HDGBurgersModel, 1 HDGnGammaModel, 2 scalar model at m = dim DOFs
  int num_cu = 0;
  hfx::DevBuf<double> gws;  // cell workspace of the orders above the LDS (3-D p = 3)
  bool graphs = false;      // hfx_burgers_set_solver_graphs
  double ts_dt = 0.0;       // implicit Euler time step (0: steady)
  hfx::DevBuf<double> old;  // "OldSolution" of the time step
  double sound_speed = 1.0;  // nGammaParams::soundSpeed
  hfx::DevBuf<double> tabs, coords, Dn, tau, f_ip, g, G, bfield;
  hfx::DevBuf<double> sol, buf, flux, trace;  // Solution, BufferSolution, Flux, Trace
  hfx::DevBuf<double> S, S0, U, Q, U0, Q0, rhs, dinv, red;
  hfx::DevBuf<int> faceNodes, cells, c2f, f2c, bfaces, bpos;
  hfx::DevBuf<int8_t> perm, invperm, perm_m, invperm_m;
  // HDG-type boundary model (hfx_burgers_set_boundary_hdg): faces, local
  // matrices, rhs, Set / Add; bfaces then lists the other boundary faces
  hfx::DevBuf<int> hdg_faces;
  hfx::DevBuf<double> hdg_B, hdg_r;
  int n_hdg = 0, n_dir = 0, hdg_mset = 1, hdg_rset = 1;
  // HDGBohmModel on the HDG-type faces (hfx_ngamma_set_boundary_bohm): B and r
  // rebuilt on the GPU before every assembly
  bool bohm = false;
  int bohm_kind = HFX_BOHM_HEAVISIDE;
  hfx::DevBuf<double> bohm_normals;
  hfx::DevBuf<uint8_t> is_dir;
  hfx::BsrPattern pat;
  hfx::BsrDev A;
  std::unique_ptr<hfx::Ilu> ilu;
  void set_device() const { HFX_HIP(hipSetDevice(device)); }
  int bs() const { return mesh.nNf * m; }
  long long n_trace() const { return (long long)mesh.nFaces * bs(); }
  long long n_sol() const { return (long long)mesh.nCells * mesh.nN * m; }
};

namespace {

void need(bool c, const char* msg) {
  if (!c) throw std::runtime_error(msg);
}

void assemble(hfx_burgers* h) {
  need(h->has_mesh && h->has_fields, "hfx_burgers_assemble: set the mesh and the fields first");
  need(h->model != 1 || h->bfield.p != nullptr, "hfx_burgers_assemble: the nGamma b field is not set");
  const hfx::Mesh& m = h->mesh;
  hipStream_t s = h->stream;
  hfx::BurgersArgs a;
  a.nEl = m.nCells;
  a.tau_sides = h->tau_sides;
  a.tabs = h->tabs.p;
  a.faceNodes = h->faceNodes.p;
  a.coords = h->coords.p;
  a.cells = h->cells.p;
  a.cell2face = h->c2f.p;
  a.face2cell = h->f2c.p;
  a.face_perm = h->perm.p;
  a.Dn = h->Dn.p;
  a.tau = h->tau.p;
  a.sol = h->buf.p;
  a.trace = h->trace.p;
  a.f_ip = h->f_ip.p;
  a.model = h->model;
  if (h->ts_dt > 0.0) {
    a.ts_idt = 1.0 / h->ts_dt;
    a.u_old = h->old.p;
  }
  a.G = h->G.p;
  a.bfield = h->bfield.p;
  a.sound_speed = h->sound_speed;
  a.S = h->S.p; a.S0 = h->S0.p; a.U = h->U.p; a.Q = h->Q.p; a.U0 = h->U0.p; a.Q0 = h->Q0.p;
  const size_t wsd = hfx::burgers_workspace_doubles(h->dim, h->order, h->model);
  if (wsd > 0) {
    // one slab per workgroup, as many as fill the GPU twice over
    if (h->num_cu <= 0) HFX_HIP(hipDeviceGetAttribute(&h->num_cu, hipDeviceAttributeMultiprocessorCount, h->device));
    const int slots = (int)std::min<long long>(m.nCells, 2LL * h->num_cu);
    if (h->gws.n < (size_t)slots * wsd) h->gws.alloc((size_t)slots * wsd);
    a.gws = h->gws.p;
    a.gws_slots = (int)(h->gws.n / wsd);
  }
  hfx::launch_burgers(h->dim, h->order, a, s);
  // trace system: HDGSolver's scatter (Dirichlet rows identity / g), then the
  // IntegratedDirichletModel rows M (x) I and (M (x) I) g
  hfx::launch_scatter(h->A, m.nFpc, h->S.p, h->S0.p, h->invperm_m.p, h->is_dir.p, h->g.p, h->rhs.p, s);
  if (h->bc_kind == HFX_BC_INTEGRATED_DIRICHLET && h->n_dir > 0) {
    const hfx::RefEl& re = *h->re;
    const int nN = re.nN, nNf = re.nNf(), nip = re.nip(), nipf = re.face->nip(), d = h->dim;
    const size_t t_fphi = (size_t)nip * nN * (1 + d), t_fdphi = t_fphi + (size_t)nipf * nNf;
    const size_t t_fw = t_fdphi + (size_t)nipf * nNf * (d - 1) + nip;
    hfx::launch_integrated_dirichlet(h->A, h->n_dir, h->bfaces.p, d, nN, nipf, m.nFpc, h->faceNodes.p, h->perm.p,
                                     h->cells.p, h->coords.p, h->tabs.p + t_fphi, h->tabs.p + t_fdphi,
                                     h->tabs.p + t_fw, h->g.p, h->rhs.p, s, h->m);
  }
  // HDGBohmModel::compute on its faces, from the current state
  if (h->bohm && h->n_hdg > 0) {
    need(h->Dn.p && h->G.p, "hfx_burgers_assemble: the Bohm model needs the D and G fields");
    const hfx::RefEl& re = *h->re;
    const int nN = re.nN, nNf = re.nNf(), nip = re.nip(), nipf = re.face->nip(), d = h->dim;
    const size_t t_fphi = (size_t)nip * nN * (1 + d), t_fdphi = t_fphi + (size_t)nipf * nNf;
    const size_t t_fw = t_fdphi + (size_t)nipf * nNf * (d - 1) + nip;
    hfx::BohmArgs ba;
    ba.nNf = nNf;
    ba.nipf = nipf;
    ba.nN = nN;
    ba.nFpc = m.nFpc;
    ba.tau_sides = h->tau_sides;
    ba.kind = h->bohm_kind;
    ba.faces = h->hdg_faces.p;
    ba.normals = h->bohm_normals.p;
    ba.face2cell = h->f2c.p;
    ba.cell2face = h->c2f.p;
    ba.faceNodes = h->faceNodes.p;
    ba.invperm = h->invperm.p;
    ba.cells = h->cells.p;
    ba.coords = h->coords.p;
    ba.fphi = h->tabs.p + t_fphi;
    ba.fdphi = h->tabs.p + t_fdphi;
    ba.fw = h->tabs.p + t_fw;
    ba.sol = h->sol.p;
    ba.flux = h->flux.p;
    ba.trace = h->trace.p;
    ba.tau = h->tau.p;
    ba.D = h->Dn.p;
    ba.G = h->G.p;
    ba.bfield = h->bfield.p;
    ba.sound_speed = h->sound_speed;
    ba.B = h->hdg_B.p;
    ba.r = h->hdg_r.p;
    hfx::launch_bohm(ba, h->n_hdg, s);
  }
  // HDG-type boundary rows (HDGSolver.cpp:503-518) on the faces of that model
  if (h->n_hdg > 0)
    hfx::launch_hdg_boundary(h->A, h->n_hdg, h->hdg_faces.p, h->dim, m.nN, m.nFpc, h->faceNodes.p, h->perm.p,
                             h->invperm.p, h->bpos.p, h->U.p, h->Q.p, h->U0.p, h->Q0.p, h->hdg_B.p, h->hdg_r.p,
                             h->hdg_mset, h->hdg_rset, h->rhs.p, s, h->m);
  h->assembled = true;
}

std::function<void(const double*, double*)> make_pc(hfx_burgers* h, int pc) {
  const long long n = h->A.nrows();
  hipStream_t s = h->stream;
  if (pc == HFX_PC_NONE) return {};
  if (pc == HFX_PC_JACOBI) {
    h->dinv.alloc(n);
    hfx::bsr_inv_point_diag(h->A, h->dinv.p, s);
    return [h, n, s](const double* in, double* out) { hfx::apply_point_diag(n, h->dinv.p, in, out, s); };
  }
  if (pc == HFX_PC_BJACOBI) {
    h->dinv.alloc((size_t)h->A.nb * h->A.bs * h->A.bs);
    hfx::bsr_invert_diag(h->A, h->dinv.p, s);
    return [h, s](const double* in, double* out) { hfx::apply_block_diag(h->A.nb, h->A.bs, h->dinv.p, in, out, s); };
  }
  if (pc == HFX_PC_ILU0 || pc == HFX_PC_ILU0_F32) {
    if (!h->ilu) {  // kept only once its set-up succeeded
      std::unique_ptr<hfx::Ilu> f(new hfx::Ilu());
      f->setup(h->A.nb, h->A.bs, h->pat.rowptr.data(), h->pat.col.data(), hfx::ILU_MULTICOLOUR, s);
      h->ilu = std::move(f);
    }
    h->ilu->low = pc == HFX_PC_ILU0_F32;
    h->ilu->factor(h->A, s);
    hfx::Ilu* f = h->ilu.get();
    return [f, s](const double* in, double* out) { f->apply(in, out, s); };
  }
  throw std::runtime_error("hfx_burgers_solve: unsupported preconditioner");
}

void solve(hfx_burgers* h, int pc, double rtol, int maxits, int restart, hfx::KrylovInfo& info) {
  need(h->assembled, "hfx_burgers_solve: assemble first");
  hipStream_t s = h->stream;
  const hfx::Mesh& m = h->mesh;
  auto M = make_pc(h, pc);
  HFX_HIP(hipMemsetAsync(h->trace.p, 0, h->n_trace() * sizeof(double), s));
  hfx::KrylovOpts o;
  o.rtol = rtol;
  o.maxits = maxits;
  o.restart = restart;
  o.graph = h->graphs;
  if (h->krylov == HFX_KRYLOV_BICGSTAB) hfx::bicgstab(h->A, h->rhs.p, h->trace.p, M, o, info, s);
  else hfx::fgmres(h->A, h->rhs.p, h->trace.p, M, o, info, s);
  const int mu = m.nN * h->m;
  hfx::launch_reconstruct(m.nCells, mu, mu * h->dim, m.nFpc, h->bs(), h->U.p, h->U0.p, h->Q.p, h->Q0.p, h->c2f.p,
                          h->perm_m.p, h->trace.p, h->sol.p, h->flux.p, s);
  HFX_HIP(hipStreamSynchronize(s));
}

}  // namespace

namespace {
hfx_burgers* create(int device, int dim, int order, int model) {
  int n = 0;
  HFX_HIP(hipGetDeviceCount(&n));
  need(device >= 0 && device < n, "hfx_burgers_create: no such GPU");
  auto h = std::unique_ptr<hfx_burgers>(new hfx_burgers());
  h->model = model;
  h->device = device;
  h->dim = dim;
  h->order = order;
  h->m = dim;
  h->re.reset(new hfx::RefEl(dim, order));
  h->set_device();
  HFX_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  std::vector<double> t;
  hfx::pack_tables(*h->re, t);
  h->tabs.upload(t, h->stream);
  std::vector<int> fnodes;
  for (auto& f : h->re->faceNodes) fnodes.insert(fnodes.end(), f.begin(), f.end());
  h->faceNodes.upload(fnodes, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  return h.release();
}
}  // namespace

extern "C" int hfx_burgers_create(int device, int dim, int order, hfx_burgers** out) {
  HFX_TRY
  need(out != nullptr, "hfx_burgers_create: null output");
  need(hfx::burgers_supported(dim, order),
       "hfx_burgers_create: the Burgers kernel covers 2-D orders 1-5 and 3-D orders 1-3");
  *out = create(device, dim, order, 0);
  HFX_END
}

extern "C" int hfx_multi_create(int device, int dim, int order, hfx_burgers** out) {
  HFX_TRY
  need(out != nullptr, "hfx_multi_create: null output");
  need(hfx::burgers_supported(dim, order), "hfx_multi_create: the multi-DOF kernel covers 2-D orders 1-5 and 3-D orders 1-3");
  *out = create(device, dim, order, 2);
  HFX_END
}

extern "C" int hfx_ngamma_create(int device, int order, hfx_burgers** out) {
  HFX_TRY
  need(out != nullptr, "hfx_ngamma_create: null output");
  need(order >= 1 && order <= 5, "hfx_ngamma_create: the nGamma kernel covers 2-D orders 1-5");
  *out = create(device, 2, order, 1);
  HFX_END
}

extern "C" void hfx_burgers_destroy(hfx_burgers* h) {
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

extern "C" int hfx_burgers_set_mesh(hfx_burgers* h, const hfx_mesh* mm) {
  HFX_TRY
  need(h && mm, "hfx_burgers_set_mesh: null handle");
  need(mm->m.dim == h->dim && mm->re->order == h->order, "hfx_burgers_set_mesh: mesh dimension/order mismatch");
  h->set_device();
  h->mesh = mm->m;
  h->ilu.reset();
  h->has_fields = h->assembled = false;
  const hfx::Mesh& m = h->mesh;
  hipStream_t s = h->stream;
  const int md = h->m, bs = m.nNf * md;
  h->coords.upload(m.coords, s);
  h->cells.upload(m.cells, s);
  h->c2f.upload(m.cell2face, s);
  h->f2c.upload(m.face2cell, s);
  h->perm.upload(m.face_perm, s);
  h->bfaces.upload(m.boundary, s);
  std::vector<int8_t> inv(m.face_perm.size());
  for (size_t blk = 0; blk < m.face_perm.size() / m.nNf; ++blk)
    for (int j = 0; j < m.nNf; ++j) inv[blk * m.nNf + m.face_perm[blk * m.nNf + j]] = (int8_t)j;
  h->invperm.upload(inv, s);
  h->n_hdg = 0;
  h->n_dir = (int)m.boundary.size();
  const long long nblk = (long long)m.nCells * m.nFpc;
  h->perm_m.alloc((size_t)nblk * bs);
  h->invperm_m.alloc((size_t)nblk * bs);
  hfx::launch_expand_perm(nblk, m.nNf, md, h->perm.p, h->perm_m.p, s);
  hfx::launch_expand_perm(nblk, m.nNf, md, h->invperm.p, h->invperm_m.p, s);
  h->pat = hfx::build_bsr_pattern(m.nCells, m.nFaces, m.nFpc, bs, m.cell2face.data(), m.face2cell.data());
  h->A.nb = h->pat.nb;
  h->A.bs = h->pat.bs;
  h->A.nnzb = (int)h->pat.col.size();
  h->A.rowptr.upload(h->pat.rowptr, s);
  h->A.col.upload(h->pat.col, s);
  h->A.rowf.upload(h->pat.rowf, s);
  h->A.diag.upload(h->pat.diag, s);
  h->A.contrib.upload(h->pat.contrib, s);
  h->A.face_cells.upload(h->pat.face_cells, s);
  h->A.vals.alloc((size_t)h->A.nnzb * bs * bs);
  h->bpos.upload(hfx::cell_block_positions(h->pat, m.nCells, m.nFpc, m.cell2face.data()), s);
  std::vector<uint8_t> dir(m.nFaces, 0);
  for (int f : m.boundary) dir[f] = 1;
  h->is_dir.upload(dir, s);
  const size_t mu = (size_t)m.nN * md, ml = (size_t)m.nFpc * bs;
  h->S.alloc(m.nCells * ml * ml);
  h->S0.alloc(m.nCells * ml);
  h->U.alloc(m.nCells * mu * ml);
  h->Q.alloc(m.nCells * mu * h->dim * ml);
  h->U0.alloc(m.nCells * mu);
  h->Q0.alloc(m.nCells * mu * h->dim);
  h->rhs.alloc(h->n_trace());
  h->sol.alloc(h->n_sol());
  h->buf.alloc(h->n_sol());
  h->flux.alloc(h->n_sol() * h->dim);
  h->trace.alloc(h->n_trace());
  h->red.alloc(2 * 256 + 2);
  h->old.alloc(h->n_sol());
  HFX_HIP(hipMemsetAsync(h->old.p, 0, h->old.n * sizeof(double), s));
  HFX_HIP(hipMemsetAsync(h->sol.p, 0, h->sol.n * sizeof(double), s));
  HFX_HIP(hipMemsetAsync(h->buf.p, 0, h->buf.n * sizeof(double), s));
  HFX_HIP(hipMemsetAsync(h->flux.p, 0, h->flux.n * sizeof(double), s));
  HFX_HIP(hipMemsetAsync(h->trace.p, 0, h->trace.n * sizeof(double), s));
  HFX_HIP(hipStreamSynchronize(s));
  h->has_mesh = true;
  HFX_END
}

namespace {
// The fields both models take: Tau, the first diffusion coefficient, the
// Dirichlet values and the source values.
void set_common_fields(hfx_burgers* h, const double* tau, int tau_sides, const double* D, const double* dirichlet,
                       const double* f_ip) {
  need(h && h->has_mesh, "hfx_burgers_set_fields: set the mesh first");
  need(tau != nullptr, "hfx_burgers_set_fields: the Tau field is required");
  need(tau_sides == 1 || tau_sides == 2,
       "hfx_burgers_set_fields: the Tau field must have the same or twice the number of values per object as the "
       "Solution field");
  h->set_device();
  const hfx::Mesh& m = h->mesh;
  hipStream_t s = h->stream;
  const size_t mm = (size_t)h->m * h->m;
  h->tau_sides = tau_sides;
  h->tau.upload(tau, (size_t)m.nFaces * m.nNf * tau_sides * mm, s);
  if (D) h->Dn.upload(D, (size_t)m.nNodes, s);
  else if (h->model == 2) h->Dn.upload(std::vector<double>((size_t)m.nNodes, 1.0), s);  // HDGLaplaceModel: D = I
  else h->Dn.release();
  h->g.alloc(h->n_trace());
  if (dirichlet) HFX_HIP(hipMemcpyAsync(h->g.p, dirichlet, h->n_trace() * sizeof(double), hipMemcpyHostToDevice, s));
  else HFX_HIP(hipMemsetAsync(h->g.p, 0, h->n_trace() * sizeof(double), s));
  if (f_ip) h->f_ip.upload(f_ip, (size_t)m.nCells * h->re->nip() * h->m, s);
  else h->f_ip.release();
  HFX_HIP(hipStreamSynchronize(s));
  h->has_fields = true;
  h->assembled = false;
}
}  // namespace

extern "C" int hfx_burgers_set_fields(hfx_burgers* h, const double* tau, int tau_sides, const double* D,
                                      const double* dirichlet, const double* f_ip) {
  HFX_TRY
  need(h != nullptr, "hfx_burgers_set_fields: null handle");
  need(h->model != 1, "hfx_burgers_set_fields: an nGamma problem takes hfx_ngamma_set_fields");
  set_common_fields(h, tau, tau_sides, D, dirichlet, f_ip);
  HFX_END
}

extern "C" int hfx_ngamma_set_fields(hfx_burgers* h, const double* tau, int tau_sides, const double* D,
                                     const double* G, const double* b, double sound_speed, const double* dirichlet,
                                     const double* f_ip) {
  HFX_TRY
  need(h && h->has_mesh, "hfx_ngamma_set_fields: set the mesh first");
  need(h->model == 1, "hfx_ngamma_set_fields: not an nGamma problem");
  need(b != nullptr, "hfx_ngamma_set_fields: must provide a b vector field to the nGamma model");
  need(std::isfinite(sound_speed), "hfx_ngamma_set_fields: the sound speed must be finite");
  set_common_fields(h, tau, tau_sides, D, dirichlet, f_ip);
  const hfx::Mesh& m = h->mesh;
  if (G) h->G.upload(G, (size_t)m.nNodes, h->stream);
  else h->G.release();
  h->bfield.upload(b, (size_t)m.nNodes * 2, h->stream);
  h->sound_speed = sound_speed;
  HFX_HIP(hipStreamSynchronize(h->stream));
  HFX_END
}

extern "C" int hfx_burgers_set_boundary_model(hfx_burgers* h, int kind) {
  HFX_TRY
  need(h != nullptr, "hfx_burgers_set_boundary_model: null handle");
  need(kind == HFX_BC_DIRICHLET || kind == HFX_BC_INTEGRATED_DIRICHLET,
       "hfx_burgers_set_boundary_model: unknown boundary model");
  h->bc_kind = kind;
  h->assembled = false;
  HFX_END
}

namespace {

// The faces of the HDG-type boundary model (checked: boundary faces, each
// once); the Dirichlet models keep the other boundary faces, and the scatter
// leaves the cell's own rows on these.  B and r are left to the caller.
void set_hdg_faces(hfx_burgers* h, const char* who, const int* faces, int n, int matrix_set, int rhs_set) {
  const hfx::Mesh& m = h->mesh;
  std::vector<char> bnd(m.nFaces, 0), seen(m.nFaces, 0);
  for (int f : m.boundary) bnd[f] = 1;
  for (int i = 0; i < n; ++i) {
    need(faces[i] >= 0 && faces[i] < m.nFaces && bnd[faces[i]], (std::string(who) + ": not a boundary face").c_str());
    need(!seen[faces[i]], (std::string(who) + ": a face is listed twice").c_str());
    seen[faces[i]] = 1;
  }
  h->set_device();
  hipStream_t s = h->stream;
  h->hdg_faces.upload(std::vector<int>(faces, faces + n), s);
  h->n_hdg = n;
  h->hdg_mset = matrix_set != 0;
  h->hdg_rset = rhs_set != 0;
  std::vector<int> dirl;
  std::vector<uint8_t> dir(m.nFaces, 0);
  for (int f : m.boundary)
    if (!seen[f]) {
      dirl.push_back(f);
      dir[f] = 1;
    }
  h->bfaces.upload(dirl, s);
  h->n_dir = (int)dirl.size();
  h->is_dir.upload(dir, s);
  h->bohm = false;
  h->assembled = false;
}

}  // namespace

extern "C" int hfx_burgers_set_boundary_hdg(hfx_burgers* h, const int* faces, int n, const double* B,
                                            const double* r, int matrix_set, int rhs_set) {
  HFX_TRY
  need(h && h->has_mesh, "hfx_burgers_set_boundary_hdg: set the mesh first");
  need(n >= 0 && (n == 0 || (faces && B && r)), "hfx_burgers_set_boundary_hdg: null argument");
  const int bs = h->bs(), nc = bs * (2 + h->dim);
  for (size_t i = 0; i < (size_t)n * bs * nc; ++i)
    need(std::isfinite(B[i]), "hfx_burgers_set_boundary_hdg: local matrices must be finite");
  set_hdg_faces(h, "hfx_burgers_set_boundary_hdg", faces, n, matrix_set, rhs_set);
  hipStream_t s = h->stream;
  h->hdg_B.upload(std::vector<double>(B, B + (size_t)n * bs * nc), s);
  h->hdg_r.upload(std::vector<double>(r, r + (size_t)n * bs), s);
  HFX_HIP(hipStreamSynchronize(s));
  HFX_END
}

extern "C" int hfx_ngamma_set_boundary_bohm(hfx_burgers* h, const int* faces, int n, const double* normals,
                                            int transfer) {
  HFX_TRY
  need(h && h->has_mesh, "hfx_ngamma_set_boundary_bohm: set the mesh first");
  need(h->model == 1, "hfx_ngamma_set_boundary_bohm: the Bohm model belongs to the nGamma problem");
  need(n >= 0 && (n == 0 || (faces && normals)), "hfx_ngamma_set_boundary_bohm: null argument");
  need(transfer >= HFX_BOHM_NEUMANN && transfer <= HFX_BOHM_TANH, "hfx_ngamma_set_boundary_bohm: unknown transfer function");
  const int nNf = h->mesh.nNf, bs = h->bs(), nc = bs * (2 + h->dim);
  for (size_t i = 0; i < (size_t)n * nNf * 2; ++i)
    need(std::isfinite(normals[i]), "hfx_ngamma_set_boundary_bohm: the normals must be finite");
  set_hdg_faces(h, "hfx_ngamma_set_boundary_bohm", faces, n, 1, 1);  // Set / Set (HDGBohmModel.cpp:89-92)
  hipStream_t s = h->stream;
  h->hdg_B.alloc((size_t)std::max(n, 1) * bs * nc);
  h->hdg_r.alloc((size_t)std::max(n, 1) * bs);
  if (n > 0) h->bohm_normals.upload(std::vector<double>(normals, normals + (size_t)n * nNf * 2), s);
  h->bohm = n > 0;
  h->bohm_kind = transfer;
  HFX_HIP(hipStreamSynchronize(s));
  HFX_END
}

extern "C" int hfx_burgers_boundary_rows(hfx_burgers* h, double* B, double* r) {
  HFX_TRY
  need(h && h->has_mesh, "hfx_burgers_boundary_rows: set the mesh first");
  h->set_device();
  const int bs = h->bs(), nc = bs * (2 + h->dim);
  if (B && h->n_hdg > 0) h->hdg_B.download(B, (size_t)h->n_hdg * bs * nc, h->stream);
  if (r && h->n_hdg > 0) h->hdg_r.download(r, (size_t)h->n_hdg * bs, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  HFX_END
}

extern "C" int hfx_burgers_set_time_scheme(hfx_burgers* h, int scheme, double dt) {
  HFX_TRY
  need(h != nullptr, "hfx_burgers_set_time_scheme: null handle");
  need(scheme == HFX_TS_NONE || scheme == HFX_TS_BEULER,
       "hfx_burgers_set_time_scheme: the multi-DOF problems take implicit Euler (BEuler) only");
  need(scheme == HFX_TS_NONE || (dt > 0.0 && std::isfinite(dt)),
       "hfx_burgers_set_time_scheme: the time step needs to be set and it should not be 0");
  h->ts_dt = scheme == HFX_TS_NONE ? 0.0 : dt;
  h->assembled = false;
  HFX_END
}

extern "C" int hfx_burgers_start_step(hfx_burgers* h) {
  HFX_TRY
  need(h && h->has_mesh, "hfx_burgers_start_step: set the mesh first");
  h->set_device();
  hipStream_t s = h->stream;
  HFX_HIP(hipMemcpyAsync(h->old.p, h->sol.p, h->n_sol() * sizeof(double), hipMemcpyDeviceToDevice, s));
  HFX_HIP(hipMemcpyAsync(h->buf.p, h->sol.p, h->n_sol() * sizeof(double), hipMemcpyDeviceToDevice, s));
  HFX_HIP(hipStreamSynchronize(s));
  h->assembled = false;
  HFX_END
}

extern "C" int hfx_burgers_set_dirichlet(hfx_burgers* h, const double* dirichlet) {
  HFX_TRY
  need(h && h->has_fields, "hfx_burgers_set_dirichlet: set the fields first");
  h->set_device();
  if (dirichlet)
    HFX_HIP(hipMemcpyAsync(h->g.p, dirichlet, h->n_trace() * sizeof(double), hipMemcpyHostToDevice, h->stream));
  else
    HFX_HIP(hipMemsetAsync(h->g.p, 0, h->n_trace() * sizeof(double), h->stream));
  HFX_HIP(hipStreamSynchronize(h->stream));
  h->assembled = false;
  HFX_END
}

extern "C" int hfx_burgers_set_solver_graphs(hfx_burgers* h, int on) {
  HFX_TRY
  need(h != nullptr, "hfx_burgers_set_solver_graphs: null handle");
  h->graphs = on != 0;
  HFX_END
}

extern "C" int hfx_burgers_set_krylov(hfx_burgers* h, int krylov) {
  HFX_TRY
  need(h != nullptr, "hfx_burgers_set_krylov: null handle");
  need(krylov == HFX_KRYLOV_FGMRES || krylov == HFX_KRYLOV_BICGSTAB, "hfx_burgers_set_krylov: unknown method");
  h->krylov = krylov;
  HFX_END
}

extern "C" int hfx_burgers_set_state(hfx_burgers* h, const double* sol, const double* buffer, const double* trace) {
  HFX_TRY
  need(h && h->has_mesh, "hfx_burgers_set_state: set the mesh first");
  h->set_device();
  hipStream_t s = h->stream;
  if (sol) HFX_HIP(hipMemcpyAsync(h->sol.p, sol, h->sol.n * sizeof(double), hipMemcpyHostToDevice, s));
  if (buffer) HFX_HIP(hipMemcpyAsync(h->buf.p, buffer, h->buf.n * sizeof(double), hipMemcpyHostToDevice, s));
  if (trace) HFX_HIP(hipMemcpyAsync(h->trace.p, trace, h->trace.n * sizeof(double), hipMemcpyHostToDevice, s));
  HFX_HIP(hipStreamSynchronize(s));
  h->assembled = false;
  HFX_END
}

extern "C" int hfx_burgers_get_state(hfx_burgers* h, double* sol, double* buffer, double* flux, double* trace) {
  HFX_TRY
  need(h && h->has_mesh, "hfx_burgers_get_state: set the mesh first");
  h->set_device();
  hipStream_t s = h->stream;
  if (sol) h->sol.download(sol, h->sol.n, s);
  if (buffer) h->buf.download(buffer, h->buf.n, s);
  if (flux) h->flux.download(flux, h->flux.n, s);
  if (trace) h->trace.download(trace, h->trace.n, s);
  HFX_HIP(hipStreamSynchronize(s));
  HFX_END
}

extern "C" int hfx_burgers_sizes(const hfx_burgers* h, long long* out) {
  HFX_TRY
  need(h && h->has_mesh && out, "hfx_burgers_sizes: set the mesh first");
  out[0] = h->A.nb;
  out[1] = h->A.bs;
  out[2] = h->A.nnzb;
  out[3] = (long long)h->mesh.nN * h->m;                // u rows per cell
  out[4] = (long long)h->mesh.nFpc * h->A.bs;           // trace entries per cell
  HFX_END
}

extern "C" int hfx_burgers_assemble(hfx_burgers* h) {
  HFX_TRY
  need(h != nullptr, "hfx_burgers_assemble: null handle");
  h->set_device();
  assemble(h);
  HFX_HIP(hipStreamSynchronize(h->stream));
  HFX_END
}

extern "C" int hfx_burgers_condensed(hfx_burgers* h, double* S, double* S0, double* U, double* Q, double* U0,
                                     double* Q0) {
  HFX_TRY
  need(h && h->assembled, "hfx_burgers_condensed: assemble first");
  h->set_device();
  hipStream_t s = h->stream;
  if (S) h->S.download(S, h->S.n, s);
  if (S0) h->S0.download(S0, h->S0.n, s);
  if (U) h->U.download(U, h->U.n, s);
  if (Q) h->Q.download(Q, h->Q.n, s);
  if (U0) h->U0.download(U0, h->U0.n, s);
  if (Q0) h->Q0.download(Q0, h->Q0.n, s);
  HFX_HIP(hipStreamSynchronize(s));
  HFX_END
}

extern "C" int hfx_burgers_trace_system(hfx_burgers* h, int* rowptr, int* col, double* vals, double* rhs) {
  HFX_TRY
  need(h && h->assembled, "hfx_burgers_trace_system: assemble first");
  h->set_device();
  if (rowptr) std::copy(h->pat.rowptr.begin(), h->pat.rowptr.end(), rowptr);
  if (col) std::copy(h->pat.col.begin(), h->pat.col.end(), col);
  if (vals) h->A.vals.download(vals, h->A.vals.n, h->stream);
  if (rhs) h->rhs.download(rhs, h->rhs.n, h->stream);
  HFX_HIP(hipStreamSynchronize(h->stream));
  HFX_END
}

extern "C" int hfx_burgers_solve(hfx_burgers* h, int pc, double rtol, int maxits, int restart, int* iters, int* reason,
                                 double* relres) {
  HFX_TRY
  need(h != nullptr, "hfx_burgers_solve: null handle");
  h->set_device();
  hfx::KrylovInfo info;
  solve(h, pc, rtol, maxits, restart, info);
  if (iters) *iters = info.iters;
  if (reason) *reason = info.reason;
  if (relres) *relres = info.true_relres;
  HFX_END
}

extern "C" int hfx_burgers_nonlinear_solve(hfx_burgers* h, int max_iters, double res_tol, double damp, int pc,
                                           double rtol, int maxits, int restart, int* iters, double* residual,
                                           int* linear_iters) {
  HFX_TRY
  need(h != nullptr, "hfx_burgers_nonlinear_solve: null handle");
  need(max_iters >= 1, "hfx_burgers_nonlinear_solve: at least one iteration");
  h->set_device();
  hipStream_t s = h->stream;
  double res = 0.0;
  int it = 0, lin = 0;
  for (it = 1; it <= max_iters; ++it) {
    assemble(h);
    hfx::KrylovInfo info;
    solve(h, pc, rtol, maxits, restart, info);
    lin += info.iters;
    if (info.reason < 0)
      throw std::runtime_error("hfx_burgers_nonlinear_solve: the linear solve of iteration " + std::to_string(it) +
                               " did not converge (reason " + std::to_string(info.reason) + " after " +
                               std::to_string(info.iters) + " iterations, relative residual " +
                               std::to_string(info.true_relres) + ")");
    hfx::burgers_residual(h->n_sol(), h->sol.p, h->buf.p, h->red.p, h->red.p + 512, s);
    double sums[2];
    HFX_HIP(hipMemcpyAsync(sums, h->red.p + 512, sizeof(sums), hipMemcpyDeviceToHost, s));
    HFX_HIP(hipStreamSynchronize(s));
    res = sums[1] != 0.0 ? std::sqrt(sums[0] / sums[1]) : std::sqrt(sums[0]);
    if (res < res_tol) break;
    hfx::burgers_damp(h->n_sol(), damp, h->sol.p, h->buf.p, s);
  }
  if (it > max_iters) it = max_iters;
  HFX_HIP(hipStreamSynchronize(s));
  if (iters) *iters = it;
  if (residual) *residual = res;
  if (linear_iters) *linear_iters = lin;
  HFX_END
}
