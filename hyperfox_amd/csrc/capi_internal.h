// Handle types of the C ABI, shared by capi.cpp (mesh, HDG bulk path) and
// capi_lpc.cpp (learned preconditioner).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <stdexcept>
#include <vector>

#include "../../include/hfx_hip.h"
#include "hdg_element.h"
#include "hfx_internal.h"
#include "ilu.h"
#include "krylov.h"
#include "lpc.h"
#include "mesh.h"
#include "refel.h"
#include "time_scheme.h"
#include "trace_system.h"

namespace hfx {
// Inference state of the learned preconditioner for one problem's pattern.
struct LpcInfer {
  LpcGraph g;
  LpcSystem sys;
  LpcActs acts;
  DevBuf<float> mvals;
  bool graph_built = false;
};
}  // namespace hfx

struct hfx_hdg;
namespace hfx {
// Packed reference tables of the general element kernels (ESz layout).
void pack_tables(const RefEl& re, std::vector<double>& t);
// Build the learned preconditioner matrix for the problem's current trace
// matrix (network attached with hfx_hdg_set_lpc); defined in capi_lpc.cpp.
void lpc_prepare(hfx_hdg* h);
}  // namespace hfx

struct hfx_mesh {
  std::unique_ptr<hfx::RefEl> re;
  hfx::Mesh m;
};

struct hfx_hdg {
  int device = 0;
  hipStream_t stream = nullptr;
  int dim = 0, order = 0;
  std::unique_ptr<hfx::RefEl> re;
  hfx::Mesh mesh;
  bool has_mesh = false, has_fields = false, assembled = false, solved = false;
  int kind = 1, ndD = 0, tau_sides = 1;
  bool hasD = false, hasB = false, hasF = false, hasR = false;
  int nt = 0, grid_cap = 0;
  bool affine = false;     // every cell straight-sided (constant Jacobian)
  int kernel_choice = 0;   // 0 auto, 1 general, 2 affine
  // device data
  hfx::DevBuf<double> tabs, tabs_aff, coords, Dn, bn, tau, f_ip, r_ip, g;
  hfx::DevBuf<int> faceNodes, cells, c2f, f2c;
  hfx::DevBuf<int8_t> perm, invperm;
  hfx::DevBuf<uint8_t> is_dir;
  hfx::DevBuf<double> S, S0, U, Q, U0, Q0, rhs, lambda, dinv, u, q;
  hfx::BsrPattern pat;
  // fused trace scatter (affine kernel): per-cell BSR block index of each face
  // pair, and the two cell-side contributions to each diagonal face block
  hfx::DevBuf<int> bpos;
  hfx::DevBuf<double> dside;
  bool fuse_scatter = true;  // hfx_hdg_set_fused_scatter
  // time scheme on the u rows (hfx_hdg_set_time_scheme); u_old is the
  // problem's current "Solution" field u
  int ts_scheme = HFX_TS_NONE;
  double ts_dt = 0.0;
  // Runge-Kutta (hfx_hdg_set_butcher, hfx_hdg_rk_*): Butcher table
  // (rk_n x rk_n, row-major), stage counter, Old* fields and stage fields of
  // Solution / Flux / Trace, and the cell-local vectors of the element kernels
  std::vector<double> rk_table;
  int rk_n = 0, rk_stage = 0;
  bool rk_started = false;
  hfx::DevBuf<double> rk_old_u, rk_old_q, rk_old_l, rk_ku, rk_kq, rk_kl, ts_w;
  // boundary model per face (hfx_hdg_set_boundary_model): 0 DirichletModel,
  // 1 IntegratedDirichletModel; the device list of the latter
  std::vector<char> bc_kind;
  hfx::DevBuf<int> bc_int_faces;
  int n_bc_int = 0;
  // HDG-type boundary model (hfx_hdg_set_boundary_hdg): faces, local
  // matrices, rhs, Set / Add
  hfx::DevBuf<int> bc_hdg_faces;
  hfx::DevBuf<double> bc_hdg_B, bc_hdg_r;
  int n_bc_hdg = 0, bc_hdg_mset = 1, bc_hdg_rset = 1;
  // HDGSolverOpts::type (hfx_hdg_set_solver_type); the explicit modes read the
  // problem's "Solution" u and "Flux" q fields at assembly
  int hdg_type = HFX_HDG_IMPLICIT;
  // bumped by every call that changes the element kernel's arguments, so
  // that a batch holding a copy of them rebuilds it
  unsigned long long args_version = 0;
  bool last_fused = false;   // the last element launch wrote the BSR, not S
  hfx::BsrDev A;
  // learned preconditioner (attached network, per-problem inference state)
  unsigned long long mesh_serial = 0;  // distinct for every set_mesh in the process
  hfx_lpc* lpc = nullptr;
  int lpc_bf16 = 1;
  std::unique_ptr<hfx::LpcInfer> lpc_state;
  double last_pc_setup_ms = 0.0;
  std::unique_ptr<hfx::Ilu> ilu;  // ILU(0) schedule + factors for the current pattern
  int krylov = HFX_KRYLOV_FGMRES;
  bool solver_graphs = false;  // hfx_hdg_set_solver_graphs
  // last solve options
  int last_pc = HFX_PC_BJACOBI, last_maxits = 10000, last_restart = 30;
  double last_rtol = 1e-10;

  void set_device() const { HFX_HIP(hipSetDevice(device)); }
  hfx::ElemArgs elem_args() {
    hfx::ElemArgs a;
    a.nEl = mesh.nCells;
    a.kind = kind;
    a.useDiff = (kind != HFX_MODEL_CDR_SOURCE) || hasD;
    a.ndD = hasD ? ndD : 0;
    a.hasB = hasB;
    a.hasF = hasF;
    a.hasR = hasR;
    a.tau_sides = tau_sides;
    a.tabs = tabs.p;
    a.faceNodes = faceNodes.p;
    a.coords = coords.p;
    a.cells = cells.p;
    a.cell2face = c2f.p;
    a.face2cell = f2c.p;
    a.face_perm = perm.p;
    a.Dn = Dn.p;
    a.bn = bn.p;
    a.tau = tau.p;
    a.f_ip = f_ip.p;
    a.r_ip = r_ip.p;
    a.S = S.p; a.S0 = S0.p; a.U = U.p; a.Q = Q.p; a.U0 = U0.p; a.Q0 = Q0.p;
    if (hdg_type != HFX_HDG_IMPLICIT) {
      a.hdg_explicit = 1;
      a.sol_old = u.p;
      a.flux_old = q.p;
    }
    if (ts_scheme == HFX_TS_BEULER) {
      a.ts_dt = ts_dt;
      a.ts_idt = 1.0 / ts_dt;
      a.u_old = u.p;
    }
    if (ts_scheme == HFX_TS_FEULER || ts_scheme == HFX_TS_RK) {
      // the kernels read w (nCells x NLOC) and u_old for every cell
      const size_t nloc = (size_t)mesh.nN * (dim + 1) + (size_t)mesh.nFpc * mesh.nNf;
      const hfx::DevBuf<double>& uo = ts_scheme == HFX_TS_RK ? rk_old_u : u;
      if (!ts_w.p || ts_w.n != (size_t)mesh.nCells * nloc || !uo.p || uo.n != (size_t)mesh.nCells * mesh.nN)
        throw std::runtime_error("time scheme: the local vectors are not prepared for this mesh");
    }
    if (ts_scheme == HFX_TS_FEULER) {
      // Euler, isExplicit = true (Euler.cpp:30-35), rows divided by dt:
      // Suu = M / dt, [Suq | Sul] kept, F_u + M u / dt - Suu u
      a.ts_gen = 1;
      a.ts_auu = 0.0;
      a.ts_arest = 1.0;
      a.ts_mu = 1.0 / ts_dt;
      a.ts_phi = 1.0;
      a.ts_dt = ts_dt;
      a.u_old = u.p;
      a.ts_w = ts_w.p;
    } else if (ts_scheme == HFX_TS_RK) {
      // RungeKutta::apply (RungeKutta.cpp:113-159) at stage s with
      // a = a_ss: rows divided by a dt (a != 0) or dt (a = 0, where the
      // model's rows drop out):  Suu + M / (a dt), F_u / a + M u_old / (a dt)
      // - [Suu | Suq | Sul] w  resp.  M / dt, F_u + M u_old / dt - ... w
      if (rk_n < 2 || rk_stage >= rk_stages()) throw std::runtime_error("RungeKutta: no stage to assemble");
      const double ass = rk_diag();
      a.ts_gen = 1;
      a.ts_auu = a.ts_arest = ass != 0.0 ? 1.0 : 0.0;
      a.ts_mu = ass != 0.0 ? 1.0 / (ass * ts_dt) : 1.0 / ts_dt;
      a.ts_phi = ass != 0.0 ? 1.0 / ass : 1.0;
      a.ts_dt = ass != 0.0 ? ass * ts_dt : ts_dt;
      a.u_old = rk_old_u.p;
      a.ts_w = ts_w.p;
    }
    // orders whose element working set is kept in device memory: one slab
    // per workgroup, as many slabs as workgroups fill the GPU twice over
    const size_t wsd = hfx::element_workspace_doubles(dim, order);
    if (wsd > 0 && mesh.nCells > 0) {
      if (num_cu <= 0) HFX_HIP(hipDeviceGetAttribute(&num_cu, hipDeviceAttributeMultiprocessorCount, device));
      const int slots = (int)std::min<long long>(mesh.nCells, 2LL * num_cu);
      if (gws.n < (size_t)slots * wsd) gws.alloc((size_t)slots * wsd);
      a.gws = gws.p;
      a.gws_slots = (int)(gws.n / wsd);
    }
    return a;
  }
  int num_cu = 0;
  hfx::DevBuf<double> gws;  // element workspace (element_workspace_doubles per slab)
  int rk_stages() const { return rk_n - 1; }
  double rk_entry(int r, int c) const { return rk_table[(size_t)r * rk_n + c]; }
  double rk_diag() const { return rk_entry(rk_stage, 1 + rk_stage); }
  // cell-local vectors w of the general time transform, before an element
  // launch: explicit Euler w = [u | 0 | 0]; Runge-Kutta stage s with
  // a = a_ss, ut the Old* fields and uj = dt sum_{j<s} a_sj k_j:
  // w = ((1 - a) ut + uj) / a (a != 0), ut + uj (a = 0)
  void prepare_time_scheme() {
    if (ts_scheme != HFX_TS_FEULER && ts_scheme != HFX_TS_RK) return;
    const hfx::Mesh& m = mesh;
    const int nloc = m.nN * (dim + 1) + m.nFpc * m.nNf;
    ts_w.alloc((size_t)m.nCells * nloc);
    hfx::RkCoef ck{};
    if (ts_scheme == HFX_TS_FEULER) {
      hfx::launch_ts_local(m.nCells, m.nN, dim, m.nFpc, m.nNf, c2f.p, perm.p, 1.0, u.p, nullptr, nullptr, 0, ck,
                           nullptr, nullptr, nullptr, 0, 0, 0, ts_w.p, stream);
      return;
    }
    if (rk_n < 2) throw std::runtime_error("RungeKutta: set a Butcher table before assembling");
    if (!rk_started)
      throw std::runtime_error("RungeKutta: start the time step (hfx_hdg_rk_start_step) before assembling a stage");
    if (rk_stage >= rk_stages())
      throw std::runtime_error("RungeKutta: every stage is computed; compute the solution");
    const double ass = rk_diag();
    const double sc = ass != 0.0 ? 1.0 / ass : 1.0;
    for (int j = 0; j < rk_stage; ++j) ck.c[j] = ts_dt * rk_entry(rk_stage, 1 + j) * sc;
    hfx::launch_ts_local(m.nCells, m.nN, dim, m.nFpc, m.nNf, c2f.p, perm.p, ass != 0.0 ? (1.0 - ass) / ass : 1.0,
                         rk_old_u.p, rk_old_q.p, rk_old_l.p, rk_stage, ck, rk_ku.p, rk_kq.p, rk_kl.p, u.n, q.n,
                         lambda.n, ts_w.p, stream);
  }
  int threads() const {
    if (nt) return nt;
    return (dim == 3 && order >= 3) ? 256 : (dim == 3 || order >= 4 ? 128 : 64);
  }
  bool use_affine() const {
    if (kernel_choice == 1) return false;
    // the affine kernel keeps a cell's working set in LDS: 2-D p <= 5, 3-D p <= 3
    const bool ok = affine && !(hasD && ndD == dim * dim) && order <= (dim == 2 ? 5 : 3);
    if (kernel_choice == 2 && !ok)
      throw std::runtime_error("affine element kernel needs straight-sided cells, a scalar diffusion field and "
                               "2-D p <= 5 or 3-D p <= 3");
    return ok;
  }
  int affine_waves() const {
    for (int w : {4, 3, 2, 1})
      if (hfx::affine_lds_bytes(dim, order, false, w) + 4096 <= 160 * 1024) return w;
    return 1;
  }
  // arguments of the affine kernel; with the fused scatter unless S must be stored
  hfx::ElemArgs affine_args(bool store_S) {
    hfx::ElemArgs a = elem_args();
    a.tabs = tabs_aff.p;
    if (fuse_scatter && !store_S) {
      a.bvals = A.vals.p;
      a.dside = dside.p;
      a.bpos = bpos.p;
      a.is_dir = is_dir.p;
    }
    return a;
  }
  // element kernel; store_S forces the stored-S path (exports)
  void run_element(bool store_S = false) {
    prepare_time_scheme();
    if (use_affine()) {
      const hfx::ElemArgs a = affine_args(store_S);
      hfx::launch_affine(dim, order, false, affine_waves(), a, grid_cap, stream);
      last_fused = a.bvals != nullptr;
    } else {
      hfx::launch_element(dim, order, elem_args(), threads(), grid_cap, stream);
      last_fused = false;
    }
  }
  // the trace-system pass that completes the last element launch, then the
  // IntegratedDirichletModel rows
  void scatter_on(hipStream_t s, bool fused) {
    if (fused)
      hfx::launch_scatter_diag(A, mesh.nFpc, dside.p, S0.p, invperm.p, is_dir.p, g.p, rhs.p, s);
    else
      hfx::launch_scatter(A, mesh.nFpc, S.p, S0.p, invperm.p, is_dir.p, g.p, rhs.p, s);
    if (n_bc_int > 0) {
      const int nN = re->nN, nNf = re->nNf(), nip = re->nip(), nipf = re->face->nip();
      const size_t t_fphi = (size_t)nip * nN * (1 + dim), t_fdphi = t_fphi + (size_t)nipf * nNf;
      const size_t t_fw = t_fdphi + (size_t)nipf * nNf * (dim - 1) + nip;
      hfx::launch_integrated_dirichlet(A, n_bc_int, bc_int_faces.p, dim, nN, nipf, mesh.nFpc, faceNodes.p, perm.p,
                                       cells.p, coords.p, tabs.p + t_fphi, tabs.p + t_fdphi, tabs.p + t_fw, g.p,
                                       rhs.p, s);
    }
    if (n_bc_hdg > 0)
      hfx::launch_hdg_boundary(A, n_bc_hdg, bc_hdg_faces.p, dim, re->nN, mesh.nFpc, faceNodes.p, perm.p, invperm.p,
                               bpos.p, U.p, Q.p, U0.p, Q0.p, bc_hdg_B.p, bc_hdg_r.p, bc_hdg_mset, bc_hdg_rset, rhs.p, s);
  }
  // Dirichlet flags of the boundary faces: every boundary face but those of
  // an HDG-type model (their rows are scattered as the cell's own)
  void upload_boundary_flags() {
    std::vector<uint8_t> dir(mesh.nFaces, 0);
    for (int f : mesh.boundary) dir[f] = (bc_kind.empty() || bc_kind[f] != 2) ? 1 : 0;
    is_dir.upload(dir, stream);
    HFX_HIP(hipStreamSynchronize(stream));
  }
  void run_scatter() { scatter_on(stream, last_fused); }
};

