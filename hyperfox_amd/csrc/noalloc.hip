// hfx_hdg_solve_noalloc: the trace solve of hfx_hdg_solve (BiCGStab, point or
// block Jacobi) with no device allocation, for callers that cannot call
// hipMalloc (code under HIP graph capture, serving loops, harnesses that
// forbid allocation).
//
// Its workspace (the Krylov vectors, device scalars, reduction partials and
// the preconditioner's inverse diagonal) is the file-scope device array
// `scratch` below, one copy on every GPU the library runs on, held until the
// process ends.  Host code never passes `scratch` itself to a kernel (in host
// code the name is a stand-in, not a device address): its device address on
// each GPU is read once with hipGetSymbolAddress, with that GPU current, and
// kept per device.
//
// Every call on one GPU shares that GPU's array, which keeps the contents of
// the last call; everything the solve reads is written by the same call
// first.  Calls that overlap on one GPU (two host threads, same device) race
// on the array and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>

#include "capi_internal.h"

namespace hfx {
void set_last_error(const std::string& msg);
}

namespace {

// 64 MiB per GPU: a BiCGStab solve needs about (8 + nNf) doubles per trace DOF
// with block Jacobi, so systems of up to ~0.6M trace DOFs (every 2-D system of
// the benchmarked sizes up to 256 x 256 cells at p = 3) fit.
constexpr size_t SCRATCH_BYTES = (size_t)64 << 20;
constexpr size_t SCRATCH_DOUBLES = SCRATCH_BYTES / sizeof(double);
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

std::mutex g_mu;
double* g_scratch[MAX_DEVICES] = {};

// Device address of `scratch` on `device`, read on the first call for that
// device; the caller has made `device` current.
double* scratch_on(int device) {
  if (device < 0 || device >= MAX_DEVICES) throw std::runtime_error("hfx_hdg_solve_noalloc: device id out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      throw std::runtime_error(std::string("hfx_hdg_solve_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                               hipGetErrorString(e));
    g_scratch[device] = static_cast<double*>(p);
  }
  return g_scratch[device];
}

inline size_t al32(size_t k) { return (k + 31) & ~(size_t)31; }

void require(bool c, const char* msg) {
  if (!c) throw std::runtime_error(msg);
}

}  // namespace

extern "C" int hfx_hdg_solve_noalloc(hfx_hdg* h, int pc, double rtol, int maxits, int* info, double* relres) {
  try {
    require(h && h->assembled, "hfx_hdg_solve_noalloc: system must be assembled before solving");
    require(rtol > 0 && maxits > 0, "hfx_hdg_solve_noalloc: rtol and maxits must be positive");
    require(pc == HFX_PC_NONE || pc == HFX_PC_JACOBI || pc == HFX_PC_BJACOBI,
            "hfx_hdg_solve_noalloc: the preconditioner must be HFX_PC_NONE, HFX_PC_JACOBI or HFX_PC_BJACOBI");
    require(h->hdg_type != HFX_HDG_SEXPLICIT,
            "hfx_hdg_solve_noalloc: SEXPLICIT uses face-local direct solves; use hfx_hdg_solve");
    h->set_device();
    const hfx::BsrDev& A = h->A;
    const long long n = A.nrows();
    const size_t need_k = al32(hfx::bicgstab_ws_doubles(n));
    const size_t need_pc = pc == HFX_PC_JACOBI    ? al32((size_t)n)
                           : pc == HFX_PC_BJACOBI ? al32((size_t)A.nb * A.bs * A.bs)
                                                  : 0;
    if (need_k + need_pc > SCRATCH_DOUBLES) {
      char msg[256];
      std::snprintf(msg, sizeof(msg),
                    "hfx_hdg_solve_noalloc: a trace system of %lld DOFs needs %.1f MiB of workspace, above the "
                    "library's fixed %zu MiB; use hfx_hdg_solve",
                    n, (need_k + need_pc) * sizeof(double) / 1048576.0, SCRATCH_BYTES >> 20);
      throw std::runtime_error(msg);
    }
    double* ws = scratch_on(h->device);
    double* dinv = ws + need_k;
    hipStream_t s = h->stream;
    // the preconditioners of hfx_hdg_solve, their inverse diagonal in the workspace
    std::function<void(const double*, double*)> M;
    if (pc == HFX_PC_JACOBI) {
      hfx::bsr_inv_point_diag(A, dinv, s);
      M = [n, dinv, s](const double* in, double* out) { hfx::apply_point_diag(n, dinv, in, out, s); };
    } else if (pc == HFX_PC_BJACOBI) {
      hfx::bsr_invert_diag(A, dinv, s);
      M = [&A, dinv, s](const double* in, double* out) { hfx::apply_block_diag(A.nb, A.bs, dinv, in, out, s); };
    }
    HFX_HIP(hipMemsetAsync(h->lambda.p, 0, n * sizeof(double), s));  // zero initial guess, as PETSc
    hfx::KrylovOpts o;
    o.rtol = rtol;
    o.maxits = maxits;
    o.graph = false;
    hfx::KrylovInfo ki;
    hfx::bicgstab(A, h->rhs.p, h->lambda.p, M, o, ki, s, ws);
    const hfx::Mesh& m = h->mesh;
    hfx::launch_reconstruct(m.nCells, m.nN, m.nN * h->dim, m.nFpc, m.nNf, h->U.p, h->U0.p, h->Q.p, h->Q0.p, h->c2f.p,
                            h->perm.p, h->lambda.p, h->u.p, h->q.p, s);
    HFX_HIP(hipStreamSynchronize(s));
    if (info) {
      info[0] = ki.iters;
      info[1] = ki.reason;
    }
    if (relres) relres[0] = ki.true_relres;
    h->last_pc = pc;
    h->last_rtol = rtol;
    h->last_maxits = maxits;
    h->solved = true;
  } catch (const std::exception& e) {
    hfx::set_last_error(e.what());
    return -1;
  }
  return 0;
}
