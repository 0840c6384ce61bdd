// C ABI of the LinAlgebraInterface mirror (hfx_la_*), declared in
// include/hfx_hip.h.
//
// Host-side semantics follow PetscInterface (src/resolution/PetscInterface.cpp:5-265):
// the state machine construct -> initialize -> configure -> allocate ->
// set/add -> assemble -> solve with the same misuse errors; ADD and INSERT
// values may not be mixed between two flushes (PETSc's MatSetValues rule);
// value blocks are row-major (MatSetValues' default); zeroOutRows keeps the
// pattern; clearSystem zeroes values and keeps the pattern; solve starts from
// a zero guess.  Set/add calls and assembly are host-only (a stash of
// (row, col, value) triples reduced into CSR at assembly), so the state
// machine needs no GPU.  The first solve after an assembly uploads the CSR to
// the GPU (a BSR of 1x1 blocks) and solves there with FP64 FGMRES.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hfx_hip.h"
#include "hfx_internal.h"
#include "ilu.h"
#include "krylov.h"

namespace hfx {
void set_last_error(const std::string& msg);
}

namespace {

struct Triple {
  long long key;  // row * n + col
  long long seq;  // arrival order (INSERT keeps the last value)
  double v;
};

struct LaError : std::runtime_error {
  LaError(const char* fn, const char* msg) : std::runtime_error(std::string("HipLinAlgebra::") + fn + ": " + msg) {}
};

}  // namespace

struct hfx_la {
  int device = 0;
  hipStream_t stream = nullptr;
  bool initialized = false, configured = false, allocated = false, assembled = false;
  int n = 0;
  int pc = HFX_PC_JACOBI;
  double rtol = 1e-6;
  int maxits = 1000, restart = 30;
  int solver = HFX_LA_GMRES;
  // assembled matrix: CSR with sorted columns, and the RHS
  std::vector<int> rowptr = std::vector<int>(1, 0), col;
  std::vector<double> val, rhs;
  // values not yet assembled: mode 0 none, 1 ADD, 2 INSERT
  int mode = 0;
  std::vector<Triple> stash;
  std::vector<Triple> rstash;
  long long seq = 0;
  // device copy (uploaded by the first solve after a change)
  hfx::BsrDev A;
  hfx::DevBuf<double> db, dx, dinv;
  bool dev_dirty = true;
  std::unique_ptr<hfx::Ilu> ilu;  // point ILU(0) of the uploaded matrix
  bool ilu_stale = true;
  int sol_low = 0, sol_high = 0;

  // Reduce the stash into the CSR: ADD sums every contribution, INSERT keeps
  // the last one; entries outside the pattern are created.
  void apply_stash() {
    if (mode == 0) return;
    auto by_key = [](const Triple& a, const Triple& b) { return a.key != b.key ? a.key < b.key : a.seq < b.seq; };
    std::sort(stash.begin(), stash.end(), by_key);
    std::vector<Triple> red;
    red.reserve(stash.size());
    for (size_t i = 0; i < stash.size();) {
      size_t j = i;
      double v = 0.0;
      for (; j < stash.size() && stash[j].key == stash[i].key; ++j) v = mode == 2 ? stash[j].v : v + stash[j].v;
      red.push_back({stash[i].key, 0, v});
      i = j;
    }
    std::vector<int> nrp(n + 1, 0), ncol;
    std::vector<double> nval;
    ncol.reserve(col.size() + red.size());
    nval.reserve(col.size() + red.size());
    size_t k = 0;
    for (int r = 0; r < n; ++r) {
      int a = rowptr[r];
      const int b = rowptr[r + 1];
      while (a < b || (k < red.size() && red[k].key / n == r)) {
        const bool new_here = k < red.size() && red[k].key / n == r;
        const int cnew = new_here ? (int)(red[k].key % n) : 0;
        if (a < b && (!new_here || col[a] < cnew)) {
          ncol.push_back(col[a]);
          nval.push_back(val[a]);
          ++a;
        } else if (a < b && col[a] == cnew) {
          ncol.push_back(cnew);
          nval.push_back(mode == 2 ? red[k].v : val[a] + red[k].v);
          ++a;
          ++k;
        } else {
          ncol.push_back(cnew);
          nval.push_back(red[k].v);
          ++k;
        }
      }
      nrp[r + 1] = (int)ncol.size();
    }
    rowptr.swap(nrp);
    col.swap(ncol);
    val.swap(nval);
    std::sort(rstash.begin(), rstash.end(), by_key);
    for (const Triple& t : rstash) {
      if (mode == 2) rhs[t.key] = t.v;
      else rhs[t.key] += t.v;
    }
    stash.clear();
    rstash.clear();
    mode = 0;
    dev_dirty = true;
  }

  void begin(int m, const char* fn) {
    if (mode != 0 && mode != m)
      throw LaError(fn, "ADD and INSERT values cannot be mixed before assembleFlush or assemble.");
    mode = m;
  }

  void put(int m, const char* fn, int nr, const int* rows, int nc, const int* cols, const double* vals) {
    if (!allocated)
      throw LaError(fn, m == 1 ? "the matrix must be allocated before adding values."
                               : "the matrix must be allocated before setting values.");
    if (nr < 0 || nc < 0 || ((long long)nr * nc > 0 && (!rows || !cols || !vals)))
      throw LaError(fn, "bad index or value arrays");
    for (int i = 0; i < nr; ++i)
      if (rows[i] < 0 || rows[i] >= n) throw LaError(fn, "row index out of range");
    for (int j = 0; j < nc; ++j)
      if (cols[j] < 0 || cols[j] >= n) throw LaError(fn, "column index out of range");
    begin(m, fn);
    for (int i = 0; i < nr; ++i)
      for (int j = 0; j < nc; ++j) stash.push_back({(long long)rows[i] * n + cols[j], seq++, vals[(size_t)i * nc + j]});
  }

  void put_rhs(int m, const char* fn, int nr, const int* rows, const double* vals) {
    if (!allocated)
      throw LaError(fn, m == 1 ? "the RHS must be allocated before adding values."
                               : "the RHS must be allocated before setting values.");
    if (nr < 0 || (nr > 0 && (!rows || !vals))) throw LaError(fn, "bad index or value arrays");
    for (int i = 0; i < nr; ++i)
      if (rows[i] < 0 || rows[i] >= n) throw LaError(fn, "row index out of range");
    begin(m, fn);
    for (int i = 0; i < nr; ++i) rstash.push_back({rows[i], seq++, vals[i]});
  }

  void upload() {
    HFX_HIP(hipSetDevice(device));
    if (!stream) HFX_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    if (!dev_dirty) return;
    std::vector<int> rowf(col.size());
    std::vector<double> di(n, 1.0);  // point Jacobi; rows without a nonzero diagonal use 1
    for (int r = 0; r < n; ++r)
      for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
        rowf[k] = r;
        if (col[k] == r && val[k] != 0.0) di[r] = 1.0 / val[k];
      }
    A.nb = n;
    A.bs = 1;
    A.nnzb = (int)col.size();
    A.rowptr.upload(rowptr, stream);
    A.col.upload(col, stream);
    A.rowf.upload(rowf, stream);
    A.vals.upload(val, stream);
    db.upload(rhs, stream);
    dinv.upload(di, stream);
    dx.alloc(n);
    HFX_HIP(hipStreamSynchronize(stream));
    dev_dirty = false;
    ilu_stale = true;
  }
};

#define LA_TRY try {
#define LA_END                           \
  }                                      \
  catch (const std::exception& e) {      \
    hfx::set_last_error(e.what());       \
    return -1;                           \
  }                                      \
  return 0;

#define LA_HANDLE(h) \
  if (!(h)) throw std::runtime_error("HipLinAlgebra: null handle");

extern "C" int hfx_la_create(int device, hfx_la** out) {
  LA_TRY
  if (!out) throw std::runtime_error("HipLinAlgebra: null output pointer");
  if (device < 0) throw std::runtime_error("HipLinAlgebra: negative device");
  auto h = std::unique_ptr<hfx_la>(new hfx_la());
  h->device = device;
  *out = h.release();
  LA_END
}

extern "C" void hfx_la_free(hfx_la* h) {
  if (!h) return;
  if (h->stream) {
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    (void)hipStreamDestroy(h->stream);
  }
  delete h;
}

extern "C" int hfx_la_initialize(hfx_la* h) {
  LA_TRY
  LA_HANDLE(h);
  h->initialized = true;
  LA_END
}

extern "C" int hfx_la_configure(hfx_la* h, int solver_type, int pc, double rtol, int maxits, int restart) {
  LA_TRY
  LA_HANDLE(h);
  if (!h->initialized) throw LaError("configure", "the object was not initialized before being configured.");
  if (solver_type != HFX_LA_GMRES && solver_type != HFX_LA_BCGS)
    throw LaError("configure", "the solver type must be GMRES or BiCGStab");
  if (pc != HFX_PC_NONE && pc != HFX_PC_JACOBI && pc != HFX_PC_ILU0 && pc != HFX_PC_ILU0_NATURAL)
    throw LaError("configure", "the preconditioner must be none, Jacobi or ILU(0)");
  if (!(rtol > 0.0) || maxits <= 0 || restart < 1 || restart > 63)
    throw LaError("configure", "rtol must be > 0, maxits > 0 and restart in [1, 63]");
  h->solver = solver_type;
  h->pc = pc;
  h->rtol = rtol;
  h->maxits = maxits;
  h->restart = restart;
  h->configured = true;
  LA_END
}

extern "C" int hfx_la_allocate(hfx_la* h, int ndofs, const int* diag_nnz, const int* off_nnz) {
  LA_TRY
  LA_HANDLE(h);
  if (!(h->initialized && h->configured))
    throw LaError("allocate", "the object was not initialized and configured before being allocated.");
  if (ndofs < 0) throw LaError("allocate", "negative number of dofs");
  if ((diag_nnz == nullptr) != (off_nnz == nullptr))
    throw LaError("allocate", "diagonal and off-diagonal patterns must be passed together");
  h->n = ndofs;
  h->rowptr.assign((size_t)ndofs + 1, 0);
  h->col.clear();
  h->val.clear();
  h->stash.clear();
  h->rstash.clear();
  if (diag_nnz) {
    size_t tot = 0;
    for (int i = 0; i < ndofs; ++i) tot += (size_t)std::max(0, diag_nnz[i]) + (size_t)std::max(0, off_nnz[i]);
    h->stash.reserve(tot);
  }
  h->rhs.assign(ndofs, 0.0);
  h->mode = 0;
  h->allocated = true;
  h->assembled = false;
  h->dev_dirty = true;
  LA_END
}

extern "C" int hfx_la_add_vals(hfx_la* h, int nr, const int* rows, int nc, const int* cols, const double* vals) {
  LA_TRY
  LA_HANDLE(h);
  h->put(1, "addValsMatrix", nr, rows, nc, cols, vals);
  LA_END
}

extern "C" int hfx_la_set_vals(hfx_la* h, int nr, const int* rows, int nc, const int* cols, const double* vals) {
  LA_TRY
  LA_HANDLE(h);
  h->put(2, "setValsMatrix", nr, rows, nc, cols, vals);
  LA_END
}

extern "C" int hfx_la_add_rhs(hfx_la* h, int nr, const int* rows, const double* vals) {
  LA_TRY
  LA_HANDLE(h);
  h->put_rhs(1, "addValsRHS", nr, rows, vals);
  LA_END
}

extern "C" int hfx_la_set_rhs(hfx_la* h, int nr, const int* rows, const double* vals) {
  LA_TRY
  LA_HANDLE(h);
  h->put_rhs(2, "setValsRHS", nr, rows, vals);
  LA_END
}

extern "C" int hfx_la_zero_rows(hfx_la* h, int nr, const int* rows) {
  LA_TRY
  LA_HANDLE(h);
  if (!h->allocated) throw LaError("zeroOutRows", "the matrix must be allocated before zeroing out rows.");
  if (nr < 0 || (nr > 0 && !rows)) throw LaError("zeroOutRows", "bad row array");
  for (int i = 0; i < nr; ++i)
    if (rows[i] < 0 || rows[i] >= h->n) throw LaError("zeroOutRows", "row index out of range");
  h->apply_stash();
  for (int i = 0; i < nr; ++i)
    for (int k = h->rowptr[rows[i]]; k < h->rowptr[rows[i] + 1]; ++k) h->val[k] = 0.0;
  h->dev_dirty = true;
  LA_END
}

extern "C" int hfx_la_assemble_flush(hfx_la* h) {
  LA_TRY
  LA_HANDLE(h);
  if (!h->allocated) throw LaError("assembleFlush", "the matrix should at least be allocated before assembling.");
  h->apply_stash();
  h->assembled = true;
  LA_END
}

extern "C" int hfx_la_assemble(hfx_la* h) {
  LA_TRY
  LA_HANDLE(h);
  if (!h->allocated) throw LaError("assemble", "the matrix should at least be allocated before assembling.");
  h->apply_stash();
  h->assembled = true;
  LA_END
}

extern "C" int hfx_la_solve(hfx_la* h, double* x, int n, int* info, double* relres) {
  LA_TRY
  LA_HANDLE(h);
  if (!h->assembled) throw LaError("solve", "the matrix needs to be assembled before solving.");
  if (n < h->n || (h->n > 0 && !x)) throw LaError("solve", "the solution buffer is shorter than the number of dofs");
  h->apply_stash();
  hfx::KrylovInfo ki;
  if (h->n == 0) {
    ki.reason = hfx::CONVERGED_ATOL;
  } else {
    h->upload();
    hipStream_t s = h->stream;
    HFX_HIP(hipMemsetAsync(h->dx.p, 0, (size_t)h->n * sizeof(double), s));
    std::function<void(const double*, double*)> M;
    if (h->pc == HFX_PC_JACOBI) {
      const double* dinv = h->dinv.p;
      const long long nn = h->n;
      M = [dinv, nn, s](const double* in, double* out) { hfx::apply_point_diag(nn, dinv, in, out, s); };
    } else if (h->pc == HFX_PC_ILU0 || h->pc == HFX_PC_ILU0_NATURAL) {
      const int ord = h->pc == HFX_PC_ILU0 ? hfx::ILU_MULTICOLOUR : hfx::ILU_NATURAL;
      if (h->ilu_stale || !h->ilu || h->ilu->ordering != ord) {
        h->ilu.reset(new hfx::Ilu());
        try {
          h->ilu->setup(h->n, 1, h->rowptr.data(), h->col.data(), ord, s);
          h->ilu->factor(h->A, s);
        } catch (const std::exception& e) {
          h->ilu.reset();
          throw LaError("solve", e.what());
        }
        h->ilu_stale = false;
      }
      hfx::Ilu* f = h->ilu.get();
      M = [f, s](const double* in, double* out) { f->apply(in, out, s); };
    }
    hfx::KrylovOpts o;
    o.rtol = h->rtol;
    o.maxits = h->maxits;
    o.restart = h->restart;
    if (h->solver == HFX_LA_BCGS) hfx::bicgstab(h->A, h->db.p, h->dx.p, M, o, ki, s);
    else hfx::fgmres(h->A, h->db.p, h->dx.p, M, o, ki, s);
    h->dx.download(x, h->n, s);
    HFX_HIP(hipStreamSynchronize(s));
  }
  h->sol_low = 0;
  h->sol_high = h->n;
  if (info) {
    info[0] = ki.iters;
    info[1] = ki.reason;
  }
  if (relres) relres[0] = ki.true_relres;
  LA_END
}

extern "C" int hfx_la_get_values(hfx_la* h, int nr, const int* rows, int nc, const int* cols, double* out) {
  LA_TRY
  LA_HANDLE(h);
  if (!h->assembled || h->mode != 0)
    throw LaError("getValues", "the matrix needs to be assembled before reading values.");
  for (int i = 0; i < nr; ++i)
    if (rows[i] < 0 || rows[i] >= h->n) throw LaError("getValues", "row index out of range");
  for (int j = 0; j < nc; ++j)
    if (cols[j] < 0 || cols[j] >= h->n) throw LaError("getValues", "column index out of range");
  for (int i = 0; i < nr; ++i) {
    const int* b = h->col.data() + h->rowptr[rows[i]];
    const int* e = h->col.data() + h->rowptr[rows[i] + 1];
    for (int j = 0; j < nc; ++j) {
      const int* p = std::lower_bound(b, e, cols[j]);
      out[(size_t)i * nc + j] = (p != e && *p == cols[j]) ? h->val[p - h->col.data()] : 0.0;
    }
  }
  LA_END
}

extern "C" int hfx_la_get_rhs(hfx_la* h, int nr, const int* rows, double* out) {
  LA_TRY
  LA_HANDLE(h);
  if (!h->assembled || h->mode != 0) throw LaError("getRHS", "the RHS needs to be assembled before reading values.");
  for (int i = 0; i < nr; ++i)
    if (rows[i] < 0 || rows[i] >= h->n) throw LaError("getRHS", "row index out of range");
  for (int i = 0; i < nr; ++i) out[i] = h->rhs[rows[i]];
  LA_END
}

extern "C" int hfx_la_get_csr(hfx_la* h, long long* rowptr, int* cols, double* vals, double* rhs) {
  LA_TRY
  LA_HANDLE(h);
  if (!h->assembled || h->mode != 0) throw LaError("getCSR", "the matrix needs to be assembled first.");
  if (rowptr)
    for (int r = 0; r <= h->n; ++r) rowptr[r] = h->rowptr[r];
  if (cols) std::copy(h->col.begin(), h->col.end(), cols);
  if (vals) std::copy(h->val.begin(), h->val.end(), vals);
  if (rhs) std::copy(h->rhs.begin(), h->rhs.end(), rhs);
  LA_END
}

extern "C" int hfx_la_nnz(hfx_la* h, long long* nnz) {
  LA_TRY
  LA_HANDLE(h);
  if (!h->assembled || h->mode != 0) throw LaError("nnz", "the matrix needs to be assembled first.");
  *nnz = (long long)h->col.size();
  LA_END
}

extern "C" int hfx_la_ownership(hfx_la* h, int* low, int* high) {
  LA_TRY
  LA_HANDLE(h);
  if (!h->assembled)
    throw LaError("getSolutionOwnership", "the matrix needs to be assembled before querying ownership.");
  *low = h->sol_low;
  *high = h->sol_high;
  LA_END
}

extern "C" int hfx_la_clear(hfx_la* h) {
  LA_TRY
  LA_HANDLE(h);
  if (!h->assembled) throw LaError("clearSystem", "the matrix needs to be assembled at least once before clearing.");
  h->stash.clear();
  h->rstash.clear();
  h->mode = 0;
  std::fill(h->val.begin(), h->val.end(), 0.0);
  std::fill(h->rhs.begin(), h->rhs.end(), 0.0);
  h->dev_dirty = true;
  LA_END
}

extern "C" int hfx_la_destroy_system(hfx_la* h) {
  LA_TRY
  LA_HANDLE(h);
  h->initialized = h->configured = h->allocated = h->assembled = false;
  h->n = 0;
  h->rowptr.assign(1, 0);
  h->col.clear();
  h->val.clear();
  h->rhs.clear();
  h->stash.clear();
  h->rstash.clear();
  h->mode = 0;
  h->dev_dirty = true;
  h->sol_low = h->sol_high = 0;
  LA_END
}

extern "C" int hfx_la_num_dofs(hfx_la* h, int* n) {
  LA_TRY
  LA_HANDLE(h);
  *n = h->n;
  LA_END
}
