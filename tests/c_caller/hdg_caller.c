/* A plain C caller of the bulk HDG path of libhfx_hip.so, making its calls in
 * the order a HipHDGSolver adapter would (HDGSolver::setMesh, setFieldMap,
 * allocate, assemble and solve; src/solver/HDGSolver.cpp:5-779): structured
 * mesh -> problem -> model -> fields -> IP coordinates -> source and
 * reaction -> assemble -> BiCGStab + block Jacobi solve -> solution fields.
 *
 * Writes the inputs it chose and the results it got to a binary file (ints,
 * then doubles, native endianness) so that the test can rebuild the same
 * problem in the FP64 oracle and compare.  Exit status 0 on success. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "hfx_hip.h"

#define CHECK(call)                                                          \
  do {                                                                       \
    if ((call) != 0) {                                                       \
      fprintf(stderr, "%s failed: %s\n", #call, hfx_last_error());           \
      return 1;                                                              \
    }                                                                        \
  } while (0)

static int put(FILE* f, const double* a, size_t n) { return fwrite(a, sizeof(double), n, f) == n ? 0 : 1; }

int main(int argc, char** argv) {
  const char* path = argc > 1 ? argv[1] : "hdg_caller.bin";
  const int dim = 3, order = 2, n = 3;
  const unsigned long long seed = 77ULL;
  const double jitter = 0.05;
  hfx_mesh* m = NULL;
  hfx_hdg* h = NULL;
  long long sz[5];
  int rs[7];
  CHECK(hfx_mesh_structured(dim, order, n, jitter, seed, &m));
  CHECK(hfx_mesh_sizes(m, sz));
  CHECK(hfx_refel_sizes(dim, order, rs));
  const int nNodes = (int)sz[0], nCells = (int)sz[1], nFaces = (int)sz[2];
  const long long ntrace = sz[4];
  const int nN = rs[0], nNf = rs[2], nIP = rs[3];
  double* coords = malloc(sizeof(double) * nNodes * dim);
  int* faces = malloc(sizeof(int) * nFaces * nNf);
  CHECK(hfx_mesh_get(m, coords, NULL, faces, NULL, NULL, NULL, NULL));

  /* HDGConvectionDiffusionReactionSource with Tau, DiffusionTensor,
   * Velocity and Dirichlet fields, a source and a reaction */
  double* tau = malloc(sizeof(double) * nFaces * nNf);
  double* D = malloc(sizeof(double) * nNodes);
  double* b = malloc(sizeof(double) * nNodes * dim);
  double* g = malloc(sizeof(double) * nFaces * nNf);
  for (int F = 0; F < nFaces; ++F)
    for (int j = 0; j < nNf; ++j) {
      const double* x = coords + (size_t)faces[F * nNf + j] * dim;
      tau[F * nNf + j] = 1.0 + 0.25 * (F % 3);
      g[F * nNf + j] = x[0] + 2.0 * x[1] - x[2] + 0.5 * x[0] * x[1];
    }
  for (int i = 0; i < nNodes; ++i) {
    const double* x = coords + (size_t)i * dim;
    D[i] = 0.5 + 0.1 * sin(3.0 * x[0]);
    b[i * dim + 0] = 0.3 + 0.1 * x[1];
    b[i * dim + 1] = -0.2 + 0.1 * x[2];
    b[i * dim + 2] = 0.1 * cos(x[0]);
  }
  CHECK(hfx_hdg_create(0, dim, order, &h));
  CHECK(hfx_hdg_set_mesh(h, m));
  CHECK(hfx_hdg_set_model(h, HFX_MODEL_CDR_SOURCE));
  CHECK(hfx_hdg_set_fields(h, tau, 1, D, 1, b, g));
  double* xip = malloc(sizeof(double) * nCells * nIP * dim);
  double* fip = malloc(sizeof(double) * nCells * nIP);
  double* rip = malloc(sizeof(double) * nCells * nIP);
  CHECK(hfx_hdg_ip_coords(h, xip));
  for (int q = 0; q < nCells * nIP; ++q) {
    const double* x = xip + (size_t)q * dim;
    fip[q] = cos(x[0] + x[1]) * exp(0.5 * x[2]);
    rip[q] = 0.3 + 0.2 * x[0] * x[2];
  }
  CHECK(hfx_hdg_set_source(h, fip, rip));
  CHECK(hfx_hdg_assemble(h));
  CHECK(hfx_hdg_set_krylov(h, HFX_KRYLOV_BICGSTAB));
  int info[2];
  double relres = 0.0;
  CHECK(hfx_hdg_solve(h, HFX_PC_BJACOBI, 1e-13, 20000, 30, info, &relres));
  double* u = malloc(sizeof(double) * nCells * nN);
  double* q = malloc(sizeof(double) * nCells * nN * dim);
  double* lam = malloc(sizeof(double) * ntrace);
  CHECK(hfx_hdg_get_solution(h, u, q, lam));

  FILE* f = fopen(path, "wb");
  if (!f) {
    perror(path);
    return 1;
  }
  const int head[12] = {dim, order, n, (int)seed, nNodes, nCells, nFaces, (int)ntrace, nN, nIP, info[0], info[1]};
  int bad = fwrite(head, sizeof(int), 12, f) != 12;
  bad |= put(f, &jitter, 1) | put(f, &relres, 1) | put(f, tau, (size_t)nFaces * nNf) | put(f, D, nNodes) |
         put(f, b, (size_t)nNodes * dim) | put(f, g, (size_t)nFaces * nNf) | put(f, fip, (size_t)nCells * nIP) |
         put(f, rip, (size_t)nCells * nIP) | put(f, u, (size_t)nCells * nN) | put(f, q, (size_t)nCells * nN * dim) |
         put(f, lam, (size_t)ntrace);
  bad |= fclose(f) != 0;
  hfx_hdg_destroy(h);
  hfx_mesh_free(m);
  free(coords); free(faces); free(tau); free(D); free(b); free(g); free(xip); free(fip); free(rip);
  free(u); free(q); free(lam);
  if (bad) {
    fprintf(stderr, "writing %s failed\n", path);
    return 1;
  }
  printf("hdg_caller: %d cells, %lld trace DOFs, %d iterations, reason %d, true relres %.2e\n", nCells, ntrace,
         info[0], info[1], relres);
  return 0;
}
