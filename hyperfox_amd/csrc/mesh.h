// Simplex mesh topology for the HDG trace system, and the synthetic
// structured meshes the benchmarks run on.  Host-side set-up code.
//
// Conventions follow the reference's Mesh (src/mesh/Mesh.cpp:183-537):
//  * cells hold global node ids in the reference element's node order;
//  * cell2face[e][lf] is the face matching the reference local face lf;
//  * face2cell[f] = {lowest cell, other cell or -1 on the boundary};
//  * a face's node list follows its first cell's local face-node order.
#pragma once

#include <cstdint>
#include <vector>

#include "refel.h"

namespace hfx {

struct Mesh {
  int dim = 0, order = 0;
  int nN = 0, nNf = 0, nFpc = 0;       // nodes per cell, nodes per face, faces per cell
  int nNodes = 0, nCells = 0, nFaces = 0;
  std::vector<double> coords;          // nNodes x dim
  std::vector<int> cells;              // nCells x nN
  std::vector<int> faces;              // nFaces x nNf
  std::vector<int> cell2face;          // nCells x nFpc
  std::vector<int> face2cell;          // nFaces x 2
  std::vector<int> boundary;           // boundary face ids, ascending
  // face_perm[(e*nFpc + lf)*nNf + j] = position, in the face's node list, of
  // the cell's j-th local node on local face lf (face2CellMap of
  // src/solver/HDGSolver.cpp:258-275).
  std::vector<int8_t> face_perm;
};

// Derive faces, cell2face, face2cell, boundary and face_perm from cells.
// Throws on a non-conforming input (a face shared by more than two cells).
void build_topology(const RefEl& re, Mesh& m);

// Unit square / cube split into simplices with seeded jitter of the interior
// vertices by U(-jit*h, jit*h) per coordinate (splitmix64 stream from seed).
// dim 2: n x n squares, two triangles each (diagonal (i,j)-(i+1,j+1));
// dim 3: n^3 cubes, six tetrahedra around the (i,j,k)-(i+1,j+1,k+1) diagonal.
// High-order nodes are placed by the affine map of the reference nodes
// (tools/convertGmsh2H5HO.cpp:367-397) and shared between cells.
// topology = false leaves the face arrays for build_topology_gpu.
Mesh structured_mesh(const RefEl& re, int n, double jit, uint64_t seed, bool topology = true);

// Mesh from caller-supplied high-order cells and node coordinates.
Mesh mesh_from_cells(const RefEl& re, int nNodes, const double* coords, int nCells, const int* cells,
                     bool topology = true);

uint64_t splitmix64(uint64_t& state);
double splitmix_uniform(uint64_t& state);  // U[0,1)

}  // namespace hfx
