// Mesh topology (faces, cell2face, face2cell, boundary, face_perm) built on
// the GPU, in place of the host hash-map pass of build_topology (the
// reference's MOAB computeFaces, src/mesh/Mesh.cpp:183-537).  The arrays are
// the host builder's exactly: faces numbered by first appearance over (cell,
// local face), face2cell = {lowest cell, other or -1}, a face's node list in
// its first cell's local order, boundary ascending.
#pragma once

#include "mesh.h"
#include "refel.h"

namespace hfx {

// Fill m.faces, cell2face, face2cell, boundary, face_perm and nFaces from
// m.cells on GPU `device`.  3-D meshes need fewer than 2^21 vertices.
// Throws on a face shared by more than two cells.
void build_topology_gpu(const RefEl& re, Mesh& m, int device);

}  // namespace hfx
