// Reference simplex element for the HDG hot path: nodes, face-node maps,
// cubature and Lagrange shape-function tables.  Host-side set-up code; the
// tables are uploaded once per (dim, order) and read by the element kernels.
//
// Follows ReferenceElement (src/element/ReferenceElement.cpp:5-26) and
// Cubature (src/element/Cubature.cpp:52-59) of the reference.
#pragma once

#include <memory>
#include <vector>

namespace hfx {

// Cubature rule on the bi-unit reference simplex of dimension dim.
struct Cubature {
  int dim = 0, order = 0, nip = 0;
  std::vector<double> pts;  // nip x dim
  std::vector<double> w;    // nip
};

// Simplex rule exact to polynomial degree `order`.  dim 1: Gauss-Legendre with
// ceil((order+1)/2) points (Cubature.cpp:106-112); dim 2/3: the
// Witherden-Vincent rule the reference maps `order` to (Cubature.cpp:292-2731).
Cubature make_cubature(int dim, int order);

// 1-D Gauss-Lobatto-Legendre points, in the reference's node order
// [-1, 1, interior ascending] (ReferenceElement.cpp:614-760).
std::vector<double> gll_nodes_ref_order(int p);

// Node set of the order-p simplex in the reference ordering: vertices, then
// the nodes of each sub-entity, then interior nodes (ReferenceElement.cpp:1045-1145).
std::vector<double> simplex_nodes(int dim, int p);

struct RefEl {
  int dim = 0, order = 0, nN = 0, nF = 0;
  std::vector<double> nodes;                 // nN x dim
  std::vector<std::vector<int>> faceNodes;   // nF x nNf, reference face order
  Cubature cub;                              // order 2p (ReferenceElement.cpp:102-106)
  std::vector<double> phi;                   // nip x nN
  std::vector<double> dphi;                  // nip x nN x dim  (d/dxi)
  std::unique_ptr<RefEl> face;               // (dim-1)-simplex of the same order

  RefEl(int dim, int order);
  int nip() const { return cub.nip; }
  int nNf() const { return face ? face->nN : 1; }

  // Lagrange basis (and reference gradient) at an arbitrary reference point.
  void eval(const double* xi, double* phi_out, double* dphi_out /* nN x dim, may be null */) const;

 private:
  std::vector<double> vinv_;  // inverse Vandermonde of the modal basis, nN x nN
  std::vector<int> modes_;    // nN x dim mode multi-indices
};

// Number of nodes of the order-p simplex (ReferenceElement.cpp:69-94).
int simplex_num_nodes(int dim, int p);

}  // namespace hfx
