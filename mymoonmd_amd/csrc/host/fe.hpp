// Finite-element tables: conforming Lagrange elements on the reference cells
// and the Gauss rules the reference selects for them.
//
// Reference behaviour restated here:
//  * local DOF order of Q_k (tensor-product, xi fastest) and P1: nodal points
//    NF_C_Q_Q1_2D.h:10-11, NF_C_H_Q1_3D.h:13-15, NF_C_H_Q2_3D.h:13-21,
//    NF_C_T_P1_2D.h:10-11;
//  * reference maps: x = sum of vertex shape functions with v0 at (-1,-1[,-1]),
//    v1 at (+1,-1..), v2 at (+1,+1..), v3 at (-1,+1..), v4.. with zeta=+1
//    (THexaTrilinear::SetCell, src/FE/HexaTrilinear.C:806-813; affine maps use
//    v0, v1, v3, v4: src/FE/HexaAffin.C:858-890); triangles x = v0 +
//    (v1-v0) xi + (v2-v0) eta (TTriaAffin);
//  * quadrature chosen from degree N_terms * polynomial_degree with N_terms = 3
//    (src/FE/FEDatabase2D.C:1771-1777, src/FE/FEDatabase3D.C:1149-1163):
//    Q1 quads -> Gauss3Quad, P1 triangles -> Gauss3Tria (7 points),
//    Q1 hexa -> Gauss3Hexa, Q2 hexa -> Gauss4Hexa (QFHexaFromDegree,
//    src/FE/FEDatabase3D.C:976-995); INTERNAL_QUAD_RULE=2 forces Gauss3Hexa
//    (src/FE/FEDatabase3D.C:1201-1206).
#pragma once
#include <vector>

#include "mesh.hpp"

namespace pmx {

struct Element {
  Shape shape;
  int order;        // polynomial order (1 or 2)
  int dim;
  int ndof;
  std::vector<double> ref_nodes;   // [ndof][dim] nodal points in reference coordinates
  // discontinuous P1 on hexahedra (D_P1_3D_H_A): basis 1, xi, eta, zeta
  // (include/FE3D/BF_D_H_P1_3D.h), every DOF inner to its cell, nodal
  // functionals from the 8 points (+-1/sqrt 3)^3 (include/FE3D/NF_D_H_P1_3D.h)
  bool disc = false;

  // basis values at one reference point:
  //  v[ndof], d1[dim][ndof] (d/dxi, d/deta, d/dzeta),
  //  d2[ncomb][ndof] with ncomb = 3 (2D: xx, xy, yy) or 6 (3D: xx, xy, xz, yy, yz, zz)
  void eval(const double* xi, double* v, double* d1, double* d2) const;

  // local DOFs lying on joint j (edge in 2D, face in 3D)
  std::vector<int> joint_dofs(int j) const;
};

Element make_element(Shape shape, int order);
Element make_element_p1disc_hexa();

// Nodal functionals of the discontinuous P1 hexahedron: F0 = mean of the
// values at the 8 points (+-1/sqrt 3)^3 (xi fastest), F1..F3 = 0.375 sum of
// value x the point's xi / eta / zeta (NF_D_H_P1_3D_EvalAll).
void p1disc_nf_points(double* xi24);
void p1disc_nf_apply(const double* values8, double* out4);

struct QuadRule {
  int n = 0;
  int dim = 0;
  std::vector<double> xi;   // [n][dim]
  std::vector<double> w;    // weights on the reference cell
};

// Gauss rule with npts points per direction on [-1,1]^dim (tensor, xi fastest)
QuadRule gauss_tensor(int dim, int npts);
// 7-point degree-5 rule on the reference triangle (Radon's formula)
QuadRule gauss3_tria();
// Rule the reference assembly uses for this element; hexa_rule = 0 follows
// the degree table, 2 forces Gauss3Hexa (INTERNAL_QUAD_RULE == 2).
QuadRule assembly_rule(const Element& e, int hexa_rule);

// 1D Gauss-Legendre nodes/weights on [-1,1] for n = 1..5
void gauss_legendre(int n, double* x, double* w);

// Affine reference map of a cell: x = x0 + J * xi.  For quadrilaterals and
// hexahedra this is the parallelogram/parallelepiped map through v0 and the
// midpoint convention (xi in [-1,1]); for triangles xi in the unit triangle.
struct AffineMap {
  int dim;
  double x0[3];
  double J[9];      // row-major J[r*dim+c] = d x_r / d xi_c
  double Jinv[9];
  double detJ;
};
AffineMap affine_map(const Mesh& m, int level, int cell);
void map_to_phys(const AffineMap& a, const double* xi, double* x);
void map_to_ref(const AffineMap& a, const double* x, double* xi);

// Cell diameter as TBaseCell::GetDiameter (longest vertex-vertex distance;
// THexahedron::GetDiameter src/Geometry/Hexahedron.C:78-104 uses the space
// diagonals, which are the longest distances on the parallelepipeds built here).
double cell_diameter(const Mesh& m, int level, int cell);

}  // namespace pmx
