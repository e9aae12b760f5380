// Mesh hierarchy of the default ParMooN geometries, reproducing the cell order
// of the reference's refinement tree.
//
// Reference behaviour restated here:
//  * initial macro meshes: TDomain::UnitSquare (src/Geometry/Tests.C:449-482),
//    TDomain::TwoTriangles (src/Geometry/Tests.C:95-170) and the default
//    unit-cube hexahedron read by TDomain::initialize_cube_hexa_mesh
//    (src/Geometry/Tests.C:1815-1856) whose parametrised vertices are mapped by
//    the planes of initializeDefaultCubeBdry (src/Geometry/Tests.C:1766-1812,
//    TBdPlane::GetXYZofTS src/Geometry/BdPlane.C:46-54);
//  * regular refinement with the child-vertex tables of TRefQuadRegDesc
//    (src/Refinement/RefQuadRegDesc.C:20-21), TRefTriRegDesc
//    (src/Refinement/RefTriRegDesc.C:20-21) and TRefHexaRegDesc
//    (src/Refinement/RefHexaRegDesc.C:31-35);
//  * the finest collection is produced by the depth-first iterator It_Finest
//    (src/Geometry/It_LE.C:17-56), so the cells of level k+1 are the children
//    of the level-k cells, parent by parent, in child order.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace pmx {

enum Shape : int { SHAPE_TRI = 0, SHAPE_QUAD = 1, SHAPE_HEXA = 2 };

enum Geometry : int {
  GEO_UNIT_SQUARE_QUAD = 0,   // "UnitSquare"
  GEO_UNIT_SQUARE_TRI = 1,    // "TwoTriangles"
  GEO_UNIT_CUBE_HEXA = 2      // "Default_UnitCube_Hexa"
};

struct MeshLevel {
  int n_cells = 0;
  std::vector<double> xyz;     // [n_cells][nv][dim], local vertex order of each cell
  std::vector<int> parent;     // index of the parent cell in the previous level
  std::vector<int> child_no;   // child index within the parent
};

struct Mesh {
  int geometry = 0;
  int dim = 0;
  Shape shape = SHAPE_QUAD;
  int nv = 0;                  // vertices per cell
  int n_faces = 0;             // joints per cell (edges in 2D, faces in 3D)
  int base_n = 1;              // level-0 vertices lie on the grid (1/base_n) Z^d
  std::vector<MeshLevel> levels;   // levels[0] = macro mesh

  const double* vertex(int level, int cell, int v) const {
    return levels[level].xyz.data() + ((size_t)cell * nv + v) * dim;
  }
};

// Build the macro mesh and refine it n_refine times uniformly.
Mesh build_mesh(int geometry, int n_refine);

// Single-level triangulation of the unit square with n x n squares, each cut
// along its (0,0)-(1,1) diagonal into (v00, v10, v11) and (v00, v11, v01) as
// the two macro triangles of TwoTriangles; squares row by row from y = 0.
// For n = 2^k the mesh covers the same triangles as TwoTriangles refined k
// times, but in another cell order (so another DOF numbering).
Mesh build_structured_tri_mesh(int n);

// Local vertices of joint j of a cell (ParMooN shape descriptors).  Returns
// the number of vertices of that joint.
int joint_vertices(Shape s, int j, int* verts);

// Boundary detection for the unit square / unit cube: returns the boundary
// component of joint j of a cell, or -1 for an interior joint.  2D components
// follow Default_UnitSquare (0: y=0, 1: x=1, 2: y=1, 3: x=0); in 3D the
// component is the face of the cube (0: z=0, 1: y=0, 2: x=1, 3: y=1, 4: x=0,
// 5: z=1).
int joint_boundary_component(const Mesh& m, int level, int cell, int j);

// Parameter t of a point on 2D boundary component comp
// (TBdLine of Default_UnitSquare: comp 0 (t,0), 1 (1,t), 2 (1-t,1), 3 (0,1-t)).
double boundary_param_2d(int comp, double x, double y);

}  // namespace pmx
