#include "mesh.hpp"

#include <cmath>
#include <stdexcept>

namespace pmx {

namespace {

// Refined-vertex construction: each new vertex is the mean of a set of parent
// vertices (edge midpoints, face and cell centres).  Indices follow the
// refinement descriptors (RefQuadRegDesc.C, RefTriRegDesc.C, RefHexaRegDesc.C).
struct RefRule {
  int n_new;                 // number of refined vertices
  int n_children;
  int nv;                    // vertices per child
  const int (*mean_of)[8];   // parent vertices averaged for refined vertex i (-1 terminated)
  const int (*child_vertex)[8];
};

// quadrilateral: 0-3 old, 4..7 edge midpoints, 8 centre
const int kQuadMean[9][8] = {
    {0, -1}, {1, -1}, {2, -1}, {3, -1},
    {0, 1, -1}, {1, 2, -1}, {2, 3, -1}, {3, 0, -1},
    {0, 1, 2, 3, -1}};
const int kQuadChild[4][8] = {{0, 4, 8, 7}, {1, 5, 8, 4}, {2, 6, 8, 5}, {3, 7, 8, 6}};

// triangle: 0-2 old, 3 = mid(0,1), 4 = mid(1,2), 5 = mid(2,0)
const int kTriMean[6][8] = {{0, -1}, {1, -1}, {2, -1}, {0, 1, -1}, {1, 2, -1}, {2, 0, -1}};
const int kTriChild[4][8] = {{0, 3, 5}, {1, 4, 3}, {2, 5, 4}, {3, 4, 5}};

// hexahedron: 0-7 old; edge midpoints, face centres and the cell centre as
// given by the vertex-child incidence of RefHexaRegDesc.C:37-48
const int kHexaMean[27][8] = {
    {0, -1}, {1, -1}, {2, -1}, {3, -1}, {4, -1}, {5, -1}, {6, -1}, {7, -1},
    {0, 1, -1},          // 8
    {1, 2, -1},          // 9
    {2, 3, -1},          // 10
    {0, 3, -1},          // 11
    {0, 1, 2, 3, -1},    // 12 bottom face
    {1, 5, -1},          // 13
    {4, 5, -1},          // 14
    {0, 4, -1},          // 15
    {0, 1, 5, 4, -1},    // 16 face y-
    {2, 6, -1},          // 17
    {5, 6, -1},          // 18
    {1, 2, 6, 5, -1},    // 19 face x+
    {3, 7, -1},          // 20
    {6, 7, -1},          // 21
    {2, 3, 7, 6, -1},    // 22 face y+
    {4, 7, -1},          // 23
    {0, 3, 7, 4, -1},    // 24 face x-
    {4, 5, 6, 7, -1},    // 25 top face
    {0, 1, 2, 3, 4, 5, 6, 7}};  // 26 centre (8 entries, no terminator needed)
const int kHexaChild[8][8] = {
    {0, 8, 12, 11, 15, 16, 26, 24}, {1, 9, 12, 8, 13, 19, 26, 16},
    {2, 10, 12, 9, 17, 22, 26, 19}, {3, 11, 12, 10, 20, 24, 26, 22},
    {4, 23, 25, 14, 15, 24, 26, 16}, {5, 14, 25, 18, 13, 16, 26, 19},
    {6, 18, 25, 21, 17, 19, 26, 22}, {7, 21, 25, 23, 20, 22, 26, 24}};

RefRule rule_for(Shape s) {
  switch (s) {
    case SHAPE_QUAD: return {9, 4, 4, kQuadMean, kQuadChild};
    case SHAPE_TRI: return {6, 4, 3, kTriMean, kTriChild};
    case SHAPE_HEXA: return {27, 8, 8, kHexaMean, kHexaChild};
  }
  throw std::runtime_error("unknown shape");
}

MeshLevel refine(const Mesh& m, const MeshLevel& coarse) {
  const RefRule r = rule_for(m.shape);
  MeshLevel fine;
  fine.n_cells = coarse.n_cells * r.n_children;
  fine.xyz.resize((size_t)fine.n_cells * m.nv * m.dim);
  fine.parent.resize(fine.n_cells);
  fine.child_no.resize(fine.n_cells);
  std::vector<double> nv((size_t)r.n_new * 3);
  for (int c = 0; c < coarse.n_cells; ++c) {
    const double* pv = coarse.xyz.data() + (size_t)c * m.nv * m.dim;
    for (int i = 0; i < r.n_new; ++i) {
      int cnt = 0;
      double s[3] = {0, 0, 0};
      for (int k = 0; k < 8 && r.mean_of[i][k] >= 0; ++k) {
        const double* p = pv + r.mean_of[i][k] * m.dim;
        for (int d = 0; d < m.dim; ++d) s[d] += p[d];
        ++cnt;
      }
      for (int d = 0; d < m.dim; ++d) nv[i * 3 + d] = s[d] / cnt;
    }
    for (int ch = 0; ch < r.n_children; ++ch) {
      const int fc = c * r.n_children + ch;
      fine.parent[fc] = c;
      fine.child_no[fc] = ch;
      double* out = fine.xyz.data() + (size_t)fc * m.nv * m.dim;
      for (int v = 0; v < m.nv; ++v)
        for (int d = 0; d < m.dim; ++d) out[v * m.dim + d] = nv[r.child_vertex[ch][v] * 3 + d];
    }
  }
  return fine;
}

}  // namespace

int joint_vertices(Shape s, int j, int* v) {
  static const int quad[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
  static const int tri[3][2] = {{0, 1}, {1, 2}, {2, 0}};
  static const int hexa[6][4] = {{0, 1, 2, 3}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                 {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 7, 6, 5}};
  switch (s) {
    case SHAPE_QUAD: v[0] = quad[j][0]; v[1] = quad[j][1]; return 2;
    case SHAPE_TRI: v[0] = tri[j][0]; v[1] = tri[j][1]; return 2;
    case SHAPE_HEXA: for (int k = 0; k < 4; ++k) v[k] = hexa[j][k]; return 4;
  }
  return 0;
}

int joint_boundary_component(const Mesh& m, int level, int cell, int j) {
  int v[4];
  const int n = joint_vertices(m.shape, j, v);
  if (m.dim == 2) {
    const double* a = m.vertex(level, cell, v[0]);
    const double* b = m.vertex(level, cell, v[1]);
    if (a[1] == 0.0 && b[1] == 0.0) return 0;
    if (a[0] == 1.0 && b[0] == 1.0) return 1;
    if (a[1] == 1.0 && b[1] == 1.0) return 2;
    if (a[0] == 0.0 && b[0] == 0.0) return 3;
    return -1;
  }
  // 3D: all four face vertices share a cube-face coordinate
  static const int axis[6] = {2, 1, 0, 1, 0, 2};
  static const double val[6] = {0, 0, 1, 1, 0, 1};
  for (int f = 0; f < 6; ++f) {
    bool on = true;
    for (int k = 0; k < n && on; ++k) on = (m.vertex(level, cell, v[k])[axis[f]] == val[f]);
    if (on) return f;
  }
  return -1;
}

double boundary_param_2d(int comp, double x, double y) {
  switch (comp) {
    case 0: return x;
    case 1: return y;
    case 2: return 1.0 - x;
    case 3: return 1.0 - y;
  }
  return 0.0;
}

Mesh build_mesh(int geometry, int n_refine) {
  if (n_refine < 0 || n_refine > 12) throw std::runtime_error("n_refine out of range [0,12]");
  Mesh m;
  m.geometry = geometry;
  MeshLevel l0;
  switch (geometry) {
    case GEO_UNIT_SQUARE_QUAD:
      m.dim = 2; m.shape = SHAPE_QUAD; m.nv = 4; m.n_faces = 4;
      l0.n_cells = 1;
      l0.xyz = {0, 0, 1, 0, 1, 1, 0, 1};
      break;
    case GEO_UNIT_SQUARE_TRI:
      // triangle 0: (0,0),(1,0),(1,1); triangle 1: (0,0),(1,1),(0,1)
      m.dim = 2; m.shape = SHAPE_TRI; m.nv = 3; m.n_faces = 3;
      l0.n_cells = 2;
      l0.xyz = {0, 0, 1, 0, 1, 1, 0, 0, 1, 1, 0, 1};
      break;
    case GEO_UNIT_CUBE_HEXA:
      // parametrised vertices (comp,t,s) of initialize_cube_hexa_mesh mapped
      // through the planes of initializeDefaultCubeBdry
      m.dim = 3; m.shape = SHAPE_HEXA; m.nv = 8; m.n_faces = 6;
      l0.n_cells = 1;
      l0.xyz = {0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0, 0, 0, 1, 1, 0, 1, 1, 1, 1, 0, 1, 1};
      break;
    default:
      throw std::runtime_error("unknown geometry");
  }
  l0.parent.assign(l0.n_cells, -1);
  l0.child_no.assign(l0.n_cells, 0);
  m.levels.push_back(std::move(l0));
  for (int r = 0; r < n_refine; ++r) m.levels.push_back(refine(m, m.levels.back()));
  return m;
}

Mesh build_structured_tri_mesh(int n) {
  if (n < 1 || n > 4096) throw std::runtime_error("structured triangle mesh: n out of range [1,4096]");
  Mesh m;
  m.geometry = GEO_UNIT_SQUARE_TRI;
  m.dim = 2; m.shape = SHAPE_TRI; m.nv = 3; m.n_faces = 3;
  m.base_n = n;
  MeshLevel l0;
  l0.n_cells = 2 * n * n;
  l0.xyz.resize((size_t)l0.n_cells * 6);
  const double h = 1.0 / n;
  size_t k = 0;
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) {
      const double x0 = i * h, y0 = j * h;
      const double x1 = (i + 1 == n) ? 1.0 : (i + 1) * h, y1 = (j + 1 == n) ? 1.0 : (j + 1) * h;
      const double t0[6] = {x0, y0, x1, y0, x1, y1};
      const double t1[6] = {x0, y0, x1, y1, x0, y1};
      for (int a = 0; a < 6; ++a) l0.xyz[k + a] = t0[a];
      for (int a = 0; a < 6; ++a) l0.xyz[k + 6 + a] = t1[a];
      k += 12;
    }
  }
  l0.parent.assign(l0.n_cells, -1);
  l0.child_no.assign(l0.n_cells, 0);
  m.levels.push_back(std::move(l0));
  return m;
}

}  // namespace pmx
