#include "space.hpp"

#include <algorithm>
#include <unordered_map>
#include <map>
#include <cmath>
#include <stdexcept>

namespace pmx {

namespace {

// Lattice resolution: nodal points of order-k elements on a mesh refined L
// times from a macro mesh on the grid (1/n) Z^d lie on (1/(k n 2^L)) Z^d.
int64_t lattice_size(const Mesh& m, int level, int order) {
  return ((int64_t)order * m.base_n) << level;
}

int64_t lattice_key(const double* x, int dim, int64_t S) {
  int64_t k = 0;
  for (int d = dim - 1; d >= 0; --d) {
    const int64_t i = std::llround(x[d] * (double)S);
    if (i < 0 || i > S) throw std::runtime_error("nodal point outside the unit domain");
    k = k * (S + 1) + i;
  }
  return k;
}

int class_of_cond(int cond) {
  if (cond == BC_DIRICHLET) return 10;
  if (cond < 1 || cond > kNBoundTypes) throw std::runtime_error("invalid boundary condition code");
  return 10 - cond;   // NEUMANN (1) -> 9 ... DIRICHLET_WEAK (9) -> 1
}

}  // namespace

FESpace build_space(const Mesh& m, int level, int order, const BoundaryConditions& bc) {
  if (level < 0 || level >= (int)m.levels.size()) throw std::runtime_error("mesh level out of range");
  FESpace s;
  s.mesh = &m;
  s.level = level;
  s.el = make_element(m.shape, order);
  s.n_cells = m.levels[level].n_cells;
  s.ndof_cell = s.el.ndof;
  const int nd = s.ndof_cell, dim = m.dim;
  const int64_t S = lattice_size(m, level, order);
  const int64_t n_keys = (dim == 3) ? (S + 1) * (S + 1) * (S + 1) : (S + 1) * (S + 1);

  // nodal point key of every (cell, local dof)
  std::vector<int64_t> key((size_t)s.n_cells * nd);
  std::vector<double> xloc((size_t)s.n_cells * nd * 3, 0.0);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < s.n_cells; ++c) {
    const AffineMap a = affine_map(m, level, c);
    for (int j = 0; j < nd; ++j) {
      double x[3] = {0, 0, 0};
      map_to_phys(a, &s.el.ref_nodes[j * dim], x);
      key[(size_t)c * nd + j] = lattice_key(x, dim, S);
      for (int d = 0; d < 3; ++d) xloc[((size_t)c * nd + j) * 3 + d] = x[d];
    }
  }

  // class of each nodal point: highest class of the boundary joints it lies on
  std::vector<uint8_t> cls(n_keys, 0);
  std::vector<int> last_comp(dim == 2 ? n_keys : 0, -1);
  std::vector<std::vector<int>> jdofs(m.n_faces);
  for (int j = 0; j < m.n_faces; ++j) jdofs[j] = s.el.joint_dofs(j);
  for (int c = 0; c < s.n_cells; ++c) {
    for (int j = 0; j < m.n_faces; ++j) {
      const int comp = joint_boundary_component(m, level, c, j);
      if (comp < 0) continue;
      int vv[4];
      const int nvj = joint_vertices(m.shape, j, vv);
      int cond;
      if (dim == 2) {
        if (!bc.cond2d) throw std::runtime_error("2D space needs a BoundCondFunct2D");
        const double* a = m.vertex(level, c, vv[0]);
        const double* b = m.vertex(level, c, vv[1]);
        double t0 = boundary_param_2d(comp, a[0], a[1]);
        double t1 = boundary_param_2d(comp, b[0], b[1]);
        const double eps = 1e-6;   // FESpace2D.C:1284
        int c0, c1;
        if (t0 < t1) { c0 = bc.cond2d(comp, t0 + eps); c1 = bc.cond2d(comp, t1 - eps); }
        else { c0 = bc.cond2d(comp, t0 - eps); c1 = bc.cond2d(comp, t1 + eps); }
        if (c0 != c1) throw std::runtime_error("different boundary condition on one edge are not allowed");
        cond = c0;
      } else {
        if (!bc.cond3d) throw std::runtime_error("3D space needs a BoundCondFunct3D");
        double X = 0, Y = 0, Z = 0;
        for (int k = 0; k < nvj; ++k) {
          const double* p = m.vertex(level, c, vv[k]);
          X += p[0]; Y += p[1]; Z += p[2];
        }
        X /= nvj; Y /= nvj; Z /= nvj;
        cond = bc.cond3d(X, Y, Z);
      }
      const uint8_t pcls = (uint8_t)class_of_cond(cond);
      for (int ld : jdofs[j]) {
        const int64_t k = key[(size_t)c * nd + ld];
        if (pcls > cls[k]) cls[k] = pcls;
        if (dim == 2 && cond == BC_DIRICHLET) last_comp[k] = comp;
      }
    }
  }

  // first-occurrence numbering per class
  std::vector<int> tmp(n_keys, -1);
  int count[11] = {0};
  std::vector<int> first_occ;   // (cell*nd + j) of first occurrence, in order of temporary id
  s.global_numbers.resize((size_t)s.n_cells * nd);
  for (size_t i = 0; i < key.size(); ++i) {
    const int64_t k = key[i];
    if (tmp[k] < 0) tmp[k] = count[cls[k]]++;
    s.global_numbers[i] = tmp[k];
  }
  s.n_inner = count[0];
  int offset[11];
  offset[0] = 0;
  int acc = count[0];
  for (int t = 1; t <= kNBoundTypes; ++t) {
    // boundary type t (NEUMANN = 1) has class 10 - t
    s.n_bound[t - 1] = count[10 - t];
    offset[10 - t] = acc;
    acc += count[10 - t];
  }
  s.active_bound = acc;
  s.n_dirichlet = count[10];
  offset[10] = acc;
  s.n_dof = acc + count[10];

  s.dof_xyz.assign((size_t)s.n_dof * 3, 0.0);
  s.dof_class.assign(s.n_dof, 0);
  if (dim == 2) { s.dof_comp.assign(s.n_dof, -1); s.dof_param.assign(s.n_dof, 0.0); }
  for (size_t i = 0; i < key.size(); ++i) {
    const int64_t k = key[i];
    const int g = s.global_numbers[i] + offset[cls[k]];
    s.global_numbers[i] = g;
    s.dof_class[g] = cls[k];
    for (int d = 0; d < 3; ++d) s.dof_xyz[(size_t)g * 3 + d] = xloc[i * 3 + d];
    if (dim == 2 && last_comp[k] >= 0) {
      s.dof_comp[g] = last_comp[k];
      s.dof_param[g] = boundary_param_2d(last_comp[k], xloc[i * 3], xloc[i * 3 + 1]);
    }
  }
  return s;
}

FESpace build_space_p1disc(const Mesh& m, int level) {
  if (level < 0 || level >= (int)m.levels.size()) throw std::runtime_error("mesh level out of range");
  if (m.shape != SHAPE_HEXA) throw std::runtime_error("the discontinuous P1 space is built on hexahedra");
  FESpace s;
  s.mesh = &m;
  s.level = level;
  s.el = make_element_p1disc_hexa();
  s.n_cells = m.levels[level].n_cells;
  s.ndof_cell = 4;
  s.n_dof = 4 * s.n_cells;
  s.n_inner = s.n_dof;
  s.active_bound = s.n_dof;
  s.global_numbers.resize((size_t)s.n_dof);
  for (int k = 0; k < s.n_dof; ++k) s.global_numbers[k] = k;
  s.dof_xyz.assign((size_t)s.n_dof * 3, 0.0);
  s.dof_class.assign(s.n_dof, 0);
  for (int c = 0; c < s.n_cells; ++c) {
    const AffineMap a = affine_map(m, level, c);
    const double xi[3] = {0, 0, 0};
    double x[3];
    map_to_phys(a, xi, x);
    for (int k = 0; k < 4; ++k)
      for (int d = 0; d < 3; ++d) s.dof_xyz[((size_t)c * 4 + k) * 3 + d] = x[d];
  }
  return s;
}

CSRPattern build_rect_structure(const FESpace& t, const FESpace& a) {
  if (t.n_cells != a.n_cells) throw std::runtime_error("spaces are not on the same cells");
  const int nt = t.ndof_cell, na = a.ndof_cell;
  CSRPattern p;
  p.n_rows = t.n_dof;
  p.n_cols = a.n_dof;
  p.active_bound = t.active_bound;
  std::vector<int> inc_ptr(t.n_dof + 1, 0);
  for (size_t i = 0; i < t.global_numbers.size(); ++i) inc_ptr[t.global_numbers[i] + 1]++;
  for (int r = 0; r < t.n_dof; ++r) inc_ptr[r + 1] += inc_ptr[r];
  std::vector<int> inc(inc_ptr[t.n_dof]);
  {
    std::vector<int> fill(inc_ptr.begin(), inc_ptr.end() - 1);
    for (int c = 0; c < t.n_cells; ++c)
      for (int j = 0; j < nt; ++j) inc[fill[t.global_numbers[(size_t)c * nt + j]]++] = c;
  }
  std::vector<std::vector<int>> rows(t.n_dof);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int r = 0; r < t.n_dof; ++r) {
    std::vector<int>& cols = rows[r];
    for (int q = inc_ptr[r]; q < inc_ptr[r + 1]; ++q) {
      const int* dof = &a.global_numbers[(size_t)inc[q] * na];
      cols.insert(cols.end(), dof, dof + na);
    }
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
  }
  p.row_ptr.assign(t.n_dof + 1, 0);
  for (int r = 0; r < t.n_dof; ++r) p.row_ptr[r + 1] = p.row_ptr[r] + (int)rows[r].size();
  p.col.resize(p.row_ptr[t.n_dof]);
  for (int r = 0; r < t.n_dof; ++r) std::copy(rows[r].begin(), rows[r].end(), p.col.begin() + p.row_ptr[r]);
  return p;
}

CSRPattern build_structure(const FESpace& s) {
  const int nd = s.ndof_cell;
  CSRPattern p;
  p.n_rows = p.n_cols = s.n_dof;
  p.active_bound = s.active_bound;
  // row -> incident cells
  std::vector<int> inc_ptr(s.n_dof + 1, 0);
  for (size_t i = 0; i < s.global_numbers.size(); ++i) inc_ptr[s.global_numbers[i] + 1]++;
  for (int r = 0; r < s.n_dof; ++r) inc_ptr[r + 1] += inc_ptr[r];
  std::vector<int> inc(inc_ptr[s.n_dof]);
  {
    std::vector<int> fill(inc_ptr.begin(), inc_ptr.end() - 1);
    for (int c = 0; c < s.n_cells; ++c)
      for (int j = 0; j < nd; ++j) inc[fill[s.global_numbers[(size_t)c * nd + j]]++] = c;
  }
  std::vector<int> len(s.n_dof, 1);
  std::vector<std::vector<int>> rows(s.n_dof);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int r = 0; r < s.active_bound; ++r) {
    std::vector<int>& cols = rows[r];
    cols.reserve((inc_ptr[r + 1] - inc_ptr[r]) * nd);
    for (int q = inc_ptr[r]; q < inc_ptr[r + 1]; ++q) {
      const int* dof = &s.global_numbers[(size_t)inc[q] * nd];
      cols.insert(cols.end(), dof, dof + nd);
    }
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    len[r] = (int)cols.size();
  }
  p.row_ptr.assign(s.n_dof + 1, 0);
  for (int r = 0; r < s.n_dof; ++r) p.row_ptr[r + 1] = p.row_ptr[r] + len[r];
  p.col.resize(p.row_ptr[s.n_dof]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int r = 0; r < s.n_dof; ++r) {
    if (r < s.active_bound) {
      std::copy(rows[r].begin(), rows[r].end(), p.col.begin() + p.row_ptr[r]);
      std::vector<int>().swap(rows[r]);
    } else {
      p.col[p.row_ptr[r]] = r;
    }
  }
  return p;
}

void build_prolongation(const FESpace& coarse, const FESpace& fine, CSRPattern& P,
                        std::vector<double>& vals) {
  const Mesh& m = *fine.mesh;
  if (coarse.mesh != fine.mesh || fine.level != coarse.level + 1)
    throw std::runtime_error("prolongation needs consecutive levels of one mesh");
  const int ndf = fine.ndof_cell, ndc = coarse.ndof_cell, dim = m.dim;
  const MeshLevel& fl = m.levels[fine.level];
  // each fine DOF's row comes from the first fine cell (in cell order) that
  // holds it: the coarse basis of its parent evaluated at the DOF's node
  std::vector<int> own((size_t)fine.n_dof, -1);
  for (int fc = 0; fc < fine.n_cells; ++fc)
    for (int k = 0; k < ndf; ++k) {
      const int g = fine.global_numbers[(size_t)fc * ndf + k];
      if (own[g] < 0) own[g] = fc * ndf + k;
    }
  if (fine.el.disc != coarse.el.disc) throw std::runtime_error("prolongation between different elements");
  auto row_of = [&](int g, std::vector<std::pair<int, double>>& row) {
    row.clear();
    const int fc = own[g] / ndf, k = own[g] % ndf;
    const int pc = fl.parent[fc];
    const AffineMap af = affine_map(m, fine.level, fc);
    const AffineMap ac = affine_map(m, coarse.level, pc);
    double x[3], xi[3], v[64];
    if (fine.el.disc) {
      // fine nodal functional k of every coarse basis function
      double pts[24], pv[8][4];
      p1disc_nf_points(pts);
      for (int i = 0; i < 8; ++i) {
        map_to_phys(af, pts + 3 * i, x);
        map_to_ref(ac, x, xi);
        coarse.el.eval(xi, pv[i], nullptr, nullptr);
      }
      for (int l = 0; l < ndc; ++l) {
        double vals8[8], f[4];
        for (int i = 0; i < 8; ++i) vals8[i] = pv[i][l];
        p1disc_nf_apply(vals8, f);
        v[l] = f[k];
      }
    } else {
      map_to_phys(af, &fine.el.ref_nodes[k * dim], x);
      map_to_ref(ac, x, xi);
      coarse.el.eval(xi, v, nullptr, nullptr);
    }
    for (int l = 0; l < ndc; ++l) {
      // basis values at refined nodal points are exact binary fractions;
      // snap round-off of the inverse map
      double val = std::nearbyint(v[l] * 1024.0) / 1024.0;
      if (std::fabs(v[l] - val) > 1e-12) val = v[l];
      if (val != 0.0) row.emplace_back(coarse.global_numbers[(size_t)pc * ndc + l], val);
    }
    std::sort(row.begin(), row.end());
  };
  if (ndc > 64) throw std::runtime_error("prolongation: too many local DOFs");
  P.n_rows = fine.n_dof;
  P.n_cols = coarse.n_dof;
  P.active_bound = fine.n_dof;
  P.row_ptr.assign(fine.n_dof + 1, 0);
#pragma omp parallel
  {
    std::vector<std::pair<int, double>> row;
#pragma omp for schedule(static)
    for (int g = 0; g < fine.n_dof; ++g) {
      row_of(g, row);
      P.row_ptr[g + 1] = (int)row.size();
    }
  }
  for (int r = 0; r < fine.n_dof; ++r) P.row_ptr[r + 1] += P.row_ptr[r];
  P.col.resize(P.row_ptr.back());
  vals.resize(P.row_ptr.back());
#pragma omp parallel
  {
    std::vector<std::pair<int, double>> row;
#pragma omp for schedule(static)
    for (int g = 0; g < fine.n_dof; ++g) {
      row_of(g, row);
      int o = P.row_ptr[g];
      for (auto& e : row) { P.col[o] = e.first; vals[o] = e.second; ++o; }
    }
  }
}

void build_function_restriction(const FESpace& coarse, const FESpace& fine, CSRPattern& RF,
                                std::vector<double>& vals) {
  const Mesh& m = *fine.mesh;
  if (coarse.mesh != fine.mesh || fine.level != coarse.level + 1)
    throw std::runtime_error("restriction needs consecutive levels of one mesh");
  if (m.shape != SHAPE_QUAD && m.shape != SHAPE_HEXA)
    throw std::runtime_error("function restriction is built for quadrilaterals and hexahedra");
  const int ndf = fine.ndof_cell, ndc = coarse.ndof_cell, dim = m.dim;
  const QuadRule q = gauss_tensor(dim, 3);   // exact for the products of Q2 functions
  const double child_det = 1.0 / (double)(1 << dim);
  // G = coarse mass matrix on the reference cell and its inverse
  std::vector<double> phc((size_t)q.n * ndc), phf((size_t)q.n * ndf);
  for (int k = 0; k < q.n; ++k) {
    coarse.el.eval(&q.xi[(size_t)k * dim], &phc[(size_t)k * ndc], nullptr, nullptr);
    fine.el.eval(&q.xi[(size_t)k * dim], &phf[(size_t)k * ndf], nullptr, nullptr);
  }
  std::vector<double> G((size_t)ndc * ndc, 0.0), Ginv((size_t)ndc * ndc, 0.0);
  for (int k = 0; k < q.n; ++k)
    for (int i = 0; i < ndc; ++i)
      for (int j = 0; j < ndc; ++j) G[(size_t)i * ndc + j] += q.w[k] * phc[(size_t)k * ndc + i] * phc[(size_t)k * ndc + j];
  for (int i = 0; i < ndc; ++i) Ginv[(size_t)i * ndc + i] = 1.0;
  for (int c = 0; c < ndc; ++c) {   // Gauss-Jordan with partial pivoting
    int p = c;
    for (int r = c + 1; r < ndc; ++r)
      if (std::fabs(G[(size_t)r * ndc + c]) > std::fabs(G[(size_t)p * ndc + c])) p = r;
    for (int j = 0; j < ndc; ++j) {
      std::swap(G[(size_t)c * ndc + j], G[(size_t)p * ndc + j]);
      std::swap(Ginv[(size_t)c * ndc + j], Ginv[(size_t)p * ndc + j]);
    }
    const double d = 1.0 / G[(size_t)c * ndc + c];
    for (int j = 0; j < ndc; ++j) { G[(size_t)c * ndc + j] *= d; Ginv[(size_t)c * ndc + j] *= d; }
    for (int r = 0; r < ndc; ++r) {
      if (r == c) continue;
      const double f = G[(size_t)r * ndc + c];
      if (f == 0.0) continue;
      for (int j = 0; j < ndc; ++j) {
        G[(size_t)r * ndc + j] -= f * G[(size_t)c * ndc + j];
        Ginv[(size_t)r * ndc + j] -= f * Ginv[(size_t)c * ndc + j];
      }
    }
  }
  // RQ of a child depends only on its position in the parent (uniform
  // refinement of affine cells): cache by the child centre in parent coordinates
  std::map<std::vector<long long>, std::vector<double>> cache;
  const MeshLevel& fl = m.levels[fine.level];
  std::vector<std::vector<int>> children(coarse.n_cells);
  for (int fc = 0; fc < fine.n_cells; ++fc) children[fl.parent[fc]].push_back(fc);
  std::vector<std::vector<std::pair<int, double>>> rows(coarse.n_dof);
  std::vector<int> count(coarse.n_dof, 0);
  std::vector<double> v(ndc);
  for (int pc = 0; pc < coarse.n_cells; ++pc) {
    const AffineMap ac = affine_map(m, coarse.level, pc);
    for (int fc : children[pc]) {
      const AffineMap af = affine_map(m, fine.level, fc);
      double zero[3] = {0, 0, 0}, x[3], xc[3];
      map_to_phys(af, zero, x);
      map_to_ref(ac, x, xc);
      std::vector<long long> key(dim);
      for (int d = 0; d < dim; ++d) key[d] = std::llround(xc[d] * 1024.0);
      auto it = cache.find(key);
      if (it == cache.end()) {
        std::vector<double> R((size_t)ndc * ndf, 0.0);
        for (int k = 0; k < q.n; ++k) {
          double xp[3], xi[3];
          map_to_phys(af, &q.xi[(size_t)k * dim], xp);
          map_to_ref(ac, xp, xi);
          coarse.el.eval(xi, v.data(), nullptr, nullptr);
          const double w = q.w[k] * child_det;
          for (int a = 0; a < ndc; ++a)
            for (int l = 0; l < ndf; ++l) R[(size_t)a * ndf + l] += w * v[a] * phf[(size_t)k * ndf + l];
        }
        std::vector<double> RQ((size_t)ndc * ndf, 0.0);
        for (int a = 0; a < ndc; ++a)
          for (int b = 0; b < ndc; ++b)
            for (int l = 0; l < ndf; ++l) RQ[(size_t)a * ndf + l] += Ginv[(size_t)a * ndc + b] * R[(size_t)b * ndf + l];
        it = cache.emplace(key, std::move(RQ)).first;
      }
      const std::vector<double>& RQ = it->second;
      for (int a = 0; a < ndc; ++a) {
        auto& row = rows[coarse.global_numbers[(size_t)pc * ndc + a]];
        for (int l = 0; l < ndf; ++l) {
          const double val = RQ[(size_t)a * ndf + l];
          if (val != 0.0) row.emplace_back(fine.global_numbers[(size_t)fc * ndf + l], val);
        }
      }
    }
    for (int a = 0; a < ndc; ++a) count[coarse.global_numbers[(size_t)pc * ndc + a]]++;
  }
  RF.n_rows = coarse.n_dof;
  RF.n_cols = fine.n_dof;
  RF.active_bound = coarse.n_dof;
  RF.row_ptr.assign(coarse.n_dof + 1, 0);
  RF.col.clear();
  vals.clear();
  for (int r = 0; r < coarse.n_dof; ++r) {
    auto& row = rows[r];
    std::sort(row.begin(), row.end(), [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
      return a.first < b.first;
    });
    for (size_t k = 0; k < row.size();) {
      const int c = row[k].first;
      double sum = 0.0;
      for (; k < row.size() && row[k].first == c; ++k) sum += row[k].second;
      RF.col.push_back(c);
      vals.push_back(sum / (double)count[r]);
    }
    RF.row_ptr[r + 1] = (int)RF.col.size();
  }
}

void transpose_csr(const CSRPattern& A, const std::vector<double>& av, CSRPattern& T,
                   std::vector<double>& tv) {
  T.n_rows = A.n_cols;
  T.n_cols = A.n_rows;
  T.active_bound = T.n_rows;
  T.row_ptr.assign(T.n_rows + 1, 0);
  for (int c : A.col) T.row_ptr[c + 1]++;
  for (int r = 0; r < T.n_rows; ++r) T.row_ptr[r + 1] += T.row_ptr[r];
  T.col.resize(A.col.size());
  tv.resize(A.col.size());
  std::vector<int> fill(T.row_ptr.begin(), T.row_ptr.end() - 1);
  for (int r = 0; r < A.n_rows; ++r)
    for (int q = A.row_ptr[r]; q < A.row_ptr[r + 1]; ++q) {
      const int o = fill[A.col[q]]++;
      T.col[o] = r;
      tv[o] = av[q];
    }
}

void color_cells(const FESpace& s, std::vector<int>& order, std::vector<int>& color_ptr) {
  const Mesh& m = *s.mesh;
  const int nc = s.n_cells, dim = m.dim;
  std::vector<int> color(nc, -1);
  int ncolors = 0;
  if (m.shape == SHAPE_QUAD || m.shape == SHAPE_HEXA) {
    // structured lattice: cells sharing a vertex differ by <= 1 in every index,
    // so the parity of the cell index vector is a valid colouring
    const double n = (double)(1 << s.level);
    for (int c = 0; c < nc; ++c) {
      double ctr[3] = {0, 0, 0};
      for (int v = 0; v < m.nv; ++v)
        for (int d = 0; d < dim; ++d) ctr[d] += m.vertex(s.level, c, v)[d] / m.nv;
      int col = 0;
      for (int d = 0; d < dim; ++d) col |= ((int)std::floor(ctr[d] * n) & 1) << d;
      color[c] = col;
    }
    ncolors = 1 << dim;
  } else {
    // greedy colouring over the DOF -> cell incidence
    std::vector<int> inc_ptr(s.n_dof + 1, 0);
    for (int g : s.global_numbers) inc_ptr[g + 1]++;
    for (int r = 0; r < s.n_dof; ++r) inc_ptr[r + 1] += inc_ptr[r];
    std::vector<int> inc(inc_ptr.back());
    std::vector<int> fill(inc_ptr.begin(), inc_ptr.end() - 1);
    for (int c = 0; c < nc; ++c)
      for (int j = 0; j < s.ndof_cell; ++j) inc[fill[s.global_numbers[(size_t)c * s.ndof_cell + j]]++] = c;
    std::vector<int> mark(64, -1);
    for (int c = 0; c < nc; ++c) {
      for (int j = 0; j < s.ndof_cell; ++j) {
        const int g = s.global_numbers[(size_t)c * s.ndof_cell + j];
        for (int q = inc_ptr[g]; q < inc_ptr[g + 1]; ++q)
          if (color[inc[q]] >= 0) mark[color[inc[q]]] = c;
      }
      int col = 0;
      while (mark[col] == c) ++col;
      color[c] = col;
      ncolors = std::max(ncolors, col + 1);
    }
  }
  color_ptr.assign(ncolors + 1, 0);
  for (int c = 0; c < nc; ++c) color_ptr[color[c] + 1]++;
  for (int k = 0; k < ncolors; ++k) color_ptr[k + 1] += color_ptr[k];
  order.resize(nc);
  std::vector<int> fill(color_ptr.begin(), color_ptr.end() - 1);
  for (int c = 0; c < nc; ++c) order[fill[color[c]]++] = c;
}

}  // namespace pmx

namespace pmx {

SellPattern build_sell(const CSRPattern& p, int sigma, bool rel16, int pack, bool patterns, int max_patterns) {
  constexpr int C = 64;
  SellPattern s;
  if (!rel16 || pack != 4) pack = 1;
  if (!rel16) patterns = false;
  s.pack = pack;
  const int na = p.active_bound;
  s.n_active = na;
  s.n_slices = (na + C - 1) / C;
  // slot entries of a row: all of them, or (rel16) those in active columns,
  // which precede the Dirichlet columns in a row sorted by column
  std::vector<int> nact(na);
  for (int r = 0; r < na; ++r) {
    int c = 0;
    for (int q = p.row_ptr[r]; q < p.row_ptr[r + 1]; ++q) c += p.col[q] < na;
    nact[r] = c;
  }
  if (rel16) {
    bool fits = true;
    for (int r = 0; r < na && fits; ++r)
      for (int q = p.row_ptr[r]; q < p.row_ptr[r + 1]; ++q) {
        const int c = p.col[q];
        if (c >= na) { fits = fits && (q - p.row_ptr[r]) >= nact[r]; continue; }   // Dirichlet entries last
        if ((q - p.row_ptr[r]) >= nact[r] || c - r > 32767 || c - r < -32768) { fits = false; break; }
      }
    if (!fits) return build_sell(p, sigma, false, 1);
  }
  auto len = [&](int r) { return rel16 ? nact[r] : p.row_ptr[r + 1] - p.row_ptr[r]; };
  if (patterns) {
    // pattern id of every active row: its active-column offsets, compared
    // exactly (a 64-bit hash picks the candidate list)
    std::vector<uint64_t> h(na);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < na; ++r) {
      uint64_t x = 1469598103934665603ull ^ (uint64_t)nact[r];
      for (int q = p.row_ptr[r]; q < p.row_ptr[r] + nact[r]; ++q) {
        x ^= (uint64_t)(uint32_t)(p.col[q] - r);
        x *= 1099511628211ull;
      }
      h[r] = x;
    }
    std::unordered_map<uint64_t, std::vector<int>> cand;   // hash -> pattern ids
    std::vector<int> rep;                                  // representative row of each pattern
    std::vector<int> pid(na);
    bool too_many = false;
    for (int r = 0; r < na && !too_many; ++r) {
      auto& list = cand[h[r]];
      int id = -1;
      for (int c : list) {
        const int r0 = rep[c];
        if (nact[r0] != nact[r]) continue;
        bool same = true;
        for (int k = 0; k < nact[r] && same; ++k)
          same = p.col[p.row_ptr[r0] + k] - r0 == p.col[p.row_ptr[r] + k] - r;
        if (same) { id = c; break; }
      }
      if (id < 0) {
        id = (int)rep.size();
        rep.push_back(r);
        list.push_back(id);
        too_many = (int)rep.size() > max_patterns;
      }
      pid[r] = id;
    }
    if (too_many) return build_sell(p, sigma, rel16, 1, false);
    const int np = (int)rep.size();
    // rows of each pattern in row order, cut into slices of 64; slices
    // ordered by their first row
    std::vector<int> cnt(np + 1, 0);
    for (int r = 0; r < na; ++r) cnt[pid[r] + 1]++;
    for (int k = 0; k < np; ++k) cnt[k + 1] += cnt[k];
    std::vector<int> by_pat(na), fill(cnt.begin(), cnt.end() - 1);
    for (int r = 0; r < na; ++r) by_pat[fill[pid[r]]++] = r;
    std::vector<std::pair<int, int>> slices;   // (first position in by_pat, pattern)
    for (int k = 0; k < np; ++k)
      for (int q = cnt[k]; q < cnt[k + 1]; q += C) slices.push_back({q, k});
    std::stable_sort(slices.begin(), slices.end(),
                     [&](const std::pair<int, int>& a, const std::pair<int, int>& b) {
                       return by_pat[a.first] < by_pat[b.first];
                     });
    s.n_slices = (int)slices.size();
    s.n_patterns = np;
    // pattern widths padded to a multiple of pack (padding offsets 0: the
    // row itself, value 0)
    auto pwidth = [&](int k) { return (nact[rep[k]] + pack - 1) / pack * pack; };
    std::vector<int> pat_ptr(np + 1, 0);
    for (int k = 0; k < np; ++k) pat_ptr[k + 1] = pat_ptr[k] + pwidth(k);
    s.pat_off.assign(pat_ptr[np], 0);
    for (int k = 0; k < np; ++k)
      for (int j = 0; j < nact[rep[k]]; ++j) s.pat_off[pat_ptr[k] + j] = p.col[p.row_ptr[rep[k]] + j] - rep[k];
    s.row.assign((size_t)s.n_slices * C, -1);
    s.slice_ptr.assign(s.n_slices + 1, 0);
    s.slice_pat.resize(s.n_slices);
    for (int sl = 0; sl < s.n_slices; ++sl) {
      const int q0 = slices[sl].first, k = slices[sl].second;
      for (int l = 0; l < C && q0 + l < cnt[k + 1]; ++l) s.row[(size_t)sl * C + l] = by_pat[q0 + l];
      s.slice_pat[sl] = pat_ptr[k];
      s.slice_ptr[sl + 1] = s.slice_ptr[sl] + pwidth(k) * C;
    }
  } else {
    std::vector<int> order(na);
    for (int r = 0; r < na; ++r) order[r] = r;
    for (int w = 0; w < na; w += sigma) {
      const int e = std::min(na, w + sigma);
      std::stable_sort(order.begin() + w, order.begin() + e, [&](int a, int b) { return len(a) > len(b); });
    }
    s.row.assign((size_t)s.n_slices * C, -1);
    s.slice_ptr.assign(s.n_slices + 1, 0);
    for (int sl = 0; sl < s.n_slices; ++sl) {
      int width = 0;
      for (int l = 0; l < C; ++l) {
        const int k = sl * C + l;
        if (k < na) {
          s.row[k] = order[k];
          width = std::max(width, len(order[k]));
        }
      }
      width = (width + pack - 1) / pack * pack;
      s.slice_ptr[sl + 1] = s.slice_ptr[sl] + width * C;
    }
  }
  const size_t total = s.slice_ptr.back();
  if (rel16) s.off.assign(total, 0);
  else s.col.assign(total, 0);
  s.src.assign(total, -1);
#pragma omp parallel for schedule(static)
  for (int sl = 0; sl < s.n_slices; ++sl) {
    const int width = (s.slice_ptr[sl + 1] - s.slice_ptr[sl]) / C;
    for (int l = 0; l < C; ++l) {
      const int r = s.row[(size_t)sl * C + l];
      for (int j = 0; j < width; ++j) {
        const size_t pos = sell_pos((size_t)s.slice_ptr[sl], j, l, pack);
        if (r >= 0 && j < len(r)) {
          const int c = p.col[p.row_ptr[r] + j];
          if (rel16) s.off[pos] = (int16_t)(c - r);
          else s.col[pos] = c;
          s.src[pos] = p.row_ptr[r] + j;
        } else if (!rel16) {
          s.col[pos] = (r >= 0) ? r : 0;   // padding: (row, 0); rel16 padding is offset 0
        }
      }
    }
  }
  // Dirichlet-column entries of every active row, in row order
  s.bd_ptr.assign(na + 1, 0);
  for (int r = 0; r < na; ++r) s.bd_ptr[r + 1] = s.bd_ptr[r] + (p.row_ptr[r + 1] - p.row_ptr[r] - nact[r]);
  s.bd_src.resize(s.bd_ptr[na]);
  s.bd_col.resize(s.bd_ptr[na]);
#pragma omp parallel for schedule(static)
  for (int r = 0; r < na; ++r) {
    int o = s.bd_ptr[r];
    for (int q = p.row_ptr[r]; q < p.row_ptr[r + 1]; ++q) {
      if (p.col[q] < na) continue;
      s.bd_src[o] = q;
      s.bd_col[o] = p.col[q];
      ++o;
    }
  }
  return s;
}

}  // namespace pmx

namespace pmx {

std::vector<int> lexicographic_order(const FESpace& s) {
  const int n = s.n_dof, na = s.active_bound;
  // coordinates quantised on a grid far finer than any mesh width here, so
  // that nodes on one plane/line compare equal despite round-off
  std::vector<long long> key((size_t)3 * n);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) key[(size_t)3 * i + d] = std::llround(s.dof_xyz[(size_t)3 * i + d] * 1073741824.0);
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  auto less = [&](int a, int b) {
    for (int d = 2; d >= 0; --d) {
      const long long qa = key[(size_t)3 * a + d], qb = key[(size_t)3 * b + d];
      if (qa != qb) return qa < qb;
    }
    return a < b;
  };
  std::sort(perm.begin(), perm.begin() + na, less);
  std::sort(perm.begin() + na, perm.end(), less);
  return perm;
}

void permute_csr(const CSRPattern& A, const std::vector<double>* av, const std::vector<int>& rperm,
                 const std::vector<int>& cinv, CSRPattern& out, std::vector<double>* ov,
                 std::vector<int>* opos, bool sort_rows) {
  const int n = A.n_rows;
  out.n_rows = A.n_rows;
  out.n_cols = A.n_cols;
  out.active_bound = A.active_bound;
  out.row_ptr.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    const int r = rperm.empty() ? i : rperm[i];
    out.row_ptr[i + 1] = out.row_ptr[i] + (A.row_ptr[r + 1] - A.row_ptr[r]);
  }
  const size_t nnz = A.col.size();
  out.col.resize(nnz);
  if (ov) ov->resize(nnz);
  if (opos) opos->resize(nnz);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const int r = rperm.empty() ? i : rperm[i];
    const int b = A.row_ptr[r], len = A.row_ptr[r + 1] - b, o = out.row_ptr[i];
    // order of the row's entries in out: as in A, or by new column
    int idx[512];
    std::vector<int> big;
    int* ord = idx;
    if (len > 512) { big.resize(len); ord = big.data(); }
    for (int k = 0; k < len; ++k) ord[k] = k;
    auto newcol = [&](int k) { return cinv.empty() ? A.col[b + k] : cinv[A.col[b + k]]; };
    if (sort_rows) std::sort(ord, ord + len, [&](int x, int y) { return newcol(x) < newcol(y); });
    for (int k = 0; k < len; ++k) {
      out.col[o + k] = newcol(ord[k]);
      if (ov) (*ov)[o + k] = (*av)[b + ord[k]];
      if (opos) (*opos)[o + k] = b + ord[k];
    }
  }
}

}  // namespace pmx
