#include "fe.hpp"

#include <cmath>
#include <stdexcept>

namespace pmx {

namespace {

// 1D Lagrange basis on nodes {-1,1} (order 1) or {-1,0,1} (order 2):
// value, first and second derivative of basis a at t.
void lagrange1d(int order, int a, double t, double& v, double& d, double& dd) {
  if (order == 1) {
    if (a == 0) { v = 0.5 * (1 - t); d = -0.5; dd = 0; }
    else { v = 0.5 * (1 + t); d = 0.5; dd = 0; }
    return;
  }
  switch (a) {
    case 0: v = 0.5 * t * (t - 1); d = t - 0.5; dd = 1; break;
    case 1: v = 1 - t * t; d = -2 * t; dd = -2; break;
    default: v = 0.5 * t * (t + 1); d = t + 0.5; dd = 1; break;
  }
}

}  // namespace

Element make_element(Shape shape, int order) {
  if (order != 1 && order != 2) throw std::runtime_error("only order 1 and 2 are supported");
  Element e;
  e.shape = shape;
  e.order = order;
  if (shape == SHAPE_TRI) {
    if (order != 1) throw std::runtime_error("triangles: only P1 is supported");
    e.dim = 2;
    e.ndof = 3;
    e.ref_nodes = {0, 0, 1, 0, 0, 1};
    return e;
  }
  e.dim = (shape == SHAPE_HEXA) ? 3 : 2;
  const int n1 = order + 1;
  e.ndof = (e.dim == 3) ? n1 * n1 * n1 : n1 * n1;
  e.ref_nodes.resize((size_t)e.ndof * e.dim);
  for (int k = 0; k < e.ndof; ++k) {
    int idx[3] = {k % n1, (k / n1) % n1, k / (n1 * n1)};
    for (int d = 0; d < e.dim; ++d)
      e.ref_nodes[k * e.dim + d] = -1.0 + 2.0 * idx[d] / order;
  }
  return e;
}

Element make_element_p1disc_hexa() {
  Element e;
  e.shape = SHAPE_HEXA;
  e.order = 1;
  e.dim = 3;
  e.ndof = 4;
  e.disc = true;
  e.ref_nodes.assign(12, 0.0);   // no nodal points: the cell centre for every DOF
  return e;
}

void p1disc_nf_points(double* xi24) {
  const double a = 0.5773502691896257645091489;
  for (int i = 0; i < 8; ++i) {
    xi24[3 * i] = (i & 1) ? a : -a;
    xi24[3 * i + 1] = (i & 2) ? a : -a;
    xi24[3 * i + 2] = (i & 4) ? a : -a;
  }
}

void p1disc_nf_apply(const double* v, double* out) {
  double xi[24];
  p1disc_nf_points(xi);
  double s = 0;
  for (int i = 0; i < 8; ++i) s += v[i];
  out[0] = s * 0.125;
  for (int d = 0; d < 3; ++d) {
    double t = 0;
    for (int i = 0; i < 8; ++i) t += v[i] * xi[3 * i + d];
    out[1 + d] = t * 0.375;
  }
}

void Element::eval(const double* xi, double* v, double* d1, double* d2) const {
  if (disc) {
    v[0] = 1; v[1] = xi[0]; v[2] = xi[1]; v[3] = xi[2];
    if (d1)
      for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 4; ++k) d1[a * 4 + k] = (k == a + 1) ? 1.0 : 0.0;
    if (d2) for (int i = 0; i < 24; ++i) d2[i] = 0;
    return;
  }
  if (shape == SHAPE_TRI) {
    v[0] = 1 - xi[0] - xi[1]; v[1] = xi[0]; v[2] = xi[1];
    if (d1) { d1[0] = -1; d1[1] = 1; d1[2] = 0; d1[3] = -1; d1[4] = 0; d1[5] = 1; }
    if (d2) for (int i = 0; i < 9; ++i) d2[i] = 0;
    return;
  }
  const int n1 = order + 1;
  for (int k = 0; k < ndof; ++k) {
    int idx[3] = {k % n1, (k / n1) % n1, k / (n1 * n1)};
    double f[3], g[3], h[3];
    for (int d = 0; d < dim; ++d) lagrange1d(order, idx[d], xi[d], f[d], g[d], h[d]);
    if (dim == 2) {
      v[k] = f[0] * f[1];
      if (d1) { d1[k] = g[0] * f[1]; d1[ndof + k] = f[0] * g[1]; }
      if (d2) {
        d2[k] = h[0] * f[1];
        d2[ndof + k] = g[0] * g[1];
        d2[2 * ndof + k] = f[0] * h[1];
      }
    } else {
      v[k] = f[0] * f[1] * f[2];
      if (d1) {
        d1[k] = g[0] * f[1] * f[2];
        d1[ndof + k] = f[0] * g[1] * f[2];
        d1[2 * ndof + k] = f[0] * f[1] * g[2];
      }
      if (d2) {
        d2[k] = h[0] * f[1] * f[2];             // xx
        d2[ndof + k] = g[0] * g[1] * f[2];      // xy
        d2[2 * ndof + k] = g[0] * f[1] * g[2];  // xz
        d2[3 * ndof + k] = f[0] * h[1] * f[2];  // yy
        d2[4 * ndof + k] = f[0] * g[1] * g[2];  // yz
        d2[5 * ndof + k] = f[0] * f[1] * h[2];  // zz
      }
    }
  }
}

std::vector<int> Element::joint_dofs(int j) const {
  std::vector<int> out;
  if (disc) return out;
  for (int k = 0; k < ndof; ++k) {
    const double* p = &ref_nodes[k * dim];
    bool on = false;
    if (shape == SHAPE_TRI) {
      if (j == 0) on = (p[1] == 0.0);
      else if (j == 1) on = (p[0] + p[1] == 1.0);
      else on = (p[0] == 0.0);
    } else if (shape == SHAPE_QUAD) {
      static const int ax[4] = {1, 0, 1, 0};
      static const double val[4] = {-1, 1, 1, -1};
      on = (p[ax[j]] == val[j]);
    } else {
      static const int ax[6] = {2, 1, 0, 1, 0, 2};
      static const double val[6] = {-1, -1, 1, 1, -1, 1};
      on = (p[ax[j]] == val[j]);
    }
    if (on) out.push_back(k);
  }
  return out;
}

void gauss_legendre(int n, double* x, double* w) {
  switch (n) {
    case 1: x[0] = 0; w[0] = 2; return;
    case 2: {
      const double a = 1.0 / std::sqrt(3.0);
      x[0] = -a; x[1] = a; w[0] = w[1] = 1; return;
    }
    case 3: {
      const double a = std::sqrt(0.6);
      x[0] = -a; x[1] = 0; x[2] = a;
      w[0] = w[2] = 5.0 / 9.0; w[1] = 8.0 / 9.0; return;
    }
    case 4: {
      const double s = 2.0 / 7.0 * std::sqrt(6.0 / 5.0);
      const double a = std::sqrt(3.0 / 7.0 - s), b = std::sqrt(3.0 / 7.0 + s);
      const double wa = (18.0 + std::sqrt(30.0)) / 36.0, wb = (18.0 - std::sqrt(30.0)) / 36.0;
      x[0] = -b; x[1] = -a; x[2] = a; x[3] = b;
      w[0] = wb; w[1] = wa; w[2] = wa; w[3] = wb; return;
    }
    case 5: {
      const double s = 2.0 * std::sqrt(10.0 / 7.0);
      const double a = std::sqrt(5.0 - s) / 3.0, b = std::sqrt(5.0 + s) / 3.0;
      const double wa = (322.0 + 13.0 * std::sqrt(70.0)) / 900.0;
      const double wb = (322.0 - 13.0 * std::sqrt(70.0)) / 900.0;
      x[0] = -b; x[1] = -a; x[2] = 0; x[3] = a; x[4] = b;
      w[0] = wb; w[1] = wa; w[2] = 128.0 / 225.0; w[3] = wa; w[4] = wb; return;
    }
  }
  throw std::runtime_error("gauss_legendre: n out of range");
}

QuadRule gauss_tensor(int dim, int npts) {
  double x[5], w[5];
  gauss_legendre(npts, x, w);
  QuadRule q;
  q.dim = dim;
  q.n = (dim == 3) ? npts * npts * npts : npts * npts;
  q.xi.resize((size_t)q.n * dim);
  q.w.resize(q.n);
  for (int k = 0; k < q.n; ++k) {
    int i = k % npts, j = (k / npts) % npts, l = k / (npts * npts);
    q.xi[k * dim + 0] = x[i];
    q.xi[k * dim + 1] = x[j];
    if (dim == 3) q.xi[k * dim + 2] = x[l];
    q.w[k] = (dim == 3) ? w[i] * w[j] * w[l] : w[i] * w[j];
  }
  return q;
}

QuadRule gauss3_tria() {
  // Radon's 7-point rule, exact for degree 5 (weights sum to the area 1/2)
  const double r = std::sqrt(15.0);
  const double a = (6.0 - r) / 21.0, b = (9.0 + 2.0 * r) / 21.0;
  const double c = (6.0 + r) / 21.0, d = (9.0 - 2.0 * r) / 21.0;
  const double wa = (155.0 - r) / 2400.0, wc = (155.0 + r) / 2400.0;
  QuadRule q;
  q.dim = 2;
  q.n = 7;
  q.xi = {1.0 / 3.0, 1.0 / 3.0, b, a, a, b, a, a, d, c, c, d, c, c};
  q.w = {9.0 / 80.0, wa, wa, wa, wc, wc, wc};
  return q;
}

QuadRule assembly_rule(const Element& e, int hexa_rule) {
  if (e.shape == SHAPE_TRI) return gauss3_tria();
  if (e.shape == SHAPE_QUAD) {
    // QFQuadFromDegree[3*order]: 3 -> Gauss3, 6 -> Gauss4
    return gauss_tensor(2, e.order == 1 ? 3 : 4);
  }
  if (hexa_rule == 2) return gauss_tensor(3, 3);
  return gauss_tensor(3, e.order == 1 ? 3 : 4);
}

AffineMap affine_map(const Mesh& m, int level, int cell) {
  AffineMap a{};
  a.dim = m.dim;
  const double* v0 = m.vertex(level, cell, 0);
  for (int d = 0; d < m.dim; ++d) a.x0[d] = v0[d];
  if (m.shape == SHAPE_TRI) {
    const double* v1 = m.vertex(level, cell, 1);
    const double* v2 = m.vertex(level, cell, 2);
    for (int r = 0; r < 2; ++r) {
      a.J[r * 2 + 0] = v1[r] - v0[r];
      a.J[r * 2 + 1] = v2[r] - v0[r];
    }
  } else {
    // parallelogram / parallelepiped through v0: xi = -1 at v0
    const int e_idx[3] = {1, 3, 4};
    for (int c = 0; c < m.dim; ++c) {
      const double* vc = m.vertex(level, cell, e_idx[c]);
      for (int r = 0; r < m.dim; ++r) a.J[r * m.dim + c] = 0.5 * (vc[r] - v0[r]);
    }
    // x0 = v0 + J*(1,1,1): the image of xi = 0
    for (int r = 0; r < m.dim; ++r) {
      double s = v0[r];
      for (int c = 0; c < m.dim; ++c) s += a.J[r * m.dim + c];
      a.x0[r] = s;
    }
  }
  if (m.dim == 2) {
    a.detJ = a.J[0] * a.J[3] - a.J[1] * a.J[2];
    const double id = 1.0 / a.detJ;
    a.Jinv[0] = a.J[3] * id; a.Jinv[1] = -a.J[1] * id;
    a.Jinv[2] = -a.J[2] * id; a.Jinv[3] = a.J[0] * id;
  } else {
    const double* J = a.J;
    a.detJ = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
             J[2] * (J[3] * J[7] - J[4] * J[6]);
    const double id = 1.0 / a.detJ;
    a.Jinv[0] = (J[4] * J[8] - J[5] * J[7]) * id;
    a.Jinv[1] = (J[2] * J[7] - J[1] * J[8]) * id;
    a.Jinv[2] = (J[1] * J[5] - J[2] * J[4]) * id;
    a.Jinv[3] = (J[5] * J[6] - J[3] * J[8]) * id;
    a.Jinv[4] = (J[0] * J[8] - J[2] * J[6]) * id;
    a.Jinv[5] = (J[2] * J[3] - J[0] * J[5]) * id;
    a.Jinv[6] = (J[3] * J[7] - J[4] * J[6]) * id;
    a.Jinv[7] = (J[1] * J[6] - J[0] * J[7]) * id;
    a.Jinv[8] = (J[0] * J[4] - J[1] * J[3]) * id;
  }
  return a;
}

void map_to_phys(const AffineMap& a, const double* xi, double* x) {
  for (int r = 0; r < a.dim; ++r) {
    double s = a.x0[r];
    for (int c = 0; c < a.dim; ++c) s += a.J[r * a.dim + c] * xi[c];
    x[r] = s;
  }
}

void map_to_ref(const AffineMap& a, const double* x, double* xi) {
  double dx[3];
  for (int r = 0; r < a.dim; ++r) dx[r] = x[r] - a.x0[r];
  for (int r = 0; r < a.dim; ++r) {
    double s = 0;
    for (int c = 0; c < a.dim; ++c) s += a.Jinv[r * a.dim + c] * dx[c];
    xi[r] = s;
  }
}

double cell_diameter(const Mesh& m, int level, int cell) {
  auto d2 = [&](int i, int j) {
    const double* a = m.vertex(level, cell, i);
    const double* b = m.vertex(level, cell, j);
    double s = 0;
    for (int d = 0; d < m.dim; ++d) s += (a[d] - b[d]) * (a[d] - b[d]);
    return s;
  };
  double diam = 0;
  if (m.shape == SHAPE_HEXA) {
    const int pairs[4][2] = {{0, 6}, {1, 7}, {2, 4}, {3, 5}};
    for (auto& p : pairs) { double l = d2(p[0], p[1]); if (l > diam) diam = l; }
  } else if (m.shape == SHAPE_QUAD) {
    diam = std::fmax(d2(0, 2), d2(1, 3));
  } else {
    const int pairs[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    for (auto& p : pairs) { double l = d2(p[0], p[1]); if (l > diam) diam = l; }
  }
  return std::sqrt(diam);
}

}  // namespace pmx
