/*
 * ORACLE — test infrastructure only, never part of the product.
 *
 * Plain-C, FP64, sequential restatement of the reference's CPU solve path
 * (nahmedwias/MyMooNMD), used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  It keeps the reference's loop
 * and summation orders:
 *   or_assemble_cd   Assemble3D/Assemble2D cell loop (src/AssembleRoutines/
 *                    Assemble3D.C:234-667, Assemble2D.C:7100-7570): reference map
 *                    per cell (THexaTrilinear/TQuadBilinear/TTriaAffin), per
 *                    quadrature point BilinearAssembleGalerkin<d> then
 *                    BilinearAssemble_SD<d> (ConvDiff.C:1794-2011), scatter by a
 *                    linear search of the CSR row (Assemble3D.C:354-371),
 *                    Dirichlet diagonal set to 1 (:393-405), Dirichlet rhs set to
 *                    the boundary value at the nodal points (:593-667), Neumann
 *                    line integrals in 2D (Assemble2D.C:7529-7570)
 *   or_assemble_tcd  the same cell loop with the TCD forms TCDStiff, TCDMass,
 *                    TCDRhs and their 2D SUPG variants (CD_local_assembling_
 *                    routines.C:6-330), matrices selected as in
 *                    LocalAssembling<d>::set_parameters_for_tcd (LocalAssembling.C:842-928)
 *   or_interpolate   TFEFunction2D::Interpolate for Lagrange elements
 *   or_errors        conv_diff_l2_h1_linf_error<d> (ConvDiff.C:2134-2185)
 *   or_apply_scaled_add  FEMatrix::multiplyActive + addScaledNonActive
 *                    (FEMatrix.C:186-199, BlockVector.C:212-228)
 *   or_prolongate / or_restrict  GridTransfer::Prolongate / DefectRestriction
 *                    (GridTransfer.C:665-1230): per child QQ*u, multiplicity
 *                    averaging; restriction divides by multiplicity, applies QQ^T
 *   or_solve         IterativeMethod::converged, flexible_gmres (Iteration_gmres.C:
 *                    374-532), Iteration_cg (Iteration_cg.C:22-79), Richardson,
 *                    Jacobi, Multigrid::cycle/cycle_step (Multigrid.C:157-415)
 *                    with CycleControl (CycleControl.C:17-111) and JacobiSmoother
 *
 * Parity pins (tests/): cd2d_test golden errors (Q1 quads, P1 triangles),
 * cd3d_solvers_test zero-error examples, matrix_test SpMV values,
 * iterative_solver_test 7x7 system, closed forms (Kronecker Q2 Laplacian,
 * partition of unity of the transfer, mass = volume).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "cd_oracle.h"

#define PI 3.14159265358979323846

/* ----------------------------------------------------------- basis tables */
/* 1D Lagrange factors as in BF_C_H_Q2_3D.h / BF_C_Q_Q1_2D.h */
static void lag1(int order, int a, double t, double* v, double* d, double* dd) {
  if (order == 1) {
    if (a == 0) { *v = 0.5 * (1 - t); *d = -0.5; *dd = 0; }
    else { *v = 0.5 * (1 + t); *d = 0.5; *dd = 0; }
    return;
  }
  if (a == 0) { *v = 0.5 * t * (t - 1); *d = t - 0.5; *dd = 1; }
  else if (a == 1) { *v = 1 - t * t; *d = -2 * t; *dd = -2; }
  else { *v = 0.5 * t * (t + 1); *d = t + 0.5; *dd = 1; }
}

int or_ndof(int shape, int order) {
  if (shape == OR_TRI) return 3;
  if (shape == OR_QUAD) return (order + 1) * (order + 1);
  return (order + 1) * (order + 1) * (order + 1);
}

/* values, reference gradients [dim][nd], second derivatives [nc][nd] */
void or_basis(int shape, int order, const double* xi, double* v, double* g, double* h) {
  const int nd = or_ndof(shape, order);
  if (shape == OR_TRI) {
    v[0] = 1 - xi[0] - xi[1]; v[1] = xi[0]; v[2] = xi[1];
    if (g) { g[0] = -1; g[1] = 1; g[2] = 0; g[3] = -1; g[4] = 0; g[5] = 1; }
    if (h) memset(h, 0, 9 * sizeof(double));
    return;
  }
  const int n1 = order + 1;
  for (int k = 0; k < nd; ++k) {
    int ix = k % n1, iy = (k / n1) % n1, iz = k / (n1 * n1);
    double fx, gx, hx, fy, gy, hy, fz = 1, gz = 0, hz = 0;
    lag1(order, ix, xi[0], &fx, &gx, &hx);
    lag1(order, iy, xi[1], &fy, &gy, &hy);
    if (shape == OR_HEXA) lag1(order, iz, xi[2], &fz, &gz, &hz);
    v[k] = fx * fy * fz;
    if (shape == OR_QUAD) {
      if (g) { g[k] = gx * fy; g[nd + k] = fx * gy; }
      if (h) { h[k] = hx * fy; h[nd + k] = gx * gy; h[2 * nd + k] = fx * hy; }
    } else {
      if (g) { g[k] = gx * fy * fz; g[nd + k] = fx * gy * fz; g[2 * nd + k] = fx * fy * gz; }
      if (h) {
        h[k] = hx * fy * fz; h[nd + k] = gx * gy * fz; h[2 * nd + k] = gx * fy * gz;
        h[3 * nd + k] = fx * hy * fz; h[4 * nd + k] = fx * gy * gz; h[5 * nd + k] = fx * fy * hz;
      }
    }
  }
}

/* reference nodal point of local dof k */
static void node_ref(int shape, int order, int k, double* xi) {
  if (shape == OR_TRI) {
    static const double p[3][2] = {{0, 0}, {1, 0}, {0, 1}};
    xi[0] = p[k][0]; xi[1] = p[k][1];
    return;
  }
  const int n1 = order + 1;
  int idx[3] = {k % n1, (k / n1) % n1, k / (n1 * n1)};
  for (int d = 0; d < 3; ++d) xi[d] = -1.0 + 2.0 * idx[d] / order;
}

/* --------------------------------------------------------- quadrature */
/* Gauss-Legendre on [-1,1] with the literal values of the reference tables
 * (QuadFormula1D.C, QuadFormulaQuad.C, QuadFormulaHexa.C) */
static int gauss1d(int n, double* x, double* w) {
  if (n == 2) {
    x[0] = -0.57735026918962576450914878050196; x[1] = -x[0];
    w[0] = w[1] = 1.0;
  } else if (n == 3) {
    x[0] = -0.77459666924148337703585307995647992; x[1] = 0; x[2] = -x[0];
    w[0] = w[2] = 0.55555555555555555555555555555556; w[1] = 0.88888888888888888888888888888889;
  } else {
    x[0] = -0.861136311594052575223946488893; x[1] = -0.339981043584856264802665759103;
    x[2] = -x[1]; x[3] = -x[0];
    w[0] = w[3] = 0.347854845137453857373063949; w[1] = w[2] = 0.652145154862546142626936051;
  }
  return n;
}

/* rule of the reference assembly: Q1 quad Gauss3, Q2 quad Gauss4, P1 tri
 * Gauss3Tria (7 points), Q1 hexa Gauss3, Q2 hexa Gauss4 (hexa_rule 2 -> Gauss3) */
int or_quad_rule(int shape, int order, int hexa_rule, double* xi, double* w) {
  if (shape == OR_TRI) {
    static const double ww[7] = {0.1125, 0.0629695902724135762978419727500906,
                                 0.0629695902724135762978419727500906,
                                 0.0629695902724135762978419727500906,
                                 0.0661970763942530903688246939165759,
                                 0.0661970763942530903688246939165759,
                                 0.0661970763942530903688246939165759};
    static const double xx[7] = {0.333333333333333333333333333333333, 0.797426985353087322398025276169754,
                                 0.101286507323456338800987361915123, 0.101286507323456338800987361915123,
                                 0.059715871789769820459117580973106, 0.470142064105115089770441209513447,
                                 0.470142064105115089770441209513447};
    static const double ee[7] = {0.333333333333333333333333333333333, 0.101286507323456338800987361915123,
                                 0.797426985353087322398025276169754, 0.101286507323456338800987361915123,
                                 0.470142064105115089770441209513447, 0.059715871789769820459117580973106,
                                 0.470142064105115089770441209513447};
    for (int k = 0; k < 7; ++k) { xi[2 * k] = xx[k]; xi[2 * k + 1] = ee[k]; w[k] = ww[k]; }
    return 7;
  }
  int n1;
  if (shape == OR_QUAD) n1 = (order == 1) ? 3 : 4;
  else n1 = (hexa_rule == 2) ? 3 : ((order == 1) ? 3 : 4);
  double x[4], ww[4];
  gauss1d(n1, x, ww);
  const int dim = (shape == OR_HEXA) ? 3 : 2;
  const int n = (dim == 3) ? n1 * n1 * n1 : n1 * n1;
  for (int k = 0; k < n; ++k) {
    int i = k % n1, j = (k / n1) % n1, l = k / (n1 * n1);
    xi[k * dim] = x[i]; xi[k * dim + 1] = x[j];
    if (dim == 3) { xi[k * dim + 2] = x[l]; w[k] = ww[i] * ww[j] * ww[l]; }
    else w[k] = ww[i] * ww[j];
  }
  return n;
}

/* ------------------------------------------------------- reference maps */
/* maps a reference point to physical coordinates and returns the Jacobian
 * (J[r][c] = dx_r / dxi_c) of the trilinear / bilinear / affine map */
static void ref_map(int shape, const double* v, const double* xi, double* x, double* J) {
  if (shape == OR_TRI) {
    for (int r = 0; r < 2; ++r) {
      J[r * 2] = v[2 + r] - v[r];
      J[r * 2 + 1] = v[4 + r] - v[r];
      x[r] = v[r] + J[r * 2] * xi[0] + J[r * 2 + 1] * xi[1];
    }
    return;
  }
  if (shape == OR_QUAD) {
    /* TQuadBilinear with v0 (-1,-1), v1 (1,-1), v2 (1,1), v3 (-1,1) */
    const double s[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
    for (int r = 0; r < 2; ++r) { x[r] = 0; J[r * 2] = 0; J[r * 2 + 1] = 0; }
    for (int a = 0; a < 4; ++a) {
      const double f = 0.25 * (1 + s[a][0] * xi[0]) * (1 + s[a][1] * xi[1]);
      const double fx = 0.25 * s[a][0] * (1 + s[a][1] * xi[1]);
      const double fy = 0.25 * (1 + s[a][0] * xi[0]) * s[a][1];
      for (int r = 0; r < 2; ++r) {
        x[r] += f * v[a * 2 + r];
        J[r * 2] += fx * v[a * 2 + r];
        J[r * 2 + 1] += fy * v[a * 2 + r];
      }
    }
    return;
  }
  /* THexaTrilinear: xc0..xc7 (HexaTrilinear.C:806-813) */
  double c[3][8];
  for (int r = 0; r < 3; ++r) {
    const double x0 = v[0 * 3 + r], x1 = v[1 * 3 + r], x2 = v[2 * 3 + r], x3 = v[3 * 3 + r];
    const double x4 = v[4 * 3 + r], x5 = v[5 * 3 + r], x6 = v[6 * 3 + r], x7 = v[7 * 3 + r];
    c[r][0] = (x0 + x1 + x3 + x2 + x4 + x5 + x7 + x6) * 0.125;
    c[r][1] = (-x0 + x1 - x3 + x2 - x4 + x5 - x7 + x6) * 0.125;
    c[r][2] = (-x0 - x1 + x3 + x2 - x4 - x5 + x7 + x6) * 0.125;
    c[r][3] = (-x0 - x1 - x3 - x2 + x4 + x5 + x7 + x6) * 0.125;
    c[r][4] = (x0 - x1 - x3 + x2 + x4 - x5 - x7 + x6) * 0.125;
    c[r][5] = (x0 - x1 + x3 - x2 - x4 + x5 - x7 + x6) * 0.125;
    c[r][6] = (x0 + x1 - x3 - x2 - x4 - x5 + x7 + x6) * 0.125;
    c[r][7] = (-x0 + x1 + x3 - x2 + x4 - x5 - x7 + x6) * 0.125;
  }
  const double a = xi[0], b = xi[1], g = xi[2];
  for (int r = 0; r < 3; ++r) {
    x[r] = c[r][0] + c[r][1] * a + c[r][2] * b + c[r][3] * g + c[r][4] * a * b + c[r][5] * a * g +
           c[r][6] * b * g + c[r][7] * a * b * g;
    J[r * 3 + 0] = c[r][1] + c[r][4] * b + c[r][5] * g + c[r][7] * b * g;
    J[r * 3 + 1] = c[r][2] + c[r][4] * a + c[r][6] * g + c[r][7] * a * g;
    J[r * 3 + 2] = c[r][3] + c[r][5] * a + c[r][6] * b + c[r][7] * a * b;
  }
}

static double inv_jac(int dim, const double* J, double* Ji) {
  if (dim == 2) {
    const double d = J[0] * J[3] - J[1] * J[2];
    Ji[0] = J[3] / d; Ji[1] = -J[1] / d; Ji[2] = -J[2] / d; Ji[3] = J[0] / d;
    return d;
  }
  const double d = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                   J[2] * (J[3] * J[7] - J[4] * J[6]);
  Ji[0] = (J[4] * J[8] - J[5] * J[7]) / d; Ji[1] = (J[2] * J[7] - J[1] * J[8]) / d;
  Ji[2] = (J[1] * J[5] - J[2] * J[4]) / d; Ji[3] = (J[5] * J[6] - J[3] * J[8]) / d;
  Ji[4] = (J[0] * J[8] - J[2] * J[6]) / d; Ji[5] = (J[2] * J[3] - J[0] * J[5]) / d;
  Ji[6] = (J[3] * J[7] - J[4] * J[6]) / d; Ji[7] = (J[1] * J[6] - J[0] * J[7]) / d;
  Ji[8] = (J[0] * J[4] - J[1] * J[3]) / d;
  return d;
}

/* physical point -> reference point of a parallelogram / parallelepiped /
 * triangle cell (Newton step of the map, exact for affine cells) */
static void phys_to_ref(int shape, const double* v, const double* x, double* xi) {
  const int dim = (shape == OR_HEXA) ? 3 : 2;
  double x0[3], J[9], Ji[9];
  for (int d = 0; d < 3; ++d) xi[d] = (shape == OR_TRI) ? 0.0 : 0.0;
  ref_map(shape, v, xi, x0, J);
  inv_jac(dim, J, Ji);
  double dx[3];
  for (int d = 0; d < dim; ++d) dx[d] = x[d] - x0[d];
  for (int r = 0; r < dim; ++r) {
    double s = 0;
    for (int c = 0; c < dim; ++c) s += Ji[r * dim + c] * dx[c];
    xi[r] = s;
  }
}

static double cell_diameter(int shape, const double* v) {
  const int dim = (shape == OR_HEXA) ? 3 : 2;
#define D2(i, j) ({ double s_ = 0; for (int d_ = 0; d_ < dim; ++d_) { double t_ = v[(i) * dim + d_] - v[(j) * dim + d_]; s_ += t_ * t_; } s_; })
  double diam = 0, len;
  if (shape == OR_HEXA) {
    len = D2(0, 6); if (len > diam) diam = len;
    len = D2(1, 7); if (len > diam) diam = len;
    len = D2(2, 4); if (len > diam) diam = len;
    len = D2(3, 5); if (len > diam) diam = len;
  } else if (shape == OR_QUAD) {
    const double a = D2(0, 2), b = D2(1, 3);
    diam = a > b ? a : b;
  } else {
    len = D2(0, 1); if (len > diam) diam = len;
    len = D2(1, 2); if (len > diam) diam = len;
    len = D2(2, 0); if (len > diam) diam = len;
  }
#undef D2
  return sqrt(diam);
}

/* ------------------------------------------------------------ examples */
/* v[0] = u, v[1..3] = grad u, v[4] = laplace u, v[5] = du/dt at time t */
static void exact(int ex, double x, double y, double z, double t, double* v) {
  v[0] = v[1] = v[2] = v[3] = v[4] = v[5] = 0;
  switch (ex) {
    case OR_EX_TCD_EXP: {
      /* include/Examples/TCD_2D/exp.h: u = e^t sin(2 pi x) sin(2 pi y) */
      const double e = exp(t), sx = sin(2 * PI * x), sy = sin(2 * PI * y);
      v[0] = e * (sx * sy);
      v[1] = e * 2 * PI * cos(2 * PI * x) * sy;
      v[2] = e * 2 * PI * sx * cos(2 * PI * y);
      v[4] = -8 * PI * PI * v[0];
      v[5] = v[0];
      break;
    }
    case OR_EX_TCD_LINEAR:
      /* include/Examples/TCD_2D/linear_space_time.h: u = 1 + 2x + 3ty */
      v[0] = 1 + 2 * x + 3 * t * y; v[1] = 2; v[2] = 3 * t; v[5] = 3 * y;
      break;
    case OR_EX_SINE_LAPLACE_2D:
    case OR_EX_SINE_2D:
      v[0] = sin(PI * x) * sin(PI * y);
      v[1] = PI * cos(PI * x) * sin(PI * y);
      v[2] = PI * sin(PI * x) * cos(PI * y);
      v[4] = -2 * PI * PI * sin(PI * x) * sin(PI * y);
      break;
    case OR_EX_P0_3D: v[0] = 1; break;
    case OR_EX_P1_3D: v[0] = 1 + 2 * x + 3 * y + 4 * z; v[1] = 2; v[2] = 3; v[3] = 4; break;
    case OR_EX_P2_3D: v[0] = x * x + y * z; v[1] = 2 * x; v[2] = z; v[3] = y; v[4] = 2; break;
    case OR_EX_EXP_SIN_3D:
      v[0] = exp(-x) * sin(PI * y) * sin(PI * z);
      v[1] = -v[0];
      v[2] = PI * exp(-x) * cos(PI * y) * sin(PI * z);
      v[3] = PI * exp(-x) * sin(PI * y) * cos(PI * z);
      v[4] = (1 - 2 * PI * PI) * v[0];
      break;
    default: break;
  }
}

/* coefficients [eps, b1, b2, b3, c, f] of the example's BilinearCoeffs */
static void coeffs(const or_cd* p, double x, double y, double z, double* c) {
  double v[6];
  c[0] = p->eps; c[1] = p->b[0]; c[2] = p->b[1]; c[3] = p->b[2]; c[4] = p->c;
  if (p->example == OR_EX_CONST) { c[5] = p->f0; return; }
  exact(p->example, x, y, z, p->time, v);
  c[5] = v[5] - c[0] * v[4];
  c[5] += c[1] * v[1] + c[2] * v[2] + c[3] * v[3];
  c[5] += c[4] * v[0];
}

static double boundary_value(const or_cd* p, double x, double y, double z) {
  double v[6];
  if (p->example == OR_EX_SINE_LAPLACE_2D || p->example == OR_EX_CONST ||
      p->example == OR_EX_TCD_EXP)
    return 0.0;
  exact(p->example, x, y, z, p->time, v);
  return v[0];
}

/* Compute_SDFEM_delta, SDFEM_TYPE 2 (ConvDiff.C:273-345) */
static double sdfem_delta(double hK, double eps, const double* b, int dim, const or_cd* p) {
  if (fabs(eps) < 1e-20) eps = 1e-20;
  double nb = b[0] * b[0] + b[1] * b[1];
  if (dim == 3) nb += b[2] * b[2];
  nb = sqrt(nb);
  const double nu = 1.0 / p->poly_degree;
  double delta = 0;
  if (nb > 0) {
    const double alpha = nu * nb * hK / (2 * eps);
    delta = nu * hK * (1 / tanh(alpha) - 1 / alpha) / (2 * nb);
  }
  return delta * p->delta0;
}

/* ------------------------------------------------------------ assembly */
static int nfaces(int shape) { return shape == OR_TRI ? 3 : (shape == OR_QUAD ? 4 : 6); }

static int face_vertices(int shape, int j, int* fv) {
  static const int quad[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
  static const int tri[3][2] = {{0, 1}, {1, 2}, {2, 0}};
  static const int hexa[6][4] = {{0, 1, 2, 3}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 7, 6, 5}};
  if (shape == OR_QUAD) { fv[0] = quad[j][0]; fv[1] = quad[j][1]; return 2; }
  if (shape == OR_TRI) { fv[0] = tri[j][0]; fv[1] = tri[j][1]; return 2; }
  for (int k = 0; k < 4; ++k) fv[k] = hexa[j][k];
  return 4;
}

/* is local dof k on joint j (nodal point on the face of the reference cell) */
static int dof_on_joint(int shape, int order, int k, int j) {
  double xi[3];
  node_ref(shape, order, k, xi);
  if (shape == OR_TRI) {
    if (j == 0) return xi[1] == 0.0;
    if (j == 1) return xi[0] + xi[1] == 1.0;
    return xi[0] == 0.0;
  }
  if (shape == OR_QUAD) {
    static const int ax[4] = {1, 0, 1, 0};
    static const double val[4] = {-1, 1, 1, -1};
    return xi[ax[j]] == val[j];
  }
  static const int ax[6] = {2, 1, 0, 1, 0, 2};
  static const double val[6] = {-1, -1, 1, 1, -1, 1};
  return xi[ax[j]] == val[j];
}

static double bparam(int comp, double x, double y) {
  switch (comp) {
    case 0: return x;
    case 1: return y;
    case 2: return 1 - x;
    default: return 1 - y;
  }
}

int or_quad_points(const or_space* s, int hexa_rule, double* xyz) {
  const int shape = s->shape, dim = (shape == OR_HEXA) ? 3 : 2;
  double qxi[64 * 3], qw[64];
  const int nq = or_quad_rule(shape, s->order, hexa_rule, qxi, qw);
  for (int cell = 0; cell < s->n_cells; ++cell) {
    const double* v = s->verts + (size_t)cell * s->nv * dim;
    for (int q = 0; q < nq; ++q) {
      double x[3] = {0, 0, 0}, J[9];
      ref_map(shape, v, qxi + q * dim, x, J);
      for (int k = 0; k < 3; ++k) xyz[((size_t)cell * nq + q) * 3 + k] = x[k];
    }
  }
  return nq;
}

void or_assemble_cd(const or_space* s, const or_cd* p, double* val, double* rhs) {
  const int shape = s->shape, dim = (shape == OR_HEXA) ? 3 : 2, nd = s->ndof_cell;
  const int nc = (dim == 3) ? 6 : 3;
  double qxi[64 * 3], qw[64];
  const int nq = or_quad_rule(shape, s->order, p->hexa_rule, qxi, qw);
  double* u = malloc(sizeof(double) * nd);
  double* gr = malloc(sizeof(double) * dim * nd);
  double* hr = malloc(sizeof(double) * nc * nd);
  double* ux = malloc(sizeof(double) * 3 * nd);
  double* lap = malloc(sizeof(double) * nd);
  double* A = malloc(sizeof(double) * nd * nd);
  double* F = malloc(sizeof(double) * nd);
  for (int cell = 0; cell < s->n_cells; ++cell) {
    const double* v = s->verts + (size_t)cell * s->nv * dim;
    const int* dof = s->gnum + (size_t)cell * nd;
    const double hK = cell_diameter(shape, v);
    memset(A, 0, sizeof(double) * nd * nd);
    memset(F, 0, sizeof(double) * nd);
    for (int q = 0; q < nq; ++q) {
      double x[3] = {0, 0, 0}, J[9], Ji[9];
      ref_map(shape, v, qxi + q * dim, x, J);
      const double det = inv_jac(dim, J, Ji);
      or_basis(shape, s->order, qxi + q * dim, u, gr, hr);
      for (int k = 0; k < nd; ++k) {
        for (int a = 0; a < dim; ++a) {
          double t = 0;
          for (int r = 0; r < dim; ++r) t += gr[r * nd + k] * Ji[r * dim + a];
          ux[a * nd + k] = t;
        }
        if (dim == 2) ux[2 * nd + k] = 0;
        /* physical Laplacian of an affine map: trace(Ji^T H Ji) */
        double l = 0;
        for (int a = 0; a < dim; ++a)
          for (int r = 0; r < dim; ++r)
            for (int t = 0; t < dim; ++t) {
              int comp;
              if (dim == 3) {
                static const int idx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
                comp = idx[r][t];
              } else {
                static const int idx2[2][2] = {{0, 1}, {1, 2}};
                comp = idx2[r][t];
              }
              l += hr[comp * nd + k] * Ji[r * dim + a] * Ji[t * dim + a];
            }
        lap[k] = l;
      }
      double cf[6];
      if (p->table) {
        const double* t = p->table + ((size_t)cell * nq + q) * 8;
        for (int k = 0; k < 6; ++k) cf[k] = t[k];
        if (dim == 2) cf[3] = 0.0;
      } else {
        coeffs(p, x[0], x[1], dim == 3 ? x[2] : 0.0, cf);
      }
      const double Mult = qw[q] * fabs(det);
      const double eps = cf[0], b1 = cf[1], b2 = cf[2], b3 = (dim == 3) ? cf[3] : 0.0;
      const double c = cf[4], f = cf[5];
      /* BilinearAssembleGalerkin<d> */
      for (int i = 0; i < nd; ++i) {
        const double t0 = u[i], t1 = ux[i], t2 = ux[nd + i], t3 = ux[2 * nd + i];
        F[i] += Mult * t0 * f;
        for (int j = 0; j < nd; ++j) {
          const double a1 = ux[j], a2 = ux[nd + j], a3 = ux[2 * nd + j], a0 = u[j];
          double vv = eps * (t1 * a1 + t2 * a2 + t3 * a3);
          vv += (b1 * a1 + b2 * a2 + b3 * a3) * t0;
          vv += c * a0 * t0;
          A[i * nd + j] += vv * Mult;
        }
      }
      /* BilinearAssemble_SD<d> */
      if (p->disc == OR_SUPG) {
        const double bb[3] = {b1, b2, b3};
        const double delta = sdfem_delta(hK, eps, bb, dim, p);
        for (int i = 0; i < nd; ++i) {
          const double bgradv = b1 * ux[i] + b2 * ux[nd + i] + b3 * ux[2 * nd + i];
          F[i] += Mult * (delta * bgradv) * f;
          for (int j = 0; j < nd; ++j) {
            double vv = -eps * lap[j];
            vv += b1 * ux[j] + b2 * ux[nd + j] + b3 * ux[2 * nd + j] + c * u[j];
            A[i * nd + j] += vv * delta * bgradv * Mult;
          }
        }
      }
      /* BilinearAssemble_GLS<d> (ConvDiff.C:2012-2131): test functions
         L phi_i = -eps lap phi_i + b.grad phi_i + c phi_i, same delta */
      if (p->disc == OR_GLS) {
        const double bb[3] = {b1, b2, b3};
        const double delta = sdfem_delta(hK, eps, bb, dim, p);
        for (int i = 0; i < nd; ++i) {
          double Lu;
          if (dim == 3) {
            Lu = -eps * lap[i];
            Lu += b1 * ux[i] + b2 * ux[nd + i] + b3 * ux[2 * nd + i] + c * u[i];
          } else {
            Lu = -eps * lap[i] + b1 * ux[i] + b2 * ux[nd + i] + c * u[i];
          }
          F[i] += Mult * delta * Lu * f;
          for (int j = 0; j < nd; ++j) {
            if (dim == 3) {
              double vv = -eps * lap[j];
              vv += b1 * ux[j] + b2 * ux[nd + j] + b3 * ux[2 * nd + j] + c * u[j];
              A[i * nd + j] += vv * (Mult * delta * Lu);
            } else {
              const double vv = delta * (-eps * lap[j] + b1 * ux[j] + b2 * ux[nd + j] + c * u[j]) * Lu;
              A[i * nd + j] += vv * Mult;
            }
          }
        }
      }
    }
    /* scatter with a linear search of the row */
    for (int m = 0; m < nd; ++m) {
      const int l = dof[m];
      if (l < s->active) {
        for (int k = 0; k < nd; ++k)
          for (int n = s->rp[l]; n < s->rp[l + 1]; ++n)
            if (s->col[n] == dof[k]) { val[n] += A[m * nd + k]; break; }
        rhs[l] += F[m];
      } else {
        for (int n = s->rp[l]; n < s->rp[l + 1]; ++n)
          if (s->col[n] == l) { val[n] = 1.0; break; }
      }
    }
    /* boundary joints: Dirichlet values at nodal points, Neumann integrals */
    for (int j = 0; j < nfaces(shape); ++j) {
      const int comp = s->bd_comp[cell * nfaces(shape) + j];
      if (comp < 0) continue;
      int fv[4];
      const int nfv = face_vertices(shape, j, fv);
      int cond = OR_DIRICHLET;
      if (dim == 2 && p->bc_sine_laplace && comp == 1) cond = OR_NEUMANN;
      if (cond == OR_DIRICHLET) {
        for (int k = 0; k < nd; ++k) {
          if (!dof_on_joint(shape, s->order, k, j)) continue;
          double xi[3], x[3] = {0, 0, 0}, J[9];
          node_ref(shape, s->order, k, xi);
          ref_map(shape, v, xi, x, J);
          rhs[dof[k]] = boundary_value(p, x[0], x[1], dim == 3 ? x[2] : 0.0);
        }
      } else if (dim == 2) {
        /* Neumann: Gauss line rule of degree 2*order */
        const int npts = (s->order == 1) ? 2 : 3;
        double zx[4], zw[4];
        gauss1d(npts, zx, zw);
        const double* a = v + fv[0] * 2;
        const double* b = v + fv[1] * 2;
        const double t0 = bparam(comp, a[0], a[1]), t1 = bparam(comp, b[0], b[1]);
        const double hE = sqrt((b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1])) / 2;
        for (int l = 0; l < npts; ++l) {
          const double t = t0 + 0.5 * (t1 - t0) * (zx[l] + 1);
          double ev[6];
          exact(p->example, 1.0, t, 0.0, p->time, ev);
          double g = p->eps * ev[1];
          g *= hE * zw[l];
          double x[2] = {a[0] + 0.5 * (zx[l] + 1) * (b[0] - a[0]), a[1] + 0.5 * (zx[l] + 1) * (b[1] - a[1])};
          double xi[3];
          phys_to_ref(shape, v, x, xi);
          or_basis(shape, s->order, xi, u, NULL, NULL);
          for (int k = 0; k < nd; ++k)
            if (dof[k] < s->active) rhs[dof[k]] += g * u[k];
        }
      }
      (void)nfv;
    }
  }
  free(u); free(gr); free(hr); free(ux); free(lap); free(A); free(F);
}

void or_errors(const or_space* s, const or_cd* p, const double* sol, double* err) {
  const int shape = s->shape, dim = (shape == OR_HEXA) ? 3 : 2, nd = s->ndof_cell;
  double qxi[64 * 3], qw[64];
  const int nq = or_quad_rule(shape, s->order, p->hexa_rule, qxi, qw);
  double* u = malloc(sizeof(double) * nd);
  double* gr = malloc(sizeof(double) * dim * nd);
  double e0s = 0, e1s = 0, sds = 0, linf = 0;
  for (int cell = 0; cell < s->n_cells; ++cell) {
    const double* v = s->verts + (size_t)cell * s->nv * dim;
    const int* dof = s->gnum + (size_t)cell * nd;
    const double hK = cell_diameter(shape, v);
    double le[4] = {0, 0, 0, 0};
    for (int q = 0; q < nq; ++q) {
      double x[3] = {0, 0, 0}, J[9], Ji[9];
      ref_map(shape, v, qxi + q * dim, x, J);
      const double det = inv_jac(dim, J, Ji);
      or_basis(shape, s->order, qxi + q * dim, u, gr, NULL);
      double val = 0, g[3] = {0, 0, 0};
      for (int k = 0; k < nd; ++k) {
        val += sol[dof[k]] * u[k];
        for (int a = 0; a < dim; ++a) {
          double t = 0;
          for (int r = 0; r < dim; ++r) t += gr[r * nd + k] * Ji[r * dim + a];
          g[a] += sol[dof[k]] * t;
        }
      }
      double ev[6], cf[6];
      exact(p->example, x[0], x[1], dim == 3 ? x[2] : 0.0, p->time, ev);
      coeffs(p, x[0], x[1], dim == 3 ? x[2] : 0.0, cf);
      const double bb[3] = {cf[1], cf[2], dim == 3 ? cf[3] : 0.0};
      const double delta = sdfem_delta(hK, cf[0], bb, dim, p);
      const double w = qw[q] * fabs(det);
      const double e0 = val - ev[0], e1 = g[0] - ev[1], e2 = g[1] - ev[2];
      const double e3 = (dim == 3) ? g[2] - ev[3] : 0.0;
      le[0] += w * e0 * e0;
      if (fabs(e0) > le[3]) le[3] = fabs(e0);
      le[1] += w * e1 * e1;
      le[1] += w * e2 * e2;
      le[1] += w * e3 * e3;
      const double e4 = bb[0] * e1 + bb[1] * e2 + bb[2] * e3;
      le[2] += w * (cf[0] * (e1 * e1 + e2 * e2 + e3 * e3) + cf[4] * e0 * e0 + delta * e4 * e4);
    }
    e0s += le[0]; e1s += le[1]; sds += le[2];
    if (le[3] > linf) linf = le[3];
  }
  err[0] = sqrt(e0s); err[1] = sqrt(e1s); err[2] = sqrt(sds); err[3] = linf;
  free(u); free(gr);
}

/* Physical basis data at one reference point of an affine cell: values u,
 * gradients ux[3][nd] (third row zero in 2D), Laplacians lap (if non-NULL),
 * the physical point x and |det J|. */
static double phys_basis(int shape, int order, const double* v, const double* xi, double* u,
                         double* gr, double* hr, double* ux, double* lap, double* x) {
  const int dim = (shape == OR_HEXA) ? 3 : 2, nd = or_ndof(shape, order);
  double J[9], Ji[9];
  x[0] = x[1] = x[2] = 0;
  ref_map(shape, v, xi, x, J);
  const double det = inv_jac(dim, J, Ji);
  or_basis(shape, order, xi, u, gr, hr);
  for (int k = 0; k < nd; ++k) {
    for (int a = 0; a < dim; ++a) {
      double t = 0;
      for (int r = 0; r < dim; ++r) t += gr[r * nd + k] * Ji[r * dim + a];
      ux[a * nd + k] = t;
    }
    if (dim == 2) ux[2 * nd + k] = 0;
    if (lap) {
      double l = 0;
      for (int a = 0; a < dim; ++a)
        for (int r = 0; r < dim; ++r)
          for (int t = 0; t < dim; ++t) {
            static const int idx3[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
            static const int idx2[2][2] = {{0, 1}, {1, 2}};
            const int comp = (dim == 3) ? idx3[r][t] : idx2[r][t];
            l += hr[comp * nd + k] * Ji[r * dim + a] * Ji[t * dim + a];
          }
      lap[k] = l;
    }
  }
  return fabs(det);
}

/* scatter of one local matrix: active rows add (linear search of the row),
 * Dirichlet rows get their diagonal set to 1 */
static void scatter_matrix(const or_space* s, const int* dof, const double* A, double* val) {
  const int nd = s->ndof_cell;
  for (int m = 0; m < nd; ++m) {
    const int l = dof[m];
    if (l < s->active) {
      for (int k = 0; k < nd; ++k)
        for (int n = s->rp[l]; n < s->rp[l + 1]; ++n)
          if (s->col[n] == dof[k]) { val[n] += A[m * nd + k]; break; }
    } else {
      for (int n = s->rp[l]; n < s->rp[l + 1]; ++n)
        if (s->col[n] == l) { val[n] = 1.0; break; }
    }
  }
}

/* Time-dependent convection-diffusion forms, per quadrature point in the
 * order LocalAssembling<d> (TCDStiffMassRhs) calls them
 * (src/AssembleRoutines/LocalAssembling.C:862-895):
 *   TCDStiff, TCDMass, TCDRhs (CD_local_assembling_routines.C:6-40, 83-93, 118-128)
 *   and for SUPG in 2D TCDStiffSUPG, TCDMassSUPG, TCDRhsSUPG (:130-330), whose
 *   coefficients are eps*theta1*tau, b*theta1*tau, 1 + c*theta1*tau, with the
 *   mass-term delta multiplied by theta1*tau.
 * Scatter and Dirichlet handling as Assemble2D/3D (every square matrix gets
 * diagonal 1 in Dirichlet rows; rhs gets the boundary value). */
void or_assemble_tcd(const or_space* s, const or_cd* p, double* aval, double* mval, double* rhs) {
  const int shape = s->shape, dim = (shape == OR_HEXA) ? 3 : 2, nd = s->ndof_cell;
  const int nc = (dim == 3) ? 6 : 3;
  const int supg = (p->disc == OR_SUPG) && dim == 2;
  double qxi[64 * 3], qw[64];
  const int nq = or_quad_rule(shape, s->order, p->hexa_rule, qxi, qw);
  double* u = malloc(sizeof(double) * nd);
  double* gr = malloc(sizeof(double) * dim * nd);
  double* hr = malloc(sizeof(double) * nc * nd);
  double* ux = malloc(sizeof(double) * 3 * nd);
  double* lap = malloc(sizeof(double) * nd);
  double* A = malloc(sizeof(double) * nd * nd);
  double* M = malloc(sizeof(double) * nd * nd);
  double* F = malloc(sizeof(double) * nd);
  for (int cell = 0; cell < s->n_cells; ++cell) {
    const double* v = s->verts + (size_t)cell * s->nv * dim;
    const int* dof = s->gnum + (size_t)cell * nd;
    const double hK = cell_diameter(shape, v);
    memset(A, 0, sizeof(double) * nd * nd);
    memset(M, 0, sizeof(double) * nd * nd);
    memset(F, 0, sizeof(double) * nd);
    for (int q = 0; q < nq; ++q) {
      double x[3];
      const double det = phys_basis(shape, s->order, v, qxi + q * dim, u, gr, hr, ux, lap, x);
      double cf[6];
      coeffs(p, x[0], x[1], dim == 3 ? x[2] : 0.0, cf);
      const double Mult = qw[q] * det;
      const double eps = cf[0], b1 = cf[1], b2 = cf[2], b3 = (dim == 3) ? cf[3] : 0.0;
      const double c = cf[4], f = cf[5];
      /* TCDStiff */
      for (int i = 0; i < nd; ++i) {
        const double t0 = u[i], tx = ux[i], ty = ux[nd + i], tz = ux[2 * nd + i];
        for (int j = 0; j < nd; ++j) {
          double val = eps * (tx * ux[j] + ty * ux[nd + j] + tz * ux[2 * nd + j]);
          val += (b1 * ux[j] + b2 * ux[nd + j] + b3 * ux[2 * nd + j]) * t0;
          val += c * t0 * u[j];
          A[i * nd + j] += Mult * val;
        }
      }
      /* TCDMass */
      for (int i = 0; i < nd; ++i)
        for (int j = 0; j < nd; ++j) M[i * nd + j] += Mult * u[i] * u[j];
      /* TCDRhs */
      for (int i = 0; i < nd; ++i) F[i] += Mult * u[i] * f;
      if (supg) {
        const double t1 = p->theta1, tau = p->tau;
        const double se = eps * t1 * tau, sb1 = b1 * t1 * tau, sb2 = b2 * t1 * tau;
        const double sc = 1. + c * t1 * tau;
        const double sbv[3] = {sb1, sb2, 0.0};
        const double delta = sdfem_delta(hK, se, sbv, 2, p);
        /* TCDStiffSUPG */
        for (int i = 0; i < nd; ++i) {
          const double bgradv = delta * (sb1 * ux[i] + sb2 * ux[nd + i]);
          for (int j = 0; j < nd; ++j) {
            double val = (sb1 * ux[j] + sb2 * ux[nd + j]) * bgradv;
            val += sc * bgradv * u[j];
            val -= se * lap[j] * bgradv;
            A[i * nd + j] += Mult * val;
          }
        }
        /* TCDMassSUPG */
        const double dm = delta * t1 * tau;
        for (int i = 0; i < nd; ++i) {
          const double bgradv = dm * (sb1 * ux[i] + sb2 * ux[nd + i]);
          for (int j = 0; j < nd; ++j) M[i * nd + j] += Mult * u[j] * bgradv;
        }
        /* TCDRhsSUPG */
        for (int i = 0; i < nd; ++i) {
          const double bgradv = delta * (sb1 * ux[i] + sb2 * ux[nd + i]);
          F[i] += Mult * f * bgradv;
        }
      }
    }
    if (aval) scatter_matrix(s, dof, A, aval);
    if (mval) scatter_matrix(s, dof, M, mval);
    if (rhs) {
      for (int m = 0; m < nd; ++m)
        if (dof[m] < s->active) rhs[dof[m]] += F[m];
      for (int j = 0; j < nfaces(shape); ++j) {
        if (s->bd_comp[cell * nfaces(shape) + j] < 0) continue;
        for (int k = 0; k < nd; ++k) {
          if (!dof_on_joint(shape, s->order, k, j)) continue;
          double xi[3], x[3] = {0, 0, 0}, J[9];
          node_ref(shape, s->order, k, xi);
          ref_map(shape, v, xi, x, J);
          rhs[dof[k]] = boundary_value(p, x[0], x[1], dim == 3 ? x[2] : 0.0);
        }
      }
    }
  }
  free(u); free(gr); free(hr); free(ux); free(lap); free(A); free(M); free(F);
}

/* TFEFunction2D/3D::Interpolate of the exact solution (Lagrange nodal points) */
void or_interpolate(const or_space* s, const or_cd* p, double* out) {
  const int shape = s->shape, dim = (shape == OR_HEXA) ? 3 : 2, nd = s->ndof_cell;
  for (int cell = 0; cell < s->n_cells; ++cell) {
    const double* v = s->verts + (size_t)cell * s->nv * dim;
    const int* dof = s->gnum + (size_t)cell * nd;
    for (int k = 0; k < nd; ++k) {
      double xi[3], x[3] = {0, 0, 0}, J[9], ev[6];
      node_ref(shape, s->order, k, xi);
      ref_map(shape, v, xi, x, J);
      exact(p->example, x[0], x[1], dim == 3 ? x[2] : 0.0, p->time, ev);
      out[dof[k]] = ev[0];
    }
  }
}

void or_multiply_active(int active, const int* rp, const int* col, const double* val,
                        const double* x, double* y, double a) {
  for (int i = 0; i < active; ++i) {
    double s = 0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) s += val[k] * x[col[k]];
    y[i] += a * s;
  }
}

/* test_u_2_p_1.h (Example_NSE3D -3): u = (y^2, xz, -y^2), p = x + y + z - 1.5;
 * ue[k] = u_k, grad u_k (1..3), lap u_k (4); pe = p, grad p */
static void nse_exact(double x, double y, double z, double ue[3][5], double pe[4]) {
  for (int k = 0; k < 3; ++k) for (int a = 0; a < 5; ++a) ue[k][a] = 0;
  ue[0][0] = y * y; ue[0][2] = 2 * y; ue[0][4] = 2;
  ue[1][0] = x * z; ue[1][1] = z; ue[1][3] = x;
  ue[2][0] = -y * y; ue[2][2] = -2 * y; ue[2][4] = -2;
  pe[0] = x + y + z - 1.5; pe[1] = 1; pe[2] = 1; pe[3] = 1;
}

/* test_u_0_p_0.h (Example_NSE3D -1): u = (1, 0, 0), p = 0 */
static void nse_exact_u0p0(double x, double y, double z, double ue[3][5], double pe[4]) {
  (void)x; (void)y; (void)z;
  for (int k = 0; k < 3; ++k) for (int a = 0; a < 5; ++a) ue[k][a] = 0;
  ue[0][0] = 1;
  for (int a = 0; a < 4; ++a) pe[a] = 0;
}

/* discontinuous P1 on hexahedra (include/FE3D/BF_D_H_P1_3D.h): 1, xi, eta, zeta */
static void p1disc_basis(const double* xi, double* v) {
  v[0] = 1; v[1] = xi[0]; v[2] = xi[1]; v[3] = xi[2];
}

static void add_entry(const int* rp, const int* col, int row, int c, double* val, double v) {
  for (int n = rp[row]; n < rp[row + 1]; ++n)
    if (col[n] == c) { val[n] += v; return; }
}

/* Assemble3D with LocalAssembling<3> NavierStokesAll + NavierStokesNL for NSTYPE 1
 * (LocalAssembling.C:931-1185): per quadrature point NSLaplaceGradGradSingle,
 * NSNonlinearTerm_convective_Single, NSDivergenceBlocks (sign 1),
 * NSRightHandSide (sign -1) (NSE_local_assembling_routines.C:52-283); the
 * coefficient function is LinCoeffs of test_u_2_p_1.h. */
void or_assemble_nse2(const or_space* vs, const or_space* ps, int np_cell, int example, double nu,
                      const double* w, double* A, const int* brp, const int* bcol, double* B1, double* B2,
                      double* B3, double* rhs) {
  const int nu_d = vs->n_dof;
  const int npc = np_cell;   /* 8: Q1 pressure, 4: discontinuous P1 */
  void (*exact)(double, double, double, double[3][5], double[4]) = example == 1 ? nse_exact_u0p0 : nse_exact;
  double qxi[64 * 3], qw[64];
  const int nq = or_quad_rule(OR_HEXA, 2, 0, qxi, qw);
  double u[27], gr[81], hr[162], ux[81], pv[8], pg[24];
  double Al[27 * 27], Bl[3][8 * 27], Fl[3][27], Gl[8];
  for (int cell = 0; cell < vs->n_cells; ++cell) {
    const double* v = vs->verts + (size_t)cell * 8 * 3;
    const int* vd = vs->gnum + (size_t)cell * 27;
    const int* pd = ps->gnum + (size_t)cell * npc;
    memset(Al, 0, sizeof Al); memset(Bl, 0, sizeof Bl); memset(Fl, 0, sizeof Fl); memset(Gl, 0, sizeof Gl);
    for (int q = 0; q < nq; ++q) {
      double x[3];
      const double det = phys_basis(OR_HEXA, 2, v, qxi + 3 * q, u, gr, hr, ux, NULL, x);
      if (npc == 4) p1disc_basis(qxi + 3 * q, pv);
      else or_basis(OR_HEXA, 1, qxi + 3 * q, pv, pg, NULL);
      const double Mult = qw[q] * det;
      double ue[3][5], pe[4], f[3];
      exact(x[0], x[1], x[2], ue, pe);
      for (int k = 0; k < 3; ++k)
        f[k] = -nu * ue[k][4] + (ue[0][0] * ue[k][1] + ue[1][0] * ue[k][2] + ue[2][0] * ue[k][3]) + pe[1 + k];
      const double g = 0.0;
      double wq[3] = {0, 0, 0};
      if (w)
        for (int k = 0; k < 3; ++k)
          for (int j = 0; j < 27; ++j) wq[k] += w[(size_t)k * nu_d + vd[j]] * u[j];
      /* NSLaplaceGradGradSingle */
      for (int i = 0; i < 27; ++i)
        for (int j = 0; j < 27; ++j)
          Al[i * 27 + j] += Mult * nu * (ux[i] * ux[j] + ux[27 + i] * ux[27 + j] + ux[54 + i] * ux[54 + j]);
      /* NSNonlinearTerm_convective_Single */
      if (w)
        for (int i = 0; i < 27; ++i)
          for (int j = 0; j < 27; ++j)
            Al[i * 27 + j] += Mult * (wq[0] * ux[j] + wq[1] * ux[27 + j] + wq[2] * ux[54 + j]) * u[i];
      /* NSDivergenceBlocks, sign 1 */
      for (int i = 0; i < npc; ++i)
        for (int j = 0; j < 27; ++j) {
          Bl[0][i * 27 + j] -= 1 * Mult * pv[i] * ux[j];
          Bl[1][i * 27 + j] -= 1 * Mult * pv[i] * ux[27 + j];
          Bl[2][i * 27 + j] -= Mult * 1 * pv[i] * ux[54 + j];
        }
      /* NSRightHandSide, sign -1 */
      for (int i = 0; i < 27; ++i)
        for (int k = 0; k < 3; ++k) Fl[k][i] += Mult * u[i] * f[k];
      for (int i = 0; i < npc; ++i) Gl[i] -= Mult * (-1) * pv[i] * g;
    }
    for (int m = 0; m < 27; ++m) {
      const int r = vd[m];
      if (r < vs->active) {
        for (int k = 0; k < 27; ++k) add_entry(vs->rp, vs->col, r, vd[k], A, Al[m * 27 + k]);
        for (int c = 0; c < 3; ++c) rhs[(size_t)c * nu_d + r] += Fl[c][m];
      } else {
        for (int n = vs->rp[r]; n < vs->rp[r + 1]; ++n)
          if (vs->col[n] == r) { A[n] = 1.0; break; }
      }
    }
    for (int i = 0; i < npc; ++i) {
      for (int j = 0; j < 27; ++j) {
        add_entry(brp, bcol, pd[i], vd[j], B1, Bl[0][i * 27 + j]);
        add_entry(brp, bcol, pd[i], vd[j], B2, Bl[1][i * 27 + j]);
        add_entry(brp, bcol, pd[i], vd[j], B3, Bl[2][i * 27 + j]);
      }
      rhs[(size_t)3 * nu_d + pd[i]] += Gl[i];
    }
    /* Dirichlet velocity rhs: boundary values at the nodal points */
    for (int k = 0; k < 27; ++k) {
      if (vd[k] < vs->active) continue;
      double xi[3], xx[3] = {0, 0, 0}, J[9], ue[3][5], pe[4];
      node_ref(OR_HEXA, 2, k, xi);
      ref_map(OR_HEXA, v, xi, xx, J);
      exact(xx[0], xx[1], xx[2], ue, pe);
      for (int c = 0; c < 3; ++c) rhs[(size_t)c * nu_d + vd[k]] = ue[c][0];
    }
  }
}

void or_assemble_nse(const or_space* vs, const or_space* ps, double nu, const double* w, double* A,
                     const int* brp, const int* bcol, double* B1, double* B2, double* B3, double* rhs) {
  or_assemble_nse2(vs, ps, 8, 0, nu, w, A, brp, bcol, B1, B2, B3, rhs);
}

/* ------------------------------------------------------ linear algebra */
void or_apply_scaled_add(int n, int active, const int* rp, const int* col, const double* val,
                         const double* x, double* y, double a) {
  for (int i = 0; i < active; ++i) {
    double s = 0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) s += val[k] * x[col[k]];
    y[i] += a * s;
  }
  for (int i = active; i < n; ++i) y[i] += a * x[i];
}

static double ddot(int n, const double* x, const double* y) {
  double s = 0;
  for (int i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}
static double dnorm(int n, const double* x) { return sqrt(ddot(n, x, x)); }

/* ------------------------------------------------------- grid transfer */
/* Prolongate: fine = average over fine cells of QQ_child * coarse_local */
void or_prolongate(const or_space* cs, const or_space* fs, const int* parent, const double* cu,
                   double* fu) {
  const int shape = fs->shape, dim = (shape == OR_HEXA) ? 3 : 2;
  const int ndf = fs->ndof_cell, ndc = cs->ndof_cell;
  double* aux = calloc(fs->n_dof, sizeof(double));
  double* cv = malloc(sizeof(double) * ndc);
  double* val2 = malloc(sizeof(double) * ndf);
  memset(fu, 0, sizeof(double) * fs->n_dof);
  for (int i = 0; i < fs->n_cells; ++i)
    for (int j = 0; j < ndf; ++j) aux[fs->gnum[(size_t)i * ndf + j]] += 1;
  for (int fc = 0; fc < fs->n_cells; ++fc) {
    const int pc = parent[fc];
    const double* vc = cs->verts + (size_t)pc * cs->nv * dim;
    const double* vf = fs->verts + (size_t)fc * fs->nv * dim;
    const int* cdof = cs->gnum + (size_t)pc * ndc;
    const int* fdof = fs->gnum + (size_t)fc * ndf;
    double Val[64];
    for (int l = 0; l < ndc; ++l) Val[l] = cu[cdof[l]];
    for (int k = 0; k < ndf; ++k) {
      /* QQ[k][l] = coarse basis l at fine nodal point k */
      double xi[3], x[3] = {0, 0, 0}, J[9], xc[3];
      node_ref(shape, fs->order, k, xi);
      ref_map(shape, vf, xi, x, J);
      phys_to_ref(shape, vc, x, xc);
      or_basis(shape, cs->order, xc, cv, NULL, NULL);
      double sum = 0;
      for (int l = 0; l < ndc; ++l) sum += cv[l] * Val[l];
      val2[k] = sum;
    }
    for (int k = 0; k < ndf; ++k) fu[fdof[k]] += val2[k];
  }
  for (int i = 0; i < fs->n_dof; ++i) fu[i] /= aux[i];
  free(aux); free(cv); free(val2);
}

/* DefectRestriction: coarse += QQ^T (fine / multiplicity) */
void or_restrict(const or_space* cs, const or_space* fs, const int* parent, double* cu,
                 const double* fd) {
  const int shape = fs->shape, dim = (shape == OR_HEXA) ? 3 : 2;
  const int ndf = fs->ndof_cell, ndc = cs->ndof_cell;
  double* aux = calloc(fs->n_dof, sizeof(double));
  double* cv = malloc(sizeof(double) * ndc);
  memset(cu, 0, sizeof(double) * cs->n_dof);
  for (int i = 0; i < fs->n_cells; ++i)
    for (int j = 0; j < ndf; ++j) aux[fs->gnum[(size_t)i * ndf + j]] += 1;
  for (int fc = 0; fc < fs->n_cells; ++fc) {
    const int pc = parent[fc];
    const double* vc = cs->verts + (size_t)pc * cs->nv * dim;
    const double* vf = fs->verts + (size_t)fc * fs->nv * dim;
    const int* cdof = cs->gnum + (size_t)pc * ndc;
    const int* fdof = fs->gnum + (size_t)fc * ndf;
    double Val2[64];
    for (int l = 0; l < ndc; ++l) Val2[l] = 0;
    for (int k = 0; k < ndf; ++k) {
      double xi[3], x[3] = {0, 0, 0}, J[9], xc[3];
      node_ref(shape, fs->order, k, xi);
      ref_map(shape, vf, xi, x, J);
      phys_to_ref(shape, vc, x, xc);
      or_basis(shape, cs->order, xc, cv, NULL, NULL);
      const double d = fd[fdof[k]] / aux[fdof[k]];
      for (int l = 0; l < ndc; ++l) Val2[l] += cv[l] * d;
    }
    for (int l = 0; l < ndc; ++l) cu[cdof[l]] += Val2[l];
  }
  free(aux); free(cv);
}

/* --------------------------------------------------------------- solvers */
typedef struct {
  const or_solver_params* p;
  double initial, previous;
  double* hist;
  int max_hist, n_hist;
  int diverged;
} conv_state;

static int converged(conv_state* c, double res, int it) {
  const or_solver_params* p = c->p;
  if (it == 0) { c->initial = res; c->previous = res; }
  if (c->n_hist < c->max_hist) c->hist[c->n_hist] = res;
  c->n_hist++;
  const int small = res < p->residual_tolerance;
  const int reduced = res < p->residual_reduction * c->initial;
  const int enough = it > p->min_n_iterations || p->min_n_iterations == 0;
  const int stag = res > 1e10 * c->previous;
  const int div = res > p->divergence_factor * c->initial;
  int conv = (small || reduced || stag || div) && enough;
  if (it == p->max_n_iterations) conv = 1;
  c->previous = res;
  if (div) { c->diverged = 1; conv = 1; }
  return conv;
}

typedef struct {
  int nl;
  const or_level* lv;
  const or_solver_params* p;
  double **rhs, **sol, **def, **dinv, **tmp;
  int* cc;
  int ncc;
  double* lu;   /* coarse LU factors (row-major, unit lower part below the diagonal) */
  int* piv;
} mg_state;

static void fill_rec(int* rec, int level, int nl, int cycle, int* cc, int* n) {
  if (level == 0) { cc[(*n)++] = 2; return; }
  for (int j = 0; j < rec[level]; ++j) {
    cc[(*n)++] = 1;
    fill_rec(rec, level - 1, nl, cycle, cc, n);
  }
  if (cycle == 2) rec[level] = 1;
  cc[(*n)++] = (level == nl - 1) ? 3 : 2;
}

/* LU factorisation with partial pivoting of the dense coarsest matrix (once per
 * solve), then forward/back substitution per application: the exact solve of
 * DirectSmoother (src/Multigrid/DirectSmoother.C:25-42) */
static void coarse_factor(mg_state* m) {
  const or_level* L = &m->lv[0];
  const int n = L->s.n_dof;
  double* a = calloc((size_t)n * n, sizeof(double));
  int* piv = malloc(sizeof(int) * n);
  for (int r = 0; r < n; ++r)
    for (int k = L->s.rp[r]; k < L->s.rp[r + 1]; ++k) a[(size_t)r * n + L->s.col[k]] = L->val[k];
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(a[(size_t)i * n + k]) > fabs(a[(size_t)p * n + k])) p = i;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) { double t = a[(size_t)k * n + j]; a[(size_t)k * n + j] = a[(size_t)p * n + j]; a[(size_t)p * n + j] = t; }
    const double d = a[(size_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const double f = a[(size_t)i * n + k] / d;
      a[(size_t)i * n + k] = f;
      if (f == 0) continue;
      for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= f * a[(size_t)k * n + j];
    }
  }
  m->lu = a;
  m->piv = piv;
}

static void mg_smooth(mg_state* m, int l) {
  const or_level* L = &m->lv[l];
  const int n = L->s.n_dof;
  if (l == 0 && m->p->mg_smoother_coarse == 1) {
    if (!m->lu) coarse_factor(m);
    double* y = m->tmp[0];
    memcpy(y, m->rhs[0], sizeof(double) * n);
    for (int k = 0; k < n; ++k) {
      const int p = m->piv[k];
      if (p != k) { double t = y[k]; y[k] = y[p]; y[p] = t; }
    }
    for (int i = 0; i < n; ++i) {
      double s = y[i];
      for (int j = 0; j < i; ++j) s -= m->lu[(size_t)i * n + j] * y[j];
      y[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = y[i];
      for (int j = i + 1; j < n; ++j) s -= m->lu[(size_t)i * n + j] * m->sol[0][j];
      m->sol[0][i] = s / m->lu[(size_t)i * n + i];
    }
    return;
  }
  /* JacobiSmoother::smooth: defect = rhs - A sol; defect /= diag; sol += damp*defect */
  double* d = m->tmp[l];
  memcpy(d, m->rhs[l], sizeof(double) * n);
  or_apply_scaled_add(n, L->s.active, L->s.rp, L->s.col, L->val, m->sol[l], d, -1.0);
  for (int i = 0; i < n; ++i) d[i] /= m->dinv[l][i];
  for (int i = 0; i < n; ++i) m->sol[l][i] += m->p->jacobi_damp * d[i];
}

static double mg_defect(mg_state* m, int l) {
  const or_level* L = &m->lv[l];
  const int n = L->s.n_dof;
  memcpy(m->def[l], m->rhs[l], sizeof(double) * n);
  or_apply_scaled_add(n, L->s.active, L->s.rp, L->s.col, L->val, m->sol[l], m->def[l], -1.0);
  return sqrt(ddot(n, m->def[l], m->def[l]));
}

static void mg_up(mg_state* m, int l) {
  const or_level* C = &m->lv[l];
  const or_level* F = &m->lv[l + 1];
  const int nf = F->s.n_dof;
  double* copy = malloc(sizeof(double) * nf);
  or_prolongate(&C->s, &F->s, F->parent, m->sol[l], copy);
  for (int j = 0; j < F->s.active; ++j) m->sol[l + 1][j] += m->p->mg_correction_damp * copy[j];
  free(copy);
}

static void mg_down_rhs(mg_state* m, int l) {
  const or_level* C = &m->lv[l - 1];
  const or_level* F = &m->lv[l];
  or_restrict(&C->s, &F->s, F->parent, m->rhs[l - 1], m->def[l]);
  for (int j = C->s.active; j < C->s.n_dof; ++j) m->rhs[l - 1][j] = 0.0;
}

static int mg_step(mg_state* m, int step, int level) {
  if (level == 0) {
    double res = 1e10;
    for (int i = 0; i < m->p->mg_coarse_max_n_iterations && res > m->p->mg_coarse_residual; ++i) {
      mg_smooth(m, 0);
      res = mg_defect(m, 0);
    }
    mg_up(m, 0);
    return 1;
  }
  const int from_below = m->cc[step - 1] == 2, up = m->cc[step] == 2, down = m->cc[step] == 1;
  if (from_below)
    for (int i = 0; i < m->p->mg_n_post_smooth; ++i) mg_smooth(m, level);
  if (down) {
    for (int i = 0; i < m->p->mg_n_pre_smooth; ++i) mg_smooth(m, level);
    mg_defect(m, level);
    memset(m->sol[level - 1], 0, sizeof(double) * m->lv[level - 1].s.n_dof);
    mg_down_rhs(m, level);
    return level - 1;
  } else if (up) {
    mg_up(m, level);
    return level + 1;
  }
  return -1;
}

static void mg_apply(mg_state* m, const double* v, double* z) {
  const int L = m->nl - 1;
  const or_level* F = &m->lv[L];
  const int n = F->s.n_dof;
  memcpy(m->rhs[L], v, sizeof(double) * n);
  memset(m->sol[L], 0, sizeof(double) * n);
  for (int i = F->s.active; i < n; ++i) m->sol[L][i] = m->rhs[L][i];
  int level = L;
  for (int step = 1; step < m->ncc; ++step) level = mg_step(m, step, level);
  memcpy(z, m->sol[L], sizeof(double) * n);
}

static void precond(mg_state* m, const double* dinv, const double* v, double* z, int n) {
  if (m->p->preconditioner == 2) mg_apply(m, v, z);
  else if (m->p->preconditioner == 1) for (int i = 0; i < n; ++i) z[i] = v[i] / dinv[i];
  else memcpy(z, v, sizeof(double) * n);
}

static void gen_rot(double dx, double dy, double* cs, double* sn) {
  if (dy == 0.0) { *cs = 1.0; *sn = 0.0; }
  else if (fabs(dy) > fabs(dx)) { double t = dx / dy; *sn = 1.0 / sqrt(1.0 + t * t); *cs = t * *sn; }
  else { double t = dy / dx; *cs = 1.0 / sqrt(1.0 + t * t); *sn = t * *cs; }
}
static void app_rot(double* dx, double* dy, double cs, double sn) {
  double t = cs * *dx + sn * *dy;
  *dy = -sn * *dx + cs * *dy;
  *dx = t;
}

int or_solve(const or_solver_params* p, int nl, const or_level* lv, const double* b, double* x,
             int* iterations, double* hist, int max_hist, int* n_hist) {
  const or_level* F = &lv[nl - 1];
  const int n = F->s.n_dof, active = F->s.active;
  const int* rp = F->s.rp;
  const int* col = F->s.col;
  const double* val = F->val;
  conv_state cst = {p, 0, 0, hist, max_hist, 0, 0};
  mg_state m;
  memset(&m, 0, sizeof(m));
  m.nl = nl; m.lv = lv; m.p = p;
  m.rhs = calloc(nl, sizeof(double*)); m.sol = calloc(nl, sizeof(double*));
  m.def = calloc(nl, sizeof(double*)); m.dinv = calloc(nl, sizeof(double*)); m.tmp = calloc(nl, sizeof(double*));
  for (int l = 0; l < nl; ++l) {
    const int nn = lv[l].s.n_dof;
    m.rhs[l] = calloc(nn, sizeof(double)); m.sol[l] = calloc(nn, sizeof(double));
    m.def[l] = calloc(nn, sizeof(double)); m.dinv[l] = calloc(nn, sizeof(double));
    m.tmp[l] = calloc(nn, sizeof(double));
    for (int r = 0; r < nn; ++r)
      for (int k = lv[l].s.rp[r]; k < lv[l].s.rp[r + 1]; ++k)
        if (lv[l].s.col[k] == r) m.dinv[l][r] = lv[l].val[k];
  }
  int cc[4096], ncc = 0;
  int rec[16];
  for (int l = 0; l < nl; ++l) rec[l] = (p->mg_cycle == 0) ? 1 : 2;
  rec[nl - 1] = 1;
  cc[ncc++] = 0;
  if (p->preconditioner == 2) fill_rec(rec, nl - 1, nl, p->mg_cycle, cc, &ncc);
  m.cc = cc; m.ncc = ncc;
  double* dinv = m.dinv[nl - 1];
  double* r = calloc(n, sizeof(double));
  int it = 0, status = 0;

  if (p->iterative_solver_type == 0) {          /* flexible GMRES */
    int mr = p->gmres_restart < p->max_n_iterations ? p->gmres_restart : p->max_n_iterations;
    if (mr < 1) mr = 1;
    double** V = malloc(sizeof(double*) * (mr + 1));
    double** Z = malloc(sizeof(double*) * mr);
    for (int k = 0; k <= mr; ++k) V[k] = calloc(n, sizeof(double));
    for (int k = 0; k < mr; ++k) Z[k] = calloc(n, sizeof(double));
    double* H = calloc((size_t)(mr + 1) * mr, sizeof(double));
    double* s = calloc(mr + 1, sizeof(double));
    double* cs = calloc(mr + 1, sizeof(double));
    double* sn = calloc(mr + 1, sizeof(double));
    double* y = calloc(mr + 1, sizeof(double));
#define HH(i, j) H[(i) + (size_t)(j) * (mr + 1)]
    memcpy(r, b, sizeof(double) * n);
    or_apply_scaled_add(n, active, rp, col, val, x, r, -1.0);
    double beta = dnorm(n, r), resid = beta;
    if (converged(&cst, resid, 0)) goto gm_done;
    it = 1;
    while (it <= p->max_n_iterations) {
      for (int i = 0; i < n; ++i) V[0][i] = r[i] * (1.0 / beta);
      memset(s, 0, sizeof(double) * (mr + 1));
      s[0] = beta;
      int i;
      for (i = 0; i < mr && it <= p->max_n_iterations; i++, it++) {
        memset(Z[i], 0, sizeof(double) * n);
        precond(&m, dinv, V[i], Z[i], n);
        memset(r, 0, sizeof(double) * n);
        or_apply_scaled_add(n, active, rp, col, val, Z[i], r, 1.0);
        for (int k = 0; k <= i; ++k) {
          HH(k, i) = ddot(n, r, V[k]);
          for (int q = 0; q < n; ++q) r[q] += -HH(k, i) * V[k][q];
        }
        HH(i + 1, i) = dnorm(n, r);
        if (HH(i + 1, i) == 0) { status = 5; goto gm_done; }
        for (int q = 0; q < n; ++q) V[i + 1][q] = r[q] * (1.0 / HH(i + 1, i));
        for (int k = 0; k < i; ++k) app_rot(&HH(k, i), &HH(k + 1, i), cs[k], sn[k]);
        gen_rot(HH(i, i), HH(i + 1, i), &cs[i], &sn[i]);
        app_rot(&HH(i, i), &HH(i + 1, i), cs[i], sn[i]);
        app_rot(&s[i], &s[i + 1], cs[i], sn[i]);
        resid = fabs(s[i + 1]);
        if (converged(&cst, resid, it)) {
          memcpy(y, s, sizeof(double) * (i + 1));
          for (int a = i; a >= 0; a--) { y[a] /= HH(a, a); for (int c = a - 1; c >= 0; c--) y[c] -= HH(c, a) * y[a]; }
          for (int c = 0; c <= i; c++) for (int q = 0; q < n; ++q) x[q] += y[c] * Z[c][q];
          goto gm_done;
        }
      }
      memcpy(y, s, sizeof(double) * mr);
      for (int a = mr - 1; a >= 0; a--) { y[a] /= HH(a, a); for (int c = a - 1; c >= 0; c--) y[c] -= HH(c, a) * y[a]; }
      for (int c = 0; c < mr; c++) for (int q = 0; q < n; ++q) x[q] += y[c] * Z[c][q];
      memcpy(r, b, sizeof(double) * n);
      or_apply_scaled_add(n, active, rp, col, val, x, r, -1.0);
      beta = dnorm(n, r);
      if (converged(&cst, resid, it)) goto gm_done;
    }
  gm_done:
#undef HH
    for (int k = 0; k <= mr; ++k) free(V[k]);
    for (int k = 0; k < mr; ++k) free(Z[k]);
    free(V); free(Z); free(H); free(s); free(cs); free(sn); free(y);
  } else if (p->iterative_solver_type == 1) {   /* preconditioned CG */
    double* z = calloc(n, sizeof(double));
    double* pp = calloc(n, sizeof(double));
    double* q = calloc(n, sizeof(double));
    double rho_1 = 0;
    memcpy(r, b, sizeof(double) * n);
    or_apply_scaled_add(n, active, rp, col, val, x, r, -1.0);
    double normb = dnorm(n, b);
    if (normb == 0.0) normb = 1;
    double resid = dnorm(n, r) / normb;
    if (!converged(&cst, resid, 0)) {
      for (it = 1; it <= p->max_n_iterations; ++it) {
        precond(&m, dinv, r, z, n);
        const double rho = ddot(n, r, z);
        if (it == 1) memcpy(pp, z, sizeof(double) * n);
        else {
          const double beta = rho / rho_1;
          for (int i = 0; i < n; ++i) pp[i] *= beta;
          for (int i = 0; i < n; ++i) pp[i] += z[i];
        }
        memset(q, 0, sizeof(double) * n);
        or_apply_scaled_add(n, active, rp, col, val, pp, q, 1.0);
        const double alpha = rho / ddot(n, pp, q);
        for (int i = 0; i < n; ++i) x[i] += alpha * pp[i];
        for (int i = 0; i < n; ++i) r[i] += -alpha * q[i];
        resid = dnorm(n, r) / normb;
        if (converged(&cst, resid, it)) break;
        rho_1 = rho;
      }
    } else it = 0;
    free(z); free(pp); free(q);
  } else {                                      /* Richardson / Jacobi */
    double* z = calloc(n, sizeof(double));
    double normb = dnorm(n, b);
    if (normb == 0.0) normb = 1;
    memcpy(r, b, sizeof(double) * n);
    or_apply_scaled_add(n, active, rp, col, val, x, r, -1.0);
    double resid = dnorm(n, r) / normb;
    if (!converged(&cst, resid, 0)) {
      for (it = 1; it <= p->max_n_iterations; ++it) {
        if (p->iterative_solver_type == 3) {
          /* Iteration_jacobi::iterate: r *= damping / diag; solution += r */
          for (int i = 0; i < n; ++i) x[i] += r[i] * (p->damping / dinv[i]);
        } else {
          precond(&m, dinv, r, z, n);
          for (int i = 0; i < n; ++i) x[i] += p->damping * z[i];
        }
        memcpy(r, b, sizeof(double) * n);
        or_apply_scaled_add(n, active, rp, col, val, x, r, -1.0);
        resid = dnorm(n, r) / normb;
        if (converged(&cst, resid, it)) break;
      }
    } else it = 0;
    free(z);
  }
  if (it > p->max_n_iterations) it = p->max_n_iterations;
  if (iterations) *iterations = it;
  if (n_hist) *n_hist = cst.n_hist;
  if (cst.diverged) status = 4;
  free(r);
  free(m.lu); free(m.piv);
  for (int l = 0; l < nl; ++l) { free(m.rhs[l]); free(m.sol[l]); free(m.def[l]); free(m.dinv[l]); free(m.tmp[l]); }
  free(m.rhs); free(m.sol); free(m.def); free(m.dinv); free(m.tmp);
  return status;
}

/* one multigrid cycle as preconditioner (Iteration_multigrid::apply) */
void or_mg_apply(const or_solver_params* p, int nl, const or_level* lv, const double* v, double* z) {
  mg_state m;
  memset(&m, 0, sizeof(m));
  m.nl = nl; m.lv = lv; m.p = p;
  m.rhs = calloc(nl, sizeof(double*)); m.sol = calloc(nl, sizeof(double*));
  m.def = calloc(nl, sizeof(double*)); m.dinv = calloc(nl, sizeof(double*)); m.tmp = calloc(nl, sizeof(double*));
  for (int l = 0; l < nl; ++l) {
    const int nn = lv[l].s.n_dof;
    m.rhs[l] = calloc(nn, sizeof(double)); m.sol[l] = calloc(nn, sizeof(double));
    m.def[l] = calloc(nn, sizeof(double)); m.dinv[l] = calloc(nn, sizeof(double));
    m.tmp[l] = calloc(nn, sizeof(double));
    for (int r = 0; r < nn; ++r)
      for (int k = lv[l].s.rp[r]; k < lv[l].s.rp[r + 1]; ++k)
        if (lv[l].s.col[k] == r) m.dinv[l][r] = lv[l].val[k];
  }
  int cc[4096], ncc = 0, rec[16];
  for (int l = 0; l < nl; ++l) rec[l] = (p->mg_cycle == 0) ? 1 : 2;
  rec[nl - 1] = 1;
  cc[ncc++] = 0;
  fill_rec(rec, nl - 1, nl, p->mg_cycle, cc, &ncc);
  m.cc = cc; m.ncc = ncc;
  mg_apply(&m, v, z);
  free(m.lu); free(m.piv);
  for (int l = 0; l < nl; ++l) { free(m.rhs[l]); free(m.sol[l]); free(m.def[l]); free(m.dinv[l]); free(m.tmp[l]); }
  free(m.rhs); free(m.sol); free(m.def); free(m.dinv); free(m.tmp);
}
