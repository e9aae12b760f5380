// Built-in convection-diffusion examples (exact solution u, right-hand side
// f = -eps*lap(u) + b.grad(u) + c*u and boundary values), usable on host and
// device.  They restate the examples the reference tests use:
//   EX_SINE_LAPLACE_2D  include/Examples/CD_2D/SineLaplace.h (Dirichlet value 0,
//                       Neumann value eps*du/dx on component 1)
//   EX_P0_3D / EX_P1_3D / EX_P2_3D  include/Examples/CD_3D/test_p0.h, test_p1.h,
//                       test_p2.h (solutions 1, 1+2x+3y+4z, x^2+yz)
//   EX_EXP_SIN_3D       u = exp(-x) sin(pi y) sin(pi z) (CD3D benchmark problem)
//   EX_SINE_2D          u = sin(pi x) sin(pi y) with Dirichlet value u everywhere
//   EX_TCD_EXP          include/Examples/TCD_2D/exp.h: u = e^t sin(2 pi x) sin(2 pi y),
//                       Dirichlet value 0 (the reference's example 0 of Example_TimeCD2D)
//   EX_TCD_LINEAR       include/Examples/TCD_2D/linear_space_time.h: u = 1 + 2x + 3ty
//                       (example -1)
// Time-dependent examples take the right-hand side f = du/dt - eps*lap(u) +
// b.grad(u) + c*u at the given time.
//   EX_NSE_U2P1_*       include/Examples/NSE_3D/test_u_2_p_1.h (NSE3D example -3):
//                       u = (y^2, xz, -y^2), p = x + y + z - 1.5; one id per
//                       component (U1, U2, U3, P) for values and errors
//   EX_NSE_U0P0_*       include/Examples/NSE_3D/test_u_0_p_0.h (NSE3D example -1):
//                       u = (1, 0, 0), p = 0; homogeneous Neumann data
//                       (nu grad u - p I) n = 0 on the interior of the face x = 1
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace pmx {

enum Example : int {
  EX_CONST = 0,
  EX_SINE_LAPLACE_2D = 1,
  EX_P0_3D = 2,
  EX_P1_3D = 3,
  EX_P2_3D = 4,
  EX_EXP_SIN_3D = 5,
  EX_SINE_2D = 6,
  EX_TCD_EXP = 7,
  EX_TCD_LINEAR = 8,
  EX_NSE_U2P1_U1 = 9,
  EX_NSE_U2P1_U2 = 10,
  EX_NSE_U2P1_U3 = 11,
  EX_NSE_U2P1_P = 12,
  EX_NSE_U0P0_U1 = 13,
  EX_NSE_U0P0_U2 = 14,
  EX_NSE_U0P0_U3 = 15,
  EX_NSE_U0P0_P = 16,
};

constexpr double kPi = 3.14159265358979323846;

// v[0] = u, v[1..3] = grad u, v[4] = laplace u, v[5] = du/dt at time t
__host__ __device__ inline void example_exact(int ex, double x, double y, double z, double* v,
                                              double t = 0.0) {
  v[0] = v[1] = v[2] = v[3] = v[4] = v[5] = 0.0;
  switch (ex) {
    case EX_TCD_EXP: {
      const double e = exp(t), sx = sin(2 * kPi * x), sy = sin(2 * kPi * y);
      v[0] = e * (sx * sy);
      v[1] = e * 2 * kPi * cos(2 * kPi * x) * sy;
      v[2] = e * 2 * kPi * sx * cos(2 * kPi * y);
      v[4] = -8 * kPi * kPi * v[0];
      v[5] = v[0];
      break;
    }
    case EX_TCD_LINEAR:
      v[0] = 1 + 2 * x + 3 * t * y; v[1] = 2; v[2] = 3 * t; v[5] = 3 * y;
      break;
    case EX_NSE_U2P1_U1: v[0] = y * y; v[2] = 2 * y; v[4] = 2; break;
    case EX_NSE_U2P1_U2: v[0] = x * z; v[1] = z; v[3] = x; break;
    case EX_NSE_U2P1_U3: v[0] = -y * y; v[2] = -2 * y; v[4] = -2; break;
    case EX_NSE_U2P1_P: v[0] = x + y + z - 1.5; v[1] = 1; v[2] = 1; v[3] = 1; break;
    case EX_NSE_U0P0_U1: v[0] = 1; break;
    case EX_NSE_U0P0_U2:
    case EX_NSE_U0P0_U3:
    case EX_NSE_U0P0_P: break;
    case EX_SINE_LAPLACE_2D:
    case EX_SINE_2D: {
      const double sx = sin(kPi * x), sy = sin(kPi * y);
      v[0] = sx * sy;
      v[1] = kPi * cos(kPi * x) * sy;
      v[2] = kPi * sx * cos(kPi * y);
      v[4] = -2 * kPi * kPi * sx * sy;
      break;
    }
    case EX_P0_3D: v[0] = 1; break;
    case EX_P1_3D: v[0] = 1 + 2 * x + 3 * y + 4 * z; v[1] = 2; v[2] = 3; v[3] = 4; break;
    case EX_P2_3D: v[0] = x * x + y * z; v[1] = 2 * x; v[2] = z; v[3] = y; v[4] = 2; break;
    case EX_EXP_SIN_3D: {
      const double e = exp(-x), sy = sin(kPi * y), sz = sin(kPi * z);
      v[0] = e * sy * sz;
      v[1] = -v[0];
      v[2] = kPi * e * cos(kPi * y) * sz;
      v[3] = kPi * e * sy * cos(kPi * z);
      v[4] = (1.0 - 2.0 * kPi * kPi) * v[0];
      break;
    }
    default: break;
  }
}

__host__ __device__ inline double example_rhs(int ex, double x, double y, double z, double eps,
                                              const double* b, double c, double f0, double t = 0.0) {
  if (ex == EX_CONST) return f0;
  double v[6];
  example_exact(ex, x, y, z, v, t);
  double f = v[5] - eps * v[4];
  f += b[0] * v[1] + b[1] * v[2] + b[2] * v[3];
  f += c * v[0];
  return f;
}

// Dirichlet boundary value at a boundary point
__host__ __device__ inline double example_dirichlet(int ex, double x, double y, double z,
                                                    double t = 0.0) {
  if (ex == EX_SINE_LAPLACE_2D || ex == EX_CONST || ex == EX_TCD_EXP) return 0.0;
  double v[6];
  example_exact(ex, x, y, z, v, t);
  return v[0];
}

// Navier-Stokes examples: f = -nu lap(u) + (u.grad) u + grad p, g = div u = 0
// (LinCoeffs of test_u_2_p_1.h); out = f1, f2, f3, g.  `ex` is the first
// component id (EX_NSE_U2P1_U1).
__host__ __device__ inline void nse_example_rhs(int ex, double x, double y, double z, double nu,
                                                double* out) {
  double u[3][6], p[6];
  for (int k = 0; k < 3; ++k) example_exact(ex + k, x, y, z, u[k]);
  example_exact(ex + 3, x, y, z, p);
  for (int k = 0; k < 3; ++k)
    out[k] = -nu * u[k][4] + (u[0][0] * u[k][1] + u[1][0] * u[k][2] + u[2][0] * u[k][3]) + p[1 + k];
  out[3] = 0.0;
}

// SUPG parameter of Compute_SDFEM_delta with SDFEM_TYPE 2
// (src/AssembleRoutines/ConvDiff.C:273-345): 1D Green's-function formula with
// nu = 1 / INTERNAL_POLYNOMIAL_DEGREE and h_K = cell diameter.
__host__ __device__ inline double sdfem_delta(double hK, double eps, const double* b, int dim,
                                              double c, double delta0, double poly_degree) {
  (void)c;
  if (fabs(eps) < 1e-20) eps = 1e-20;
  double nb = b[0] * b[0] + b[1] * b[1];
  if (dim == 3) nb += b[2] * b[2];
  nb = sqrt(nb);
  const double nu = 1.0 / poly_degree;
  double delta = 0.0;
  if (nb > 0) {
    const double alpha = nu * nb * hK / (2 * eps);
    delta = nu * hK * (1 / tanh(alpha) - 1 / alpha) / (2 * nb);
  }
  return delta * delta0;
}

}  // namespace pmx
