/* ORACLE — test infrastructure only (see cd_oracle.c). */
#ifndef CD_ORACLE_H
#define CD_ORACLE_H

#define OR_TRI 0
#define OR_QUAD 1
#define OR_HEXA 2
#define OR_GALERKIN 0
#define OR_SUPG 1
#define OR_GLS 2
#define OR_DIRICHLET 0
#define OR_NEUMANN 1
#define OR_EX_CONST 0
#define OR_EX_SINE_LAPLACE_2D 1
#define OR_EX_P0_3D 2
#define OR_EX_P1_3D 3
#define OR_EX_P2_3D 4
#define OR_EX_EXP_SIN_3D 5
#define OR_EX_SINE_2D 6
#define OR_EX_TCD_EXP 7      /* TCD_2D/exp.h (reference example 0) */
#define OR_EX_TCD_LINEAR 8   /* TCD_2D/linear_space_time.h (reference example -1) */

typedef struct {
  int shape, order, nv, n_cells, ndof_cell, n_dof, active;
  const double* verts;   /* [n_cells][nv][dim] */
  const int* gnum;       /* [n_cells][ndof_cell] */
  const int* rp;         /* CSR pattern */
  const int* col;
  const int* bd_comp;    /* [n_cells][n_joints] boundary component or -1 */
} or_space;

typedef struct {
  int disc, example, hexa_rule, bc_sine_laplace;
  double eps, b[3], c, f0, delta0, poly_degree;
  double time;     /* TimeDB->CURRENTTIME (time-dependent examples) */
  double theta1;   /* TimeDB->THETA1 (TCD SUPG) */
  double tau;      /* TimeDB->CURRENTTIMESTEPLENGTH (TCD SUPG) */
  /* optional [n_cells][n_quad][8] per-point coefficients (eps, b1, b2, b3, c,
     f, -, -) in the order of or_quad_points: the values a CoeffFct would
     return there (LocalAssembling::GetParameters, LocalAssembling.C:766-840);
     NULL: constants + example */
  const double* table;
} or_cd;

typedef struct {
  int iterative_solver_type;   /* 0 fgmres, 1 cg, 2 richardson, 3 jacobi */
  int preconditioner;          /* 0 none, 1 jacobi, 2 multigrid */
  int max_n_iterations, min_n_iterations;
  double residual_tolerance, residual_reduction, divergence_factor;
  int gmres_restart;
  double damping;
  int mg_cycle;                /* 0 V, 1 W, 2 F */
  int mg_smoother_coarse;      /* 0 jacobi, 1 direct */
  double mg_correction_damp;
  int mg_n_pre_smooth, mg_n_post_smooth;
  double mg_coarse_residual;
  int mg_coarse_max_n_iterations;
  double jacobi_damp;
} or_solver_params;

typedef struct {
  or_space s;
  const double* val;     /* matrix values on the pattern */
  const int* parent;     /* fine cell -> coarse cell (NULL on the coarsest level) */
} or_level;

int or_ndof(int shape, int order);
void or_basis(int shape, int order, const double* xi, double* v, double* g, double* h);
int or_quad_rule(int shape, int order, int hexa_rule, double* xi, double* w);
void or_assemble_cd(const or_space* s, const or_cd* p, double* val, double* rhs);
/* physical quadrature points [n_cells][n_quad][3] of or_assemble_cd's loop */
int or_quad_points(const or_space* s, int hexa_rule, double* xyz);
void or_errors(const or_space* s, const or_cd* p, const double* sol, double* err4);
/* TCD forms; aval, mval, rhs may each be NULL */
void or_assemble_tcd(const or_space* s, const or_cd* p, double* aval, double* mval, double* rhs);
/* y[i] += a * (A x)[i] for active rows only (FEMatrix::multiplyActive) */
void or_multiply_active(int active, const int* rp, const int* col, const double* val,
                        const double* x, double* y, double a);
/* interpolation of the example's exact solution at time p->time at the nodal points */
void or_interpolate(const or_space* s, const or_cd* p, double* u);
/* Navier-Stokes NSTYPE 1 on hexahedra (velocity Q2 space vs, pressure Q1
 * space ps on the same cells): A (pattern of vs), B1..B3 (pattern brp/bcol:
 * pressure rows, velocity columns), rhs [3 n_u + n_p]; w = convection field
 * [3][n_u] or NULL; example test_u_2_p_1 with viscosity nu */
void or_assemble_nse(const or_space* vs, const or_space* ps, double nu, const double* w, double* A,
                     const int* brp, const int* bcol, double* B1, double* B2, double* B3, double* rhs);
/* the same for np_cell = 8 (Q1) or 4 (discontinuous P1, basis 1, xi, eta,
 * zeta, ps->gnum [n_cells][4]) pressure and example 0 (test_u_2_p_1) or 1
 * (test_u_0_p_0: u = (1,0,0), p = 0, f = 0) */
void or_assemble_nse2(const or_space* vs, const or_space* ps, int np_cell, int example, double nu,
                      const double* w, double* A, const int* brp, const int* bcol, double* B1, double* B2,
                      double* B3, double* rhs);
void or_apply_scaled_add(int n, int active, const int* rp, const int* col, const double* val,
                         const double* x, double* y, double a);
void or_prolongate(const or_space* cs, const or_space* fs, const int* parent, const double* cu, double* fu);
void or_restrict(const or_space* cs, const or_space* fs, const int* parent, double* cu, const double* fd);
int or_solve(const or_solver_params* p, int nl, const or_level* lv, const double* b, double* x,
             int* iterations, double* hist, int max_hist, int* n_hist);
void or_mg_apply(const or_solver_params* p, int nl, const or_level* lv, const double* v, double* z);

#endif
