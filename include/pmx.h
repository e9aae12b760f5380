/*
 * pmx — C ABI of the MI355X-native solve path for ParMooN-style
 * convection-diffusion problems (assembly into the reference CSR pattern,
 * FE-matrix operator, Krylov / geometric-multigrid solvers).
 *
 * Conventions
 *  - every function returns PMX_OK (0) or a PMX_E_* code; pmx_last_error()
 *    gives the message of the last failure of the calling thread;
 *  - "dev" pointers are device (HBM) addresses, "host" pointers host memory;
 *    sizes are element counts;
 *  - all device work of a context runs on the context's HIP stream and is
 *    asynchronous unless stated otherwise; calls that return host results
 *    synchronise that stream;
 *  - objects are not thread-safe; use one pmx_ctx per host thread.
 *
 * Each entry point names the reference interface (nahmedwias/MyMooNMD) it
 * replaces.
 */
#ifndef PMX_H
#define PMX_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ errors */
#define PMX_OK 0
#define PMX_E_ARG 1         /* invalid argument */
#define PMX_E_HIP 2         /* HIP runtime error */
#define PMX_E_ZERO_DIAG 3   /* zero diagonal entry (Iteration_jacobi.C:8-15) */
#define PMX_E_DIVERGED 4    /* IterativeMethod::converged divergence (IterativeMethod.h:105-131) */
#define PMX_E_BREAKDOWN 5   /* FGMRES unhappy breakdown (Iteration_gmres.C:470-477) */
#define PMX_E_STATE 6       /* object used in a wrong state */
#define PMX_E_NOMEM 7       /* allocation failed */
#define PMX_E_INTERNAL 8

/* Message of the last error on the calling thread (never NULL). */
const char* pmx_last_error(void);
/* ABI version (major*100 + minor). */
int pmx_version(void);

/* ----------------------------------------------------------------- context */
typedef struct pmx_ctx pmx_ctx;
/* Binds HIP device `device` and creates the context's stream, a
 * non-blocking one: every call on the context runs there, ordered among
 * themselves but not with the caller's own HIP work.  Device buffers handed
 * in must be complete (synchronous copies, or hipDeviceSynchronize after
 * asynchronous ones); copies between buffers the library uses go through
 * pmx_vector_copy; results are complete when a call that returns host values
 * returns, or after pmx_ctx_synchronize. */
int pmx_ctx_create(int device, pmx_ctx** out);
/* Drops the caller's reference: every object created on the context holds
 * one too, so objects may be destroyed before or after the context (in any
 * order); the context's stream and buffers are freed with the last reference.
 * The handle must not be used for new calls afterwards. */
int pmx_ctx_destroy(pmx_ctx* ctx);
int pmx_ctx_synchronize(pmx_ctx* ctx);

/* ------------------------------------------------------------ domain/mesh */
/* Replaces TDomain(db) + TDomain::Init(PRM, GEO) + refine_and_get_hierarchy_
 * of_collections (src/Geometry/Domain.C:270-461, 3065-3173) for the default
 * geometries; level k holds the collection after k regular refinements in
 * It_Finest order. */
#define PMX_GEO_UNIT_SQUARE_QUAD 0   /* "UnitSquare" */
#define PMX_GEO_UNIT_SQUARE_TRI 1    /* "TwoTriangles" */
#define PMX_GEO_UNIT_CUBE_HEXA 2     /* "Default_UnitCube_Hexa" */
typedef struct pmx_domain pmx_domain;
int pmx_domain_create(int geometry, int n_refine, pmx_domain** out);
/* Single-level n x n triangulation of the unit square (the mesh of the
 * batched solves for n != 2^k; no reference counterpart): squares row by row,
 * each cut along its (0,0)-(1,1) diagonal. */
int pmx_domain_create_structured_tri(int n_side, pmx_domain** out);
int pmx_domain_destroy(pmx_domain* d);
/* number of cells of a level; vertex coordinates [n_cells][nv][dim] */
int pmx_domain_level_info(const pmx_domain* d, int level, int* n_cells, int* n_vertices_per_cell,
                          int* dim);
int pmx_domain_get_vertices(const pmx_domain* d, int level, double* host_xyz);

/* ------------------------------------------------------------------ spaces */
/* Boundary-condition callbacks, replacing BoundCondFunct2D / BoundCondFunct3D
 * (include/General/Constants.h); return 0 DIRICHLET, 1 NEUMANN, ... */
typedef int (*pmx_bc2d_fn)(int component, double t, void* user);
typedef int (*pmx_bc3d_fn)(double x, double y, double z, void* user);
/* Built-in boundary conditions (bc callbacks may then be NULL) */
#define PMX_BC_ALL_DIRICHLET 0
#define PMX_BC_SINE_LAPLACE 1        /* component 1 Neumann, others Dirichlet */

typedef struct pmx_space pmx_space;
/* Replaces TFESpace2D/TFESpace3D(collection, ..., BoundaryCondition, order)
 * + ConstructSpace (src/FE/FESpace3D.C:58-229, 506-2080).  order 1 or 2
 * (Q1/Q2 on quads and hexahedra, P1 on triangles).  When bc2d/bc3d are NULL
 * the built-in `builtin_bc` is used.  ctx may be NULL for a host-only space
 * (numbering and pattern queries, no assembly). */
int pmx_space_create(pmx_ctx* ctx, const pmx_domain* d, int level, int order, pmx_bc2d_fn bc2d,
                     pmx_bc3d_fn bc3d, void* user, int builtin_bc, pmx_space** out);
int pmx_space_destroy(pmx_space* s);
typedef struct {
  int n_cells, ndof_cell, n_dof, active_bound, n_inner, n_dirichlet;
  int n_bound[9];   /* NEUMANN, ROBIN, SLIP, ... DIRICHLET_WEAK */
  int dim;
} pmx_space_info;
int pmx_space_get_info(const pmx_space* s, pmx_space_info* out);
/* GlobalNumbers [n_cells][ndof_cell] (TFESpace::GetGlobalNumbers) */
int pmx_space_get_global_numbers(const pmx_space* s, int* host_out);
/* nodal points of all DOFs [n_dof][3] */
int pmx_space_get_dof_coords(const pmx_space* s, double* host_out);
/* TStructure(FESpace) pattern on the host: *nnz always set; row_ptr [n_dof+1]
 * and col [nnz] filled when not NULL.  Works without a device (ctx NULL). */
int pmx_space_get_structure(const pmx_space* s, int* nnz, int* host_row_ptr, int* host_col);
/* Grid-transfer matrix P (fine x coarse) between consecutive levels:
 * prolongation = P*u (GridTransfer::Prolongate, src/Multigrid/GridTransfer.C:665-941),
 * defect restriction = P^T*d (GridTransfer::DefectRestriction, :943-1230). */
int pmx_space_get_prolongation(const pmx_space* coarse, const pmx_space* fine, int* nnz,
                               int* host_row_ptr, int* host_col, double* host_val);

/* ---------------------------------------------------------------- matrices */
/* Square FE matrix on device: TStructure(FESpace) pattern
 * (src/Matrix/Structure.C:829-1277) + FEMatrix values. */
typedef struct pmx_matrix pmx_matrix;
int pmx_matrix_create(pmx_ctx* ctx, const pmx_space* s, pmx_matrix** out);
/* Square matrix from a host CSR array set (BlockMatrix::get_combined_matrix
 * style use, src/Matrix/BlockMatrix.C); rows >= active_bound act as identity
 * rows in the operator. */
int pmx_matrix_create_csr(pmx_ctx* ctx, int n, int nnz, const int* host_row_ptr, const int* host_col,
                          const double* host_val, int active_bound, pmx_matrix** out);
int pmx_matrix_destroy(pmx_matrix* A);
int pmx_matrix_get_info(const pmx_matrix* A, int* n_rows, int* nnz, int* active_bound);
int pmx_matrix_get_structure(const pmx_matrix* A, int* host_row_ptr, int* host_col);
int pmx_matrix_get_entries(pmx_ctx* ctx, const pmx_matrix* A, double* host_values);
int pmx_matrix_set_entries(pmx_ctx* ctx, pmx_matrix* A, const double* host_values);
/* device address of the value array (nnz doubles) */
int pmx_matrix_entries_device(pmx_matrix* A, double** dev_values);
/* TMatrix::reset: all entries to zero */
int pmx_matrix_reset(pmx_ctx* ctx, pmx_matrix* A);
/* Lanes per row of the SpMV kernels (4, 8, 16, 32, 64); 0 = automatic */
int pmx_matrix_set_spmv_lanes(pmx_matrix* A, int lanes);

/* BlockFEMatrix::apply_scaled_add (src/Matrix/BlockFEMatrix.C:863-924):
 * y += a*A*x on active rows, y += a*x on the non-active (Dirichlet) rows. */
int pmx_matrix_apply_scaled_add(pmx_ctx* ctx, const pmx_matrix* A, const double* dev_x,
                                double* dev_y, double a);
/* BlockFEMatrix::apply (src/Matrix/BlockFEMatrix.C:855-860): y = A*x with
 * identity on non-active rows. */
int pmx_matrix_apply(pmx_ctx* ctx, const pmx_matrix* A, const double* dev_x, double* dev_y);

/* Times `reps` back-to-back launches of one operator kernel on the context's
 * stream with HIP events; *avg_ms = mean duration per launch.
 * kernel: PMX_KERNEL_APPLY_FP64 (y = A x), PMX_KERNEL_JACOBI_FP64 (fused damped
 * Jacobi step), PMX_KERNEL_APPLY_BF16 / PMX_KERNEL_JACOBI_BF16 (bf16 values with
 * FP32 row scale and FP32 vectors; the bf16 copy is made internally).
 * Input vectors are internal buffers filled with ones. */
#define PMX_KERNEL_APPLY_FP64 0
#define PMX_KERNEL_JACOBI_FP64 1
#define PMX_KERNEL_APPLY_BF16 2
#define PMX_KERNEL_JACOBI_BF16 3
int pmx_matrix_time_kernel(pmx_ctx* ctx, const pmx_matrix* A, int kernel, int reps, double* avg_ms);

/* ---------------------------------------------------------------- assembly */
#define PMX_DISC_GALERKIN 0
#define PMX_DISC_SUPG 1
#define PMX_DISC_GLS 2   /* BilinearAssemble_GLS (ConvDiff.C:2012-2131), stationary CD only */
/* built-in examples (coefficients + exact solution + boundary values) */
#define PMX_EX_CONST 0
#define PMX_EX_SINE_LAPLACE_2D 1   /* include/Examples/CD_2D/SineLaplace.h */
#define PMX_EX_P0_3D 2             /* include/Examples/CD_3D/test_p0.h */
#define PMX_EX_P1_3D 3             /* include/Examples/CD_3D/test_p1.h */
#define PMX_EX_P2_3D 4             /* include/Examples/CD_3D/test_p2.h */
#define PMX_EX_EXP_SIN_3D 5        /* u = exp(-x) sin(pi y) sin(pi z) */
#define PMX_EX_SINE_2D 6           /* u = sin(pi x) sin(pi y), Dirichlet u */
#define PMX_EX_TCD_EXP 7           /* include/Examples/TCD_2D/exp.h (Example_TimeCD2D 0):
                                      u = e^t sin(2 pi x) sin(2 pi y), Dirichlet 0 */
#define PMX_EX_TCD_LINEAR 8        /* include/Examples/TCD_2D/linear_space_time.h (example -1):
                                      u = 1 + 2x + 3ty */
typedef struct {
  int disc;            /* space_discretization_type: galerkin / supg / gls */
  int example;         /* PMX_EX_* : right-hand side from the exact solution */
  int hexa_quad_rule;  /* 0: degree table (Q2 hexa -> Gauss4), 2: Gauss3 (INTERNAL_QUAD_RULE=2) */
  double eps;          /* diffusion coefficient (coeff[0]) */
  double b[3];         /* convection (coeff[1..d]) */
  double c;            /* reaction */
  double f0;           /* constant rhs for PMX_EX_CONST */
  double delta0;       /* DELTA0 (default 1) */
  double poly_degree;  /* INTERNAL_POLYNOMIAL_DEGREE used by SDFEM delta (reference default 1) */
  const double* dev_coeff_table; /* optional [n_cells][n_quad][8] coefficients (eps,b1,b2,b3,c,f,-,-)
                                    replacing the CoeffFct callback; NULL: constants + example */
  const double* host_dirichlet;  /* optional Dirichlet values [n_dirichlet]; NULL: example values */
  double time;         /* TDatabase::TimeDB->CURRENTTIME for time-dependent examples */
  double theta1;       /* TimeDB->THETA1 (TCD SUPG forms) */
  double tau;          /* TimeDB->CURRENTTIMESTEPLENGTH (TCD SUPG forms) */
  /* Q2 hexahedra, pmx_assemble_cd.  0 (default): where the coefficients are
     constant on each cell (no dev_coeff_table) and disc is Galerkin or SUPG,
     the element matrices are one GEMM of per-cell coefficients with
     reference matrices of the quadrature rule on the matrix cores, and each
     CSR row is assembled once from its cells' element matrices (cell order);
     elsewhere the quadrature kernel.  1: always the quadrature kernel (one
     launch per cell colour). */
  int assembly_variant;
  /* with assembly_variant 0: element matrices on the FP32 matrix cores (FP32
     operands and accumulation), summed into the FP64 CSR in FP64; entries
     carry relative errors of at most about 1e-6 of their row's largest entry.
     PMX_E_ARG where the GEMM path does not apply. */
  int fp32_assembly;
  /* pmx_assemble_cd: 0 adds into A and dev_rhs (the caller zeroes them first,
     as ConvectionDiffusion::assemble does); 1 sets them: the element-GEMM
     route writes every entry once (no zeroing pass, no read of the old
     values), the other routes zero A and dev_rhs themselves first. */
  int overwrite;
} pmx_cd_params;
/* number of quadrature points per cell the assembly uses for this space */
int pmx_assembly_n_quad(const pmx_space* s, int hexa_quad_rule, int* n_quad);
/* physical quadrature points [n_cells][n_quad][3] (to evaluate user CoeffFct on host) */
int pmx_assembly_quad_points(const pmx_space* s, int hexa_quad_rule, double* host_xyz);

/* Assemble3D / Assemble2D with LocalAssembling<d>(ConvDiff)
 * (src/AssembleRoutines/Assemble3D.C:37-1132, Assemble2D.C:6940-7620):
 * ADDS the Galerkin (+SUPG) matrix into A and the load vector into dev_rhs,
 * sets Dirichlet diagonals to 1 and Dirichlet rhs entries to the boundary
 * values, and adds Neumann boundary integrals.  As in the reference the
 * caller zeroes A and rhs first (src/System/ConvectionDiffusion.C:329-340),
 * or sets p->overwrite. */
int pmx_assemble_cd(pmx_ctx* ctx, const pmx_space* s, const pmx_cd_params* p, pmx_matrix* A,
                    double* dev_rhs);

/* Assemble2D / Assemble3D with LocalAssembling<d> of type TCDStiffMassRhs /
 * TCDStiffRhs (src/AssembleRoutines/LocalAssembling.C:842-928) as
 * TimeConvectionDiffusion<d>::call_assembling_routine calls it
 * (src/System/TimeConvectionDiffusion.C:465-517): ADDS the stiffness matrix
 * (TCDStiff) into A, the mass matrix (TCDMass) into M and the load vector
 * (TCDRhs, f at p->time) into dev_rhs; with SUPG (2D only) also TCDStiffSUPG,
 * TCDMassSUPG, TCDRhsSUPG with THETA1 = p->theta1 and time step p->tau
 * (CD_local_assembling_routines.C:130-330).  Any of A, M, dev_rhs may be NULL.
 * Dirichlet rows of A and M get diagonal 1, Dirichlet rhs entries the boundary
 * value at p->time.  The caller zeroes A, M, rhs first. */
int pmx_assemble_tcd(pmx_ctx* ctx, const pmx_space* s, const pmx_cd_params* p, pmx_matrix* A,
                     pmx_matrix* M, double* dev_rhs);
/* TFEFunction2D/3D::Interpolate of the example's exact solution at p->time
 * (Lagrange nodal points of every DOF) */
int pmx_interpolate(pmx_ctx* ctx, const pmx_space* s, const pmx_cd_params* p, double* dev_u);

/* TFEFunction::GetErrors with conv_diff_l2_h1_linf_error
 * (src/AssembleRoutines/ConvDiff.C:2134-2185): errors[0..3] = L2, H1-semi,
 * SD, L_inf (at quadrature points) against the example's exact solution. */
int pmx_compute_errors(pmx_ctx* ctx, const pmx_space* s, const pmx_cd_params* p,
                       const double* dev_u, double* host_errors4);
/* No reference counterpart.  pmx_compute_errors for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops): the same kernels and results, with the per-block partial sums and
 * the 4 results in a fixed scratch array of the library (2 MiB per GPU,
 * resolved by pmx_ctx_create and held until the process ends) instead of a
 * hipMalloc per call.  Limits: at most 16,777,216 cells (PMX_E_ARG beyond);
 * the space's quadrature tables for p->hexa_quad_rule must already be on the
 * device (any earlier pmx_assemble_cd / pmx_assemble_tcd / pmx_compute_errors
 * on the space; else PMX_E_STATE).  All calls on one GPU share the scratch:
 * calls that overlap on one GPU (several host threads or contexts) race. */
int pmx_compute_errors_noalloc(pmx_ctx* ctx, const pmx_space* s, const pmx_cd_params* p,
                               const double* dev_u, double* host_errors4);

/* ---------------------------------------------------------- vector (BLAS-1) */
/* BlockVector::dot / norm (src/Matrix/BlockVector.C:293-386, 491-530) */
int pmx_vector_dot(pmx_ctx* ctx, int n, const double* dev_x, const double* dev_y, double* host_out);
int pmx_vector_norm(pmx_ctx* ctx, int n, const double* dev_x, double* host_out);
/* BlockVector::add_scaled: y += a*x */
int pmx_vector_add_scaled(pmx_ctx* ctx, int n, double a, const double* dev_x, double* dev_y);
/* BlockVector::scale / scaleActive (n = active bound): x *= a (src/Matrix/BlockVector.C:138-181) */
int pmx_vector_scale(pmx_ctx* ctx, int n, double a, double* dev_x);
/* BlockVector::copy / copy_nonactive (with offsets): y = x (src/Matrix/BlockVector.C:244-268) */
int pmx_vector_copy(pmx_ctx* ctx, int n, const double* dev_x, double* dev_y);

/* ------------------------------------------- matrix operations on active rows */
/* FEMatrix::scaleActive (src/Matrix/FEMatrix.C:139-150): entries of the active
 * rows *= factor */
int pmx_matrix_scale_active(pmx_ctx* ctx, pmx_matrix* A, double factor);
/* BlockFEMatrix::add_blockfe_matrix / add_matrix_actives
 * (src/Matrix/BlockFEMatrix.C:842-852, 1829-1862): active entries of A +=
 * factor * active entries of B (same sparsity pattern) */
int pmx_matrix_add_scaled_active(pmx_ctx* ctx, pmx_matrix* A, const pmx_matrix* B, double factor);
/* FEMatrix::multiplyActive (src/Matrix/FEMatrix.C:186-199): y[i] += a (A x)[i]
 * for the active rows, the other entries of y untouched */
int pmx_matrix_multiply_active(pmx_ctx* ctx, const pmx_matrix* A, const double* dev_x, double* dev_y,
                               double a);

/* ------------------------------------------------------ time discretisation */
#define PMX_TIME_BACKWARD_EULER 0
#define PMX_TIME_CRANK_NICOLSON 1
#define PMX_TIME_BDF2 2   /* step 1 is done with backward Euler (pre-stage) */
/* TimeDiscretization::prepare_rhs_from_time_disc
 * (src/TimeDiscRoutines/TimeDiscretizations.C:106-168): on entry dev_rhs is the
 * load vector of the new time level; on return the active part holds the
 * right-hand side of the time-discrete system (BE: tau f + M u_m1; CN:
 * tau/2 (f + f_old) + M u_m1 - tau/2 A u_m1; BDF2: 2tau/3 f + 4/3 M u_m1 -
 * 1/3 M u_m2).  A is the stiffness matrix before prepare_system. `step` counts
 * from 1. */
int pmx_time_prepare_rhs(pmx_ctx* ctx, int scheme, int step, double tau, const pmx_matrix* A,
                         const pmx_matrix* M, double* dev_rhs, const double* dev_old_rhs,
                         const double* dev_u_m1, const double* dev_u_m2);
/* TimeDiscretization::prepare_system_matrix (TimeDiscretizations.C:172-210):
 * active entries of A = factor * A + M (factor tau, tau/2 or 2tau/3) */
int pmx_time_prepare_system(pmx_ctx* ctx, int scheme, int step, double tau, pmx_matrix* A,
                            const pmx_matrix* M);
/* TimeDiscretization::reset_linear_matrices (TimeDiscretizations.C:194-236):
 * active entries of A = (A - M) / factor */
int pmx_time_reset_system(pmx_ctx* ctx, int scheme, int step, double tau, pmx_matrix* A,
                          const pmx_matrix* M);

/* ------------------- Oseen / Navier-Stokes (3D, Q2/Q1 and Q2/P1disc) */
/* The saddle-point system of one mesh level with the NSTYPE-1 block layout of
 * BlockFEMatrix::NSE3D_Type1 (src/Matrix/BlockFEMatrix.C:471-720):
 * ( A 0 0 B1^T ; 0 A 0 B2^T ; 0 0 A B3^T ; B1 B2 B3 0 ), vectors [u1|u2|u3|p].
 * Velocity Q2; pressure Q1 (continuous) or discontinuous P1, all pressure
 * DOFs active.  pmx_nse_create: Dirichlet conditions on the whole boundary
 * (an enclosed flow with pressure projection); pmx_nse_create_pair chooses
 * the pressure space and the example whose boundary conditions apply. */
typedef struct pmx_nse pmx_nse;
#define PMX_EX_NSE_U2P1 9   /* include/Examples/NSE_3D/test_u_2_p_1.h (Example_NSE3D -3) */
#define PMX_EX_NSE_U0P0 13  /* include/Examples/NSE_3D/test_u_0_p_0.h (Example_NSE3D -1):
                               u = (1,0,0), p = 0, Dirichlet boundary except the interior
                               of the face x = 1 (Neumann, data 0) */
/* pressure spaces (TDatabase PRESSURE_SPACE codes of
 * NavierStokes.C:get_velocity_pressure_orders): Q1 continuous, or P1
 * discontinuous (D_P1_3D_H_A, the pair of VELOCITY_SPACE 22) */
#define PMX_NSE_PRESSURE_Q1 1
#define PMX_NSE_PRESSURE_P1DISC (-110)
#define PMX_NSE_BLOCK_A 0
#define PMX_NSE_BLOCK_B 1   /* structure of B1, B2, B3 */
#define PMX_NSE_BLOCK_B1 1
#define PMX_NSE_BLOCK_B2 2
#define PMX_NSE_BLOCK_B3 3
#define PMX_NSE_COMBINED 4
/* The B^T blocks: structure (PMX_NSE_BLOCK_BT, velocity rows x pressure
 * columns) and values (B1T..B3T: Bk^T on the active velocity rows, zero on
 * the Dirichlet rows, which the block operator never applies).  With them
 * and A an NSTYPE-4 caller (BlockFEMatrix::NSE3D_Type4,
 * BlockFEMatrix.C:611-660) reads its blocks: for the Galerkin forms with
 * LAPLACETYPE 0 assembled here A11 = A22 = A33 = A, A12, A13, A21, A23, A31,
 * A32 = 0 and BkT as above (NSTYPE 4 stores the same values in separate
 * blocks). */
#define PMX_NSE_BLOCK_BT 5
#define PMX_NSE_BLOCK_B1T 5
#define PMX_NSE_BLOCK_B2T 6
#define PMX_NSE_BLOCK_B3T 7
typedef struct {
  int n_cells, n_u, n_u_active, n_p;
  long long nnz_a, nnz_b;
  int n_combined;
  long long nnz_combined;
  int pressure_space;     /* PMX_NSE_PRESSURE_Q1 or PMX_NSE_PRESSURE_P1DISC */
  int n_p_cell;           /* pressure DOFs per cell: 8 or 4 */
  int enclosed;           /* 1: no Neumann velocity DOF (pressure row correction and projection) */
} pmx_nse_info;
typedef struct {
  int example;   /* PMX_EX_NSE_U2P1 or PMX_EX_NSE_U0P0 (the example the system was
                    created for): right-hand side f = -nu lap u + (u.grad) u + grad p, g = 0 */
  double nu;     /* viscosity 1/Re (coeff[0]) */
} pmx_nse_params;
/* NavierStokes<3> spaces and matrices on one level: TFESpace3D velocity (Q2) and
 * pressure (Q1), TStructure of A (src/Matrix/Structure.C:829-1277) and of B
 * (rectangular, :1709-1924), the combined matrix pattern.  The velocity space
 * has Dirichlet conditions on the whole boundary (example PMX_EX_NSE_U2P1). */
int pmx_nse_create(pmx_ctx* ctx, const pmx_domain* d, int level, pmx_nse** out);
/* The same for a chosen pair and example, as the NavierStokes<3> constructor
 * builds its spaces (src/System/NavierStokes.C:270-360, velocity space 2 or 22;
 * TFESpace3D with the example's BoundCondition, evaluated at the centre of
 * every boundary face, Dirichlet winning on shared DOFs).  pressure_space:
 * PMX_NSE_PRESSURE_Q1 or PMX_NSE_PRESSURE_P1DISC; example: PMX_EX_NSE_U2P1 or
 * PMX_EX_NSE_U0P0.  With Neumann velocity DOFs the flow is not enclosed
 * (BlockFEMatrix.C:1869-1920): the combined matrix keeps its first pressure
 * row, and the pressure is not projected. */
int pmx_nse_create_pair(pmx_ctx* ctx, const pmx_domain* d, int level, int pressure_space, int example,
                        pmx_nse** out);
int pmx_nse_destroy(pmx_nse* s);
int pmx_nse_get_info(const pmx_nse* s, pmx_nse_info* out);
/* No reference counterpart (kernel choice): the velocity block of the block
 * operator is applied to the three components in one pass over its SELL copy
 * (one lane per row) when it has at least `min_rows` rows (default 131,072),
 * else by three CSR launches with 8 lanes per row.  Same products; the row
 * sums are rounded in another order. */
int pmx_nse_set_velocity_sell_rows(pmx_nse* s, int min_rows);
/* NavierStokes<3>::assemble_linear_terms + assemble_nonlinear_term
 * (src/System/NavierStokes.C:407-700) with the Galerkin NSTYPE-1 forms
 * NSLaplaceGradGradSingle, NSNonlinearTerm_convective_Single (convection field
 * dev_w = [w1|w2|w3] in the velocity space; NULL: Stokes), NSDivergenceBlocks,
 * NSRightHandSide (NSE_local_assembling_routines.C); Gauss4 quadrature.
 * Resets and fills A, B1..B3 and dev_rhs (NULL allowed); Dirichlet velocity
 * rows of A get diagonal 1 and their rhs the example's boundary values. */
int pmx_nse_assemble(pmx_ctx* ctx, pmx_nse* s, const pmx_nse_params* p, const double* dev_w,
                     double* dev_rhs);
/* BlockFEMatrix::apply / apply_scaled_add (src/Matrix/BlockFEMatrix.C:855-924):
 * B^T blocks act on the active velocity rows (FEMatrix::multiplyTransposedActive,
 * src/Matrix/FEMatrix.C:201-218), Dirichlet velocity rows are identity rows. */
int pmx_nse_apply(pmx_ctx* ctx, const pmx_nse* s, const double* dev_x, double* dev_y);
int pmx_nse_apply_scaled_add(pmx_ctx* ctx, const pmx_nse* s, const double* dev_x, double* dev_y,
                             double a);
/* structures and values (A, B, B1..B3, combined matrix of
 * BlockFEMatrix::get_combined_matrix, src/Matrix/BlockFEMatrix.C:1168-1345, with
 * unit Dirichlet rows and the first pressure row replaced by a unit row;
 * zero entries are kept) */
int pmx_nse_get_structure(const pmx_nse* s, int which, int* nnz, int* host_row_ptr, int* host_col);
int pmx_nse_get_values(pmx_ctx* ctx, const pmx_nse* s, int which, double* host_out);
int pmx_nse_get_global_numbers(const pmx_nse* s, int* host_velocity, int* host_pressure);
/* exact solution of the example at the nodal points of every DOF */
int pmx_nse_interpolate(pmx_ctx* ctx, const pmx_nse* s, int example, double* dev_x);
/* TFEFunction3D::project_into_L20 of the pressure block (FEFunction3D.C:1678-1725) */
int pmx_nse_project_pressure(pmx_ctx* ctx, pmx_nse* s, double* dev_x);
/* GridTransfer::RestrictFunction (src/Multigrid/GridTransfer.C:1218-1501) of the
 * three velocity components, as NavierStokes<3>::assemble_nonlinear_term calls
 * it through RestrictFunctionRepeatedly (src/System/NavierStokes.C:598-612):
 * dev_coarse_u[0 .. 3 n_u(coarse)) = R dev_fine_u[0 .. 3 n_u(fine)), where per
 * coarse cell the fine function is projected locally in L2 onto the coarse
 * element and the cell values of a DOF are averaged.  `fine` is the next finer
 * level of the same domain; the operator is built on the first call. */
int pmx_nse_restrict_velocity(pmx_ctx* ctx, pmx_nse* coarse, const pmx_nse* fine, const double* dev_fine_u,
                              double* dev_coarse_u);
/* NavierStokes<3>::get_errors layout: [0] L2 u, [1] H1-semi u, [2] unused,
 * [3] L2 p, [4] H1-semi p */
int pmx_nse_errors(pmx_ctx* ctx, pmx_nse* s, int example, const double* dev_x, double* host_err5);

/* ------------------------------------------- batched time-dependent solves */
/* Many independent TimeConvectionDiffusion<2> problems (P1 on an n x n
 * unit-square triangulation each, own coefficients) advanced together by
 * backward Euler.  No reference counterpart: the reference runs such parameter
 * sweeps one problem after another; per problem the result is that of
 * TimeConvectionDiffusion<2>::assemble/solve with the Galerkin TCD forms and
 * TimeDiscretization (src/System/TimeConvectionDiffusion.C:222-386,
 * src/TimeDiscRoutines/TimeDiscretizations.C:106-210).  For n = 2^k the mesh is
 * TwoTriangles refined k times (the reference's cell order and DOF numbering);
 * otherwise an n x n triangulation with the same diagonal direction, cells row
 * by row.  The linear solve of each step is FP64-residual iterative refinement
 * around FP32 GMRES(m) or BiCGStab + Jacobi, run for all problems in one persistent launch
 * (one workgroup per problem); each step is replayed from a HIP graph. */
typedef struct pmx_batch pmx_batch;
#define PMX_BATCH_INNER_GMRES 0     /* restarted GMRES(m) (Iteration_gmres, right preconditioned) */
#define PMX_BATCH_INNER_BICGSTAB 1  /* BiCGStab (Iteration_bicgstab, right preconditioned) */
typedef struct {
  int n_side;        /* squares per side (2 n^2 triangles) */
  double eps;        /* diffusion */
  double b[2];       /* convection */
  double c;          /* reaction */
} pmx_batch_problem;
typedef struct {
  int example;               /* PMX_EX_TCD_EXP / PMX_EX_TCD_LINEAR: f and boundary data at time t */
  int scheme;                /* PMX_TIME_BACKWARD_EULER */
  double tau;                /* time step length */
  double t0;                 /* start time */
  int inner_solver;          /* PMX_BATCH_INNER_GMRES or PMX_BATCH_INNER_BICGSTAB (FP32) */
  int gmres_restart;         /* inner GMRES(m), m <= 64 */
  int max_inner;             /* inner BiCGStab iterations per refinement step */
  int max_outer;             /* FP64 refinement steps per time step */
  double inner_reduction;    /* inner residual reduction per refinement step (default 1e-2; the
                                256-problem sweep takes 1.65 s for 20 steps with it, 2.11 s with
                                1e-4; the FP64 residual test decides convergence either way) */
  double residual_reduction; /* FP64 residual reduction per time step */
  double residual_tolerance; /* FP64 absolute residual per time step */
  int block;                 /* threads per problem: 256, 512, 1024, or 0 = tune on the first run */
  int max_groups;            /* workgroups one problem may get (1..32, default 16; 1 = one per problem).
                                The first run measures each problem's iterations and gives the
                                costliest problems more workgroups, as long as the whole grid
                                fits on the device at once */
} pmx_batch_params;
typedef struct {
  int n_problems, total_dof, total_active, total_cells;
  long long total_nnz;
  int block;
  double time;
  int n_workgroups;   /* of the solve kernel (> n_problems when problems share several) */
} pmx_batch_info;
typedef struct {
  double seconds;                 /* wall time of the replayed steps */
  long long total_inner_iterations, total_outer_iterations;
  int max_inner_iterations;       /* of one problem, summed over the steps run so far */
  int failed_solves;              /* (problem, step) pairs that missed the tolerance */
  int block;
} pmx_batch_report;
int pmx_batch_default_params(pmx_batch_params* p);
int pmx_batch_create(pmx_ctx* ctx, int n_problems, const pmx_batch_problem* problems,
                     const pmx_batch_params* p, pmx_batch** out);
int pmx_batch_destroy(pmx_batch* b);
int pmx_batch_get_info(const pmx_batch* b, pmx_batch_info* out);
int pmx_batch_problem_size(const pmx_batch* b, int problem, int* n_dof, int* active_bound, int* nnz);
/* time = t0, u = interpolant of the initial condition, assemble stiffness and
 * mass matrices and the system matrix tau A + M of every problem */
int pmx_batch_initialize(pmx_batch* b);
/* advance every problem by n_steps backward-Euler steps (blocks until done) */
int pmx_batch_run(pmx_batch* b, int n_steps, pmx_batch_report* rep);
/* solution of one problem in its own DOF numbering */
int pmx_batch_get_solution(const pmx_batch* b, int problem, double* host_u);
/* inner / outer iteration counts and failed solves per problem (summed over steps) */
int pmx_batch_get_iterations(const pmx_batch* b, int* host_inner, int* host_outer, int* host_failed);
/* CSR pattern (TStructure) and cell DOF map of one problem */
int pmx_batch_get_structure(const pmx_batch* b, int problem, int* host_row_ptr, int* host_col,
                            int* host_global_numbers);
/* stiffness, mass and system matrix values of one problem (any may be NULL) */
int pmx_batch_get_matrices(const pmx_batch* b, int problem, double* host_stiffness, double* host_mass,
                           double* host_system);
/* L2, H1-semi, SD, L_inf errors of one problem at the current time */
int pmx_batch_errors(pmx_batch* b, int problem, double* host_errors4);

/* ------------------------------------------------------------- output */
/* DataWriter<d>::writeVtk (src/FE/DataWriter.C:232-340): legacy ASCII VTK of
 * the space's mesh level with one SCALARS block per FE function (values at
 * the mesh vertices, scientific with 12 digits; coordinates with 8
 * significant digits).  Scalar Lagrange spaces.  The points are numbered in
 * order of first appearance in the cell order (the reference sorts its vertex
 * objects by address).  Reads the functions from device memory. */
int pmx_write_vtk(const pmx_space* s, int n_functions, const double* const* dev_u, const char* const* names,
                  double time, const char* path);

/* -------------------------------------------------- snapshots and POD */
/* SnapshotsCollector (include/ReducedOrderModels/SnapshotsCollector.h:22-80,
 * src/ReducedOrderModels/SnapshotsCollector.C:36-95).  The snapshots stay in
 * HBM as a column-major length x max_snaps matrix (snapshot j at dev + j *
 * length); when `path` is not NULL every stored snapshot is also appended to
 * that file in the reference's text format (one line per snapshot, each
 * entry printed with 12 significant digits followed by a space; the file is
 * truncated on creation).  steps_per_snap: db["steps_per_snap"]. */
typedef struct pmx_snapshots pmx_snapshots;
int pmx_snapshots_create(pmx_ctx* ctx, const char* path, int steps_per_snap, int length, int max_snaps,
                         pmx_snapshots** out);
int pmx_snapshots_destroy(pmx_snapshots* s);
/* SnapshotsCollector::write_data(solution, time_step): stores dev_solution
 * (length entries) when time_step % steps_per_snap == 0; *written (may be
 * NULL) tells whether it did.  More than max_snaps snapshots: PMX_E_ARG. */
int pmx_snapshots_write_data(pmx_snapshots* s, const double* dev_solution, int time_step, int* written);
/* number of stored snapshots and the device matrix (valid until destroy) */
int pmx_snapshots_get(const pmx_snapshots* s, int* n_snaps, const double** dev_matrix);

/* POD (include/ReducedOrderModels/POD.h:35-283, src/ReducedOrderModels/
 * POD.C:20-616): method of snapshots.  The mean over the snapshots is removed
 * (pod_fluctuations_only), the autocorrelation matrix S^T S (euclidean) or
 * S^T M S (L2, M = the Gramian set by pmx_pod_set_gramian) is formed on the
 * FP64 matrix cores, its eigenproblem (LAPACK dsyev in the reference) is
 * solved on the host by cyclic Jacobi rotations, and the basis
 * phi_k = S v_k / sqrt(lambda_k) of the `rank` largest eigenvalues above
 * 1e-10 is formed on the device.  Eigenvector signs are fixed so that the
 * largest-magnitude component of each v_k is positive (dsyev leaves the sign
 * to the implementation). */
#define PMX_POD_EUCLIDEAN 0
#define PMX_POD_L2 1
typedef struct pmx_pod pmx_pod;
typedef struct {
  int pod_rank;                 /* 0: all eigenvalues above the threshold */
  int pod_fluctuations_only;    /* 1 (default): subtract the mean of the snapshots */
  int pod_inner_product;        /* PMX_POD_EUCLIDEAN (default) or PMX_POD_L2 */
} pmx_pod_params;
/* POD::default_pod_database (src/ReducedOrderModels/POD.C:20-60) */
int pmx_pod_default_params(pmx_pod_params* p);
int pmx_pod_create(pmx_ctx* ctx, const pmx_pod_params* p, pmx_pod** out);
int pmx_pod_destroy(pmx_pod* pod);
/* POD::set_gramian (POD.C:238-242): the inner-product matrix of the L2 POD,
 * e.g. the mass matrix (TimeConvectionDiffusionPOD::assemble_gramian) */
int pmx_pod_set_gramian(pmx_pod* pod, const pmx_matrix* gramian);
/* POD::read_snapshots (POD.C:92-118): the snapshot file of a collector */
int pmx_pod_read_snapshots(pmx_pod* pod, const char* path);
/* snapshots from device memory (copied): column j at dev_snaps + j * ld */
int pmx_pod_set_snapshots(pmx_pod* pod, const double* dev_snaps, int length, int n_snaps, int ld);
/* POD::compute_basis (POD.C:121-235) */
int pmx_pod_compute_basis(pmx_pod* pod);
/* length, number of snapshots, rank of the basis, eigenvalues above 1e-10 */
int pmx_pod_get_info(const pmx_pod* pod, int* length, int* n_snaps, int* rank, int* valid_eigs);
/* the n_snaps eigenvalues, largest first (POD::get_pod_eigs after compute_basis) */
int pmx_pod_get_eigs(const pmx_pod* pod, double* host_eigs);
/* basis (length x rank, mode k at dev_out + k * length) and mean (length)
 * copied into device memory; POD::get_basis / get_snaps_avr */
int pmx_pod_get_basis(const pmx_pod* pod, double* dev_out);
int pmx_pod_get_mean(const pmx_pod* pod, double* dev_out);
/* POD::write_pod(basename) (POD.C:299-330, 424-498): <directory>/<basename>pod
 * (header line "POD basis: fluctuating field" or "POD basis: full field", then
 * one mode per line), <basename>mean (with fluctuations) and <basename>eigs
 * ("i \t lambda_i \t cumulative energy" per eigenvalue >= 1e-10) */
int pmx_pod_write(const pmx_pod* pod, const char* directory, const char* basename);

/* The same for every problem of a batched sweep (no reference counterpart:
 * the reference runs one TimeConvectionDiffusionPOD per problem).  Snapshots
 * of all problems are stored together (one device copy of the current
 * solutions per snapshot); the Gram matrices of all problems come from one
 * launch pair, the L2 variant with each problem's mass matrix. */
int pmx_batch_snapshots_reserve(pmx_batch* b, int max_snaps);
int pmx_batch_write_snapshot(pmx_batch* b);
int pmx_batch_compute_pod(pmx_batch* b, const pmx_pod_params* p);
/* rank, eigenvalues (n_snaps, largest first), basis (n_dof x rank, mode k at
 * k * n_dof) and mean (n_dof) of one problem; host pointers may be NULL */
int pmx_batch_get_pod(const pmx_batch* b, int problem, int* rank, double* host_eigs, double* host_basis,
                      double* host_mean);

/* ------------------------------------------------------------------ solver */
#define PMX_IT_FGMRES 0
#define PMX_IT_CG 1
#define PMX_IT_RICHARDSON 2
#define PMX_IT_JACOBI 3
#define PMX_PC_NONE 0
#define PMX_PC_JACOBI 1
#define PMX_PC_MULTIGRID 2
#define PMX_SMOOTHER_JACOBI 0
#define PMX_SMOOTHER_DIRECT 1
#define PMX_SMOOTHER_CELL_VANKA_JACOBI 2   /* VankaSmoother CELL_JACOBI (saddle-point systems) */
#define PMX_SMOOTHER_CELL_VANKA 3          /* VankaSmoother CELL (multiplicative), cells in colour-major
                                              order: colour = lattice index mod 3 per direction */
#define PMX_CYCLE_V 0
#define PMX_CYCLE_W 1
#define PMX_CYCLE_F 2
#define PMX_PRECISION_FP64 0
#define PMX_PRECISION_MIXED 1   /* FP32 V-cycle on bf16 (row-scaled) matrices */
typedef struct {
  int iterative_solver_type;   /* iterative_solver_type */
  int preconditioner;          /* preconditioner */
  int max_n_iterations;        /* default 100 */
  int min_n_iterations;        /* default 0 */
  double residual_tolerance;   /* default 1e-8 */
  double residual_reduction;   /* default 0 */
  double divergence_factor;    /* default 1e10 */
  int gmres_restart;           /* default 20 */
  double damping;              /* Richardson damping (default 1) */
  int mg_cycle;                /* multigrid_cycle_type */
  int mg_smoother;             /* multigrid_smoother */
  int mg_smoother_coarse;      /* multigrid_smoother_coarse */
  double mg_correction_damp;   /* multigrid_correction_damp_factor */
  int mg_n_pre_smooth, mg_n_post_smooth;
  double mg_coarse_residual;   /* multigrid_coarse_residual */
  int mg_coarse_max_n_iterations;
  double jacobi_damp;          /* jacobi_damp */
  int precision;               /* PMX_PRECISION_* (multigrid preconditioner only; for
                                  saddle-point systems MIXED stores the Vanka factors in FP32,
                                  or bf16 with vanka_inverse 2) */
  double vanka_damp;           /* multigrid_vanka_damp_factor (default 1) */
  int zero_start_shortcut;     /* no reference counterpart.  1 (default): a Jacobi
                                  smoothing step on a level whose solution is zero on the
                                  active DOFs (the first step after the cycle resets it)
                                  multiplies only the Dirichlet-column entries; the result
                                  is the same bit for bit.  0: always the full sweep */
  int renumber_levels;         /* no reference counterpart.  1: the multigrid levels above
                                  the coarsest (and scalar FGMRES) run in a solver-internal
                                  lexicographic DOF order with the entries of each row
                                  sorted by the new column (better gather locality; row
                                  sums and dot products are rounded in another order).
                                  2: as 1, and where every active-column offset fits in 16
                                  bits the level matrix stores col - row as int16, the
                                  Dirichlet-column entries in a per-row tail (same results
                                  as 1 bit for bit; slower).  3: as 2 with the
                                  Dirichlet-column products of each row summed once per
                                  reset of the level's solution (its Dirichlet part does
                                  not change between resets) and added first: fewer bytes
                                  per sweep and no tail; rounding differs from 1.
                                  4: as 3 with the slots of a row packed by four (one lane
                                  loads four offsets and four values per instruction); same
                                  results as 3 bit for bit, no faster on MI355X (FP64
                                  slower).  5 (default): as 3 with the rows grouped into slices of
                                  equal active-column offsets (a lexicographic order gives
                                  few such patterns on structured meshes): the offsets of
                                  a slice are read once per slot by the whole wave and only
                                  the values are streamed; same results as 3 bit for bit.
                                  Levels whose offsets do not fit keep 32-bit columns.
                                  0: the reference order */
  int transfer_sell;           /* no reference counterpart.  1 (default): the scalar
                                  multigrid's prolongation and restriction read a SELL-64
                                  copy of P or P^T (one lane per row, rows sorted by
                                  length; each row summed sequentially in column order)
                                  when the product has at least 131,072 rows, else CSR
                                  with 1-8 lanes per row (partial sums combined across
                                  lanes).  2: SELL for every level.  3 (default): as 1,
                                  and on renumbered levels the rows that share their
                                  column offsets and weights (most of them on structured
                                  hierarchies) are grouped into slices that read those
                                  once per wave, with the Dirichlet columns left out (the
                                  transfers' inputs vanish there); same sums as 1 bit for
                                  bit.  4: as 3 for every level.  0: CSR throughout */
  int vanka_inverse;           /* no reference counterpart.  1: the cell_vanka_jacobi
                                  smoother stores the inverse of each cell's local matrix
                                  (from its LU factors, in place, as dgetri) and applies it
                                  by one product; 0: it stores the LU factors and substitutes
                                  (the reference's dgetrs).  Same smoother; the rounding
                                  differs.  2 (default): as 1, and with precision mixed the
                                  local inverses and the values of the smoother's defect
                                  f - K x are bf16 instead of FP32 (rounded from FP32;
                                  FP64 vectors and sums). */
  int refinement;              /* no reference counterpart; precision mixed, FGMRES with
                                  multigrid on a scalar system whose finest level has the
                                  folded 16-bit layout (renumber_levels >= 3).  1 (default):
                                  FP64 iterative refinement: the residual b - A x and the
                                  solution update in FP64 with the FP64 matrix; each
                                  correction A d = r from FGMRES(gmres_restart) in FP32
                                  (FP32 basis vectors, dot products and Hessenberg in FP64)
                                  with the FP32 multigrid as preconditioner and an FP32 (or
                                  bf16) copy of the finest matrix as operator.  The stopping
                                  test of the reference (residual_tolerance,
                                  residual_reduction, max_n_iterations) is applied to the
                                  FP64 residual; iterations count the inner ones.  0: FGMRES
                                  in FP64 throughout, only the preconditioner in FP32. */
  double refinement_inner_reduction; /* reduction of the inner residual estimate per
                                  refinement step (default 1e-4) */
  int refinement_inner_bf16;   /* 1: the inner operator reads the bf16 row-scaled values of
                                  the multigrid's finest level instead of an FP32 copy */
  int graph_cycle;             /* no reference counterpart.  1 (default): with a direct
                                  coarse solve the scalar multigrid cycle is captured once
                                  as a HIP graph and replayed (the same kernels on the same
                                  buffers, fewer launch gaps); 0: launched kernel by kernel */
} pmx_solver_params;
/* Defaults of Solver<>::default_solver_database (src/Solver/Solver.C:20-117)
 * merged with Multigrid::default_multigrid_database (src/Multigrid/Multigrid.C:422-514). */
int pmx_solver_default_params(pmx_solver_params* p);

typedef struct pmx_solver pmx_solver;
int pmx_solver_create(pmx_ctx* ctx, const pmx_solver_params* p, pmx_solver** out);
int pmx_solver_destroy(pmx_solver* s);
/* Solver::update_matrix (src/Solver/Solver.C:260-326) + Multigrid::initialize/update:
 * `levels` are the matrices of all multigrid levels, COARSEST FIRST (the
 * order of Multigrid::initialize); `spaces` their spaces (consecutive mesh
 * levels).  Without multigrid pass n_levels = 1. */
int pmx_solver_update(pmx_solver* s, pmx_matrix* const* levels, const pmx_space* const* spaces,
                      int n_levels);
typedef struct {
  int iterations;
  double initial_residual;
  double final_residual;
  /* 1 when the stop met residual_tolerance or residual_reduction; 0 when the
   * loop ended at max_n_iterations or on stagnation (the reference's
   * converged() returns true in all these cases and only its log differs) */
  int converged;
} pmx_solve_report;
/* Solver::update_matrix for a BlockFEMatrix saddle-point system + Multigrid
 * initialize/update (src/Multigrid/Multigrid.C:157-420) with the
 * cell_vanka_jacobi smoother (src/Multigrid/VankaSmoother.C:44-300): `levels`
 * are the NSE systems of consecutive mesh levels, COARSEST FIRST (assembled);
 * the coarsest is solved with the inverse of its combined matrix
 * (direct_solve) or smoothed.  FGMRES only; without multigrid pass one level. */
int pmx_solver_update_nse(pmx_solver* s, pmx_nse* const* levels, int n_levels);
/* Solver::solve(rhs, solution) (src/Solver/Solver.C:414-488): iterates from
 * the given dev_x.  Blocks until done. */
int pmx_solver_solve(pmx_solver* s, const double* dev_b, double* dev_x, pmx_solve_report* rep);
/* residual history of the last solve (residuals as printed by LoopInfo) */
int pmx_solver_get_history(const pmx_solver* s, double* host_out, int max, int* n);
/* One multigrid cycle as preconditioner: dev_z = MG(dev_r) (Iteration_multigrid::apply) */
int pmx_solver_apply_preconditioner(pmx_solver* s, const double* dev_r, double* dev_z);
/* No reference counterpart (instrumentation): mean duration (HIP events on the
 * solver's stream, `reps` launches) of the Jacobi smoothing sweep of multigrid
 * level `level` (coarsest = 0) exactly as a cycle runs it: the level's own
 * (possibly renumbered) matrix copy and vectors.  kernel = PMX_KERNEL_JACOBI_FP64
 * for an FP64 solver, PMX_KERNEL_JACOBI_BF16 for a mixed one.  The level's
 * scratch vector is overwritten. */
int pmx_solver_time_level_kernel(pmx_solver* s, int level, int kernel, int reps, double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif /* PMX_H */
