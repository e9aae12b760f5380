// Host-callable launchers of the device kernels (all asynchronous on `st`).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace pmx {

// Sliced-ELLPACK (SELL-64-sigma) copy of the active rows of a DevCSR: one
// 64-wide wavefront per slice of 64 rows, entries column-major per slice.
struct DevSELL {
  int n_slices = 0;
  size_t total = 0;               // entries incl. padding
  int* slice_ptr = nullptr;       // [n_slices+1]
  int* row = nullptr;             // [n_slices*64] original row (-1: empty slot)
  int* col = nullptr;             // [total]
  int* src = nullptr;             // [total] CSR position (-1: padding)
  double* val = nullptr;          // FP64 values (nullptr: not built)
  uint16_t* val_bf16 = nullptr;   // bf16 values divided by the row scale
  // active-row entries in Dirichlet columns (row order): CSR position, column
  // and (filled by the owner of the values, see sell_dirichlet_fill) value
  int n_bd = 0;                   // number of such entries
  int* bd_ptr = nullptr;          // [active+1] (nullptr: not built)
  int* bd_src = nullptr;
  int* bd_col = nullptr;
  double* bd_val = nullptr;
  uint16_t* bd_val_bf16 = nullptr;
  // 16-bit layout: slots hold col - row of the active-column entries (col and
  // the Dirichlet-column entries are then only in the bd lists)
  int16_t* off = nullptr;
  // folded 16-bit layout: per active row the sum of its Dirichlet-column
  // products (FP64 or FP32 as the vectors, gsum_bytes = 8 or 4), owned by the
  // user of the layout and refreshed with sell_dirichlet_sums whenever the
  // Dirichlet part of the vector the kernels read changes
  void* gsum = nullptr;
  int gsum_bytes = 0;
  // 1: every vector this copy is applied to vanishes on the Dirichlet rows
  // (coarse multigrid levels, whose right-hand side is reset there; the
  // finest level inside an iterative refinement), so the Dirichlet-column
  // sums are zero: the sweeps from zero are elementwise, gsum (zeros) is never
  // recomputed and the pattern kernel does not read it
  int dir_zero = 0;
  int pack = 1;                   // packed slots (see SellPattern::pack)
  // pattern layout (SellPattern::slice_pat / pat_off): the 64 rows of a
  // slice share their active-column offsets, read once per slot by the wave
  int* slice_pat = nullptr;       // [n_slices] position of the slice's pattern in pat_off
  int* pat_off = nullptr;         // offsets col - row of all patterns
  int n_patterns = 0;
  // grid-transfer matrices: the SELL copy (slice_ptr, row, col, val and, when
  // the weights are bf16 numbers, val_bf16) covers rows [0, n_rows)
  int n_rows = 0;
};

// Device CSR matrix (0-based int32 row pointer / columns, sorted columns).
// Pattern slices of a grid-transfer matrix (the scalar multigrid's P or
// P^T on lexicographically renumbered levels): the rows of a slice share
// their column offsets from the row's first column and their weights, so a
// wave reads each (offset, weight) once and every lane only its first column.
// Columns of Dirichlet DOFs are left out (the transfers' inputs vanish there).
struct DevPatTransfer {
  int n_slices = 0, n_patterns = 0;
  int* row = nullptr;         // [n_slices * 64] output row, -1 for an empty lane
  int* base = nullptr;        // [n_slices * 64] first column of the row (a valid one for empty lanes)
  int* slice_pat = nullptr;   // [n_slices] pattern of the slice
  int* pat_ptr = nullptr;     // [n_patterns + 1] into pat_col / pat_val
  int* pat_col = nullptr;     // column offsets from base, ascending
  double* pat_val = nullptr;  // weights
};

struct DevCSR {
  int n = 0;            // rows
  int m = 0;            // columns
  int nnz = 0;
  int active = 0;       // ActiveBound: rows >= active are identity rows
  int lpr = 64;         // lanes per row used by the SpMV kernels
  int* row_ptr = nullptr;
  int* col = nullptr;
  double* val = nullptr;
  uint16_t* val_bf16 = nullptr;   // mixed precision copy (bf16, row-scaled)
  float* rscale = nullptr;        // per-row scale of the bf16 copy
  DevSELL sell;                   // optional SELL copy used by the FE-matrix kernels
  DevPatTransfer pt;              // optional pattern slices of a transfer matrix (the rows
                                  // not in pattern slices are in sell)
  // fe_apply_scaled_add3: matrices with fewer rows use three CSR launches
  // with 8 lanes per row instead of the one-lane-per-row SELL pass
  int sell3_min_rows = 131072;
};

// SELL values from the CSR values (FP64) / bf16 with the row scales A.rscale
void sell_fill_f64(const DevCSR& A, double* sell_val, hipStream_t st);
void sell_fill_bf16(const DevCSR& A, uint16_t* sell_bf16, hipStream_t st);

// ---- FE-matrix operators (FP64)
void fe_apply_scaled_add(const DevCSR& A, const double* x, double* y, double a, hipStream_t st);
// the same for x + k ld, y + k ld, k = 0, 1, 2 (one pass over A)
void fe_apply_scaled_add3(const DevCSR& A, const double* x, double* y, size_t ld, double a, hipStream_t st);
void fe_apply(const DevCSR& A, const double* x, double* y, hipStream_t st);
int fe_defect_partials(const DevCSR& A);
void fe_defect_f32val(const DevCSR& A, const float* val32, const double* x, const double* b, double* d,
                      double* partial, hipStream_t st);
void fe_defect_bf16val(const DevCSR& A, const uint16_t* val16, const double* x, const double* b, double* d,
                       double* partial, hipStream_t st);
void fe_defect(const DevCSR& A, const double* x, const double* b, double* d, double* partial,
               double* norm_out, hipStream_t st);
void fe_jacobi(const DevCSR& A, const double* x, const double* b, const double* dinv, double* xn,
               double omega, hipStream_t st);
// the same sweep for an x that is zero on the active rows (only the products
// with the Dirichlet entries of x remain); bitwise equal to fe_jacobi on such
// an x.  Needs A.sell.bd_val (FP64) / bd_val_bf16 (bf16) when A has
// Dirichlet couplings.  Returns false (and launches nothing) when A has no
// SELL copy.
bool fe_jacobi_from_zero(const DevCSR& A, const double* x, const double* b, const double* dinv, double* xn,
                         double omega, hipStream_t st);
// values of the Dirichlet-column entries from the CSR values: FP64 copies
// and/or bf16 with the row scales A.rscale (the rounding of sell_fill_bf16)
void sell_dirichlet_fill(const DevCSR& A, double* out64, uint16_t* out16, hipStream_t st);
// ---- FE-matrix operators (bf16 matrix, FP32 vectors)
void fe_apply_scaled_add_bf16(const DevCSR& A, const float* x, float* y, double a, hipStream_t st);
void fe_defect_bf16(const DevCSR& A, const float* x, const float* b, float* d, double* partial,
                    double* norm_out, hipStream_t st);
void fe_jacobi_bf16(const DevCSR& A, const float* x, const float* b, const float* dinv, float* xn,
                    double omega, hipStream_t st);
// folded 16-bit layout: refresh the Dirichlet-column sums for vector x
void sell_dirichlet_sums(const DevCSR& A, const double* x, hipStream_t st);
void sell_dirichlet_sums_bf16(const DevCSR& A, const float* x, hipStream_t st);
bool fe_jacobi_from_zero_bf16(const DevCSR& A, const float* x, const float* b, const float* dinv, float* xn,
                              double omega, hipStream_t st);
// ---- plain CSR (transfer matrices): y = alpha*P x + beta*y on rows < row_end;
//      rows >= row_end are set to 0 when zero_rest, else left unchanged
// divergence blocks of a saddle-point operator (values B.val, v2, v3 on one
// pattern): y += a sum_k B_k (x + k ldx) and (y + k ldy) += a B_k^T p
void csr_b3(const DevCSR& B, const double* v2, const double* v3, const double* x, size_t ldx, double* y, double a,
            int row_end, hipStream_t st);
void csr_bt3(const DevCSR& T, const double* v2, const double* v3, const double* p, double* y, size_t ldy, double a,
             int row_end, hipStream_t st);
// three vectors x + k ldx -> y + k ldy (k = 0, 1, 2) with one pass over P
void csr_spmv3(const DevCSR& P, const double* x, size_t ldx, double* y, size_t ldy, double alpha, double beta,
               int row_end, int zero_rest, hipStream_t st);
void csr_spmv(const DevCSR& P, const double* x, double* y, double alpha, double beta, int row_end,
              int zero_rest, hipStream_t st);
void csr_spmv_f32(const DevCSR& P, const float* x, float* y, double alpha, double beta,
                  int row_end, int zero_rest, hipStream_t st);
void csr_inv_diag(const DevCSR& A, double* dinv, int* zero_flag, hipStream_t st);
void csr_to_bf16(const DevCSR& A, uint16_t* vb, float* rscale, hipStream_t st);

// ---- BLAS-1 (partial buffers hold >= 1024 doubles)
void vec_dot(int n, const double* x, const double* y, double* partial, double* out, hipStream_t st);
void vec_norm(int n, const double* x, double* partial, double* out, hipStream_t st);
void vec_norm_f32(int n, const float* x, double* partial, double* out, hipStream_t st);
// r -= h[0]*v (skipped when h == nullptr); out = <r, w> (or ||r|| when w == nullptr)
void vec_axpy_dot(int n, const double* h, const double* v, double* r, const double* w,
                  double* partial, double* out, hipStream_t st);
// Gram-Schmidt chain without separate reductions: per-block partial sums
// (vec_partials() of them) of <x, y>; r -= h v with h = sum(partial_in), written
// to h_out, and the partial sums of <r, w> (or |r|^2 when w == nullptr);
// out = sum (or sqrt of the sum) of partial sums
int vec_partials();
void vec_dot_partial(int n, const double* x, const double* y, double* partial, hipStream_t st);
void vec_axpy_dot_reduced(int n, const double* partial_in, double* h_out, const double* v, double* r,
                          const double* w, double* partial_out, hipStream_t st);
void vec_reduce(const double* partial, double* out, int sqrt_of_sum, hipStream_t st);
// FP32 vectors with FP64 dot products (the inner Krylov of an iterative
// refinement): as vec_dot_partial / vec_axpy_dot_reduced, r updated in FP64
// and rounded to FP32
void vec_dot_partial_f32(int n, const float* x, const float* y, double* partial, hipStream_t st);
void vec_axpy_dot_reduced_f32(int n, const double* partial_in, double* h_out, const float* v, float* r,
                              const float* w, double* partial_out, hipStream_t st);
// y = x / a_dev[0] rounded to FP32
void vec_scale_to_f32(int n, const double* x, const double* a_dev, float* y, hipStream_t st);
// x += sum_{j<k} c_j Z_j with Z_j = Z + j ld (FP32), k <= 64, c on the host
void vec_multi_axpy_f32(int n, int k, const double* c, const float* Z, size_t ld, double* x, hipStream_t st);
// y = A x on the active rows (y = x = 0 elsewhere) for an x that vanishes on
// the Dirichlet rows, from the folded 16-bit SELL copy (pattern layout when
// present) with FP32 values `vals` (sell_fill_f32) or the bf16 ones;
// zero_gsum: A.active zeros (the Dirichlet-column sums of such an x)
void sell_fill_f32(const DevCSR& A, float* out, hipStream_t st);
void fe_apply_active_f32(const DevCSR& A, const float* vals, const float* x, float* y, const float* zero_gsum,
                         hipStream_t st);
void fe_apply_active_bf16(const DevCSR& A, const float* x, float* y, const float* zero_gsum, hipStream_t st);
// y += (a_dev ? sign*a_dev[0] : a) * x
void vec_axpy(int n, double a, const double* a_dev, double sign, const double* x, double* y,
              hipStream_t st);
void vec_axpy_f32(int n, double a, const float* x, float* y, hipStream_t st);
// y = s * x with s = a, a_dev[0] or 1/a_dev[0]
void vec_scale(int n, double a, const double* a_dev, int a_dev_inv, const double* x, double* y,
               hipStream_t st);
void vec_scale_f32(int n, double a, const double* a_dev, int a_dev_inv, const float* x, float* y,
                   hipStream_t st);
void vec_fill(int n, double v, double* y, hipStream_t st);
// y = a .* x (elementwise)
void vec_mul(int n, const double* a, const double* x, double* y, hipStream_t st);
void vec_fill_f32(int n, float v, float* y, hipStream_t st);
void vec_add_scaled_range(int from, int n, double a, const double* x, double* y, hipStream_t st);
void vec_add_scaled_range_f32(int from, int n, double a, const float* x, float* y, hipStream_t st);
void vec_f64_to_f32(int n, const double* x, float* y, hipStream_t st);
// y[i] = x[perm[i]] (gather into a renumbered vector) / y[perm[i]] = x[i]
// (scatter back); with conversion between FP64 and FP32 where the types differ
void vec_gather(int n, const int* perm, const double* x, double* y, hipStream_t st);
void vec_gather(int n, const int* perm, const double* x, float* y, hipStream_t st);
void vec_gather(int n, const int* perm, const float* x, float* y, hipStream_t st);
void vec_scatter(int n, const int* perm, const double* x, double* y, hipStream_t st);
void vec_scatter(int n, const int* perm, const float* x, double* y, hipStream_t st);
void vec_f32_to_f64(int n, const float* x, double* y, hipStream_t st);

// ---- dense in-place inverse (n x n row-major): blocked Gauss-Jordan with
//      partial pivoting; synchronises st; false when a pivot is zero
bool dense_invert(int n, double* A, hipStream_t st);
// copy cnt device doubles to host-mapped pinned memory, then set *host_flag =
// seq (the host waits for the flag instead of a stream synchronisation)
void publish_to_host(const double* src, int cnt, double* host_dst, unsigned int* host_flag, unsigned int seq,
                     hipStream_t st);
// classical Gram-Schmidt of w against the nv FP32 basis vectors V_k = V + k ld:
// h_out[k] = <w, V_k> (FP64 sums, fixed order), w -= sum h_k V_k,
// h_out[nv] = |w|.  mpartial: vec_multi_partials() doubles, partial:
// vec_partials() doubles; nv <= vec_multi_dot_max()
int vec_multi_dot_max();
int vec_multi_partials();
void vec_cgs_f32(int n, int nv, const float* V, size_t ld, float* w, double* h_out, double* mpartial,
                 double* partial, hipStream_t st);
// D (n x n row-major, n <= 20000) = the CSR matrix A
void csr_to_dense(const DevCSR& A, double* D, hipStream_t st);
// y = A x for a dense row-major n x n matrix
void dense_mv(int n, const double* A, const double* x, double* y, hipStream_t st);
void dense_mv_f32(int n, const double* A, const float* x, float* y, hipStream_t st);
void dense_mv_f32a(int n, const float* A, const float* x, float* y, hipStream_t st);

// ---- assembly
struct AssemblyData {
  int dim = 0;
  int ndof = 0;           // local DOFs per cell
  int nq = 0;             // quadrature points
  int n_cells = 0;
  const double* ref = nullptr;     // per q: w, phi[nd], dphi[dim][nd], d2phi[ncomb][nd]
  const double* geo = nullptr;     // per cell: x0[3], J[9], Jinv[9], detJ, hK, pad
  const int* dofs = nullptr;       // [n_cells][ndof] global numbers
  // Q2 hexahedra: CSR position of every element-matrix entry in the lane
  // order of the matrix-core kernel, [cell][16][64] (-1: padding or a
  // Dirichlet row off the diagonal); nullptr: positions are searched
  const int* pos = nullptr;
};
constexpr int kGeoStride = 24;
constexpr int kQ2PosPerCell = 1024;
// fill the position table of the Q2 matrix-core assembly for all cells
void q2_scatter_positions(const AssemblyData& ad, const DevCSR& A, int* pos, hipStream_t st);

struct CDCoeffs {
  int example = 0;        // built-in right-hand side / exact solution
  int supg = 0;           // 1: add BilinearAssemble_SD (CD) / the TCD SUPG forms; 2: BilinearAssemble_GLS
  double eps = 1, b[3] = {0, 0, 0}, c = 0, f0 = 0;
  double delta0 = 1, poly_degree = 1;
  double time = 0;        // TimeDB->CURRENTTIME of time-dependent examples
  double theta1 = 1, tau = 0;      // TimeDB->THETA1, CURRENTTIMESTEPLENGTH (TCD SUPG)
  const double* table = nullptr;   // optional per-(cell,q) coefficients [8]
  // batched assembly: coefficients per problem and per-problem active bounds
  const double* prob_coef = nullptr;      // [n_prob][4] eps, b1, b2, c
  const int* cell_prob = nullptr;         // problem of each cell
  const int* prob_active_end = nullptr;   // first Dirichlet row (global) of each problem
  const double* time_dev = nullptr;       // time read from device memory (graph replays)
};

// ---- Oseen / Navier-Stokes (kernels_nse.hip)
struct NseAsm {
  int nq = 0, n_u = 0, n_u_active = 0, example = 0, convection = 0;
  int np = 8;                    // pressure DOFs per cell: 8 (Q1) or 4 (discontinuous P1)
  double nu = 1;
  const double* ref = nullptr;   // per q: w, xi[3], phi[27], dphi[3][27], psi[8]
  const double* geo = nullptr;   // per cell (kGeoStride)
  const int* vdofs = nullptr;    // [cells][27]
  const int* pdofs = nullptr;    // [cells][np]
  const double* w = nullptr;     // convection field [3][n_u] (null: Stokes)
  DevCSR A{};                    // velocity block (NSTYPE 1: shared by the 3 components)
  const int* Brp = nullptr;
  const int* Bcol = nullptr;
  double* B1 = nullptr;
  double* B2 = nullptr;
  double* B3 = nullptr;
  double* rhs = nullptr;         // [3 n_u + n_p] or null
};
void assemble_nse(const NseAsm& d, const int* cells, int n, hipStream_t st);
void nse_dirichlet_rhs(int example, int n_u, int active, const double* xyz, double* rhs, hipStream_t st);
void combined_values(long long nnz, const uint8_t* kind, const int* pos, const double* A, const double* B1,
                     const double* B2, const double* B3, double* out, hipStream_t st);
void pressure_project_l20(int n_cells, int nq, int np_cell, const double* ref, const double* geo, const int* pdofs,
                          int n_p, double* p, double* partial, hipStream_t st);
// cell Vanka (Jacobi variant) on Q2 hexahedra: 89 local DOFs per cell (Q1
// pressure); with discontinuous P1 pressure the last 4 are kVankaNone
constexpr int kVankaLocal = 89;
constexpr int kVankaNone = 0x7fffffff;
// factor: LU of every cell's local matrix (gathered from the combined matrix C
// by the cell's 89 sorted DOFs `map`); pmap[cell][i] = global DOF whose
// defect enters row i of the row-interchanged local system
void vanka_factor(int n_cells, const int* map, const DevCSR& C, double* lu64, float* lu32, int* pmap,
                  int* singular, hipStream_t st, bool inverse = false, bool padded = false);
// Jacobi Vanka step with the explicit local inverses (vanka_factor with inverse)
void vanka_apply_inverse(int n_cells, const double* inv64, const float* inv32, const int* pmap, const double* r,
                         double damp, double* upd, hipStream_t st);
// the same with bf16 local inverses in the layout of vanka_inverse_to_bf16
void vanka_apply_inverse_bf16(int n_cells, const uint16_t* inv16, const int* pmap, const double* r, double damp,
                              double* upd, hipStream_t st);
// FP32 local inverses rounded to bf16, columns padded to 90 entries
// (vanka_bf16_inverse_elems(n_cells) elements)
size_t vanka_bf16_inverse_elems(int n_cells);
void vanka_inverse_to_bf16(int n_cells, const float* inv32, uint16_t* inv16, hipStream_t st);
// FP64 / FP32 local inverses in the same padded layout (vanka_factor with
// inverse and padded; vanka_padded_inverse_elems(n_cells) elements)
size_t vanka_padded_inverse_elems(int n_cells);
void vanka_apply_inverse_padded(int n_cells, const double* inv64, const float* inv32, const int* pmap,
                                const double* r, double damp, double* upd, hipStream_t st);
// b = a rounded to bf16 (nearest even)
void f32_to_bf16(long long n, const float* a, uint16_t* b, hipStream_t st);
// upd[cell][i] = damp * (local solution) for local DOF i (the order of map)
void vanka_solve(int n_cells, const double* lu64, const float* lu32, const int* pmap, const double* r,
                 double damp, double* upd, hipStream_t st);
void vanka_accumulate(int n, const int* acc_ptr, const int* acc_idx, const double* upd, double* x,
                      hipStream_t st);
// multiplicative variant over the n cells `cells` of one colour (no shared
// DOFs): local defect f - C x on the cell's rows, substitution, x += damp * local
// solution; map[cell][i] = global DOF of local DOF i
void vanka_colour(int n, const int* cells, const double* lu64, const float* lu32, const int* pmap, const int* map,
                  const DevCSR& C, const double* f, double damp, double* x, hipStream_t st);

// ---- batched time-dependent solve (kernels_batch.hip)
struct BatchDev {
  int n_prob = 0;
  const int* order = nullptr;       // workgroup -> problem (largest first)
  const int* dof_off = nullptr;     // [n_prob+1]
  const int* n_active = nullptr;    // [n_prob]
  const int* slice_off = nullptr;   // [n_prob+1]
  const long long* ws_off = nullptr;   // FP32 Krylov workspace offset per problem
  const int* sptr = nullptr;        // slice -> first slot
  const int* srow = nullptr;        // slot row (local, -1 empty) [slices*64]
  const int* scol = nullptr;        // slot column (local)
  const double* S64 = nullptr;      // system matrix tau A + M, FP64 slots
  const float* S32 = nullptr;       // the same in FP32
  const float* dinv32 = nullptr;    // 1/diag(S) per DOF
  double* x = nullptr;
  const double* rhs = nullptr;
  double* r = nullptr;
  float* w = nullptr;
  float* V = nullptr;
  int* iters = nullptr;
  int* outer = nullptr;
  double* res = nullptr;
  int* fail = nullptr;
  // several workgroups per problem (nullptr: one each, in `order`):
  // workgroup -> problem and rank within it, workgroups per problem,
  // per-problem barrier counters (zeroed before each launch), partial-sum
  // slots [n_prob][2][kBatchMaxGroups][kBatchPartStride], a flag set when a
  // barrier wait gave up
  int n_wg = 0;
  const int* wg_prob = nullptr;
  const int* wg_rank = nullptr;
  const int* groups = nullptr;
  unsigned int* bar = nullptr;
  double* part = nullptr;
  int* bar_timeout = nullptr;
};
constexpr int kBatchMaxGroups = 32;
constexpr int kBatchPartStride = 64;   // partial-sum slots per workgroup (GMRES: up to 64 sums at once)
struct BatchSolveParams {
  int method = 0;          // inner solver: 0 GMRES(m), 1 BiCGStab
  int m = 30;              // GMRES restart (<= 64)
  int max_inner = 1000;    // BiCGStab iterations per refinement step
  int max_outer = 50;      // FP64 refinement steps per time step
  double inner_red = 1e-4; // FP32 GMRES residual reduction per refinement step
  double red = 1e-10;      // FP64 residual reduction per time step
  double tol = 0.0;        // FP64 absolute residual
};
void batch_solve(const BatchDev& B, const BatchSolveParams& P, int block, hipStream_t st);
// workgroups of the BiCGStab batch kernel that can be resident at once
int batch_solve_capacity(int block, int method);
void batch_rhs(int n_slices, const int* slice_prob, const int* dof_off, const int* sptr, const int* srow,
               const int* scol, const double* M64, double tau, const double* f, const double* u,
               double* rhs, hipStream_t st);
void batch_dirichlet_copy(int n, const int* idx, const double* f, double* rhs, double* u, hipStream_t st);
void batch_dirichlet_values(int example, const double* t_dev, int n, const int* idx, const double* xyz,
                            double* out, hipStream_t st);
void batch_system(int n_rows, const int* rp, const uint8_t* row_active, double tau, const double* A,
                  const double* M, double* S, hipStream_t st);
void add_scalar(double* t, double dt, hipStream_t st);
void sell_gather(size_t total, const int* src, const double* v, double* out64, float* out32, hipStream_t st);

// TCD forms of the cells cells[0..n) (one colour): stiffness into A, mass into
// M, load vector into rhs; any of A.val, M.val, rhs may be null.
void assemble_tcd(const AssemblyData& ad, const CDCoeffs& co, const int* cells, int n, int active,
                  const DevCSR& A, const DevCSR& M, double* rhs, hipStream_t st);
// u[i] = exact solution at the nodal point xyz[i] (Lagrange interpolation)
void interpolate_example(int example, double time, int n, const double* xyz, double* u, hipStream_t st);

// Adds the cell contributions of the cells cells[0..n) (one colour) into A and rhs.
void assemble_cd(const AssemblyData& ad, const CDCoeffs& co, const int* cells, int n,
                 const DevCSR& A, double* rhs, hipStream_t st);

// Q2 hexahedra with coefficients constant on each cell: element matrices as
// E_e = sum_r alpha_r(e) R_r (kQ2Terms reference matrices of the quadrature
// rule, [kQ2Terms][kQ2N] row-major, entry n = 32 i + j, j < 27), then every
// active CSR row gathered once from the element matrices of its cells.
constexpr int kQ2Terms = 32;   // 31 reference matrices and one zero row
constexpr int kQ2N = 27 * 32;  // element rows padded to 32 entries (54 tiles of 16)
constexpr int kQ2MaxInc = 8;   // cells per DOF (a vertex of a hexahedral mesh)
constexpr int kQ2MaxRow = 128; // CSR row length the gather holds in LDS
// The incidences of the active rows, in row order and within a row in
// increasing cell order: row r owns the incidence slots inc_ptr[r] ..
// inc_ptr[r + 1]; slot[27 e + i] is the slot of (cell e, local DOF i), -1 for
// a Dirichlet row.  Per slot k, off[32 k + j] is the position of the cell's
// local DOF j in row r relative to row_ptr[r] (j < 27); the element GEMM
// writes row i of cell e's element matrix to E[32 k ..] and its load-vector
// entry to fe[k].
struct Q2Gather {
  int n_active = 0;
  int n_inc = 0;
  const int* inc_ptr = nullptr;
  const int* slot = nullptr;
  const uint8_t* off = nullptr;
};
// host: the reference matrices of a Q2 hexahedron quadrature table
// (ref layout of AssemblyData::ref), FP64, [kQ2Terms][kQ2N]
void q2_reference_matrices(const double* ref, int nq, double* R);
// off[] of all slots from inc[k] = 27 e + i (the slot's cell and local DOF);
// *bad = 1 if a position is missing or beyond the LDS row of the gather
void q2_gather_offsets(const Q2Gather& g, const int* inc, const int* dofs, const DevCSR& A, uint8_t* off,
                       int* bad, hipStream_t st);
// E[32 k + j] and fe[k] of all slots k
void q2_element_gemm(const AssemblyData& ad, const CDCoeffs& co, const Q2Gather& g, const double* R, double* E,
                     double* fe, hipStream_t st);
void q2_element_gemm(const AssemblyData& ad, const CDCoeffs& co, const Q2Gather& g, const float* R, float* E,
                     double* fe, hipStream_t st);
// A.val[row] += sum over the row's slots (cell order), rhs[row] += the
// element load-vector entries, for the active rows (= with overwrite);
// Dirichlet diagonals set to 1
void q2_row_gather(const Q2Gather& g, const double* E, const double* fe, const DevCSR& A, double* rhs,
                   bool overwrite, hipStream_t st);
void q2_row_gather(const Q2Gather& g, const float* E, const double* fe, const DevCSR& A, double* rhs,
                   bool overwrite, hipStream_t st);
// rhs[i] = g(x_i) for Dirichlet rows from the built-in example (i >= active)
void dirichlet_rhs_example(int example, int n_dir, const double* dir_xyz, double* rhs_dir,
                           const CDCoeffs& co, hipStream_t st);
// y[i] += a (A x)[i] on the active rows only (FEMatrix::multiplyActive)
void fe_multiply_active(const DevCSR& A, const double* x, double* y, double a, hipStream_t st);
// L2, H1-semi, SD errors (squared partials) and Linf over quadrature points
void cd_errors(const AssemblyData& ad, const CDCoeffs& co, const double* u, double* partial4,
               double* out4, hipStream_t st);
// The same reduction on the library's fixed per-device scratch (no
// allocation): errors_scratch_init resolves the scratch address on `device`
// (that device current; returns 0 or a HIP error code), errors_scratch gives
// it back (nullptr before the init).  cd_errors_scratch returns false (and
// launches nothing) when the space has more than 256 * kErrorsScratchBlocks
// cells; *out4 is the device address of the 4 results inside the scratch.
constexpr int kErrorsScratchBlocks = 65536;
int errors_scratch_init(int device);
double* errors_scratch(int device);
bool cd_errors_scratch(const AssemblyData& ad, const CDCoeffs& co, const double* u, double* scratch_ptr,
                       const double** out4, hipStream_t st);


// ---- snapshots / POD (kernels_pod.hip); snapshot matrices are column-major
// over global rows with leading dimension ld
constexpr int kPodChunk = 4096;   // rows per workgroup of the Gram kernel
void pod_mean(int rows, int m, long long ld, double* S, double* mean, bool fluct, hipStream_t st);
void pod_spmm(int rows, int m, long long ld, const int* rp, const int* ci, const double* va, const double* S,
              double* WS, hipStream_t st);
void pod_gram(int n_chunks, const int* chunk_row0, const int* chunk_rows, int n_jobs, const int* job_chunk0, int m,
              long long ld, const double* S, const double* WS, double* partial, double* G, hipStream_t st);
void pod_basis(int rows, int n_jobs, const int* job_row0, const int* job_rank, int m, int kmax, long long ld,
               const double* S, const double* C, double* Phi, hipStream_t st);

}  // namespace pmx
