// Krylov solvers and geometric multigrid of the solve path (C ABI part).
//
// Algorithms restated from the reference:
//  * IterativeMethod::converged (include/Solver/IterativeMethod.h:81-139)
//  * Iteration_gmres::flexible_gmres (src/Solver/Iteration_gmres.C:374-532) with
//    modified Gram-Schmidt, Givens rotations (:123-150) and the back solve (:151-168)
//  * Iteration_cg::iterate (src/Solver/Iteration_cg.C:22-79)
//  * Iteration_richardson::iterate (src/Solver/Iteration_richardson.C:28-117)
//  * Iteration_jacobi::iterate/apply (src/Solver/Iteration_jacobi.C:31-108)
//  * Multigrid::cycle / cycle_step / update_rhs_in_coarser_grid /
//    update_solution_in_finer_grid (src/Multigrid/Multigrid.C:157-415) driven by
//    the CycleControl direction sequence (src/Multigrid/CycleControl.C:17-111)
//  * JacobiSmoother::smooth (src/Multigrid/JacobiSmoother.C:22-30)
//  * DirectSmoother (src/Multigrid/DirectSmoother.C:25-42): the coarsest
//    system is solved exactly, here through its dense inverse.
// The vectors live on the GPU; the loops run on the host and read back only
// scalars (residual norms and one Hessenberg column per FGMRES iteration).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "nse_internal.hpp"

using namespace pmx_internal;

namespace {

constexpr int kSellSigmaSolver = 4096;   // SELL sorting window of the renumbered levels

enum Dir { DIR_DUMMY = 0, DIR_DOWN = 1, DIR_UP = 2, DIR_END = 3 };

void fill_recursively(std::vector<int>& rec, int level, int n_levels, int cycle, std::vector<Dir>& cc) {
  if (level == 0) { cc.push_back(DIR_UP); return; }
  for (int j = 0; j < rec[level]; ++j) {
    cc.push_back(DIR_DOWN);
    fill_recursively(rec, level - 1, n_levels, cycle, cc);
  }
  if (cycle == PMX_CYCLE_F) rec[level] = 1;
  cc.push_back(level == n_levels - 1 ? DIR_END : DIR_UP);
}

std::vector<Dir> cycle_control(int cycle, int n_levels) {
  std::vector<int> rec(n_levels, cycle == PMX_CYCLE_V ? 1 : 2);
  rec[n_levels - 1] = 1;
  std::vector<Dir> cc{DIR_DUMMY};
  fill_recursively(rec, n_levels - 1, n_levels, cycle, cc);
  return cc;
}

void free_csr(pmx::DevCSR& M) {
  dfree(M.row_ptr); dfree(M.col); dfree(M.val); dfree(M.val_bf16); dfree(M.rscale);
  dfree(M.sell.slice_ptr); dfree(M.sell.row); dfree(M.sell.col); dfree(M.sell.val); dfree(M.sell.val_bf16);
  M.sell = pmx::DevSELL{};
  dfree(M.pt.row); dfree(M.pt.base); dfree(M.pt.slice_pat); dfree(M.pt.pat_ptr); dfree(M.pt.pat_col);
  dfree(M.pt.pat_val);
  M.pt = pmx::DevPatTransfer{};
}

pmx::DevCSR upload_csr(const pmx::CSRPattern& p, const std::vector<double>& v, hipStream_t st) {
  pmx::DevCSR M;
  M.n = p.n_rows;
  M.m = p.n_cols;
  M.nnz = (int)p.col.size();
  M.active = p.active_bound;
  // lanes per row of the transfer kernel from the mean row length
  const double avg = M.n ? (double)M.nnz / M.n : 0.0;
  M.lpr = avg <= 2.5 ? 1 : avg <= 5.0 ? 2 : avg <= 10.0 ? 4 : 8;
  M.row_ptr = upload(p.row_ptr.data(), p.row_ptr.size(), st);
  M.col = upload(p.col.data(), p.col.size(), st);
  M.val = upload(v.data(), v.size(), st);
  // grid-transfer weights of nested Lagrange spaces are short binary fractions:
  // when every value is a bf16 number, the transfer kernels read a 2-byte copy
  // (same products, same result)
  std::vector<uint16_t> h(v.size());
  bool exact = !v.empty();
  for (size_t k = 0; k < v.size() && exact; ++k) {
    const float f = (float)v[k];
    uint32_t u;
    std::memcpy(&u, &f, sizeof(u));
    exact = (double)f == v[k] && (u & 0xFFFFu) == 0;
    h[k] = (uint16_t)(u >> 16);
  }
  if (exact) M.val_bf16 = upload(h.data(), h.size(), st);
  return M;
}

// SELL-64 copy of a grid-transfer matrix over its rows [0, row_end), for the
// transfer kernel with one lane per row: rows sorted by length (descending,
// stable) within windows of 4096, each row's entries in CSR column order,
// padding slots (column 0, value 0) up to the slice's longest row; the bf16
// copy exists when the CSR upload found every weight to be a bf16 number
// (only the rows listed in `rows`, when given: the rest of a pattern transfer)
void attach_transfer_sell(pmx::DevCSR& M, const pmx::CSRPattern& p, const std::vector<double>& v, int row_end,
                          hipStream_t st, const std::vector<int>* rows = nullptr) {
  constexpr int C = 64, sigma = 4096;
  require(row_end >= 0 && row_end <= p.n_rows, "transfer rows out of range");
  M.sell.n_rows = row_end;
  auto len = [&](int r) { return p.row_ptr[r + 1] - p.row_ptr[r]; };
  std::vector<int> order;
  if (rows) order = *rows;
  else
    for (int r = 0; r < row_end; ++r) order.push_back(r);
  const int nr = (int)order.size();
  if (nr == 0) return;
  for (int w = 0; w < nr; w += sigma)
    std::stable_sort(order.begin() + w, order.begin() + std::min(nr, w + sigma),
                     [&](int a, int b) { return len(a) > len(b); });
  const int n_slices = (nr + C - 1) / C;
  std::vector<int> rowmap((size_t)n_slices * C, -1), sptr(n_slices + 1, 0);
  for (int sl = 0; sl < n_slices; ++sl) {
    int width = 0;
    for (int l = 0; l < C && sl * C + l < nr; ++l) {
      rowmap[(size_t)sl * C + l] = order[sl * C + l];
      width = std::max(width, len(order[sl * C + l]));
    }
    sptr[sl + 1] = sptr[sl] + width * C;
  }
  const size_t total = sptr[n_slices];
  std::vector<int> col(total, 0);
  std::vector<double> val(total, 0.0);
  for (int sl = 0; sl < n_slices; ++sl) {
    const int width = (sptr[sl + 1] - sptr[sl]) / C;
    for (int l = 0; l < C; ++l) {
      const int r = rowmap[(size_t)sl * C + l];
      if (r < 0) continue;
      for (int j = 0; j < len(r); ++j) {
        const size_t k = (size_t)sptr[sl] + (size_t)j * C + l;
        col[k] = p.col[p.row_ptr[r] + j];
        val[k] = v[p.row_ptr[r] + j];
      }
      (void)width;
    }
  }
  M.sell.n_slices = n_slices;
  M.sell.total = total;
  M.sell.slice_ptr = upload(sptr.data(), sptr.size(), st);
  M.sell.row = upload(rowmap.data(), rowmap.size(), st);
  M.sell.col = upload(col.data(), col.size(), st);
  M.sell.val = upload(val.data(), val.size(), st);
  if (M.val_bf16) {
    std::vector<uint16_t> h(total);
    for (size_t k = 0; k < total; ++k) {
      const float f = (float)val[k];
      uint32_t u;
      std::memcpy(&u, &f, sizeof(u));
      h[k] = (uint16_t)(u >> 16);
    }
    M.sell.val_bf16 = upload(h.data(), h.size(), st);
  }
}

// Pattern slices of a grid-transfer matrix over its rows [0, row_end) with
// the columns >= col_end (Dirichlet DOFs: the transfers' inputs are zero
// there) left out.  Rows whose remaining (column - first column, weight)
// sequences agree share a pattern; patterns of at least 64 rows are cut into
// slices of 64 rows (ascending), the rows of rarer patterns go to the SELL
// copy (attach_transfer_sell with all their columns).  On lexicographically
// numbered levels of a structured hierarchy a few hundred patterns (fine node
// type x nearness to the boundary) cover nearly all rows.
void attach_pattern_transfer(pmx::DevCSR& M, const pmx::CSRPattern& p, const std::vector<double>& v, int row_end,
                             int col_end, hipStream_t st) {
  constexpr int C = 64;
  require(row_end >= 0 && row_end <= p.n_rows, "transfer rows out of range");
  std::map<std::pair<std::vector<int>, std::vector<double>>, int> ids;
  std::vector<int> pid(row_end), base(row_end, 0), count;
  std::vector<std::vector<int>> pat_cols;
  std::vector<std::vector<double>> pat_vals;
  for (int r = 0; r < row_end; ++r) {
    std::vector<int> off;
    std::vector<double> w;
    int c0 = -1;
    for (int k = p.row_ptr[r]; k < p.row_ptr[r + 1]; ++k) {
      const int c = p.col[k];
      if (c >= col_end) continue;
      if (c0 < 0) c0 = c;
      off.push_back(c - c0);
      w.push_back(v[k]);
    }
    base[r] = std::max(c0, 0);
    auto key = std::make_pair(off, w);
    auto it = ids.find(key);
    if (it == ids.end()) {
      it = ids.emplace(std::move(key), (int)pat_cols.size()).first;
      pat_cols.push_back(off);
      pat_vals.push_back(w);
      count.push_back(0);
    }
    pid[r] = it->second;
    ++count[it->second];
  }
  // kept patterns, renumbered in order of first appearance
  std::vector<int> keep_id(pat_cols.size(), -1);
  std::vector<int> pat_ptr(1, 0), pcol;
  std::vector<double> pval;
  int n_kept = 0;
  for (size_t q = 0; q < pat_cols.size(); ++q)
    if (count[q] >= C) {
      keep_id[q] = n_kept++;
      pcol.insert(pcol.end(), pat_cols[q].begin(), pat_cols[q].end());
      pval.insert(pval.end(), pat_vals[q].begin(), pat_vals[q].end());
      pat_ptr.push_back((int)pcol.size());
    }
  std::vector<std::vector<int>> rows_of(n_kept);
  std::vector<int> rest;
  for (int r = 0; r < row_end; ++r) {
    if (keep_id[pid[r]] >= 0) rows_of[keep_id[pid[r]]].push_back(r);
    else rest.push_back(r);
  }
  std::vector<int> rowmap, bmap, spat;
  for (int q = 0; q < n_kept; ++q)
    for (size_t i0 = 0; i0 < rows_of[q].size(); i0 += C) {
      const int first = rows_of[q][i0];
      for (int l = 0; l < C; ++l) {
        const size_t i = i0 + l;
        rowmap.push_back(i < rows_of[q].size() ? rows_of[q][i] : -1);
        bmap.push_back(i < rows_of[q].size() ? base[rows_of[q][i]] : base[first]);
      }
      spat.push_back(q);
    }
  if (pcol.empty()) pcol.push_back(0), pval.push_back(0.0);   // keep the arrays non-empty
  M.pt.n_slices = (int)spat.size();
  M.pt.n_patterns = n_kept;
  if (M.pt.n_slices > 0) {
    M.pt.row = upload(rowmap.data(), rowmap.size(), st);
    M.pt.base = upload(bmap.data(), bmap.size(), st);
    M.pt.slice_pat = upload(spat.data(), spat.size(), st);
    M.pt.pat_ptr = upload(pat_ptr.data(), pat_ptr.size(), st);
    M.pt.pat_col = upload(pcol.data(), pcol.size(), st);
    M.pt.pat_val = upload(pval.data(), pval.size(), st);
  }
  attach_transfer_sell(M, p, v, row_end, st, &rest);
}

// one level of the saddle-point multigrid: the system (not owned), the
// cell-Vanka data and the transfer from this level to the next finer one
struct NseLevel {
  pmx_nse* sys = nullptr;
  int n = 0, n_u = 0, n_ua = 0, n_p = 0, n_cells = 0;
  int *map = nullptr, *pmap = nullptr, *acc_ptr = nullptr, *acc_idx = nullptr;
  double* lu64 = nullptr;
  float* lu32 = nullptr;
  uint16_t* lu16 = nullptr;    // bf16 local inverses (vanka_inverse 2, mixed precision)
  float* cval32 = nullptr;     // mixed precision: FP32 copy of the combined values (smoother defects)
  uint16_t* cval16 = nullptr;  // vanka_inverse 2: bf16 copy instead
  bool inv = false;            // lu64/lu32 hold the local inverses instead of the LU factors
  bool padded = false;         // ... with columns padded to 90 entries
  double *upd = nullptr, *rhs = nullptr, *sol = nullptr, *def = nullptr, *res = nullptr;
  pmx::DevCSR Pu{}, Ru{}, Pp{}, Rp{};   // P: finer rows from this level; R = P^T (rows: this level)
  double* dinv_dense = nullptr;          // coarsest level: inverse of the combined matrix
  double* partial = nullptr;
  bool sol_zero = false;                 // sol is zero (set by the cycle's resets)
  bool multiplicative = false;           // cell_vanka: colours one after another
  int* colour_cells = nullptr;           // cells in colour-major order (27 colours)
  std::vector<int> colour_ptr;
  void release() {
    dfree(map); dfree(pmap); dfree(acc_ptr); dfree(acc_idx); dfree(lu64); dfree(lu32); dfree(lu16); dfree(upd); dfree(cval32); dfree(cval16);
    dfree(rhs); dfree(sol); dfree(def); dfree(res);
    free_csr(Pu); free_csr(Ru); free_csr(Pp); free_csr(Rp);
    dfree(dinv_dense); dfree(partial); dfree(colour_cells);
    colour_ptr.clear();
  }
};

struct Level {
  pmx::DevCSR A;             // matrix view: fp64 arrays owned by pmx_matrix, bf16 copy owned here
  int n = 0, active = 0;
  double* dinv = nullptr;
  float* dinv32 = nullptr;
  pmx::DevCSR P;             // prolongation to the next finer level (rows: finer DOFs)
  pmx::DevCSR R;             // restriction from the next finer level (rows: this level)
  double* dinv_dense = nullptr;   // dense inverse (direct coarse solve), row-major n x n
  float* dinv_dense32 = nullptr;  // its FP32 copy (mixed precision cycle)
  double *rhs = nullptr, *sol = nullptr, *def = nullptr, *tmp = nullptr;
  float *rhs32 = nullptr, *sol32 = nullptr, *def32 = nullptr, *tmp32 = nullptr;
  double* partial = nullptr;
  bool sol_zero = false;     // sol vanishes on the active DOFs (set by the cycle's resets)
  int* perm = nullptr;       // solver-internal numbering: perm[new] = reference DOF (nullptr: identity)
  double* bd64 = nullptr;    // values of A's Dirichlet-column slots (zero-start sweep), FP64 ...
  uint16_t* bd16 = nullptr;  // ... or bf16
  bool own_sell = false;     // A.sell is this level's renumbered copy
  // folded 16-bit layout: Dirichlet-column sums of the current solution
  // (valid until the level's solution is reset)
  double* g64 = nullptr;
  float* g32 = nullptr;
  bool g_valid = false;
  void release() {
    if (own_sell) {
      dfree(A.sell.slice_ptr); dfree(A.sell.row); dfree(A.sell.col); dfree(A.sell.src); dfree(A.sell.val);
      dfree(A.sell.bd_ptr); dfree(A.sell.bd_src); dfree(A.sell.bd_col); dfree(A.sell.off);
      dfree(A.sell.slice_pat); dfree(A.sell.pat_off);
      own_sell = false;
    }
    dfree(perm); dfree(bd64); dfree(bd16); dfree(g64); dfree(g32);
    dfree(dinv); dfree(dinv32);
    free_csr(P); free_csr(R); dfree(dinv_dense); dfree(dinv_dense32);
    dfree(rhs); dfree(sol); dfree(def); dfree(tmp);
    dfree(rhs32); dfree(sol32); dfree(def32); dfree(tmp32);
    dfree(partial);
    dfree(A.val_bf16); dfree(A.rscale); dfree(A.sell.val_bf16);
  }
};

}  // namespace

struct pmx_solver {
  CtxRef ctx;
  pmx_solver_params p{};
  std::vector<Level> lv;     // coarsest first
  std::vector<NseLevel> nlv; // saddle-point levels, coarsest first (pmx_solver_update_nse)
  bool saddle = false;
  const pmx_nse* nse_fine = nullptr;   // saddle-point operator without multigrid
  std::vector<pmx_matrix*> mats;   // level matrices (their SELL values are refreshed per solve)
  std::vector<uint64_t> mat_uids;        // ... and their serial numbers at the last update
  std::vector<const pmx_space*> spaces;  // the levels' spaces at the last update
  std::vector<Dir> cc;
  pmx::DevCSR A;             // finest operator (view)
  // FGMRES in the multigrid's numbering of the finest level (renumbered
  // levels): the operator is that level's SELL copy with FP64 values, b and x
  // are gathered once per solve, the preconditioner needs no permutation
  const int* kperm = nullptr;    // finest level's perm (owned by the level)
  pmx::DevCSR K;                 // view of the renumbered finest operator
  double* kval = nullptr;        // FP64 SELL values of K (mixed precision only)
  double* kbd = nullptr;         // FP64 Dirichlet-column values of K (mixed only)
  double *kb = nullptr, *kx = nullptr;   // b and x in that numbering
  bool kown = false;             // K.sell is FGMRES's own 32-bit copy (16-bit levels)
  // iterative refinement (precision mixed, refinement 1): FP32 Krylov basis
  // blocks (column j at + j n), the FP32 values of the finest level's folded
  // copy (nullptr: its bf16 values are used) and zeros for its
  // Dirichlet-column sums
  bool ir = false;
  // K is the finest level's folded layout: its FP64 Dirichlet-column sums
  // of the current operand (formed before every product)
  bool kfold = false;
  double* kg64 = nullptr;
  // the multigrid cycle captured as a HIP graph (direct coarse solve only),
  // one per state of the finest level's dir_zero flag
  hipGraph_t cycle_graph[2] = {nullptr, nullptr};
  hipGraphExec_t cycle_exec[2] = {nullptr, nullptr};
  float *V32 = nullptr, *Z32 = nullptr, *w32 = nullptr, *kval32 = nullptr, *zero32 = nullptr;
  size_t ld32 = 0;   // stride of the FP32 basis vectors (n rounded up to 64 entries)
  double* mdot = nullptr;   // block partials of the classical Gram-Schmidt dots (refinement)
  // refinement: one captured cycle per Krylov slot i, reading v_i and
  // writing z_i in place (no copies in and out of the level vectors)
  std::vector<hipGraph_t> ir_graph;
  std::vector<hipGraphExec_t> ir_exec;
  // pinned host memory the device writes each Hessenberg column to, with a
  // sequence flag the host waits for (instead of a copy + stream sync)
  double* hpin = nullptr;
  unsigned int* hflag = nullptr;
  double* hpin_dev = nullptr;
  unsigned int* hflag_dev = nullptr;
  unsigned int hseq = 0;
  int ir_outer = 0;              // refinement steps of the last solve
  const pmx::DevCSR* op = nullptr;       // operator of the running solve
  bool internal = false;                 // the running solve uses that numbering
  int n = 0;
  std::vector<double*> V, Z;
  double *r = nullptr, *w1 = nullptr, *w2 = nullptr, *w3 = nullptr, *dinv = nullptr;
  double* d_H = nullptr;
  double* d_scal = nullptr;
  double* partial = nullptr;
  std::vector<double> history;
  double initial_res = 0, previous_res = 0;
  bool reached = false;   // the last stop met residual_tolerance or residual_reduction
  bool updated = false;
  void release() {
    for (auto& l : lv) l.release();
    lv.clear();
    for (auto& l : nlv) l.release();
    nlv.clear();
    saddle = false;
    for (auto*& v : V) dfree(v);
    for (auto*& z : Z) dfree(z);
    V.clear(); Z.clear();
    dfree(r); dfree(w1); dfree(w2); dfree(w3); dfree(dinv); dfree(d_H); dfree(d_scal); dfree(partial);
    dfree(kval); dfree(kbd); dfree(kb); dfree(kx);
    dfree(V32); dfree(Z32); dfree(w32); dfree(kval32); dfree(zero32); dfree(mdot);
    ir = false;
    dfree(kg64);
    kfold = false;
    for (int k = 0; k < 2; ++k) {
      if (cycle_exec[k]) (void)hipGraphExecDestroy(cycle_exec[k]);
      if (cycle_graph[k]) (void)hipGraphDestroy(cycle_graph[k]);
      cycle_exec[k] = nullptr;
      cycle_graph[k] = nullptr;
    }
    for (auto& e : ir_exec)
      if (e) (void)hipGraphExecDestroy(e);
    for (auto& g : ir_graph)
      if (g) (void)hipGraphDestroy(g);
    ir_exec.clear();
    ir_graph.clear();
    if (hpin) (void)hipHostFree(hpin);
    hpin = nullptr;
    hflag = nullptr;
    hpin_dev = nullptr;
    hflag_dev = nullptr;
    if (kown) {
      dfree(K.sell.slice_ptr); dfree(K.sell.row); dfree(K.sell.col); dfree(K.sell.src);
      dfree(K.sell.bd_ptr); dfree(K.sell.bd_src); dfree(K.sell.bd_col);
      kown = false;
    }
    kperm = nullptr;
    K = pmx::DevCSR{};
    updated = false;
  }
  ~pmx_solver() { release(); }
};

namespace {

hipStream_t S(pmx_solver* s) { return s->ctx->stream; }

double read_scalar(pmx_solver* s, const double* d) {
  HIP_CHECK(hipGetLastError());
  double h = 0;
  HIP_CHECK(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, S(s)));
  HIP_CHECK(hipStreamSynchronize(S(s)));
  return h;
}

// IterativeMethod::converged
bool converged(pmx_solver* s, double res, int it) {
  const pmx_solver_params& p = s->p;
  if (it == 0) { s->initial_res = res; s->previous_res = res; }
  s->history.push_back(res);
  const bool small = res < p.residual_tolerance;
  const bool reduced = res < p.residual_reduction * s->initial_res;
  const bool enough = it > p.min_n_iterations || p.min_n_iterations == 0;
  const bool stagnation = res > 1e10 * s->previous_res;
  const bool diverged = res > p.divergence_factor * s->initial_res;
  bool conv = (small || reduced || stagnation || diverged) && enough;
  s->reached = (small || reduced) && enough;
  if (it == p.max_n_iterations) conv = true;
  s->previous_res = res;
  if (diverged) throw PmxError(PMX_E_DIVERGED, "Execution of the code stopped because of solver divergence.");
  return conv;
}

// ---------------------------------------------------------------- multigrid
// Folded 16-bit levels: the sums of the Dirichlet-column products of the
// current solution.  The Dirichlet part of a level's solution is set when the
// cycle resets it and is left unchanged by the smoothing steps (unit rows:
// x + w (b - x) = b when x = b) and by the coarse-grid correction (active rows
// only), so the sums are computed once per reset.
void ensure_dirichlet_sums(pmx_solver* s, Level& L, bool mixed) {
  if (L.g_valid || !L.A.sell.gsum) return;
  if (mixed) pmx::sell_dirichlet_sums_bf16(L.A, L.sol32, S(s));
  else pmx::sell_dirichlet_sums(L.A, L.sol, S(s));
  L.g_valid = true;
}

void mg_smooth(pmx_solver* s, int l, bool mixed) {
  Level& L = s->lv[l];
  if (l == 0 && s->p.mg_smoother_coarse == PMX_SMOOTHER_DIRECT) {
    L.sol_zero = false;
    if (mixed && L.dinv_dense32) pmx::dense_mv_f32a(L.n, L.dinv_dense32, L.rhs32, L.sol32, S(s));
    else if (mixed) pmx::dense_mv_f32(L.n, L.dinv_dense, L.rhs32, L.sol32, S(s));
    else pmx::dense_mv(L.n, L.dinv_dense, L.rhs, L.sol, S(s));
    return;
  }
  const bool zero = L.sol_zero && s->p.zero_start_shortcut;
  L.sol_zero = false;
  if (mixed) {
    // the sweep from zero also stores the Dirichlet-column sums of a folded level
    if (zero && pmx::fe_jacobi_from_zero_bf16(L.A, L.sol32, L.rhs32, L.dinv32, L.tmp32, s->p.jacobi_damp, S(s))) {
      L.g_valid = true;
    } else {
      ensure_dirichlet_sums(s, L, mixed);
      pmx::fe_jacobi_bf16(L.A, L.sol32, L.rhs32, L.dinv32, L.tmp32, s->p.jacobi_damp, S(s));
    }
    std::swap(L.sol32, L.tmp32);
  } else {
    if (zero && pmx::fe_jacobi_from_zero(L.A, L.sol, L.rhs, L.dinv, L.tmp, s->p.jacobi_damp, S(s))) {
      L.g_valid = true;
    } else {
      ensure_dirichlet_sums(s, L, mixed);
      pmx::fe_jacobi(L.A, L.sol, L.rhs, L.dinv, L.tmp, s->p.jacobi_damp, S(s));
    }
    std::swap(L.sol, L.tmp);
  }
}

void mg_defect(pmx_solver* s, int l, bool mixed, double* norm_out) {
  Level& L = s->lv[l];
  ensure_dirichlet_sums(s, L, mixed);
  if (mixed) pmx::fe_defect_bf16(L.A, L.sol32, L.rhs32, L.def32, L.partial, norm_out, S(s));
  else pmx::fe_defect(L.A, L.sol, L.rhs, L.def, L.partial, norm_out, S(s));
}

void mg_update_solution_in_finer_grid(pmx_solver* s, int l, bool mixed) {
  Level& C = s->lv[l];
  Level& F = s->lv[l + 1];
  const double damp = s->p.mg_correction_damp;
  F.sol_zero = false;
  if (mixed) pmx::csr_spmv_f32(C.P, C.sol32, F.sol32, damp, 1.0, F.active, 0, S(s));
  else pmx::csr_spmv(C.P, C.sol, F.sol, damp, 1.0, F.active, 0, S(s));
}

void mg_update_rhs_in_coarser_grid(pmx_solver* s, int l, bool mixed) {
  Level& F = s->lv[l];
  Level& C = s->lv[l - 1];
  if (mixed) pmx::csr_spmv_f32(C.R, F.def32, C.rhs32, 1.0, 0.0, C.active, 1, S(s));
  else pmx::csr_spmv(C.R, F.def, C.rhs, 1.0, 0.0, C.active, 1, S(s));
}

// The level's solution need not be cleared when the first operation after the
// reset overwrites all of it without reading it: the direct coarse solve, or
// a first smoothing step from zero on a level whose Dirichlet part is zero as
// well (dir_zero; k_jacobi_from_zero_nodir, the conditions of
// fe_jacobi_from_zero).
bool reset_overwritten(const pmx_solver* s, int l, bool mixed) {
  const Level& L = s->lv[l];
  if (l == 0 && s->p.mg_smoother_coarse == PMX_SMOOTHER_DIRECT) return true;
  if (!s->p.zero_start_shortcut || s->p.mg_n_pre_smooth < 1 || !L.A.sell.dir_zero) return false;
  if (L.A.lpr != 0 || !L.A.sell.bd_ptr) return false;
  const bool vals = mixed ? L.A.sell.bd_val_bf16 != nullptr : L.A.sell.bd_val != nullptr;
  return vals || L.A.sell.n_bd == 0;
}

int mg_cycle_step(pmx_solver* s, int step, int level, bool mixed) {
  const int nl = (int)s->lv.size();
  if (level == 0) {
    // the initial defect only feeds verbose output in the reference
    double res = 1e10;
    if (s->p.mg_smoother_coarse == PMX_SMOOTHER_DIRECT) {
      // an exact solve does not depend on the start value, so repeating it
      // until the defect drops below multigrid_coarse_residual (the
      // reference's loop) yields the same solution as one application
      mg_smooth(s, 0, mixed);
    } else {
      for (int i = 0; i < s->p.mg_coarse_max_n_iterations && res > s->p.mg_coarse_residual; ++i) {
        mg_smooth(s, 0, mixed);
        mg_defect(s, 0, mixed, s->d_scal);
        res = read_scalar(s, s->d_scal);
      }
    }
    if (nl > 1) mg_update_solution_in_finer_grid(s, 0, mixed);
    return 1;
  }
  const bool from_below = s->cc[step - 1] == DIR_UP;
  const bool up = s->cc[step] == DIR_UP;
  const bool down = s->cc[step] == DIR_DOWN;
  if (from_below)
    for (int i = 0; i < s->p.mg_n_post_smooth; ++i) mg_smooth(s, level, mixed);
  if (down) {
    for (int i = 0; i < s->p.mg_n_pre_smooth; ++i) mg_smooth(s, level, mixed);
    mg_defect(s, level, mixed, nullptr);
    Level& C = s->lv[level - 1];
    if (!reset_overwritten(s, level - 1, mixed)) {
      if (mixed) pmx::vec_fill_f32(C.n, 0.f, C.sol32, S(s));
      else pmx::vec_fill(C.n, 0.0, C.sol, S(s));
    }
    C.sol_zero = true;
    C.g_valid = false;
    mg_update_rhs_in_coarser_grid(s, level, mixed);
    return level - 1;
  } else if (up) {
    mg_update_solution_in_finer_grid(s, level, mixed);
    return level + 1;
  }
  return -1;
}

// Multigrid::cycle: finest solution = 0 with the rhs' non-active entries
void mg_cycle(pmx_solver* s, bool mixed) {
  Level& F = s->lv.back();
  if (s->lv.size() > 1 && reset_overwritten(s, (int)s->lv.size() - 1, mixed)) {
    // the rhs vanishes on the Dirichlet rows (dir_zero), so the reset
    // solution is zero everywhere and the first sweep does not read it
  } else if (mixed) {
    pmx::vec_fill_f32(F.active, 0.f, F.sol32, S(s));
    if (F.n > F.active)
      HIP_CHECK(hipMemcpyAsync(F.sol32 + F.active, F.rhs32 + F.active, (F.n - F.active) * sizeof(float),
                               hipMemcpyDeviceToDevice, S(s)));
  } else {
    pmx::vec_fill(F.active, 0.0, F.sol, S(s));
    if (F.n > F.active)
      HIP_CHECK(hipMemcpyAsync(F.sol + F.active, F.rhs + F.active, (F.n - F.active) * sizeof(double),
                               hipMemcpyDeviceToDevice, S(s)));
  }
  F.sol_zero = true;
  F.g_valid = false;
  int level = (int)s->lv.size() - 1;
  const int n_steps = (int)s->cc.size() - 1;
  for (int step = 1; step <= n_steps; ++step) level = mg_cycle_step(s, step, level, mixed);
}

// The cycle is a fixed sequence of kernels and device copies when the
// coarsest level is solved directly (no residual read back): it is captured
// once into a HIP graph and replayed.  The graph holds the buffers of the
// capture: the host-side pointer swaps of the smoothing steps are not redone
// on a replay, so the level vectors stay where the captured cycle left them,
// which is where the replays leave them too.
void mg_cycle_run(pmx_solver* s, bool mixed) {
  if (!s->p.graph_cycle || s->p.mg_smoother_coarse != PMX_SMOOTHER_DIRECT) {
    mg_cycle(s, mixed);
    return;
  }
  hipStream_t st = S(s);
  const int key = s->lv.back().A.sell.dir_zero ? 1 : 0;
  if (!s->cycle_exec[key]) {
    HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      mg_cycle(s, mixed);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(st, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIP_CHECK(hipStreamEndCapture(st, &s->cycle_graph[key]));
    HIP_CHECK(hipGraphInstantiate(&s->cycle_exec[key], s->cycle_graph[key], nullptr, nullptr, 0));
  }
  HIP_CHECK(hipGraphLaunch(s->cycle_exec[key], st));
}

// The refinement's cycle for Krylov slot i: z_i = MG(v_i) with the finest
// level's right-hand side and solution pointed at v_i and z_i while the cycle
// is captured (the graph keeps those addresses; a final copy puts the result
// into z_i if the smoothing steps left it in the other buffer).  Without
// graphs the level vectors are copied in and out.
void mg_cycle_slot(pmx_solver* s, int i, const float* vi, float* zi) {
  Level& F = s->lv.back();
  hipStream_t st = S(s);
  if (!s->p.graph_cycle || s->p.mg_smoother_coarse != PMX_SMOOTHER_DIRECT) {
    HIP_CHECK(hipMemcpyAsync(F.rhs32, vi, (size_t)F.n * sizeof(float), hipMemcpyDeviceToDevice, st));
    mg_cycle_run(s, true);
    HIP_CHECK(hipMemcpyAsync(zi, F.sol32, (size_t)F.n * sizeof(float), hipMemcpyDeviceToDevice, st));
    return;
  }
  if ((int)s->ir_exec.size() <= i) {
    s->ir_exec.resize(i + 1, nullptr);
    s->ir_graph.resize(i + 1, nullptr);
  }
  if (!s->ir_exec[i]) {
    float* const rhs = F.rhs32;
    float* const sol = F.sol32;
    float* const tmp = F.tmp32;
    F.rhs32 = const_cast<float*>(vi);
    F.sol32 = zi;
    F.tmp32 = tmp;
    HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      mg_cycle(s, true);
      if (F.sol32 != zi)
        HIP_CHECK(hipMemcpyAsync(zi, F.sol32, (size_t)F.n * sizeof(float), hipMemcpyDeviceToDevice, st));
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(st, &g);
      if (g) (void)hipGraphDestroy(g);
      F.rhs32 = rhs; F.sol32 = sol; F.tmp32 = tmp;
      throw;
    }
    F.rhs32 = rhs; F.sol32 = sol; F.tmp32 = tmp;
    HIP_CHECK(hipStreamEndCapture(st, &s->ir_graph[i]));
    HIP_CHECK(hipGraphInstantiate(&s->ir_exec[i], s->ir_graph[i], nullptr, nullptr, 0));
  }
  HIP_CHECK(hipGraphLaunch(s->ir_exec[i], st));
}

// Wait until the device has published sequence number seq (bounded: the
// stream is queried now and then, and a stream that finished or failed
// without the flag ends the wait with an error).
void wait_published(pmx_solver* s, unsigned int seq) {
  for (long it = 1;; ++it) {
    if (__atomic_load_n(s->hflag, __ATOMIC_ACQUIRE) == seq) return;
    if ((it & 0xFFFF) == 0) {
      const hipError_t e = hipStreamQuery(S(s));
      if (e == hipSuccess) {
        if (__atomic_load_n(s->hflag, __ATOMIC_ACQUIRE) == seq) return;
        throw PmxError(PMX_E_INTERNAL, "published Hessenberg column not visible after the stream finished");
      }
      if (e != hipErrorNotReady) HIP_CHECK(e);
    }
  }
}

// Iteration_multigrid::apply: z = MG(v)
void mg_apply(pmx_solver* s, const double* v, double* z) {
  Level& F = s->lv.back();
  if (s->p.precision == PMX_PRECISION_MIXED) {
    if (F.perm) pmx::vec_gather(F.n, F.perm, v, F.rhs32, S(s));
    else pmx::vec_f64_to_f32(F.n, v, F.rhs32, S(s));
    mg_cycle_run(s, true);
    if (F.perm) pmx::vec_scatter(F.n, F.perm, F.sol32, z, S(s));
    else pmx::vec_f32_to_f64(F.n, F.sol32, z, S(s));
  } else {
    if (F.perm) pmx::vec_gather(F.n, F.perm, v, F.rhs, S(s));
    else HIP_CHECK(hipMemcpyAsync(F.rhs, v, F.n * sizeof(double), hipMemcpyDeviceToDevice, S(s)));
    mg_cycle_run(s, false);
    if (F.perm) pmx::vec_scatter(F.n, F.perm, F.sol, z, S(s));
    else HIP_CHECK(hipMemcpyAsync(z, F.sol, F.n * sizeof(double), hipMemcpyDeviceToDevice, S(s)));
  }
}

// the same on vectors in the numbering of the finest level (no permutation)
void mg_apply_internal(pmx_solver* s, const double* v, double* z) {
  Level& F = s->lv.back();
  if (s->p.precision == PMX_PRECISION_MIXED) {
    pmx::vec_f64_to_f32(F.n, v, F.rhs32, S(s));
    mg_cycle_run(s, true);
    pmx::vec_f32_to_f64(F.n, F.sol32, z, S(s));
  } else {
    HIP_CHECK(hipMemcpyAsync(F.rhs, v, F.n * sizeof(double), hipMemcpyDeviceToDevice, S(s)));
    mg_cycle_run(s, false);
    HIP_CHECK(hipMemcpyAsync(z, F.sol, F.n * sizeof(double), hipMemcpyDeviceToDevice, S(s)));
  }
}

// ------------------------------------------------- saddle-point multigrid
// Multigrid::cycle / cycle_step (src/Multigrid/Multigrid.C:157-420) on
// BlockFEMatrix levels: defects with the block operator, cell-Vanka smoothing
// (CELL_JACOBI) on the combined matrix, velocity and pressure transfers block
// by block; the coarsest level is solved with the inverse of its combined
// matrix (DirectSmoother) or smoothed.
void nse_smooth(pmx_solver* s, int l) {
  NseLevel& L = s->nlv[l];
  if (l == 0 && s->p.mg_smoother_coarse == PMX_SMOOTHER_DIRECT) {
    L.sol_zero = false;
    pmx::dense_mv(L.n, L.dinv_dense, L.rhs, L.sol, S(s));
    return;
  }
  if (L.multiplicative) {
    // VankaSmoother CELL: each colour's cells see the updates of the earlier ones
    const pmx::DevCSR C = nse_combined(L.sys);
    for (size_t c = 0; c + 1 < L.colour_ptr.size(); ++c)
      pmx::vanka_colour(L.colour_ptr[c + 1] - L.colour_ptr[c], L.colour_cells + L.colour_ptr[c], L.lu64, L.lu32,
                        L.pmap, L.map, C, L.rhs, s->p.vanka_damp, L.sol, S(s));
    L.sol_zero = false;
    return;
  }
  // local defects come from r = rhs - K sol; for sol = 0 that is rhs itself
  // (K 0 sums signed zeros), so the product is skipped
  const double* r = L.rhs;
  if (!(L.sol_zero && s->p.zero_start_shortcut)) {
    if (L.cval16) pmx::fe_defect_bf16val(nse_combined(L.sys), L.cval16, L.sol, L.rhs, L.res, L.partial, S(s));
    else if (L.cval32) pmx::fe_defect_f32val(nse_combined(L.sys), L.cval32, L.sol, L.rhs, L.res, L.partial, S(s));
    else pmx::fe_defect(nse_combined(L.sys), L.sol, L.rhs, L.res, L.partial, nullptr, S(s));
    r = L.res;
  }
  L.sol_zero = false;
  if (L.lu16) pmx::vanka_apply_inverse_bf16(L.n_cells, L.lu16, L.pmap, r, s->p.vanka_damp, L.upd, S(s));
  else if (L.padded)
    pmx::vanka_apply_inverse_padded(L.n_cells, L.lu64, L.lu32, L.pmap, r, s->p.vanka_damp, L.upd, S(s));
  else if (L.inv) pmx::vanka_apply_inverse(L.n_cells, L.lu64, L.lu32, L.pmap, r, s->p.vanka_damp, L.upd, S(s));
  else pmx::vanka_solve(L.n_cells, L.lu64, L.lu32, L.pmap, r, s->p.vanka_damp, L.upd, S(s));
  pmx::vanka_accumulate(L.n, L.acc_ptr, L.acc_idx, L.upd, L.sol, S(s));
}

void nse_defect(pmx_solver* s, int l, double* norm_out) {
  NseLevel& L = s->nlv[l];
  HIP_CHECK(hipMemcpyAsync(L.def, L.rhs, L.n * sizeof(double), hipMemcpyDeviceToDevice, S(s)));
  nse_apply_scaled_add(L.sys, L.sol, L.def, -1.0, S(s));
  if (norm_out) pmx::vec_norm(L.n, L.def, L.partial, norm_out, S(s));
}

void nse_update_solution_in_finer_grid(pmx_solver* s, int l) {
  NseLevel& C = s->nlv[l];
  NseLevel& F = s->nlv[l + 1];
  const double damp = s->p.mg_correction_damp;
  F.sol_zero = false;
  pmx::csr_spmv3(C.Pu, C.sol, (size_t)C.n_u, F.sol, (size_t)F.n_u, damp, 1.0, F.n_ua, 0, S(s));
  pmx::csr_spmv(C.Pp, C.sol + (size_t)3 * C.n_u, F.sol + (size_t)3 * F.n_u, damp, 1.0, F.n_p, 0, S(s));
}

void nse_update_rhs_in_coarser_grid(pmx_solver* s, int l) {
  NseLevel& F = s->nlv[l];
  NseLevel& C = s->nlv[l - 1];
  pmx::csr_spmv3(C.Ru, F.def, (size_t)F.n_u, C.rhs, (size_t)C.n_u, 1.0, 0.0, C.n_ua, 1, S(s));
  pmx::csr_spmv(C.Rp, F.def + (size_t)3 * F.n_u, C.rhs + (size_t)3 * C.n_u, 1.0, 0.0, C.n_p, 1, S(s));
}

int nse_cycle_step(pmx_solver* s, int step, int level) {
  const int nl = (int)s->nlv.size();
  if (level == 0) {
    double res = 1e10;
    if (s->p.mg_smoother_coarse == PMX_SMOOTHER_DIRECT) {
      nse_smooth(s, 0);   // an exact solve: repeating it gives the same solution
    } else {
      for (int i = 0; i < s->p.mg_coarse_max_n_iterations && res > s->p.mg_coarse_residual; ++i) {
        nse_smooth(s, 0);
        nse_defect(s, 0, s->d_scal);
        res = read_scalar(s, s->d_scal);
      }
    }
    if (nl > 1) nse_update_solution_in_finer_grid(s, 0);
    return 1;
  }
  const bool from_below = s->cc[step - 1] == DIR_UP;
  const bool up = s->cc[step] == DIR_UP;
  const bool down = s->cc[step] == DIR_DOWN;
  if (from_below)
    for (int i = 0; i < s->p.mg_n_post_smooth; ++i) nse_smooth(s, level);
  if (down) {
    for (int i = 0; i < s->p.mg_n_pre_smooth; ++i) nse_smooth(s, level);
    nse_defect(s, level, nullptr);
    NseLevel& C = s->nlv[level - 1];
    pmx::vec_fill(C.n, 0.0, C.sol, S(s));
    C.sol_zero = true;
    nse_update_rhs_in_coarser_grid(s, level);
    return level - 1;
  } else if (up) {
    nse_update_solution_in_finer_grid(s, level);
    return level + 1;
  }
  return -1;
}

void nse_mg_apply(pmx_solver* s, const double* v, double* z) {
  NseLevel& F = s->nlv.back();
  HIP_CHECK(hipMemcpyAsync(F.rhs, v, F.n * sizeof(double), hipMemcpyDeviceToDevice, S(s)));
  // finest solution 0, non-actives (Dirichlet velocity entries) from the rhs
  pmx::vec_fill(F.n, 0.0, F.sol, S(s));
  for (int k = 0; k < 3; ++k) {
    const size_t o = (size_t)k * F.n_u + F.n_ua;
    if (F.n_u > F.n_ua)
      HIP_CHECK(hipMemcpyAsync(F.sol + o, F.rhs + o, (F.n_u - F.n_ua) * sizeof(double), hipMemcpyDeviceToDevice,
                               S(s)));
  }
  int level = (int)s->nlv.size() - 1;
  const int n_steps = (int)s->cc.size() - 1;
  for (int step = 1; step <= n_steps; ++step) level = nse_cycle_step(s, step, level);
  HIP_CHECK(hipMemcpyAsync(z, F.sol, F.n * sizeof(double), hipMemcpyDeviceToDevice, S(s)));
}

// dense inverse of an n x n CSR matrix (device values) on the GPU
double* dense_inverse_of(pmx_solver* s, int n, const std::vector<int>& rp, const std::vector<int>& col,
                         const double* dev_vals) {
  if (n > 20000) throw PmxError(PMX_E_ARG, "direct coarse solve supports at most 20000 DOFs");
  std::vector<double> vals(col.size());
  HIP_CHECK(hipMemcpyAsync(vals.data(), dev_vals, vals.size() * sizeof(double), hipMemcpyDeviceToHost, S(s)));
  HIP_CHECK(hipStreamSynchronize(S(s)));
  std::vector<double> a((size_t)n * n, 0.0);
  for (int r = 0; r < n; ++r)
    for (int k = rp[r]; k < rp[r + 1]; ++k) a[(size_t)r * n + col[k]] = vals[k];
  double* d = upload(a.data(), a.size(), S(s));
  const bool ok = pmx::dense_invert(n, d, S(s));
  HIP_CHECK(hipGetLastError());
  if (!ok) { dfree(d); throw PmxError(PMX_E_ZERO_DIAG, "singular coarse matrix in the direct coarse solve"); }
  return d;
}

// Dense inverse of the coarsest matrix by Gauss-Jordan elimination with
// partial pivoting on the GPU (set-up only; the coarse solve is then one dense
// matrix-vector product).
void build_coarse_inverse(pmx_solver* s, const pmx_matrix* M) {
  Level& L = s->lv[0];
  const int n = L.n;
  if (n > 20000) throw PmxError(PMX_E_ARG, "direct coarse solve supports at most 20000 DOFs");
  require(M->A.n == n, "coarse matrix size mismatch");
  // a refresh keeps the buffer (a captured cycle refers to it)
  if (!L.dinv_dense) L.dinv_dense = dmalloc<double>((size_t)n * n);
  pmx::csr_to_dense(M->A, L.dinv_dense, S(s));
  HIP_CHECK(hipGetLastError());
  const bool ok = pmx::dense_invert(n, L.dinv_dense, S(s));
  HIP_CHECK(hipGetLastError());
  if (!ok) throw PmxError(PMX_E_ZERO_DIAG, "singular coarse matrix in the direct coarse solve");
  if (s->p.precision == PMX_PRECISION_MIXED) {
    // the FP32 cycle reads an FP32 copy: half the bytes of its coarse solve
    if (!L.dinv_dense32) L.dinv_dense32 = dmalloc<float>((size_t)n * n);
    pmx::vec_f64_to_f32((int)((size_t)n * n), L.dinv_dense, L.dinv_dense32, S(s));
    HIP_CHECK(hipGetLastError());
  }
}

template <typename T>
T* dzeros(size_t n, hipStream_t st) {
  T* p = dmalloc<T>(n);
  HIP_CHECK(hipMemsetAsync(p, 0, n * sizeof(T), st));
  return p;
}

}  // namespace

namespace {

// r = b - K x and ||r|| into d_scal (scalar FE matrix or saddle-point system)
void op_defect(pmx_solver* s, const double* x, const double* b, double* r) {
  if (s->kfold && s->op == &s->K) pmx::sell_dirichlet_sums(s->K, x, S(s));
  if (s->saddle) {
    const pmx_nse* sys = s->nlv.empty() ? s->nse_fine : s->nlv.back().sys;
    HIP_CHECK(hipMemcpyAsync(r, b, s->n * sizeof(double), hipMemcpyDeviceToDevice, S(s)));
    nse_apply_scaled_add(sys, x, r, -1.0, S(s));
    pmx::vec_norm(s->n, r, s->partial, s->d_scal, S(s));
  } else {
    pmx::fe_defect(s->op ? *s->op : s->A, x, b, r, s->partial, s->d_scal, S(s));
  }
}

void op_apply(pmx_solver* s, const double* x, double* y) {
  if (s->kfold && s->op == &s->K) pmx::sell_dirichlet_sums(s->K, x, S(s));
  if (s->saddle) {
    const pmx_nse* sys = s->nlv.empty() ? s->nse_fine : s->nlv.back().sys;
    pmx::vec_fill(s->n, 0.0, y, S(s));
    nse_apply_scaled_add(sys, x, y, 1.0, S(s));
  } else {
    pmx::fe_apply(s->op ? *s->op : s->A, x, y, S(s));
  }
}

void precondition(pmx_solver* s, const double* v, double* z) {
  if (s->p.preconditioner == PMX_PC_MULTIGRID && s->saddle) nse_mg_apply(s, v, z);
  else if (s->p.preconditioner == PMX_PC_MULTIGRID && s->internal) mg_apply_internal(s, v, z);
  else if (s->p.preconditioner == PMX_PC_MULTIGRID) mg_apply(s, v, z);
  else if (s->p.preconditioner == PMX_PC_JACOBI) pmx::vec_mul(s->n, s->dinv, v, z, S(s));
  else HIP_CHECK(hipMemcpyAsync(z, v, s->n * sizeof(double), hipMemcpyDeviceToDevice, S(s)));
}

void generate_rotation(double dx, double dy, double& cs, double& sn) {
  if (dy == 0.0) { cs = 1.0; sn = 0.0; }
  else if (std::fabs(dy) > std::fabs(dx)) {
    const double t = dx / dy;
    sn = 1.0 / std::sqrt(1.0 + t * t);
    cs = t * sn;
  } else {
    const double t = dy / dx;
    cs = 1.0 / std::sqrt(1.0 + t * t);
    sn = t * cs;
  }
}
void apply_rotation(double& dx, double& dy, double cs, double sn) {
  const double t = cs * dx + sn * dy;
  dy = -sn * dx + cs * dy;
  dx = t;
}

// x += sum_k y_k z_k after the back solve with H (Update, Iteration_gmres.C:151-168)
void gmres_update(pmx_solver* s, double* x, int k, const std::vector<double>& H, int ld,
                  std::vector<double> y) {
  for (int i = k; i >= 0; --i) {
    y[i] /= H[i + (size_t)i * ld];
    for (int j = i - 1; j >= 0; --j) y[j] -= H[j + (size_t)i * ld] * y[i];
  }
  for (int j = 0; j <= k; ++j) pmx::vec_axpy(s->n, y[j], nullptr, 1.0, s->Z[j], x, S(s));
}

void solve_fgmres_run(pmx_solver* s, const double* b, double* x, pmx_solve_report* rep) {
  const int n = s->n;
  const int m = std::min(s->p.gmres_restart, s->p.max_n_iterations);
  const int ld = m + 1;
  std::vector<double> H((size_t)ld * m, 0.0), sv(m + 1), cs(m + 1), sn(m + 1);
  // r = b - A x
  op_defect(s, x, b, s->r);
  double beta = read_scalar(s, s->d_scal);
  double resid = beta;
  rep->initial_residual = beta;
  if (converged(s, resid, 0)) { rep->iterations = 0; rep->final_residual = resid; rep->converged = s->reached ? 1 : 0; return; }
  int j = 1;
  while (j <= s->p.max_n_iterations) {
    pmx::vec_scale(n, 1.0 / beta, nullptr, 0, s->r, s->V[0], S(s));
    std::fill(sv.begin(), sv.end(), 0.0);
    sv[0] = beta;
    int i = 0;
    for (i = 0; i < m && j <= s->p.max_n_iterations; ++i, ++j) {
      precondition(s, s->V[i], s->Z[i]);
      op_apply(s, s->Z[i], s->r);
      double* Hcol = s->d_H + (size_t)i * ld;
      // modified Gram-Schmidt, pipelined: H(k,i) = <r, v_k>; r -= H(k,i) v_k;
      // each step finishes the previous dot product itself (two partial-sum
      // buffers alternate)
      double* pa = s->partial;
      double* pb = s->partial + pmx::vec_partials();
      pmx::vec_dot_partial(n, s->r, s->V[0], pa, S(s));
      for (int k = 0; k < i; ++k) {
        pmx::vec_axpy_dot_reduced(n, pa, Hcol + k, s->V[k], s->r, s->V[k + 1], pb, S(s));
        std::swap(pa, pb);
      }
      pmx::vec_axpy_dot_reduced(n, pa, Hcol + i, s->V[i], s->r, nullptr, pb, S(s));
      pmx::vec_reduce(pb, Hcol + i + 1, 1, S(s));
      HIP_CHECK(hipMemcpyAsync(&H[(size_t)i * ld], Hcol, (i + 2) * sizeof(double), hipMemcpyDeviceToHost, S(s)));
      HIP_CHECK(hipStreamSynchronize(S(s)));
      const double hn = H[(i + 1) + (size_t)i * ld];
      if (hn == 0.0) throw PmxError(PMX_E_BREAKDOWN, "Unhappy breakdown in flexible gmres");
      pmx::vec_scale(n, 0.0, Hcol + i + 1, 1, s->r, s->V[i + 1], S(s));
      for (int k = 0; k < i; ++k)
        apply_rotation(H[k + (size_t)i * ld], H[k + 1 + (size_t)i * ld], cs[k], sn[k]);
      generate_rotation(H[i + (size_t)i * ld], H[i + 1 + (size_t)i * ld], cs[i], sn[i]);
      apply_rotation(H[i + (size_t)i * ld], H[i + 1 + (size_t)i * ld], cs[i], sn[i]);
      apply_rotation(sv[i], sv[i + 1], cs[i], sn[i]);
      resid = std::fabs(sv[i + 1]);
      if (converged(s, resid, j)) {
        gmres_update(s, x, i, H, ld, sv);
        HIP_CHECK(hipStreamSynchronize(S(s)));
        rep->iterations = j; rep->final_residual = resid; rep->converged = s->reached ? 1 : 0;
        return;
      }
    }
    gmres_update(s, x, m - 1, H, ld, sv);
    op_defect(s, x, b, s->r);
    beta = read_scalar(s, s->d_scal);
    if (converged(s, resid, j)) {
      rep->iterations = j; rep->final_residual = resid; rep->converged = s->reached ? 1 : 0;
      return;
    }
  }
  rep->iterations = s->p.max_n_iterations;
  rep->final_residual = resid;
  rep->converged = 0;
}

// FP64 iterative refinement around FP32 FGMRES (precision mixed, refinement
// 1), on vectors in the finest level's numbering.  Outer step: r = b - A x
// with the FP64 matrix, the reference's stopping test on ||r||; inner:
// FGMRES(m) for A d = r from d = 0 with FP32 basis vectors (dot products,
// Hessenberg matrix and Givens rotations in FP64), z_i = multigrid cycle in
// FP32, w = A z_i with the FP32 (or bf16) copy of the finest level; it stops
// at m iterations or when its residual estimate drops below
// max(refinement_inner_reduction ||r||, half the final target); then
// x += sum y_j z_j in FP64.  Every vector of the inner solve vanishes on the
// Dirichlet rows (r_D = b_D - x_D = 0 once x_D = b_D; the cycle and the unit
// rows keep zeros there), so the inner operator needs only the
// active-column entries; x_D is set to b_D first.
void solve_ir_run(pmx_solver* s, const double* b, double* x, pmx_solve_report* rep) {
  const int n = s->n;
  const int m = std::min(s->p.gmres_restart, s->p.max_n_iterations);
  const int ld = m + 1;
  Level& F = s->lv.back();
  const int na = F.active;
  hipStream_t st = S(s);
  if (n > na)
    HIP_CHECK(hipMemcpyAsync(x + na, b + na, (size_t)(n - na) * sizeof(double), hipMemcpyDeviceToDevice, st));
  std::vector<double> H((size_t)ld * m, 0.0), sv(m + 1), cs(m + 1), sn(m + 1);
  op_defect(s, x, b, s->r);
  double beta = read_scalar(s, s->d_scal);
  rep->initial_residual = beta;
  s->ir_outer = 0;
  if (converged(s, beta, 0)) { rep->iterations = 0; rep->final_residual = beta; rep->converged = s->reached ? 1 : 0; return; }
  const double target = std::max(s->p.residual_tolerance, s->p.residual_reduction * s->initial_res);
  // the cycle's input vanishes on the Dirichlet rows in here (see above)
  struct DirZero {
    Level& L;
    bool on;
    ~DirZero() { L.A.sell.dir_zero = 0; L.g_valid = false; }
  } dz{F, s->p.zero_start_shortcut != 0};
  if (dz.on) {
    F.A.sell.dir_zero = 1;
    if (F.g32) pmx::vec_fill_f32(std::max(1, F.active), 0.f, F.g32, st);
  }
  int j = 0;
  while (true) {
    ++s->ir_outer;
    pmx::vec_scale_to_f32(n, s->r, s->d_scal, s->V32, st);
    const size_t ldv = s->ld32;   // stride of the FP32 basis (ld is the Hessenberg's)
    std::fill(sv.begin(), sv.end(), 0.0);
    sv[0] = beta;
    const double inner_target = std::max(s->p.refinement_inner_reduction * beta, 0.5 * target);
    int i = 0, used = 0;
    for (i = 0; i < m && j < s->p.max_n_iterations; ++i) {
      float* vi = s->V32 + (size_t)i * ldv;
      float* zi = s->Z32 + (size_t)i * ldv;
      mg_cycle_slot(s, i, vi, zi);
      if (s->kval32) pmx::fe_apply_active_f32(F.A, s->kval32, zi, s->w32, s->zero32, st);
      else pmx::fe_apply_active_bf16(F.A, zi, s->w32, s->zero32, st);
      double* Hcol = s->d_H + (size_t)i * ld;
      if (s->mdot) {
        // classical Gram-Schmidt: all dots in one pass over the basis, the
        // update and the norm in a second (the inner solve only has to
        // reach refinement_inner_reduction; the FP64 outer residual sets the
        // accuracy)
        pmx::vec_cgs_f32(n, i + 1, s->V32, ldv, s->w32, Hcol, s->mdot, s->partial, st);
      } else {
        double* pa = s->partial;
        double* pb = s->partial + pmx::vec_partials();
        pmx::vec_dot_partial_f32(n, s->w32, s->V32, pa, st);
        for (int k = 0; k < i; ++k) {
          pmx::vec_axpy_dot_reduced_f32(n, pa, Hcol + k, s->V32 + (size_t)k * ldv, s->w32,
                                        s->V32 + (size_t)(k + 1) * ldv, pb, st);
          std::swap(pa, pb);
        }
        pmx::vec_axpy_dot_reduced_f32(n, pa, Hcol + i, vi, s->w32, nullptr, pb, st);
        pmx::vec_reduce(pb, Hcol + i + 1, 1, st);
      }
      // v_{i+1} = w / h_{i+1,i} needs no host value: queued before the wait
      pmx::vec_scale_f32(n, 0.0, Hcol + i + 1, 1, s->w32, s->V32 + (size_t)(i + 1) * ldv, st);
      if (s->hpin) {
        const unsigned int seq = ++s->hseq;
        pmx::publish_to_host(Hcol, i + 2, s->hpin_dev, s->hflag_dev, seq, st);
        HIP_CHECK(hipGetLastError());
        wait_published(s, seq);
        std::copy(s->hpin, s->hpin + i + 2, H.begin() + (size_t)i * ld);
      } else {
        HIP_CHECK(hipMemcpyAsync(&H[(size_t)i * ld], Hcol, (i + 2) * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
      }
      const double hn = H[(i + 1) + (size_t)i * ld];
      if (hn == 0.0) throw PmxError(PMX_E_BREAKDOWN, "Unhappy breakdown in flexible gmres");
      for (int k = 0; k < i; ++k) apply_rotation(H[k + (size_t)i * ld], H[k + 1 + (size_t)i * ld], cs[k], sn[k]);
      generate_rotation(H[i + (size_t)i * ld], H[i + 1 + (size_t)i * ld], cs[i], sn[i]);
      apply_rotation(H[i + (size_t)i * ld], H[i + 1 + (size_t)i * ld], cs[i], sn[i]);
      apply_rotation(sv[i], sv[i + 1], cs[i], sn[i]);
      ++j;
      used = i + 1;
      const double est = std::fabs(sv[i + 1]);
      s->history.push_back(est);
      if (est <= inner_target) break;
    }
    if (used == 0) break;   // max_n_iterations reached
    // back solve and x += sum_k y_k z_k (FP64 accumulation)
    std::vector<double> y(sv.begin(), sv.begin() + used);
    for (int q = used - 1; q >= 0; --q) {
      y[q] /= H[q + (size_t)q * ld];
      for (int k = q - 1; k >= 0; --k) y[k] -= H[k + (size_t)q * ld] * y[q];
    }
    pmx::vec_multi_axpy_f32(n, used, y.data(), s->Z32, ldv, x, st);
    op_defect(s, x, b, s->r);
    beta = read_scalar(s, s->d_scal);
    if (converged(s, beta, j) || j >= s->p.max_n_iterations) {
      rep->iterations = j; rep->final_residual = beta; rep->converged = s->reached ? 1 : 0;
      return;
    }
  }
  rep->iterations = j;
  rep->final_residual = beta;
  rep->converged = 0;
}

// Scalar systems with renumbered multigrid levels run FGMRES in the finest
// level's numbering: b and x are permuted once per solve instead of the
// preconditioner's input and output on every iteration, and the operator is
// that level's SELL copy, whose gathers are local.  Only the summation order
// of row sums and dot products differs from the reference numbering.
void solve_fgmres(pmx_solver* s, const double* b, double* x, pmx_solve_report* rep) {
  if (!s->kperm) {
    s->op = &s->A;
    s->internal = false;
    solve_fgmres_run(s, b, x, rep);
    return;
  }
  pmx::vec_gather(s->n, s->kperm, b, s->kb, S(s));
  pmx::vec_gather(s->n, s->kperm, x, s->kx, S(s));
  s->op = &s->K;
  s->internal = true;
  try {
    if (s->ir) solve_ir_run(s, s->kb, s->kx, rep);
    else solve_fgmres_run(s, s->kb, s->kx, rep);
  } catch (...) {
    s->op = &s->A;
    s->internal = false;
    throw;
  }
  s->op = &s->A;
  s->internal = false;
  pmx::vec_scatter(s->n, s->kperm, s->kx, x, S(s));
}

void solve_cg(pmx_solver* s, const double* b, double* x, pmx_solve_report* rep) {
  const int n = s->n;
  double* r = s->r;
  double* z = s->w1;
  double* pv = s->w2;
  double* q = s->w3;
  pmx::fe_defect(s->A, x, b, r, s->partial, s->d_scal, S(s));
  double rn = read_scalar(s, s->d_scal);
  pmx::vec_norm(n, b, s->partial, s->d_scal, S(s));
  double normb = read_scalar(s, s->d_scal);
  if (normb == 0.0) normb = 1;
  double resid = rn / normb;
  rep->initial_residual = resid;
  if (converged(s, resid, 0)) { rep->iterations = 0; rep->final_residual = resid; rep->converged = s->reached ? 1 : 0; return; }
  double rho_1 = 0;
  for (int i = 1; i <= s->p.max_n_iterations; ++i) {
    precondition(s, r, z);
    pmx::vec_dot(n, r, z, s->partial, s->d_scal, S(s));
    const double rho = read_scalar(s, s->d_scal);
    if (i == 1) HIP_CHECK(hipMemcpyAsync(pv, z, n * sizeof(double), hipMemcpyDeviceToDevice, S(s)));
    else {
      const double beta = rho / rho_1;
      pmx::vec_scale(n, beta, nullptr, 0, pv, pv, S(s));
      pmx::vec_axpy(n, 1.0, nullptr, 1.0, z, pv, S(s));
    }
    pmx::fe_apply(s->A, pv, q, S(s));
    pmx::vec_dot(n, pv, q, s->partial, s->d_scal, S(s));
    const double alpha = rho / read_scalar(s, s->d_scal);
    pmx::vec_axpy(n, alpha, nullptr, 1.0, pv, x, S(s));
    pmx::vec_axpy(n, -alpha, nullptr, 1.0, q, r, S(s));
    pmx::vec_norm(n, r, s->partial, s->d_scal, S(s));
    resid = read_scalar(s, s->d_scal) / normb;
    if (converged(s, resid, i)) { rep->iterations = i; rep->final_residual = resid; rep->converged = s->reached ? 1 : 0; return; }
    rho_1 = rho;
  }
  rep->iterations = s->p.max_n_iterations; rep->final_residual = resid; rep->converged = 0;
}

// Richardson (preconditioned) and the plain Jacobi method
void solve_richardson(pmx_solver* s, const double* b, double* x, pmx_solve_report* rep, bool jacobi) {
  const int n = s->n;
  double* r = s->r;
  double* z = s->w1;
  pmx::vec_norm(n, b, s->partial, s->d_scal, S(s));
  double normb = read_scalar(s, s->d_scal);
  if (normb == 0.0) normb = 1;
  pmx::fe_defect(s->A, x, b, r, s->partial, s->d_scal, S(s));
  double resid = read_scalar(s, s->d_scal) / normb;
  rep->initial_residual = resid;
  if (converged(s, resid, 0)) { rep->iterations = 0; rep->final_residual = resid; rep->converged = s->reached ? 1 : 0; return; }
  const double damp = s->p.damping;
  for (int i = 1; i <= s->p.max_n_iterations; ++i) {
    if (jacobi) pmx::vec_mul(n, s->dinv, r, z, S(s));
    else precondition(s, r, z);
    pmx::vec_axpy(n, damp, nullptr, 1.0, z, x, S(s));
    pmx::fe_defect(s->A, x, b, r, s->partial, s->d_scal, S(s));
    resid = read_scalar(s, s->d_scal) / normb;
    if (converged(s, resid, i)) { rep->iterations = i; rep->final_residual = resid; rep->converged = s->reached ? 1 : 0; return; }
  }
  rep->iterations = s->p.max_n_iterations; rep->final_residual = resid; rep->converged = 0;
}

void check_zero_diag(pmx_solver* s, const pmx::DevCSR& A, double* dinv);

// Solver::update_matrix with the structure of the last update (the same
// level matrices and spaces): only what depends on the matrix values is
// recomputed, in the buffers of the last update (the reference's
// Multigrid::update after Multigrid::initialize).  Pattern-dependent data --
// numbering, SELL layouts, transfer matrices -- is kept.
void refresh_values(pmx_solver* s) {
  hipStream_t st = S(s);
  const bool mixed = s->p.precision == PMX_PRECISION_MIXED;
  for (pmx_matrix* M : s->mats) ensure_sell_values(M, st);
  s->A = s->mats.back()->A;
  if (s->p.preconditioner == PMX_PC_JACOBI || s->p.iterative_solver_type == PMX_IT_JACOBI)
    check_zero_diag(s, s->A, s->dinv);
  const int nl = (int)s->lv.size();
  for (int l = 0; l < nl; ++l) {
    Level& L = s->lv[l];
    const pmx_matrix* M = s->mats[l];
    const pmx::DevCSR ref = M->A;   // reference-order CSR (values just assembled)
    L.sol_zero = false;
    L.g_valid = false;
    if (L.own_sell) {
      double* d = dmalloc<double>(L.n);
      check_zero_diag(s, ref, d);
      pmx::vec_gather(L.n, L.perm, d, L.dinv, st);
      if (mixed) {
        float* rs = dmalloc<float>(L.n);
        pmx::csr_to_bf16(ref, nullptr, rs, st);   // row scales only
        pmx::vec_gather(L.n, L.perm, rs, L.A.rscale, st);
        pmx::sell_fill_bf16(L.A, L.A.sell.val_bf16, st);
        HIP_CHECK(hipStreamSynchronize(st));
        dfree(rs);
      } else {
        pmx::sell_fill_f64(L.A, L.A.sell.val, st);
      }
      HIP_CHECK(hipStreamSynchronize(st));
      dfree(d);
    } else {
      check_zero_diag(s, ref, L.dinv);
      if (mixed) {
        pmx::csr_to_bf16(L.A, L.A.val_bf16, L.A.rscale, st);
        if (L.A.sell.val_bf16) pmx::sell_fill_bf16(L.A, L.A.sell.val_bf16, st);
      }
    }
    if (L.bd16) pmx::sell_dirichlet_fill(L.A, nullptr, L.bd16, st);
    if (L.bd64) pmx::sell_dirichlet_fill(L.A, L.bd64, nullptr, st);
    if (L.g32) pmx::vec_fill_f32(std::max(1, L.active), 0.f, L.g32, st);
    if (L.g64) pmx::vec_fill(std::max(1, L.active), 0.0, L.g64, st);
    if (mixed) pmx::vec_f64_to_f32(L.n, L.dinv, L.dinv32, st);
  }
  if (nl > 0 && s->p.mg_smoother_coarse == PMX_SMOOTHER_DIRECT) build_coarse_inverse(s, s->mats[0]);
  if (s->kval) {
    pmx::sell_fill_f64(s->K, s->kval, st);
    if (s->kbd) pmx::sell_dirichlet_fill(s->K, s->kbd, nullptr, st);
  }
  if (s->kval32) pmx::sell_fill_f32(s->lv.back().A, s->kval32, st);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(st));
}

void check_zero_diag(pmx_solver* s, const pmx::DevCSR& A, double* dinv) {
  int* flag = dmalloc<int>(1);
  HIP_CHECK(hipMemsetAsync(flag, 0, sizeof(int), S(s)));
  pmx::csr_inv_diag(A, dinv, flag, S(s));
  int h = 0;
  HIP_CHECK(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, S(s)));
  HIP_CHECK(hipStreamSynchronize(S(s)));
  dfree(flag);
  if (h) throw PmxError(PMX_E_ZERO_DIAG, "zero entry on the diagonal using Jacobi solver/preconditioner");
}

}  // namespace

extern "C" {

int pmx_solver_default_params(pmx_solver_params* p) {
  return guarded([&] {
    require(p != nullptr, "NULL argument");
    pmx_solver_params d{};
    d.iterative_solver_type = PMX_IT_FGMRES;
    d.preconditioner = PMX_PC_NONE;
    d.max_n_iterations = 100;
    d.min_n_iterations = 0;
    d.residual_tolerance = 1e-8;
    d.residual_reduction = 0.0;
    d.divergence_factor = 1e10;
    d.gmres_restart = 20;
    d.damping = 1.0;
    d.mg_cycle = PMX_CYCLE_V;
    d.mg_smoother = PMX_SMOOTHER_JACOBI;
    d.mg_smoother_coarse = PMX_SMOOTHER_DIRECT;
    d.mg_correction_damp = 1.0;
    d.mg_n_pre_smooth = 1;
    d.mg_n_post_smooth = 1;
    d.mg_coarse_residual = 1e-1;
    d.mg_coarse_max_n_iterations = 10;
    d.jacobi_damp = 1.0;
    d.precision = PMX_PRECISION_FP64;
    d.vanka_damp = 1.0;
    d.zero_start_shortcut = 1;
    d.renumber_levels = 5;
    d.transfer_sell = 3;
    d.vanka_inverse = 2;
    d.refinement = 1;
    d.refinement_inner_reduction = 1e-4;
    d.refinement_inner_bf16 = 0;
    d.graph_cycle = 1;
    *p = d;
  });
}

int pmx_solver_create(pmx_ctx* ctx, const pmx_solver_params* p, pmx_solver** out) {
  return guarded([&] {
    require(ctx && p && out, "NULL argument");
    require(p->iterative_solver_type >= 0 && p->iterative_solver_type <= 3, "unknown iterative solver");
    require(p->preconditioner >= 0 && p->preconditioner <= 2, "unknown preconditioner");
    require(p->max_n_iterations >= 0 && p->gmres_restart >= 1, "invalid iteration counts");
    require(p->mg_smoother == PMX_SMOOTHER_JACOBI || p->mg_smoother == PMX_SMOOTHER_CELL_VANKA_JACOBI ||
                p->mg_smoother == PMX_SMOOTHER_CELL_VANKA,
            "unknown multigrid smoother");
    require(p->mg_smoother_coarse == PMX_SMOOTHER_JACOBI || p->mg_smoother_coarse == PMX_SMOOTHER_DIRECT ||
                p->mg_smoother_coarse == PMX_SMOOTHER_CELL_VANKA_JACOBI ||
                p->mg_smoother_coarse == PMX_SMOOTHER_CELL_VANKA,
            "unknown coarse smoother");
    require(p->precision == PMX_PRECISION_FP64 || p->precision == PMX_PRECISION_MIXED, "unknown precision");
    auto s = std::make_unique<pmx_solver>();
    s->ctx = ctx;
    s->p = *p;
    if (s->p.min_n_iterations > s->p.max_n_iterations) s->p.min_n_iterations = s->p.max_n_iterations;
    *out = s.release();
  });
}

int pmx_solver_destroy(pmx_solver* s) {
  return guarded([&] {
    if (!s) return;
    (void)hipStreamSynchronize(s->ctx->stream);
    delete s;
  });
}

int pmx_solver_update(pmx_solver* s, pmx_matrix* const* levels, const pmx_space* const* spaces,
                      int n_levels) {
  return guarded([&] {
    require(s && levels && n_levels >= 1, "invalid argument");
    const bool mg = s->p.preconditioner == PMX_PC_MULTIGRID;
    require(!mg || (spaces && n_levels >= 2), "multigrid needs >= 2 levels and their spaces");
    require(mg || n_levels == 1, "without multigrid pass exactly one matrix");
    require(!mg || (s->p.mg_smoother == PMX_SMOOTHER_JACOBI &&
                    s->p.mg_smoother_coarse != PMX_SMOOTHER_CELL_VANKA_JACOBI &&
                    s->p.mg_smoother_coarse != PMX_SMOOTHER_CELL_VANKA),
            "scalar problems use the jacobi smoother (cell Vanka is for saddle-point systems)");
    for (int l = 0; l < n_levels; ++l) require(levels[l] != nullptr, "NULL matrix");
    // the same hierarchy as the last update: values only
    bool same = s->updated && !s->saddle && (int)s->mats.size() == n_levels;
    for (int l = 0; same && l < n_levels; ++l)
      same = s->mats[l] == levels[l] && s->mat_uids[l] == levels[l]->uid &&
             s->spaces[l] == (spaces ? spaces[l] : nullptr);
    if (same) {
      refresh_values(s);
      return;
    }
    s->release();
    hipStream_t st = S(s);
    const pmx_matrix* fine = levels[n_levels - 1];
    s->mats.assign(levels, levels + n_levels);
    s->mat_uids.resize(n_levels);
    s->spaces.assign(n_levels, nullptr);
    for (int l = 0; l < n_levels; ++l) {
      s->mat_uids[l] = levels[l]->uid;
      if (spaces) s->spaces[l] = spaces[l];
    }
    for (int l = 0; l < n_levels; ++l)
      if (levels[l]) ensure_sell_values(levels[l], st);
    s->A = fine->A;
    s->n = fine->A.n;
    const int n = s->n;
    s->partial = dmalloc<double>(std::max(4096, pmx::fe_defect_partials(s->A)));
    s->d_scal = dmalloc<double>(64);
    s->r = dzeros<double>(n, st);
    s->w1 = dzeros<double>(n, st);
    s->w2 = dzeros<double>(n, st);
    s->w3 = dzeros<double>(n, st);
    s->dinv = dmalloc<double>(n);
    if (s->p.preconditioner == PMX_PC_JACOBI || s->p.iterative_solver_type == PMX_IT_JACOBI)
      check_zero_diag(s, s->A, s->dinv);
    if (s->p.iterative_solver_type == PMX_IT_FGMRES) {
      const int m = std::min(s->p.gmres_restart, std::max(1, s->p.max_n_iterations));
      for (int k = 0; k <= m; ++k) s->V.push_back(dzeros<double>(n, st));
      for (int k = 0; k < m; ++k) s->Z.push_back(dzeros<double>(n, st));
      s->d_H = dzeros<double>((size_t)(m + 1) * m + 16, st);
    }
    if (mg) {
      const bool mixed = s->p.precision == PMX_PRECISION_MIXED;
      s->lv.resize(n_levels);
      // solver-internal numbering of the levels above the coarsest (which
      // keeps the reference order for its dense inverse)
      std::vector<std::vector<int>> perm(n_levels), pinv(n_levels);
      pmx::SellPattern kpat;
      bool have_kpat = false;
      for (int l = 1; l < n_levels; ++l)
        if (s->p.renumber_levels && levels[l] && levels[l]->A.sell.total && spaces[l]) {
          perm[l] = pmx::lexicographic_order(spaces[l]->fs);
          pinv[l].resize(perm[l].size());
          for (size_t i = 0; i < perm[l].size(); ++i) pinv[l][perm[l][i]] = (int)i;
        }
      for (int l = 0; l < n_levels; ++l) {
        Level& L = s->lv[l];
        const pmx_matrix* M = levels[l];
        require(M != nullptr && spaces[l] != nullptr, "NULL level");
        require(M->space == spaces[l], "matrix/space mismatch");
        require(perm[l].empty() || (int)perm[l].size() == M->A.n, "space/matrix size mismatch");
        L.A = M->A;
        L.A.val_bf16 = nullptr;
        L.A.rscale = nullptr;
        L.A.sell.val_bf16 = nullptr;
        L.n = M->A.n;
        L.active = M->A.active;
        L.dinv = dmalloc<double>(L.n);
        check_zero_diag(s, L.A, L.dinv);
        if (mixed) {
          L.A.val_bf16 = dmalloc<uint16_t>(L.A.nnz);
          L.A.rscale = dmalloc<float>(L.n);
          pmx::csr_to_bf16(L.A, L.A.val_bf16, L.A.rscale, st);
        }
        if (!perm[l].empty()) {
          // renumbered SELL copy whose slots read the reference CSR values;
          // the entries of a row are sorted by their new column, so the
          // lanes of a slice (rows of one DOF type, neighbours in the new
          // order) gather neighbouring x entries at every step.  This changes
          // the summation order within rows (a different rounding of the
          // preconditioner, not of the FGMRES operator)
          pmx::CSRPattern pp;
          std::vector<int> opos;
          pmx::permute_csr(*M->pattern, nullptr, perm[l], pinv[l], pp, nullptr, &opos, true);
          pmx::SellPattern sp = pmx::build_sell(pp, kSellSigmaSolver, s->p.renumber_levels >= 2,
                                                (s->p.renumber_levels == 4 || s->p.renumber_levels == 6) ? 4 : 1,
                                                s->p.renumber_levels >= 5);
          for (auto& k : sp.src)
            if (k >= 0) k = opos[k];
          for (auto& k : sp.bd_src) k = opos[k];
          if (l == n_levels - 1 && !sp.off.empty() && s->p.iterative_solver_type == PMX_IT_FGMRES &&
              s->p.renumber_levels < 3) {
            // FGMRES applies the finest operator to vectors whose Dirichlet
            // part changes every time: without the folded layout it keeps a
            // 32-bit copy of the layout
            kpat = pmx::build_sell(pp, kSellSigmaSolver, false);
            for (auto& k : kpat.src)
              if (k >= 0) k = opos[k];
            for (auto& k : kpat.bd_src) k = opos[k];
            have_kpat = true;
          }
          pmx::DevCSR Ap = L.A;
          Ap.lpr = 0;
          Ap.sell = upload_sell(sp, st);
          L.perm = upload(perm[l].data(), perm[l].size(), st);
          double* dp = dmalloc<double>(L.n);
          pmx::vec_gather(L.n, L.perm, L.dinv, dp, st);
          dfree(L.dinv);
          L.dinv = dp;
          if (mixed) {
            float* rs = dmalloc<float>(L.n);
            pmx::vec_gather(L.n, L.perm, L.A.rscale, rs, st);
            Ap.rscale = rs;
            Ap.val_bf16 = nullptr;
            Ap.sell.val_bf16 = dmalloc<uint16_t>(Ap.sell.total);
            pmx::sell_fill_bf16(Ap, Ap.sell.val_bf16, st);
            HIP_CHECK(hipStreamSynchronize(st));
            dfree(L.A.val_bf16);
            dfree(L.A.rscale);
          } else {
            Ap.sell.val = dmalloc<double>(Ap.sell.total);
            pmx::sell_fill_f64(Ap, Ap.sell.val, st);
          }
          // the CSR arrays of the view are the reference-order ones: never
          // used for a renumbered level (lpr = 0 selects the SELL kernels)
          L.A = Ap;
          L.own_sell = true;
        } else if (mixed && L.A.sell.total) {
          L.A.sell.val_bf16 = dmalloc<uint16_t>(L.A.sell.total);
          pmx::sell_fill_bf16(L.A, L.A.sell.val_bf16, st);
        }
        L.partial = dmalloc<double>(std::max(4096, pmx::fe_defect_partials(L.A)));
        if (L.A.sell.n_bd > 0 && L.A.sell.bd_src) {
          // values of the Dirichlet-column entries (zero-start sweep; the
          // 16-bit layout keeps them only there)
          if (mixed) {
            L.bd16 = dmalloc<uint16_t>(L.A.sell.n_bd);
            pmx::sell_dirichlet_fill(L.A, nullptr, L.bd16, st);
            L.A.sell.bd_val_bf16 = L.bd16;
          } else {
            L.bd64 = dmalloc<double>(L.A.sell.n_bd);
            pmx::sell_dirichlet_fill(L.A, L.bd64, nullptr, st);
            L.A.sell.bd_val = L.bd64;
          }
        }
        if (s->p.renumber_levels >= 3 && L.A.sell.off && L.own_sell) {
          // folded 16-bit layout: Dirichlet-column sums per reset, no per-row tail
          if (mixed) { L.g32 = dzeros<float>(std::max(1, L.active), st); L.A.sell.gsum = L.g32; L.A.sell.gsum_bytes = 4; }
          else { L.g64 = dzeros<double>(std::max(1, L.active), st); L.A.sell.gsum = L.g64; L.A.sell.gsum_bytes = 8; }
        }
        // below the finest level the cycle resets the right-hand side to zero on
        // the Dirichlet rows and the solution everywhere, so the solution
        // vanishes on the Dirichlet rows for good: no Dirichlet-column sums
        if (l + 1 < n_levels && s->p.zero_start_shortcut) L.A.sell.dir_zero = 1;
        if (mixed) {
          L.rhs32 = dzeros<float>(L.n, st); L.sol32 = dzeros<float>(L.n, st);
          L.def32 = dzeros<float>(L.n, st); L.tmp32 = dzeros<float>(L.n, st);
          L.dinv32 = dmalloc<float>(L.n);
          pmx::vec_f64_to_f32(L.n, L.dinv, L.dinv32, st);
        } else {
          L.rhs = dzeros<double>(L.n, st); L.sol = dzeros<double>(L.n, st);
          L.def = dzeros<double>(L.n, st); L.tmp = dzeros<double>(L.n, st);
        }
      }
      for (int l = 0; l + 1 < n_levels; ++l) {
        Level& L = s->lv[l];
        const pmx_space* cs = spaces[l];
        const pmx_space* fsp = spaces[l + 1];
        require(cs->dom == fsp->dom && fsp->fs.level == cs->fs.level + 1,
                "multigrid spaces must be consecutive levels of one domain");
        pmx::CSRPattern P, R, Pp, Rp;
        std::vector<double> pv, rv, ppv, rpv;
        pmx::build_prolongation(cs->fs, fsp->fs, P, pv);
        pmx::transpose_csr(P, pv, R, rv);
        pmx::permute_csr(P, &pv, perm[l + 1], pinv[l], Pp, &ppv, nullptr, true);
        pmx::permute_csr(R, &rv, perm[l], pinv[l + 1], Rp, &rpv, nullptr, true);
        L.P = upload_csr(Pp, ppv, st);
        L.R = upload_csr(Rp, rpv, st);
        // SELL pays off where there are enough rows for the lane-per-row
        // layout to fill the GPU; long restriction rows on small levels run
        // faster with several lanes per row
        const int ts = s->p.transfer_sell;
        auto use_sell = [&](int rows) { return ts == 2 || ts == 4 || ((ts == 1 || ts == 3) && rows >= 131072); };
        // on renumbered levels most rows share a few patterns (transfer_sell 3, 4)
        const bool pat = (ts == 3 || ts == 4) && !perm[l + 1].empty();
        if (use_sell(s->lv[l + 1].active)) {
          if (pat) attach_pattern_transfer(L.P, Pp, ppv, s->lv[l + 1].active, L.active, st);
          else attach_transfer_sell(L.P, Pp, ppv, s->lv[l + 1].active, st);
        }
        if (use_sell(L.active)) {
          if (pat && !perm[l].empty()) attach_pattern_transfer(L.R, Rp, rpv, L.active, s->lv[l + 1].active, st);
          else attach_transfer_sell(L.R, Rp, rpv, L.active, st);
        }
      }
      if (s->p.mg_smoother_coarse == PMX_SMOOTHER_DIRECT) build_coarse_inverse(s, levels[0]);
      s->cc = cycle_control(s->p.mg_cycle, n_levels);
      Level& F = s->lv.back();
      if (s->p.iterative_solver_type == PMX_IT_FGMRES && F.perm && F.own_sell) {
        s->kperm = F.perm;
        s->K = F.A;
        s->K.val_bf16 = nullptr;
        s->K.rscale = nullptr;
        s->K.sell.val_bf16 = nullptr;
        s->K.sell.bd_val_bf16 = nullptr;
        s->K.sell.gsum = nullptr;
        s->K.sell.gsum_bytes = 0;
        s->K.sell.dir_zero = 0;
        if (F.A.sell.off && F.A.sell.gsum) {
          // the folded layout of the finest level with FP64 values and sums
          // of its own: the Dirichlet-column sums of each operand are formed
          // before each product (op_apply / op_defect)
          if (mixed) {
            s->kval = dmalloc<double>(s->K.sell.total);
            pmx::sell_fill_f64(s->K, s->kval, st);
            s->K.sell.val = s->kval;
            if (s->K.sell.n_bd > 0) {
              s->kbd = dmalloc<double>(s->K.sell.n_bd);
              pmx::sell_dirichlet_fill(s->K, s->kbd, nullptr, st);
            }
            s->K.sell.bd_val = s->kbd;
          } else {
            s->K.sell.bd_val = F.bd64;
          }
          s->kg64 = dzeros<double>(std::max(1, F.active), st);
          s->K.sell.gsum = s->kg64;
          s->K.sell.gsum_bytes = 8;
          s->kfold = true;
        } else if (have_kpat) {
          s->K.sell = upload_sell(kpat, st);
          s->kown = true;
          s->kval = dmalloc<double>(s->K.sell.total);
          pmx::sell_fill_f64(s->K, s->kval, st);
          s->K.sell.val = s->kval;
          if (s->K.sell.n_bd > 0) {
            s->kbd = dmalloc<double>(s->K.sell.n_bd);
            pmx::sell_dirichlet_fill(s->K, s->kbd, nullptr, st);
          }
          s->K.sell.bd_val = s->kbd;
        } else if (mixed) {
          // the multigrid keeps only bf16 values: an FP64 copy for FGMRES
          s->kval = dmalloc<double>(s->K.sell.total);
          pmx::sell_fill_f64(s->K, s->kval, st);
          s->K.sell.val = s->kval;
          if (s->K.sell.n_bd > 0) {
            s->kbd = dmalloc<double>(s->K.sell.n_bd);
            pmx::sell_dirichlet_fill(s->K, s->kbd, nullptr, st);
          }
          s->K.sell.bd_val = s->kbd;
        }
        require(s->K.sell.val != nullptr, "renumbered finest level without FP64 values");
        require(s->K.sell.n_bd == 0 || s->K.sell.bd_val != nullptr, "renumbered finest level without Dirichlet values");
        s->kb = dmalloc<double>(n);
        s->kx = dmalloc<double>(n);
        if (pmx::fe_defect_partials(s->K) > std::max(4096, pmx::fe_defect_partials(s->A))) {
          dfree(s->partial);
          s->partial = dmalloc<double>(pmx::fe_defect_partials(s->K));
        }
        if (mixed && s->p.refinement == 1 && F.A.sell.off && F.A.sell.gsum &&
            (F.A.sell.pack == 1 || F.A.sell.slice_pat)) {
          // iterative refinement: FP32 Krylov blocks and the inner operator
          const int m = std::min(s->p.gmres_restart, std::max(1, s->p.max_n_iterations));
          s->ld32 = ((size_t)n + 63) / 64 * 64;
          s->V32 = dzeros<float>((size_t)(m + 1) * s->ld32, st);
          s->Z32 = dzeros<float>((size_t)m * s->ld32, st);
          s->w32 = dzeros<float>(n, st);
          s->zero32 = dzeros<float>(std::max(1, F.active), st);
          if (m + 1 <= pmx::vec_multi_dot_max()) s->mdot = dmalloc<double>(pmx::vec_multi_partials());
          {
            void* hp = nullptr;
            if (hipHostMalloc(&hp, 64 * sizeof(double) + 64, hipHostMallocMapped | hipHostMallocCoherent) ==
                hipSuccess) {
              s->hpin = static_cast<double*>(hp);
              s->hflag = reinterpret_cast<unsigned int*>(s->hpin + 64);
              *s->hflag = 0;
              s->hseq = 0;
              void* dp = nullptr;
              HIP_CHECK(hipHostGetDevicePointer(&dp, hp, 0));
              s->hpin_dev = static_cast<double*>(dp);
              s->hflag_dev = reinterpret_cast<unsigned int*>(s->hpin_dev + 64);
            } else {
              (void)hipGetLastError();   // no pinned memory: copy + synchronise instead
            }
          }
          if (!s->p.refinement_inner_bf16) {
            s->kval32 = dmalloc<float>(F.A.sell.total);
            pmx::sell_fill_f32(F.A, s->kval32, st);
          }
          s->ir = true;
          for (auto*& v : s->V) dfree(v);
          for (auto*& z : s->Z) dfree(z);
          s->V.clear();
          s->Z.clear();
        }
      }
    }
    HIP_CHECK(hipStreamSynchronize(st));
    s->updated = true;
  });
}

int pmx_solver_update_nse(pmx_solver* s, pmx_nse* const* levels, int n_levels) {
  return guarded([&] {
    require(s && levels && n_levels >= 1, "invalid argument");
    const bool mg = s->p.preconditioner == PMX_PC_MULTIGRID;
    require(s->p.iterative_solver_type == PMX_IT_FGMRES, "saddle-point systems are solved with fgmres");
    require(s->p.preconditioner != PMX_PC_JACOBI, "use multigrid (cell Vanka) or no preconditioner");
    require(mg || n_levels == 1, "without multigrid pass exactly one system");
    require(!mg || n_levels >= 2, "multigrid needs >= 2 levels");
    require(!mg || s->p.mg_smoother == PMX_SMOOTHER_CELL_VANKA_JACOBI || s->p.mg_smoother == PMX_SMOOTHER_CELL_VANKA,
            "saddle-point multigrid uses the cell_vanka_jacobi or cell_vanka smoother");
    require(!mg || s->p.mg_smoother_coarse != PMX_SMOOTHER_JACOBI,
            "coarse smoother must be direct_solve, cell_vanka_jacobi or cell_vanka");
    s->release();
    hipStream_t st = S(s);
    for (int l = 0; l < n_levels; ++l) {
      require(levels[l] != nullptr, "NULL level");
      require(levels[l]->values_ready, "assemble the system before updating the solver");
      ensure_sell_values(levels[l]->A, st);
    }
    s->saddle = true;
    s->nse_fine = levels[n_levels - 1];
    s->n = s->nse_fine->n_comb;
    const int n = s->n;
    s->partial = dmalloc<double>(4096);
    s->d_scal = dmalloc<double>(64);
    s->r = dzeros<double>(n, st);
    const int m = std::min(s->p.gmres_restart, std::max(1, s->p.max_n_iterations));
    for (int k = 0; k <= m; ++k) s->V.push_back(dzeros<double>(n, st));
    for (int k = 0; k < m; ++k) s->Z.push_back(dzeros<double>(n, st));
    s->d_H = dzeros<double>((size_t)(m + 1) * m + 16, st);
    if (mg) {
      const bool f32 = s->p.precision == PMX_PRECISION_MIXED;
      int* sing = dmalloc<int>(1);
      HIP_CHECK(hipMemsetAsync(sing, 0, sizeof(int), st));
      s->nlv.resize(n_levels);
      for (int l = 0; l < n_levels; ++l) {
        NseLevel& L = s->nlv[l];
        pmx_nse* sys = levels[l];
        L.sys = sys;
        L.n = sys->n_comb;
        L.n_u = sys->n_u;
        L.n_ua = sys->n_u_active;
        L.n_p = sys->n_p;
        L.n_cells = sys->vs.n_cells;
        L.rhs = dzeros<double>(L.n, st);
        L.sol = dzeros<double>(L.n, st);
        L.def = dzeros<double>(L.n, st);
        L.res = dzeros<double>(L.n, st);
        L.partial = dmalloc<double>(std::max(4096, pmx::fe_defect_partials(nse_combined(sys))));
        const bool vanka = !(l == 0 && s->p.mg_smoother_coarse == PMX_SMOOTHER_DIRECT);
        if (vanka) {
          // local DOF maps: [u1 | u2 | u3 | p] of every cell, sorted per block
          // (VankaSmoother.C:44-105 collects the cell's DOFs block by block);
          // a 4-DOF pressure element leaves the last 4 entries kVankaNone
          const int NLc = pmx::kVankaLocal, npc = sys->np_cell;
          require(npc == 8 || npc == 4, "cell Vanka needs 8 or 4 pressure DOFs per cell");
          std::vector<int> map((size_t)L.n_cells * NLc, pmx::kVankaNone);
          for (int c = 0; c < L.n_cells; ++c) {
            int* mc = &map[(size_t)c * NLc];
            std::vector<int> v(sys->vs.global_numbers.begin() + (size_t)c * 27,
                               sys->vs.global_numbers.begin() + (size_t)(c + 1) * 27);
            std::vector<int> p(sys->ps.global_numbers.begin() + (size_t)c * npc,
                               sys->ps.global_numbers.begin() + (size_t)(c + 1) * npc);
            std::sort(v.begin(), v.end());
            std::sort(p.begin(), p.end());
            for (int k = 0; k < 3; ++k)
              for (int j = 0; j < 27; ++j) mc[k * 27 + j] = k * L.n_u + v[j];
            for (int j = 0; j < npc; ++j) mc[81 + j] = 3 * L.n_u + p[j];
          }
          std::vector<int> acc_ptr(L.n + 1, 0), acc_idx;
          for (int d : map)
            if (d != pmx::kVankaNone) acc_ptr[d + 1]++;
          for (int i = 0; i < L.n; ++i) acc_ptr[i + 1] += acc_ptr[i];
          acc_idx.resize(acc_ptr[L.n]);
          std::vector<int> fill(acc_ptr.begin(), acc_ptr.end() - 1);
          for (size_t k = 0; k < map.size(); ++k)
            if (map[k] != pmx::kVankaNone) acc_idx[fill[map[k]]++] = (int)k;
          L.map = upload(map.data(), map.size(), st);
          L.acc_ptr = upload(acc_ptr.data(), acc_ptr.size(), st);
          L.acc_idx = upload(acc_idx.data(), acc_idx.size(), st);
          const int kind = (l == 0) ? s->p.mg_smoother_coarse : s->p.mg_smoother;
          L.inv = s->p.vanka_inverse != 0 && kind == PMX_SMOOTHER_CELL_VANKA_JACOBI;
          const bool bf16 = f32 && L.inv && s->p.vanka_inverse == 2;
          // explicit inverses kept as they are applied: columns padded to 90
          // entries (two local rows per lane and load); the bf16 copy is
          // made from unpadded FP32 inverses
          L.padded = L.inv && !bf16;
          const size_t nlu = L.padded ? pmx::vanka_padded_inverse_elems(L.n_cells) : (size_t)L.n_cells * NLc * NLc;
          if (f32) L.lu32 = dmalloc<float>(nlu);
          else L.lu64 = dmalloc<double>(nlu);
          L.pmap = dmalloc<int>((size_t)L.n_cells * NLc);
          L.upd = dmalloc<double>((size_t)L.n_cells * NLc);
          pmx::vanka_factor(L.n_cells, L.map, nse_combined(sys), L.lu64, L.lu32, L.pmap, sing, st, L.inv, L.padded);
          if (bf16) {
            L.lu16 = dmalloc<uint16_t>(pmx::vanka_bf16_inverse_elems(L.n_cells));
            pmx::vanka_inverse_to_bf16(L.n_cells, L.lu32, L.lu16, st);
            HIP_CHECK(hipStreamSynchronize(st));
            dfree(L.lu32);
            L.lu32 = nullptr;
          }
          if (f32 && kind == PMX_SMOOTHER_CELL_VANKA_JACOBI) {
            // the smoother's defects read FP32 values (FP64 vectors and sums)
            const pmx::DevCSR C = nse_combined(sys);
            L.cval32 = dmalloc<float>(std::max(1, C.nnz));
            pmx::vec_f64_to_f32(C.nnz, C.val, L.cval32, st);
            if (L.lu16) {
              L.cval16 = dmalloc<uint16_t>(std::max(1, C.nnz));
              pmx::f32_to_bf16(C.nnz, L.cval32, L.cval16, st);
              HIP_CHECK(hipStreamSynchronize(st));
              dfree(L.cval32);
              L.cval32 = nullptr;
            }
          }
          if (kind == PMX_SMOOTHER_CELL_VANKA) {
            // colour = cell lattice index mod 3 per direction: cells of one
            // colour are >= 3 cells apart, so none reads a DOF another updates
            const pmx::Mesh& mesh = *sys->vs.mesh;
            const pmx::MeshLevel& ml = mesh.levels[sys->vs.level];
            const int nv = mesh.nv, dim = mesh.dim;
            auto centre = [&](int c, int d) {
              double t = 0;
              for (int v = 0; v < nv; ++v) t += ml.xyz[((size_t)c * nv + v) * dim + d];
              return t / nv;
            };
            double h[3];
            for (int d = 0; d < dim; ++d) {
              double lo = 1e300, hi = -1e300;
              for (int v = 0; v < nv; ++v) {
                const double x = ml.xyz[(size_t)v * dim + d];
                lo = std::min(lo, x);
                hi = std::max(hi, x);
              }
              h[d] = hi - lo;
            }
            std::vector<int> colour(L.n_cells);
            for (int c = 0; c < L.n_cells; ++c) {
              int col = 0, mul = 1;
              for (int d = 0; d < dim; ++d, mul *= 3)
                col += mul * ((int)std::floor(centre(c, d) / h[d]) % 3);
              colour[c] = col;
            }
            L.colour_ptr.assign(28, 0);
            for (int c = 0; c < L.n_cells; ++c) L.colour_ptr[colour[c] + 1]++;
            for (int k = 0; k < 27; ++k) L.colour_ptr[k + 1] += L.colour_ptr[k];
            std::vector<int> fillc(L.colour_ptr.begin(), L.colour_ptr.end() - 1), cells(L.n_cells);
            for (int c = 0; c < L.n_cells; ++c) cells[fillc[colour[c]]++] = c;
            L.colour_cells = upload(cells.data(), cells.size(), st);
            L.multiplicative = true;
          }
        } else {
          L.dinv_dense = dense_inverse_of(s, L.n, sys->Cp.row_ptr, sys->Cp.col, sys->d_Cval);
        }
        if (l + 1 < n_levels) {
          const pmx_nse* fsys = levels[l + 1];
          require(fsys->dom == sys->dom && fsys->level == sys->level + 1,
                  "multigrid systems must be consecutive levels of one domain");
          pmx::CSRPattern P, R;
          std::vector<double> pv, rv;
          pmx::build_prolongation(sys->vs, fsys->vs, P, pv);
          pmx::transpose_csr(P, pv, R, rv);
          L.Pu = upload_csr(P, pv, st);
          L.Ru = upload_csr(R, rv, st);
          pmx::build_prolongation(sys->ps, fsys->ps, P, pv);
          pmx::transpose_csr(P, pv, R, rv);
          L.Pp = upload_csr(P, pv, st);
          L.Rp = upload_csr(R, rv, st);
        }
      }
      HIP_CHECK(hipGetLastError());
      int hs = 0;
      HIP_CHECK(hipMemcpyAsync(&hs, sing, sizeof(int), hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      dfree(sing);
      if (hs) throw PmxError(PMX_E_ZERO_DIAG, "singular local Vanka system");
      s->cc = cycle_control(s->p.mg_cycle, n_levels);
    }
    HIP_CHECK(hipStreamSynchronize(st));
    s->updated = true;
  });
}

int pmx_solver_solve(pmx_solver* s, const double* b, double* x, pmx_solve_report* rep) {
  return guarded([&] {
    require(s && b && x && rep, "NULL argument");
    if (!s->updated) throw PmxError(PMX_E_STATE, "pmx_solver_update must be called before solve");
    s->history.clear();
    *rep = pmx_solve_report{};
    for (pmx_matrix* M : s->mats)
      if (M && !M->sell_valid)
        throw PmxError(PMX_E_STATE, "matrix values changed since pmx_solver_update; update the solver first");
    switch (s->p.iterative_solver_type) {
      case PMX_IT_FGMRES: solve_fgmres(s, b, x, rep); break;
      case PMX_IT_CG: solve_cg(s, b, x, rep); break;
      case PMX_IT_RICHARDSON: solve_richardson(s, b, x, rep, false); break;
      case PMX_IT_JACOBI: solve_richardson(s, b, x, rep, true); break;
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(S(s)));
  });
}

int pmx_solver_get_history(const pmx_solver* s, double* out, int max, int* n) {
  return guarded([&] {
    require(s && n, "NULL argument");
    const int k = std::min<int>(max, (int)s->history.size());
    if (out) for (int i = 0; i < k; ++i) out[i] = s->history[i];
    *n = (int)s->history.size();
  });
}

int pmx_solver_apply_preconditioner(pmx_solver* s, const double* r, double* z) {
  return guarded([&] {
    require(s && r && z, "NULL argument");
    if (!s->updated) throw PmxError(PMX_E_STATE, "pmx_solver_update must be called first");
    precondition(s, r, z);
    HIP_CHECK(hipStreamSynchronize(S(s)));
  });
}

}  // extern "C"

int pmx_solver_time_level_kernel(pmx_solver* s, int level, int kernel, int reps, double* avg_ms) {
  return guarded([&] {
    require(s && avg_ms && reps >= 1, "invalid argument");
    require(s->updated && !s->lv.empty(), "a scalar multigrid solver after pmx_solver_update is needed");
    require(level >= 0 && level < (int)s->lv.size(), "level out of range");
    const bool mixed = s->p.precision == PMX_PRECISION_MIXED;
    require(kernel == (mixed ? PMX_KERNEL_JACOBI_BF16 : PMX_KERNEL_JACOBI_FP64),
            "kernel must be the solver's smoothing sweep (JACOBI_BF16 mixed, JACOBI_FP64 fp64)");
    Level& L = s->lv[level];
    hipStream_t st = S(s);
    // the finest level's sweep as the refinement's cycles run it: with zero
    // Dirichlet-column sums
    const bool ir_sweep = s->ir && level + 1 == (int)s->lv.size() && s->p.zero_start_shortcut;
    struct Restore {
      Level& L;
      bool on;
      ~Restore() { if (on) { L.A.sell.dir_zero = 0; L.g_valid = false; } }
    } restore{L, ir_sweep};
    if (ir_sweep) {
      L.A.sell.dir_zero = 1;
      if (L.g32) pmx::vec_fill_f32(std::max(1, L.active), 0.f, L.g32, st);
    }
    auto launch = [&]() {
      ensure_dirichlet_sums(s, L, mixed);
      if (mixed) pmx::fe_jacobi_bf16(L.A, L.sol32, L.rhs32, L.dinv32, L.tmp32, s->p.jacobi_damp, st);
      else pmx::fe_jacobi(L.A, L.sol, L.rhs, L.dinv, L.tmp, s->p.jacobi_damp, st);
    };
    launch();
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) launch();
    HIP_CHECK(hipEventRecord(e1, st));
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipGetLastError());
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  });
}
