// FE space (DOF numbering), CSR sparsity pattern and grid-transfer matrices.
//
// DOF numbering restates TFESpace3D::ConstructSpace (src/FE/FESpace3D.C:506-2080)
// and TFESpace2D::ConstructSpace (src/FE/FESpace2D.C:1284-2000).  The
// reference links coinciding local DOFs of neighbouring cells by a
// union-find whose root is always the lowest (cell, local dof) index
// (TFE3DMapper::MapDOF, src/FE/FE3DMapper.C:143-220) and marks each root with a
// class counter; the highest class wins (Dirichlet > Neumann > Robin > ... >
// inner).  Its renumbering pass (src/FE/FESpace3D.C:1830-1905) then walks the
// cells in collection order and local DOFs in order, and hands out the next
// number of a DOF's class at the DOF's first occurrence; the classes are laid
// out inner, Neumann, Robin, ..., hanging, Dirichlet (src/FE/FESpace3D.C:1923-1990).
// This file computes the same result directly: coinciding DOFs are
// identified by their nodal point, each DOF's class is the highest class of
// the boundary joints it lies on, and numbers are issued at first occurrence.
//
// The CSR pattern restates TStructure(FESpace) (src/Matrix/Structure.C:829-1277):
// active rows hold every DOF that shares a cell with the row DOF, ascending;
// Dirichlet rows (the last rows) hold only their diagonal.
#pragma once
#include <cstdint>
#include <functional>
#include <vector>

#include "fe.hpp"
#include "mesh.hpp"

namespace pmx {

enum BoundCond : int {
  BC_DIRICHLET = 0, BC_NEUMANN = 1, BC_ROBIN = 2, BC_SLIP = 3, BC_FREESURF = 4,
  BC_SLIP_FRICTION = 5, BC_INTERFACE = 6, BC_SUBDOMAIN_INTERFACE = 7,
  BC_SUBDOMAIN_HALOBOUND = 8, BC_DIRICHLET_WEAK = 9
};
constexpr int kNBoundTypes = 9;   // non-Dirichlet boundary classes

struct BoundaryConditions {
  std::function<int(int comp, double t)> cond2d;                 // BoundCondFunct2D
  std::function<int(double x, double y, double z)> cond3d;       // BoundCondFunct3D
};

struct FESpace {
  const Mesh* mesh = nullptr;
  int level = 0;
  Element el;
  int n_cells = 0;
  int ndof_cell = 0;
  std::vector<int> global_numbers;   // [n_cells][ndof_cell]
  int n_dof = 0;
  int n_inner = 0;
  int n_bound[kNBoundTypes] = {0};
  int n_dirichlet = 0;
  int active_bound = 0;
  std::vector<double> dof_xyz;       // [n_dof][3] nodal points
  std::vector<int> dof_comp;         // 2D: boundary component of the last Dirichlet joint (or -1)
  std::vector<double> dof_param;     // 2D: parameter on that component
  std::vector<uint8_t> dof_class;    // 0 inner, 1..9 boundary types, 10 Dirichlet
};

FESpace build_space(const Mesh& m, int level, int order, const BoundaryConditions& bc);
// Discontinuous P1 on the hexahedra of a level (TFESpace3D with D_P1_3D_H_A):
// every DOF is an inner DOF of one cell, numbered in cell order and local
// order (FESpace3D.C:1486-1492 counts them per cell; the renumbering pass of
// :1835-1880 keeps the order of first occurrence): global = 4 cell + local.
FESpace build_space_p1disc(const Mesh& m, int level);

struct CSRPattern {
  int n_rows = 0, n_cols = 0;
  int active_bound = 0;
  std::vector<int> row_ptr, col;
};

// Square FE matrix pattern of a space.
CSRPattern build_structure(const FESpace& s);

// Rectangular pattern TStructure(testspace, ansatzspace)
// (src/Matrix/Structure.C:1709-1924): every row of a test DOF holds the
// ansatz DOFs of all cells containing it, ascending (both spaces on the same
// cells); active_bound is the test space's.
CSRPattern build_rect_structure(const FESpace& test, const FESpace& ansatz);

// Interpolation matrix from the coarse space (mesh level l) to the fine space
// (level l+1): P[f][c] = phi_c(x_f).  Prolongation of GridTransfer::Prolongate
// (src/Multigrid/GridTransfer.C:665-941) equals P*u (the reference averages
// identical per-child values by multiplicity), and DefectRestriction
// (src/Multigrid/GridTransfer.C:943-1230) equals P^T*d.
// For the discontinuous P1 space the row of fine DOF k of a child is the
// child's nodal functional k applied to the parent's basis functions
// (TFEDatabase3D::GetProlongationMatrix3D evaluates the fine nodal
// functionals on the coarse basis); each fine DOF has one cell, so nothing is
// averaged.
void build_prolongation(const FESpace& coarse, const FESpace& fine, CSRPattern& P,
                        std::vector<double>& vals);
// Lexicographic (z, y, x) order of the DOFs of a space, active DOFs first and
// Dirichlet DOFs last (each group sorted on its own): perm[new] = old.  A
// solver-internal numbering whose neighbouring rows touch neighbouring
// columns, so the columns gathered by one wavefront share cache lines.
std::vector<int> lexicographic_order(const FESpace& s);
// Rows and columns renumbered: out row i = row rperm[i] of A (rperm empty:
// identity), column c -> cinv[c] (cinv empty: identity).  The entries of a
// row keep their order (and so the row's summation order), or, with
// sort_rows, are sorted by their new column.  opos (optional) receives the
// position in A of every entry of out.
void permute_csr(const CSRPattern& A, const std::vector<double>* av, const std::vector<int>& rperm,
                 const std::vector<int>& cinv, CSRPattern& out, std::vector<double>* ov,
                 std::vector<int>* opos, bool sort_rows = false);
// GridTransfer::RestrictFunction (src/Multigrid/GridTransfer.C:1218-1501) as
// a matrix: per coarse cell the local L2 projection G^-1 R_j of the fine
// function on each child j (G the coarse mass matrix, R_j the coarse/fine
// mixed mass matrix of child j, both on the parent's reference cell), summed
// over the cells around each coarse DOF and divided by their number.
// Quadrilaterals and hexahedra.
void build_function_restriction(const FESpace& coarse, const FESpace& fine, CSRPattern& RF,
                                std::vector<double>& vals);
void transpose_csr(const CSRPattern& A, const std::vector<double>& av, CSRPattern& T,
                   std::vector<double>& tv);

// Greedy colouring of the cells such that no two cells of one colour share a
// DOF (used by the device scatter).  Returns colour offsets into `order`.
void color_cells(const FESpace& s, std::vector<int>& order, std::vector<int>& color_ptr);

}  // namespace pmx

namespace pmx {

// Sliced-ELLPACK copy of the active rows of a CSR pattern (SELL-C-sigma with
// C = 64 = wavefront width).  Inside windows of `sigma` rows the rows are
// ordered by decreasing length, so the 64 rows of a slice have (nearly) equal
// length; entries are stored column-major per slice (entry j of the 64 rows is
// contiguous), padding entries point at the row itself with value 0.
struct SellPattern {
  int n_active = 0;
  int n_slices = 0;
  std::vector<int> slice_ptr;   // [n_slices+1] offsets in units of entries
  std::vector<int> row;         // [n_slices*64] original row of each slot (-1: empty)
  std::vector<int> col;         // [slice_ptr.back()]
  std::vector<int> src;         // [slice_ptr.back()] CSR position of the entry (-1: padding)
  // entries of the active rows in Dirichlet columns (col >= n_active), per
  // row in column order: their CSR position and their column
  std::vector<int> bd_ptr;      // [n_active+1]
  std::vector<int> bd_src;
  std::vector<int> bd_col;
  // rel16 layout: the slots hold only the active-column entries, as 16-bit
  // offsets col - row (col empty); the Dirichlet-column entries (the last ones
  // of every row) live only in the bd lists
  std::vector<int16_t> off;
  // packed slots (rel16 only, pack 1 or 4): slot j of lane l of a slice lies at
  // slice_ptr + (j / pack) * 64 * pack + l * pack + j % pack, so one lane
  // loads `pack` consecutive entries of its row with one instruction; the
  // slice width is a multiple of pack
  int pack = 1;
  // pattern layout (rel16 only): the active rows are grouped by their
  // sequence of active-column offsets (col - row), and the 64 rows of a slice
  // share one such pattern; slice_pat[slice] is the position of the slice's
  // pattern in pat_off (width = the slice's slot count / 64 offsets).  The
  // slots keep the 16-bit layout (off filled as well), so every 16-bit kernel
  // also runs on it.  Empty lanes (row -1) exist only in the last slice of a
  // pattern.  Slices are ordered by their first row.
  std::vector<int> slice_pat;
  std::vector<int> pat_off;
  int n_patterns = 0;
};
// position of slot j of lane l in a slice starting at s0
inline size_t sell_pos(size_t s0, int j, int l, int pack) {
  return s0 + (size_t)(j / pack) * 64 * pack + (size_t)l * pack + (size_t)(j % pack);
}
// rel16: build the 16-bit offset layout; if an offset does not fit, the
// result has an empty `off` and the 32-bit layout
// patterns: group rows by offset pattern (rel16 only); when the
// rows have more than max_patterns distinct patterns the length-sorted
// layout is built instead
SellPattern build_sell(const CSRPattern& p, int sigma, bool rel16 = false, int pack = 1, bool patterns = false,
                       int max_patterns = 65536);

}  // namespace pmx
