// Internal state of the Oseen / Navier-Stokes systems (nse.cpp,
// nse_solver.cpp).
#pragma once
#include <map>
#include <memory>
#include <vector>

#include "host/space.hpp"
#include "internal.hpp"

struct pmx_nse {
  CtxRef ctx;
  const pmx_domain* dom = nullptr;
  int level = 0;
  pmx::FESpace vs, ps;           // velocity Q2, pressure Q1 (continuous) or discontinuous P1
  int pressure_space = 1;        // PMX_NSE_PRESSURE_Q1 or PMX_NSE_PRESSURE_P1DISC
  int np_cell = 8;               // pressure DOFs per cell
  int example = 0;               // example whose boundary conditions built the velocity space
  bool enclosed = true;          // no Neumann/Robin velocity DOF: pressure row correction and projection
  int n_u = 0, n_u_active = 0, n_p = 0, n_comb = 0;
  pmx_matrix* A = nullptr;       // velocity block (+ SELL copy)
  pmx::CSRPattern Bp, BTp, Cp;   // B (p x u), B^T (u x p), combined matrix
  std::vector<int> BTsrc;        // B position of every B^T entry
  std::vector<uint8_t> Ckind;    // combined entry source: 0 zero, 1 one, 2 A, 3..5 B1..B3
  std::vector<int> Cpos;
  std::vector<int> color_ptr;
  int nq = 0;
  bool values_ready = false;
  // device
  double* d_geo = nullptr;
  int* d_vdofs = nullptr;
  int* d_pdofs = nullptr;
  int* d_cells = nullptr;
  double* d_vxyz = nullptr;
  double* d_ref = nullptr;
  int *d_Brp = nullptr, *d_Bcol = nullptr;
  double *d_B1 = nullptr, *d_B2 = nullptr, *d_B3 = nullptr;
  int *d_BTrp = nullptr, *d_BTcol = nullptr, *d_BTsrc = nullptr;
  double *d_BT1 = nullptr, *d_BT2 = nullptr, *d_BT3 = nullptr;
  int *d_Crp = nullptr, *d_Ccol = nullptr;
  double* d_Cval = nullptr;
  uint8_t* d_Ckind = nullptr;
  int* d_Cpos = nullptr;
  double* d_partial = nullptr;
  double* d_pref = nullptr;
  double* d_pgeo = nullptr;
  std::map<int, double*> err_ref;
  std::map<const pmx_nse*, pmx::DevCSR> restrict_from;   // function restriction from a finer level
  ~pmx_nse();
};

namespace pmx_internal {
pmx::DevCSR nse_b(const pmx_nse* s, int k);
pmx::DevCSR nse_bt(const pmx_nse* s, int k);
pmx::DevCSR nse_combined(const pmx_nse* s);
// y += a K x with the NSTYPE-1 block operator (identity on Dirichlet velocity rows)
void nse_apply_scaled_add(const pmx_nse* s, const double* x, double* y, double a, hipStream_t st);
// B^T and combined-matrix values from the assembled blocks
void nse_refresh(pmx_nse* s, hipStream_t st);
}  // namespace pmx_internal
