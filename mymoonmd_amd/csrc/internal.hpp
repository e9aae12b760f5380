// Internal definitions shared by the C-ABI translation units.
#pragma once
#include <atomic>
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pmx.h"
#include "hip/launch.hpp"
#include "host/fe.hpp"
#include "host/mesh.hpp"
#include "host/space.hpp"

namespace pmx_internal {

extern thread_local std::string g_last_error;

struct PmxError : std::runtime_error {
  int code;
  PmxError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIP_CHECK(expr)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw pmx_internal::PmxError(PMX_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return PMX_OK;
  } catch (const PmxError& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "out of host memory";
    return PMX_E_NOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return PMX_E_ARG;
  }
}

inline void require(bool cond, const char* msg) {
  if (!cond) throw PmxError(PMX_E_ARG, msg);
}

template <typename T>
T* dmalloc(size_t n) {
  if (n == 0) n = 1;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, n * sizeof(T));
  if (e != hipSuccess) throw PmxError(PMX_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return (T*)p;
}
template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}
template <typename T>
T* upload(const T* h, size_t n, hipStream_t st) {
  T* d = dmalloc<T>(n);
  if (n) HIP_CHECK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, st));
  HIP_CHECK(hipStreamSynchronize(st));
  return d;
}

}  // namespace pmx_internal

struct pmx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  double* partial = nullptr;   // reduction scratch (>= 1024 doubles)
  double* scal = nullptr;      // device scalars
  // one reference held by the caller (pmx_ctx_destroy drops it) and one by
  // every object created on the context, so that objects destroyed after
  // pmx_ctx_destroy (garbage collectors finalise in any order) still find
  // their stream; the last release frees the context
  std::atomic<int> refs{1};
};

// drop one reference; the last one synchronises and frees the context
void pmx_ctx_release(pmx_ctx* c);

// the context pointer an object keeps: takes a reference on assignment and
// drops it when the object is destroyed (after the object's own destructor,
// which may still use the stream)
struct CtxRef {
  pmx_ctx* p = nullptr;
  CtxRef() = default;
  CtxRef(const CtxRef&) = delete;
  CtxRef& operator=(const CtxRef&) = delete;
  CtxRef& operator=(pmx_ctx* c) {
    if (c) c->refs.fetch_add(1);
    if (p) pmx_ctx_release(p);
    p = c;
    return *this;
  }
  ~CtxRef() {
    if (p) pmx_ctx_release(p);
  }
  operator pmx_ctx*() const { return p; }
  pmx_ctx* operator->() const { return p; }
};

struct pmx_domain {
  pmx::Mesh mesh;
};

struct RefTable {
  int nq = 0;
  double* d_ref = nullptr;
  std::vector<double> h_ref;
  pmx::QuadRule rule;
  // Q2 hexahedra: reference matrices of the element GEMM (FP64, FP32 copy)
  double* d_R64 = nullptr;
  float* d_R32 = nullptr;
};

struct pmx_space {
  CtxRef ctx;
  const pmx_domain* dom = nullptr;
  pmx::FESpace fs;
  int builtin_bc = 0;
  pmx_bc2d_fn bc2d = nullptr;
  pmx_bc3d_fn bc3d = nullptr;
  void* user = nullptr;
  std::shared_ptr<pmx::CSRPattern> pattern;
  // device data
  int* d_dofs = nullptr;
  double* d_geo = nullptr;
  int* d_cells = nullptr;          // cells ordered by colour
  std::vector<int> color_ptr;
  double* d_dof_xyz = nullptr;     // nodal points of all DOFs [n_dof][3]
  double* d_dir_xyz = nullptr;     // the Dirichlet part of d_dof_xyz (not owned)
  std::map<int, RefTable> ref;     // keyed by hexa quad rule
  int* d_q2_pos = nullptr;         // CSR positions of the Q2 element entries (matrix-core assembly)
  // Q2 element GEMM + row gather: incidence slots of the active rows (state
  // 0: not built, 1: built, -1: the pattern does not fit the gather), element
  // rows (FP64 / FP32, 32 entries per slot) and load-vector entries per slot
  int q2_gather_state = 0;
  pmx::Q2Gather q2g;
  int* d_q2_inc_ptr = nullptr;
  int* d_q2_slot = nullptr;
  uint8_t* d_q2_off = nullptr;
  double* d_q2_E64 = nullptr;
  float* d_q2_E32 = nullptr;
  double* d_q2_fe = nullptr;
  ~pmx_space() {
    pmx_internal::dfree(d_dofs); pmx_internal::dfree(d_geo); pmx_internal::dfree(d_cells); pmx_internal::dfree(d_dof_xyz);
    pmx_internal::dfree(d_q2_pos);
    pmx_internal::dfree(d_q2_inc_ptr); pmx_internal::dfree(d_q2_slot); pmx_internal::dfree(d_q2_off);
    pmx_internal::dfree(d_q2_E64); pmx_internal::dfree(d_q2_E32); pmx_internal::dfree(d_q2_fe);
    for (auto& kv : ref) {
      pmx_internal::dfree(kv.second.d_ref);
      pmx_internal::dfree(kv.second.d_R64);
      pmx_internal::dfree(kv.second.d_R32);
    }
  }
};

// a process-wide serial number per object (identifies a matrix across
// solver updates even when an address is reused)
inline uint64_t pmx_next_uid() {
  static std::atomic<uint64_t> c{1};
  return c.fetch_add(1);
}

struct pmx_matrix {
  uint64_t uid = pmx_next_uid();
  CtxRef ctx;
  const pmx_space* space = nullptr;
  std::shared_ptr<pmx::CSRPattern> pattern;
  pmx::DevCSR A;
  bool sell_valid = false;   // SELL values in sync with the CSR values
  ~pmx_matrix() {
    using pmx_internal::dfree;
    dfree(A.row_ptr); dfree(A.col); dfree(A.val); dfree(A.val_bf16); dfree(A.rscale);
    dfree(A.sell.slice_ptr); dfree(A.sell.row); dfree(A.sell.col); dfree(A.sell.src);
    dfree(A.sell.val); dfree(A.sell.val_bf16);
    dfree(A.sell.bd_ptr); dfree(A.sell.bd_src); dfree(A.sell.bd_col); dfree(A.sell.off);
    dfree(A.sell.slice_pat); dfree(A.sell.pat_off);
  }
};

namespace pmx_internal {
// POD helpers (pod.cpp)
struct PodEig {
  std::vector<double> eigs;   // all eigenvalues, largest first
  int rank = 0, valid = 0;
  bool clamped = false;       // pod_rank was larger than the valid eigenvalues
  std::vector<double> C;      // m x rank coefficients of the modes (column-major)
};
void sym_eig(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& V);
PodEig pod_eigen(const double* G, int m, int pod_rank);
void pod_gram_jobs(hipStream_t st, const std::vector<int>& row0, const std::vector<int>& rows, int m, long long ld,
                   const double* S, const double* WS, std::vector<double>& G);
void put_value(std::string& out, double v);
void make_directory(const std::string& dir);
// device copy of a SELL layout (values not allocated); sizes checked
pmx::DevSELL upload_sell(const pmx::SellPattern& sp, hipStream_t st);
// upload the SELL layout of a matrix pattern (once) and mark its values stale
void attach_sell(pmx_matrix* M, hipStream_t st);
// refresh the SELL FP64 values from the CSR values when they are stale
void ensure_sell_values(pmx_matrix* M, hipStream_t st);
}  // namespace pmx_internal

