// Snapshots and proper orthogonal decomposition (C ABI pmx_snapshots_* and
// pmx_pod_* of include/pmx.h): SnapshotsCollector and POD of the reference
// (src/ReducedOrderModels/SnapshotsCollector.C, src/ReducedOrderModels/POD.C)
// with the snapshot matrix resident in HBM.
//
// Division of work: the O(n m) and O(n m^2) parts -- mean, W S, the Gram
// matrix S^T W S (FP64 matrix cores) and the basis S V Lambda^-1/2 -- run on
// the device; the m x m symmetric eigenproblem (LAPACK dsyev in the
// reference) runs on the host with cyclic Jacobi rotations.
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "hip/launch.hpp"
#include "internal.hpp"
#include "pmx.h"

using namespace pmx_internal;

namespace pmx_internal {

// Symmetric eigenproblem by the cyclic Jacobi method (Golub & Van Loan,
// Matrix Computations, 4th ed., algorithms 8.5.2 and 8.5.3): rotations
// J(p, q, theta) annihilate a_pq row by row until the off-diagonal Frobenius
// norm is below 1e-16 of the matrix norm.  A is row-major n x n.  Output:
// eigenvalues ascending (dsyev's order), V column-major with eigenvector k in
// column k, each normalised so that its largest-magnitude component (the
// first one on ties) is positive.
void sym_eig(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  auto a = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
  double frob = 0.0;
  for (double x : A) frob += x * x;
  frob = std::sqrt(frob);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = 0; q < n; ++q)
        if (p != q) off += a(p, q) * a(p, q);
    if (std::sqrt(off) <= 1e-16 * frob) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a(p, q);
        if (apq == 0.0) continue;
        // symmetric 2x2 Schur decomposition of [[a_pp, a_pq], [a_pq, a_qq]]
        const double tau = (a(q, q) - a(p, p)) / (2.0 * apq);
        const double t = tau >= 0.0 ? 1.0 / (tau + std::sqrt(1.0 + tau * tau))
                                    : -1.0 / (-tau + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < n; ++k) {   // A J
          const double kp = a(k, p), kq = a(k, q);
          a(k, p) = c * kp - s * kq;
          a(k, q) = s * kp + c * kq;
        }
        for (int k = 0; k < n; ++k) {   // J^T (A J)
          const double pk = a(p, k), qk = a(q, k);
          a(p, k) = c * pk - s * qk;
          a(q, k) = s * pk + c * qk;
        }
        for (int k = 0; k < n; ++k) {   // V J
          double& vp = V[(size_t)p * n + k];
          double& vq = V[(size_t)q * n + k];
          const double x = vp, y = vq;
          vp = c * x - s * y;
          vq = s * x + c * y;
        }
      }
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int i, int j) { return a(i, i) < a(j, j); });
  std::vector<double> Vs((size_t)n * n);
  w.resize(n);
  for (int k = 0; k < n; ++k) {
    w[k] = a(ord[k], ord[k]);
    const double* src = &V[(size_t)ord[k] * n];   // column ord[k] of V (stored as row of the transposed)
    int big = 0;
    for (int i = 1; i < n; ++i)
      if (std::fabs(src[i]) > std::fabs(src[big])) big = i;
    const double sg = src[big] < 0.0 ? -1.0 : 1.0;
    for (int i = 0; i < n; ++i) Vs[(size_t)k * n + i] = sg * src[i];
  }
  V.swap(Vs);
}

// POD::compute_basis after the autocorrelation matrix (POD.C:139-235):
// eigenvalues above the threshold, rank, coefficients C (m x rank,
// column-major) of the modes phi_k = S C(:, k) = S v_i / sqrt(lambda_i),
// i = m - 1 - k; eigenvalues returned largest first
PodEig pod_eigen(const double* G, int m, int pod_rank) {
  constexpr double threshold = 1.0e-10;   // POD::eigen_threshold (POD.C:66)
  PodEig r;
  // dsyev is called with uplo 'U' on the autocorrelation matrix (POD.C:157-
  // 178): only G(i, j), i <= j, is read.  S^T W S is not symmetric when W is
  // not (a Gramian with unit Dirichlet rows), so the lower triangle is taken
  // from the upper one.
  std::vector<double> A((size_t)m * m);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) A[(size_t)i * m + j] = G[(size_t)std::min(i, j) * m + std::max(i, j)];
  std::vector<double> w, V;
  sym_eig(m, std::move(A), w, V);
  r.valid = 0;
  for (int i = m - 1; i >= 0; --i)
    if (w[i] > threshold) r.valid++;
  r.rank = pod_rank <= 0 ? r.valid : pod_rank;
  r.clamped = r.rank > r.valid;
  if (r.clamped) r.rank = r.valid;
  r.C.assign((size_t)m * std::max(1, r.rank), 0.0);
  for (int k = 0; k < r.rank; ++k) {
    const int i = m - 1 - k;
    for (int j = 0; j < m; ++j) r.C[(size_t)k * m + j] = 1.0 / std::sqrt(w[i]) * V[(size_t)i * m + j];
  }
  r.eigs.assign(w.rbegin(), w.rend());
  return r;
}

// Gram matrices S_p^T WS_p of the row ranges [row0[p], row0[p] + rows[p]),
// row-major m x m each, into host G (n_jobs * m * m)
void pod_gram_jobs(hipStream_t st, const std::vector<int>& row0, const std::vector<int>& rows, int m, long long ld,
                   const double* S, const double* WS, std::vector<double>& G) {
  const int n_jobs = (int)row0.size();
  std::vector<int> c_row0, c_rows, job_c0(n_jobs + 1, 0);
  for (int p = 0; p < n_jobs; ++p) {
    for (int r = 0; r < rows[p]; r += pmx::kPodChunk) {
      c_row0.push_back(row0[p] + r);
      c_rows.push_back(std::min(pmx::kPodChunk, rows[p] - r));
    }
    job_c0[p + 1] = (int)c_row0.size();
  }
  const int nc = (int)c_row0.size();
  int* d_r0 = upload(c_row0.data(), c_row0.size(), st);
  int* d_nr = upload(c_rows.data(), c_rows.size(), st);
  int* d_c0 = upload(job_c0.data(), job_c0.size(), st);
  double* d_part = dmalloc<double>((size_t)std::max(1, nc) * m * m);
  double* d_G = dmalloc<double>((size_t)n_jobs * m * m);
  pmx::pod_gram(nc, d_r0, d_nr, n_jobs, d_c0, m, ld, S, WS, d_part, d_G, st);
  G.resize((size_t)n_jobs * m * m);
  HIP_CHECK(hipMemcpyAsync(G.data(), d_G, G.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  HIP_CHECK(hipGetLastError());
  dfree(d_r0); dfree(d_nr); dfree(d_c0); dfree(d_part); dfree(d_G);
}

// one value as std::ostream prints it after setprecision(12) in the default
// floating-point format, i.e. printf's %.12g
void put_value(std::string& out, double v) {
  char buf[40];
  const int k = std::snprintf(buf, sizeof buf, "%.12g", v);
  out.append(buf, (size_t)k);
}

void make_directory(const std::string& dir) {
  if (mkdir(dir.c_str(), 0777) != 0 && errno != EEXIST)
    throw PmxError(PMX_E_ARG, "cannot create directory " + dir);
}

}  // namespace pmx_internal

struct pmx_snapshots {
  CtxRef ctx;
  FILE* file = nullptr;
  int steps = 5, length = 0, max = 0, count = 0;
  double* d = nullptr;
  std::vector<double> host;
  ~pmx_snapshots() {
    if (file) std::fclose(file);
    dfree(d);
  }
};

struct pmx_pod {
  CtxRef ctx;
  pmx_pod_params p{};
  const pmx_matrix* gram = nullptr;
  int length = 0, m = 0, rank = 0, valid = 0;
  double* d_S0 = nullptr;     // snapshots as given (length x m)
  double* d_S = nullptr;      // working copy (fluctuations)
  double* d_WS = nullptr;
  double* d_mean = nullptr;
  double* d_basis = nullptr;  // length x rank
  std::vector<double> eigs;   // largest first
  bool computed = false;
  void drop_snapshots() {
    dfree(d_S0); dfree(d_S); dfree(d_WS); dfree(d_mean); dfree(d_basis);
    length = m = rank = valid = 0;
    eigs.clear();
    computed = false;
  }
  ~pmx_pod() { drop_snapshots(); }
};

extern "C" {

int pmx_snapshots_create(pmx_ctx* ctx, const char* path, int steps_per_snap, int length, int max_snaps,
                         pmx_snapshots** out) {
  return guarded([&] {
    require(ctx && out, "NULL argument");
    require(steps_per_snap >= 1 && length >= 1 && max_snaps >= 1, "invalid snapshot sizes");
    auto s = std::make_unique<pmx_snapshots>();
    s->ctx = ctx;
    s->steps = steps_per_snap;
    s->length = length;
    s->max = max_snaps;
    s->d = dmalloc<double>((size_t)length * max_snaps);
    if (path) {
      s->file = std::fopen(path, "w");
      if (!s->file) throw PmxError(PMX_E_ARG, std::string("cannot open snapshot file ") + path);
    }
    *out = s.release();
  });
}

int pmx_snapshots_destroy(pmx_snapshots* s) {
  return guarded([&] { delete s; });
}

int pmx_snapshots_write_data(pmx_snapshots* s, const double* dev_solution, int time_step, int* written) {
  return guarded([&] {
    require(s && dev_solution && time_step >= 0, "invalid argument");
    if (written) *written = 0;
    if (time_step % s->steps != 0) return;
    if (s->count >= s->max) throw PmxError(PMX_E_ARG, "snapshot collector is full");
    hipStream_t st = s->ctx->stream;
    HIP_CHECK(hipMemcpyAsync(s->d + (size_t)s->count * s->length, dev_solution, sizeof(double) * s->length,
                             hipMemcpyDeviceToDevice, st));
    if (s->file) {
      s->host.resize(s->length);
      HIP_CHECK(hipMemcpyAsync(s->host.data(), dev_solution, sizeof(double) * s->length, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      std::string line;
      line.reserve((size_t)s->length * 20);
      for (double v : s->host) {
        put_value(line, v);
        line.push_back(' ');
      }
      line.push_back('\n');
      if (std::fwrite(line.data(), 1, line.size(), s->file) != line.size() || std::fflush(s->file) != 0)
        throw PmxError(PMX_E_ARG, "writing the snapshot file failed");
    }
    s->count++;
    if (written) *written = 1;
  });
}

int pmx_snapshots_get(const pmx_snapshots* s, int* n_snaps, const double** dev_matrix) {
  return guarded([&] {
    require(s, "NULL argument");
    if (n_snaps) *n_snaps = s->count;
    if (dev_matrix) *dev_matrix = s->d;
  });
}

int pmx_pod_default_params(pmx_pod_params* p) {
  return guarded([&] {
    require(p, "NULL argument");
    p->pod_rank = 0;
    p->pod_fluctuations_only = 1;
    p->pod_inner_product = PMX_POD_EUCLIDEAN;
  });
}

int pmx_pod_create(pmx_ctx* ctx, const pmx_pod_params* p, pmx_pod** out) {
  return guarded([&] {
    require(ctx && p && out, "NULL argument");
    require(p->pod_inner_product == PMX_POD_EUCLIDEAN || p->pod_inner_product == PMX_POD_L2,
            "pod_inner_product must be euclidean or L2");
    auto pod = std::make_unique<pmx_pod>();
    pod->ctx = ctx;
    pod->p = *p;
    *out = pod.release();
  });
}

int pmx_pod_destroy(pmx_pod* pod) {
  return guarded([&] { delete pod; });
}

int pmx_pod_set_gramian(pmx_pod* pod, const pmx_matrix* gramian) {
  return guarded([&] {
    require(pod && gramian, "NULL argument");
    require(gramian->A.n == gramian->A.m, "the Gramian must be square");
    pod->gram = gramian;
    pod->computed = false;
  });
}

int pmx_pod_set_snapshots(pmx_pod* pod, const double* dev_snaps, int length, int n_snaps, int ld) {
  return guarded([&] {
    require(pod && dev_snaps && length >= 1 && n_snaps >= 1 && ld >= length, "invalid argument");
    pod->drop_snapshots();
    pod->length = length;
    pod->m = n_snaps;
    pod->d_S0 = dmalloc<double>((size_t)length * n_snaps);
    HIP_CHECK(hipMemcpy2DAsync(pod->d_S0, sizeof(double) * length, dev_snaps, sizeof(double) * ld,
                               sizeof(double) * length, n_snaps, hipMemcpyDeviceToDevice, pod->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(pod->ctx->stream));
  });
}

int pmx_pod_read_snapshots(pmx_pod* pod, const char* path) {
  return guarded([&] {
    require(pod && path, "NULL argument");
    std::ifstream in(path);
    if (!in) throw PmxError(PMX_E_ARG, std::string("File ") + path + " could not be opened");
    // POD::read_data: one snapshot per line, values separated by blanks
    std::vector<std::vector<double>> snaps;
    std::string line;
    while (std::getline(in, line)) {
      std::vector<double> v;
      std::istringstream iss(line);
      double x;
      while (iss >> x) v.push_back(x);
      snaps.push_back(std::move(v));
    }
    require(!snaps.empty() && !snaps[0].empty(), "empty snapshot file");
    const size_t n = snaps[0].size();
    for (const auto& v : snaps) require(v.size() == n, "snapshots of different lengths");
    std::vector<double> h(n * snaps.size());
    for (size_t j = 0; j < snaps.size(); ++j) std::copy(snaps[j].begin(), snaps[j].end(), h.begin() + j * n);
    pod->drop_snapshots();
    pod->length = (int)n;
    pod->m = (int)snaps.size();
    pod->d_S0 = upload(h.data(), h.size(), pod->ctx->stream);
  });
}

int pmx_pod_compute_basis(pmx_pod* pod) {
  return guarded([&] {
    require(pod, "NULL argument");
    if (!pod->d_S0) throw PmxError(PMX_E_STATE, "Snapshots are not available: read or set them first");
    const bool l2 = pod->p.pod_inner_product == PMX_POD_L2;
    if (l2) {
      if (!pod->gram) throw PmxError(PMX_E_STATE, "L2 POD needs a Gramian (pmx_pod_set_gramian)");
      require(pod->gram->A.n == pod->length, "Dimension of inner product matrix and length of snapshots must coincide");
    }
    hipStream_t st = pod->ctx->stream;
    const int n = pod->length, m = pod->m;
    dfree(pod->d_S); dfree(pod->d_WS); dfree(pod->d_mean); dfree(pod->d_basis);
    pod->d_S = dmalloc<double>((size_t)n * m);
    HIP_CHECK(hipMemcpyAsync(pod->d_S, pod->d_S0, sizeof(double) * n * m, hipMemcpyDeviceToDevice, st));
    pod->d_mean = dmalloc<double>(n);
    if (pod->p.pod_fluctuations_only) pmx::pod_mean(n, m, n, pod->d_S, pod->d_mean, true, st);
    else pmx::vec_fill(n, 0.0, pod->d_mean, st);
    const double* WS = pod->d_S;
    if (l2) {
      pod->d_WS = dmalloc<double>((size_t)n * m);
      const pmx::DevCSR& W = pod->gram->A;
      pmx::pod_spmm(n, m, n, W.row_ptr, W.col, W.val, pod->d_S, pod->d_WS, st);
      WS = pod->d_WS;
    }
    std::vector<double> G;
    pod_gram_jobs(st, {0}, {n}, m, n, pod->d_S, WS, G);
    PodEig e = pod_eigen(G.data(), m, pod->p.pod_rank);
    pod->valid = e.valid;
    pod->rank = e.rank;
    pod->eigs = e.eigs;
    pod->d_basis = dmalloc<double>((size_t)n * std::max(1, e.rank));
    if (e.rank > 0) {
      double* d_C = upload(e.C.data(), e.C.size(), st);
      const int zero = 0;
      int* d_row0 = upload(&zero, 1, st);
      int* d_rank = upload(&e.rank, 1, st);
      pmx::pod_basis(n, 1, d_row0, d_rank, m, e.rank, n, pod->d_S, d_C, pod->d_basis, st);
      HIP_CHECK(hipStreamSynchronize(st));
      dfree(d_C); dfree(d_row0); dfree(d_rank);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st));
    pod->computed = true;
  });
}

int pmx_pod_get_info(const pmx_pod* pod, int* length, int* n_snaps, int* rank, int* valid_eigs) {
  return guarded([&] {
    require(pod, "NULL argument");
    if (length) *length = pod->length;
    if (n_snaps) *n_snaps = pod->m;
    if (rank) *rank = pod->rank;
    if (valid_eigs) *valid_eigs = pod->valid;
  });
}

int pmx_pod_get_eigs(const pmx_pod* pod, double* host_eigs) {
  return guarded([&] {
    require(pod && host_eigs, "NULL argument");
    if (!pod->computed) throw PmxError(PMX_E_STATE, "POD basis not computed");
    std::copy(pod->eigs.begin(), pod->eigs.end(), host_eigs);
  });
}

int pmx_pod_get_basis(const pmx_pod* pod, double* dev_out) {
  return guarded([&] {
    require(pod && dev_out, "NULL argument");
    if (!pod->computed) throw PmxError(PMX_E_STATE, "POD basis not computed");
    if (pod->rank > 0)
      HIP_CHECK(hipMemcpyAsync(dev_out, pod->d_basis, sizeof(double) * pod->length * pod->rank,
                               hipMemcpyDeviceToDevice, pod->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(pod->ctx->stream));
  });
}

int pmx_pod_get_mean(const pmx_pod* pod, double* dev_out) {
  return guarded([&] {
    require(pod && dev_out, "NULL argument");
    if (!pod->computed) throw PmxError(PMX_E_STATE, "POD basis not computed");
    HIP_CHECK(hipMemcpyAsync(dev_out, pod->d_mean, sizeof(double) * pod->length, hipMemcpyDeviceToDevice,
                             pod->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(pod->ctx->stream));
  });
}

int pmx_pod_write(const pmx_pod* pod, const char* directory, const char* basename) {
  return guarded([&] {
    require(pod && directory && basename, "NULL argument");
    if (!pod->computed) throw PmxError(PMX_E_STATE, "Eigenvalues not found. POD not created?");
    hipStream_t st = pod->ctx->stream;
    const std::string dir(directory), base = dir + "/" + basename;
    make_directory(dir);
    const int n = pod->length;
    std::vector<double> basis((size_t)n * pod->rank), mean(n);
    if (pod->rank > 0)
      HIP_CHECK(hipMemcpyAsync(basis.data(), pod->d_basis, basis.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(mean.data(), pod->d_mean, mean.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    auto dump = [](const std::string& path, const std::string& text) {
      FILE* f = std::fopen(path.c_str(), "w");
      if (!f) throw PmxError(PMX_E_ARG, "File " + path + " could not be created.");
      const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
      std::fclose(f);
      if (!ok) throw PmxError(PMX_E_ARG, "writing " + path + " failed");
    };
    // POD::write_pod (POD.C:299-330)
    const bool fl = pod->p.pod_fluctuations_only != 0;
    std::string t = fl ? "POD basis: fluctuating field\n" : "POD basis: full field\n";
    for (int k = 0; k < pod->rank; ++k) {
      for (int i = 0; i < n; ++i) {
        put_value(t, basis[(size_t)k * n + i]);
        t.push_back(' ');
      }
      t.push_back('\n');
    }
    dump(base + "pod", t);
    // POD::write_averages (POD.C:424-449), only with fluctuations
    if (fl) {
      std::string a;
      for (int i = 0; i < n; ++i) {
        put_value(a, mean[i]);
        a.push_back(' ');
      }
      dump(base + "mean", a);
    }
    // POD::write_eigenvalues (POD.C:452-487): eigenvalues >= 1e-10 with the
    // cumulative share of the sum of all eigenvalues
    constexpr double threshold = 1.0e-10;
    int valid = 0;
    double sum = 0.0, cum = 0.0;
    for (double e : pod->eigs) {
      if (e >= threshold) valid++;
      sum += e;
    }
    std::string ev;
    const int k_end = std::min(pod->m, valid);
    for (int i = 0; i < k_end; ++i) {
      cum += pod->eigs[i] / sum;
      ev += std::to_string(i + 1) + " \t";
      put_value(ev, pod->eigs[i]);
      ev += " \t";
      put_value(ev, cum);
      ev.push_back('\n');
    }
    dump(base + "eigs", ev);
  });
}

}  // extern "C"
