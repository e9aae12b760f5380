// CPU check of the SELL layouts of the multigrid levels (built by the host code
// of the library):
// 32-bit, 16-bit and packed 16-bit layouts describe the same rows; every
// index a kernel would form stays inside its array, and a packed slice's
// four-slot groups start on 8-byte (16-bit arrays) / 32-byte (FP64 values)
// boundaries.
#include <cstdio>
#include <functional>
#include <cstdlib>
#include <vector>
#include "space.hpp"
using namespace pmx;
static int fail = 0;
#define CHECK(c) do { if (!(c)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); if (++fail > 20) std::exit(1); } } while (0)
void check_level(const FESpace& fs) {
  CSRPattern p = build_structure(fs);
  std::vector<int> perm = lexicographic_order(fs), pinv(perm.size());
  for (size_t i = 0; i < perm.size(); ++i) pinv[perm[i]] = (int)i;
  CSRPattern pp; std::vector<int> opos;
  permute_csr(p, nullptr, perm, pinv, pp, nullptr, &opos, true);
  SellPattern a = build_sell(pp, 4096, false), b = build_sell(pp, 4096, true), c4 = build_sell(pp, 4096, true, 4);
  SellPattern pt = build_sell(pp, 4096, true, 1, true), pt4 = build_sell(pp, 4096, true, 4, true);
  const int na = pp.active_bound, n = pp.n_rows;
  std::printf("n %d active %d nnz %zu slots32 %zu slots16 %zu rel16 %d n_bd %zu\n", n, na, pp.col.size(), a.src.size(),
              b.src.size(), (int)!b.off.empty(), b.bd_src.size());
  CHECK(c4.off.empty() == b.off.empty());
  CHECK(pt.off.empty() == b.off.empty());
  for (const SellPattern& pt : {std::cref(pt), std::cref(pt4)})
  if (!pt.off.empty()) {
    // pattern layout: the lanes of a slice read the slice's pattern offsets;
    // an empty lane's reads (first row of the slice + offset) stay in the
    // active rows
    std::printf("  patterns %d slices %d (length-sorted %d)\n", pt.n_patterns, pt.n_slices, b.n_slices);
    CHECK((int)pt.slice_pat.size() == pt.n_slices);
    for (int sl = 0; sl < pt.n_slices; ++sl) {
      const int s0 = pt.slice_ptr[sl], width = (pt.slice_ptr[sl + 1] - s0) / 64;
      const int r0 = pt.row[(size_t)sl * 64];
      CHECK(r0 >= 0);
      CHECK(pt.slice_pat[sl] >= 0 && (size_t)(pt.slice_pat[sl] + width) <= pt.pat_off.size());
      for (int l = 0; l < 64; ++l) {
        const int r = pt.row[(size_t)sl * 64 + l];
        for (int j = 0; j < width; ++j) {
          const int o = pt.pat_off[pt.slice_pat[sl] + j];
          const size_t k = sell_pos((size_t)s0, j, l, pt.pack);
          if (r >= 0) CHECK(pt.off[k] == o);
          if (r < 0) CHECK(pt.src[k] == -1);
          const int base = r >= 0 ? r : r0;
          CHECK(base + o >= 0 && base + o < na);
        }
      }
    }
  }
  for (const SellPattern* s : {&a, &b, &c4, &pt, &pt4}) {
    const bool rel = !s->off.empty();
    if (s->pack == 4) {
      for (int sl = 0; sl <= s->n_slices; ++sl) CHECK(s->slice_ptr[sl] % 256 == 0);
    }
    const size_t total = s->src.size();
    CHECK((size_t)s->slice_ptr.back() == total);
    CHECK(rel ? s->col.empty() && s->off.size() == total : s->col.size() == total);
    // per row: entries reached through the slots (+ bd lists for rel16)
    std::vector<std::vector<std::pair<int, int>>> rows(na);
    for (int sl = 0; sl < s->n_slices; ++sl) {
      const int s0 = s->slice_ptr[sl], width = (s->slice_ptr[sl + 1] - s0) / 64;
      for (int l = 0; l < 64; ++l) {
        const int r = s->row[(size_t)sl * 64 + l];
        CHECK(r >= -1 && r < na);
        const int rb = r >= 0 ? r : 0;
        for (int j = 0; j < width; ++j) {
          const size_t k = sell_pos((size_t)s0, j, l, s->pack);
          if (s->pack == 4 && j % 4 == 0) CHECK(k % 4 == 0);
          CHECK(k < total);
          const int c = rel ? rb + s->off[k] : s->col[k];
          CHECK(c >= 0 && c < n);
          if (s->src[k] >= 0) { CHECK(r >= 0); rows[r].push_back({c, s->src[k]}); }
        }
      }
    }
    for (int r = 0; r < na; ++r) {
      if (rel)
        for (int k = s->bd_ptr[r]; k < s->bd_ptr[r + 1]; ++k) rows[r].push_back({s->bd_col[k], s->bd_src[k]});
      const int len = pp.row_ptr[r + 1] - pp.row_ptr[r];
      CHECK((int)rows[r].size() == len);
      for (int j = 0; j < len && j < (int)rows[r].size(); ++j) {
        CHECK(rows[r][j].first == pp.col[pp.row_ptr[r] + j]);
        CHECK(rows[r][j].second == pp.row_ptr[r] + j);
      }
    }
    for (size_t k = 0; k < s->bd_src.size(); ++k) { CHECK(s->bd_col[k] >= na && s->bd_col[k] < n); CHECK(s->bd_src[k] >= 0 && (size_t)s->bd_src[k] < pp.col.size()); }
    CHECK((int)s->bd_ptr.size() == na + 1 && (size_t)s->bd_ptr[na] == s->bd_src.size());
  }
  for (int q : opos) CHECK(q >= 0 && (size_t)q < p.col.size());
}
int main() {
  BoundaryConditions bc;
  bc.cond3d = [](double, double, double) { return (int)BC_DIRICHLET; };
  bc.cond2d = [](int comp, double) { return comp == 1 ? (int)BC_NEUMANN : (int)BC_DIRICHLET; };
  {
    Mesh m = build_mesh(GEO_UNIT_CUBE_HEXA, 4);
    for (int lev = 1; lev <= 4; ++lev) check_level(build_space(m, lev, 2, bc));
  }
  {
    Mesh m = build_mesh(GEO_UNIT_SQUARE_QUAD, 3);
    for (int lev = 1; lev <= 3; ++lev) { check_level(build_space(m, lev, 1, bc)); check_level(build_space(m, lev, 2, bc)); }
  }
  {
    Mesh m = build_mesh(GEO_UNIT_SQUARE_TRI, 3);
    for (int lev = 1; lev <= 3; ++lev) check_level(build_space(m, lev, 1, bc));
  }
  std::printf(fail ? "FAILED %d\n" : "all checks passed\n", fail);
  return fail != 0;
}
