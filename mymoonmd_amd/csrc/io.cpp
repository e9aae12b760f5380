// VTK output of scalar FE functions (C ABI pmx_write_vtk of include/pmx.h):
// DataWriter<d>::writeVtk with writeMesh and computeNodeValues
// (src/FE/DataWriter.C:232-340, 1698-1800) in the legacy ASCII format.
//
// The function values come from device memory (one copy per function); the
// value at a mesh vertex is the mean over the cells containing the vertex of
// the FE function at that vertex, which for the Lagrange spaces here is the
// value of the DOF located there.  The reference numbers the points by
// sorting the vertex objects by their address (TCollection::
// createElementLists, src/Geometry/Collection.C:120-190), an order fixed by
// the allocator; here the points are numbered in order of first appearance
// in the cell order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "internal.hpp"
#include "pmx.h"

using namespace pmx_internal;

namespace {

// n_local_vertices_to_type<d> (src/FE/DataWriter.C): VTK cell type codes
int vtk_cell_type(int dim, int nv) {
  if (dim == 2) return nv == 3 ? 5 : 9;    // VTK_TRIANGLE, VTK_QUAD
  return nv == 4 ? 10 : 12;                // VTK_TETRA, VTK_HEXAHEDRON
}

void put(std::string& out, const char* fmt, double v) {
  char buf[48];
  const int k = std::snprintf(buf, sizeof buf, fmt, v);
  out.append(buf, (size_t)k);
}

}  // namespace

extern "C" {

int pmx_write_vtk(const pmx_space* s, int n_functions, const double* const* dev_u, const char* const* names,
                  double time, const char* path) {
  return guarded([&] {
    require(s && path && n_functions >= 0 && (n_functions == 0 || (dev_u && names)), "NULL argument");
    const pmx::FESpace& fs = s->fs;
    const pmx::Mesh& mesh = *fs.mesh;
    const int dim = mesh.dim, nv = mesh.nv, nc = fs.n_cells;
    // points in order of first appearance; coordinates of coinciding vertices
    // are identical (every cell's vertices come from the same refinement
    // arithmetic), matched on a 1e-9 lattice
    std::map<std::tuple<long long, long long, long long>, int> id;
    std::vector<int> cell_pts((size_t)nc * nv);
    std::vector<double> pts;
    std::vector<int> pt_dof;   // DOF at the point (-1: none)
    for (int c = 0; c < nc; ++c)
      for (int v = 0; v < nv; ++v) {
        const double* x = mesh.vertex(fs.level, c, v);
        const auto key = std::make_tuple(std::llround(x[0] * 1e9), std::llround(x[1] * 1e9),
                                         dim == 3 ? std::llround(x[2] * 1e9) : 0LL);
        auto it = id.find(key);
        int p;
        if (it == id.end()) {
          p = (int)id.size();
          id.emplace(key, p);
          for (int k = 0; k < 3; ++k) pts.push_back(k < dim ? x[k] : 0.0);
          pt_dof.push_back(-1);
        } else {
          p = it->second;
        }
        cell_pts[(size_t)c * nv + v] = p;
        if (pt_dof[p] < 0) {
          // the cell's DOF whose nodal point is this vertex
          int best = -1;
          double bd = 1e300;
          for (int i = 0; i < fs.ndof_cell; ++i) {
            const int g = fs.global_numbers[(size_t)c * fs.ndof_cell + i];
            double d2 = 0;
            for (int k = 0; k < dim; ++k) d2 += (fs.dof_xyz[(size_t)g * 3 + k] - x[k]) * (fs.dof_xyz[(size_t)g * 3 + k] - x[k]);
            if (d2 < bd) { bd = d2; best = g; }
          }
          if (bd > 1e-20) throw PmxError(PMX_E_ARG, "the space has no DOF at a mesh vertex (VTK needs Lagrange spaces)");
          pt_dof[p] = best;
        }
      }
    const int np = (int)pt_dof.size();
    std::string t;
    t.reserve((size_t)np * 60 + (size_t)nc * 80 + (size_t)n_functions * np * 24);
    // writeMesh
    t += "# vtk DataFile Version 4.0\n";
    t += "file created by ParMooN Time < ";
    put(t, "%.12g", time);
    t += " >\nASCII\nDATASET UNSTRUCTURED_GRID\n";
    t += "POINTS " + std::to_string(np) + " double\n";
    for (int p = 0; p < np; ++p) {   // writeCoord: precision 8
      put(t, "%.8g", pts[(size_t)p * 3]);
      t += " ";
      put(t, "%.8g", pts[(size_t)p * 3 + 1]);
      t += " ";
      put(t, "%.8g", pts[(size_t)p * 3 + 2]);
      t += "\n";
    }
    t += "\n";
    t += "CELLS " + std::to_string(nc) + " " + std::to_string((long long)nc + (long long)nc * nv) + "\n";
    for (int c = 0; c < nc; ++c) {
      t += std::to_string(nv) + " ";
      for (int v = 0; v < nv; ++v) t += std::to_string(cell_pts[(size_t)c * nv + v]) + " ";
      t += "\n";
    }
    t += "\n";
    t += "CELL_TYPES " + std::to_string(nc) + "\n";
    const std::string ty = std::to_string(vtk_cell_type(dim, nv)) + " ";
    for (int c = 0; c < nc; ++c) t += ty;
    // writeVtk: point data, one SCALARS block per function (scientific, 12 digits)
    t += "\n\nPOINT_DATA " + std::to_string(np) + "\n";
    std::vector<double> u(fs.n_dof);
    hipStream_t st = s->ctx->stream;
    for (int f = 0; f < n_functions; ++f) {
      require(dev_u[f] && names[f], "NULL function");
      HIP_CHECK(hipMemcpyAsync(u.data(), dev_u[f], sizeof(double) * fs.n_dof, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      t += std::string("SCALARS ") + names[f] + " double\nLOOKUP_TABLE default\n";
      for (int p = 0; p < np; ++p) {
        put(t, "%.12e", u[pt_dof[p]]);
        t += "\n";
      }
      t += "\n\n";
    }
    t += "\n";
    FILE* fp = std::fopen(path, "w");
    if (!fp) throw PmxError(PMX_E_ARG, std::string("cannot open file for output. ") + path);
    const bool ok = std::fwrite(t.data(), 1, t.size(), fp) == t.size();
    std::fclose(fp);
    if (!ok) throw PmxError(PMX_E_ARG, std::string("writing ") + path + " failed");
  });
}

}  // extern "C"
