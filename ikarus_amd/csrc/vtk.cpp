// vtk.cpp — XML VTK UnstructuredGrid output of a handle's mesh, fields and stress results: the role of
// Ikarus::Vtk::Writer (ikarus/io/vtkwriter.hh:37-146) with addInterpolation (:125-140) and addResult / the
// ResultFunction (ikarus/io/resultfunction.hh:55-110), written without DUNE: ASCII .vtu, one piece.
//
// Cells: VTK_HEXAHEDRON (12) for hex8, VTK_QUAD (9) for plane-strain Q1. Dune numbers the reference-cube corners
// lexicographically (corner a at (a&1, a>>1&1, a>>2&1)); VTK walks each face ring, so corners 2 and 3 (and 6 and
// 7) swap. Stress results are evaluated on the device at the bound requirement (ik_stress_at): as cell data at
// the element centre, as point data at each element corner and averaged over the elements sharing the node.
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "ik_internal.hh"

using ik::Error;
using ik::require;

namespace {

const char* kResultNames[3] = {"PK2Stress", "KirchhoffStress", "CauchyStress"};

struct File {
  FILE* f = nullptr;
  explicit File(const char* path) : f(std::fopen(path, "w")) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

void writeArray(FILE* f, const char* type, const char* name, int comps, const double* v, int64_t n) {
  std::fprintf(f, "        <DataArray type=\"%s\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"ascii\">\n", type, name,
               comps);
  for (int64_t i = 0; i < n; ++i) {
    std::fprintf(f, "          ");
    for (int c = 0; c < comps; ++c) std::fprintf(f, c ? " %.17g" : "%.17g", v[i * comps + c]);
    std::fprintf(f, "\n");
  }
  std::fprintf(f, "        </DataArray>\n");
}

}  // namespace

extern "C" int ik_write_vtu(ik_handle* h, const char* path, const ik_vtk_field* fields, int32_t num_fields,
                            int32_t results, int32_t result_tag) {
  if (!h) return IK_EINVAL;
  try {
    require(path != nullptr && (num_fields == 0 || fields != nullptr), IK_EINVAL, "null argument");
    require(num_fields >= 0 && (results & ~7) == 0, IK_EINVAL, "bad field count or result mask");
    require(result_tag == IK_VTK_POINT_DATA || result_tag == IK_VTK_CELL_DATA, IK_EINVAL, "bad data tag");
    const bool q1  = h->plane2d;
    const int dim  = q1 ? 2 : 3;
    const int npe  = q1 ? 4 : 8;
    const int64_t nn = q1 ? h->nn2 : h->nn, ne = h->ne;
    for (int32_t i = 0; i < num_fields; ++i) {
      const ik_vtk_field& fd = fields[i];
      require(fd.name && fd.values && fd.components > 0 && (fd.cell == 0 || fd.cell == 1), IK_EINVAL,
              "bad VTK field");
      for (const char* c = fd.name; *c; ++c)
        require(*c != '"' && *c != '<' && *c != '&', IK_EINVAL, "VTK field names may not contain \" < &");
    }
    // stresses at the bound requirement: Voigt [11,22,33,23,13,12] (3D) or [11,22,12] (plane strain)
    const int nv = q1 ? 3 : 6;
    std::vector<std::vector<double>> res(3);
    if (results) {
      require(h->bound, IK_ENOTBOUND, "stress results need a bound requirement (call ik_set_requirement first)");
      std::vector<double> tmp((size_t)ne * nv);
      for (int k = 0; k < 3; ++k) {
        if (!((results >> k) & 1)) continue;
        if (result_tag == IK_VTK_CELL_DATA) {
          const double xc[3] = {0.5, 0.5, 0.5};
          const int rc       = ik_stress_at(h, k, xc, tmp.data(), IK_HOST);
          if (rc != IK_OK) throw Error{rc, ik_last_error(h)};
          res[k] = tmp;
        } else {
          std::vector<double> sum((size_t)nn * nv, 0.0), cnt(nn, 0.0);
          for (int a = 0; a < npe; ++a) {
            const double xa[3] = {(double)(a & 1), (double)((a >> 1) & 1), (double)((a >> 2) & 1)};
            const int rc       = ik_stress_at(h, k, xa, tmp.data(), IK_HOST);
            if (rc != IK_OK) throw Error{rc, ik_last_error(h)};
            for (int64_t e = 0; e < ne; ++e) {
              const int64_t v = q1 ? h->hconn2[e * 4 + a] : h->hconn[e * 8 + a];
              for (int c = 0; c < nv; ++c) sum[v * nv + c] += tmp[e * nv + c];
              cnt[v] += 1.0;
            }
          }
          for (int64_t v = 0; v < nn; ++v)
            for (int c = 0; c < nv; ++c) sum[v * nv + c] = cnt[v] > 0 ? sum[v * nv + c] / cnt[v] : 0.0;
          res[k].swap(sum);
        }
      }
    }
    File out(path);
    require(out.f != nullptr, IK_EINVAL, std::string("cannot open ") + path + " for writing");
    FILE* f = out.f;
    std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
                    "byte_order=\"LittleEndian\">\n  <UnstructuredGrid>\n");
    std::fprintf(f, "    <Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n", (long long)nn, (long long)ne);
    // points: reference coordinates, padded to 3 components
    std::vector<double> P((size_t)nn * 3, 0.0);
    for (int64_t v = 0; v < nn; ++v)
      for (int k = 0; k < dim; ++k) P[v * 3 + k] = h->hX[v * 3 + k];
    std::fprintf(f, "      <Points>\n");
    writeArray(f, "Float64", "Points", 3, P.data(), nn);
    std::fprintf(f, "      </Points>\n      <Cells>\n");
    std::fprintf(f, "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n");
    static const int hexOrder[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    for (int64_t e = 0; e < ne; ++e) {
      std::fprintf(f, "          ");
      for (int a = 0; a < npe; ++a) {
        const int da = hexOrder[a];
        const long long v = q1 ? h->hconn2[e * 4 + da] : h->hconn[e * 8 + da];
        std::fprintf(f, a ? " %lld" : "%lld", v);
      }
      std::fprintf(f, "\n");
    }
    std::fprintf(f, "        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
    for (int64_t e = 0; e < ne; ++e) std::fprintf(f, "          %lld\n", (long long)((e + 1) * npe));
    std::fprintf(f, "        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
    for (int64_t e = 0; e < ne; ++e) std::fprintf(f, "          %d\n", q1 ? 9 : 12);
    std::fprintf(f, "        </DataArray>\n      </Cells>\n");
    for (int tag = 0; tag < 2; ++tag) {
      const int64_t n = tag ? ne : nn;
      std::fprintf(f, tag ? "      <CellData>\n" : "      <PointData>\n");
      for (int32_t i = 0; i < num_fields; ++i)
        if (fields[i].cell == tag) {
          // a field with dim components is a vector: VTK wants 3
          const int c = fields[i].components;
          if (c == dim && dim == 2) {
            std::vector<double> v3((size_t)n * 3, 0.0);
            for (int64_t j = 0; j < n; ++j)
              for (int k = 0; k < 2; ++k) v3[j * 3 + k] = fields[i].values[j * 2 + k];
            writeArray(f, "Float64", fields[i].name, 3, v3.data(), n);
          } else {
            writeArray(f, "Float64", fields[i].name, c, fields[i].values, n);
          }
        }
      if ((tag == 1) == (result_tag == IK_VTK_CELL_DATA))
        for (int k = 0; k < 3; ++k)
          if (!res[k].empty()) writeArray(f, "Float64", kResultNames[k], nv, res[k].data(), n);
      std::fprintf(f, tag ? "      </CellData>\n" : "      </PointData>\n");
    }
    std::fprintf(f, "    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n");
    require(!std::ferror(f), IK_EINVAL, std::string("write error on ") + path);
    h->err.clear();
    return IK_OK;
  } catch (const Error& e) {
    h->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    h->err = e.what();
    return IK_EINVAL;
  }
}
