// VTK output of the model state: the reference's write_vtk
// (model_single_compartment.impl.hh:250-319, model_multi_compartment.impl.hh:
// 219-300, driven by model.writer.vtk.path in src/dune_copasi.cc:406-418).
// Per compartment, one XML UnstructuredGrid file per call named
// "<path>/<stem>-<compartment>-NNNNN.vtu" (stem = last component of path)
// with the compartment's vertices and cells (conforming: one point per
// vertex) and one scalar vertex field per species, plus the time series
// "<path>/<stem>-<compartment>.pvd" listing every file written to that path
// since the last call without append.  Arrays are Float64/Int32/UInt8,
// inline base64 with a UInt32 byte-count header (VTK's "binary" format).
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>

#include "error.hh"
#include "model.hh"

namespace dcopasi {

namespace {

std::string base64(const std::vector<unsigned char>& in) {
  static const char* tab = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  std::string out;
  out.reserve((in.size() + 2) / 3 * 4);
  size_t i = 0;
  for (; i + 2 < in.size(); i += 3) {
    const unsigned v = (in[i] << 16) | (in[i + 1] << 8) | in[i + 2];
    out += tab[(v >> 18) & 63];
    out += tab[(v >> 12) & 63];
    out += tab[(v >> 6) & 63];
    out += tab[v & 63];
  }
  if (i + 1 == in.size()) {
    const unsigned v = in[i] << 16;
    out += tab[(v >> 18) & 63];
    out += tab[(v >> 12) & 63];
    out += "==";
  } else if (i + 2 == in.size()) {
    const unsigned v = (in[i] << 16) | (in[i + 1] << 8);
    out += tab[(v >> 18) & 63];
    out += tab[(v >> 12) & 63];
    out += tab[(v >> 6) & 63];
    out += '=';
  }
  return out;
}

// one DataArray body: UInt32 byte count, then the values, base64 as a whole
template <typename T>
std::string encoded(const std::vector<T>& v) {
  const uint32_t n = static_cast<uint32_t>(v.size() * sizeof(T));
  std::vector<unsigned char> raw(sizeof n + n);
  std::memcpy(raw.data(), &n, sizeof n);
  if (n) std::memcpy(raw.data() + sizeof n, v.data(), n);
  return base64(raw);
}

template <typename T>
void data_array(std::ostream& o, const char* type, const std::string& name, int ncomp, const std::vector<T>& v) {
  o << "        <DataArray type=\"" << type << "\"";
  if (!name.empty()) o << " Name=\"" << name << "\"";
  o << " NumberOfComponents=\"" << ncomp << "\" format=\"binary\">\n          " << encoded(v)
    << "\n        </DataArray>\n";
}

}  // namespace

void Model::write_vtk(const std::string& path, bool append) {
  require_device();
  namespace fs = std::filesystem;
  std::error_code ec;
  fs::create_directories(path, ec);
  if (ec) throw IOError("cannot create output directory '" + path + "': " + ec.message());
  std::vector<double> u(L_.ndof);
  double t = 0.0;
  get_state(u.data(), &t);
  auto& steps = vtk_steps_[path];
  if (!append) steps.clear();
  steps.push_back(t);
  const size_t count = steps.size() - 1;
  std::string stem = fs::path(path).filename().string();
  if (stem.empty()) stem = fs::path(path).parent_path().filename().string();
  const Grid& g = L_.grid;
  const int d = g.dim;
  // VTK corner order of a box cell (bit k = +1 along axis k)
  static const int quad[4] = {0, 1, 3, 2}, hex[8] = {0, 1, 3, 2, 4, 5, 7, 6};
  for (size_t c = 0; c < L_.comps.size(); ++c) {
    const auto& C = L_.comps[c];
    const std::string name = stem + "-" + C.name;
    std::vector<double> pts;
    pts.reserve(static_cast<size_t>(C.nslot) * 3);
    for (long l = 0; l < C.nslot; ++l) {
      double x[3] = {0, 0, 0};
      g.vertex_pos(C.nodes[l], x);
      pts.insert(pts.end(), x, x + 3);
    }
    std::vector<int32_t> conn, offs;
    std::vector<uint8_t> types;
    for (size_t i = 0; i < C.cells.size(); ++i) {
      for (int p = 0; p < g.nclass; ++p) {
        if (!((C.cell_mask[i] >> p) & 1)) continue;
        const int* slot = &C.cell_slots[i * g.ncorner];
        auto local = [&](int corner) { return static_cast<int32_t>(slot[corner] - C.slot_off); };
        if (g.cube) {
          for (int a = 0; a < g.ncorner; ++a) conn.push_back(local(d == 3 ? hex[a] : d == 2 ? quad[a] : a));
          types.push_back(d == 3 ? 12 : d == 2 ? 9 : 3);
        } else {
          for (int a = 0; a < g.nvpe; ++a) conn.push_back(local(g.corner[p][a]));
          types.push_back(d == 3 ? 10 : d == 2 ? 5 : 3);
        }
        offs.push_back(static_cast<int32_t>(conn.size()));
      }
    }
    char num[16];
    std::snprintf(num, sizeof num, "%05zu", count);
    const std::string file = name + "-" + num + ".vtu";
    std::ofstream o(fs::path(path) / file);
    if (!o) throw IOError("cannot write '" + (fs::path(path) / file).string() + "'");
    o << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt32\">\n"
      << "  <UnstructuredGrid>\n"
      << "    <Piece NumberOfPoints=\"" << C.nslot << "\" NumberOfCells=\"" << types.size() << "\">\n"
      << "      <PointData Scalars=\"" << (C.b ? spec_.species[C.species[0]].name : std::string()) << "\">\n";
    for (int i = 0; i < C.b; ++i) {
      std::vector<double> f(static_cast<size_t>(C.nslot));
      for (long l = 0; l < C.nslot; ++l) f[l] = u[C.dof_off + l * C.b + i];
      data_array(o, "Float64", spec_.species[C.species[i]].name, 1, f);
    }
    o << "      </PointData>\n      <Points>\n";
    data_array(o, "Float64", "Coordinates", 3, pts);
    o << "      </Points>\n      <Cells>\n";
    data_array(o, "Int32", "connectivity", 1, conn);
    data_array(o, "Int32", "offsets", 1, offs);
    data_array(o, "UInt8", "types", 1, types);
    o << "      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
    if (!o) throw IOError("writing '" + file + "' failed");
    std::ofstream pvd(fs::path(path) / (name + ".pvd"));
    pvd.precision(17);
    pvd << "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
        << "  <Collection>\n";
    for (size_t k = 0; k < steps.size(); ++k) {
      std::snprintf(num, sizeof num, "%05zu", k);
      pvd << "    <DataSet timestep=\"" << steps[k] << "\" group=\"\" part=\"0\" file=\"" << name << "-" << num
          << ".vtu\"/>\n";
    }
    pvd << "  </Collection>\n</VTKFile>\n";
    if (!pvd) throw IOError("writing '" + name + ".pvd' failed");
  }
}

}  // namespace dcopasi
