// config_io.cpp -- ICARAT's configuration file (TOML) and result file (legacy
// VTK) for the voxel hexa8 path: icarat_hip_create_from_toml and
// icarat_hip_write_vtk of include/icarat_hip.h.
//
// Keys read (reference io/mesh_icarat.hpp:72-139, problem/elasticity.cpp:19-33,
// io/boundary_condition_icarat.hpp:72-138, 212-264):
//   [mesh] dimension = 3, type = "hexa8", division = [nx, ny, nz], length = [lx, ly, lz]
//   [material] young = [E1, E2], poisson, penalty, designs_0 (default 1)
//   [boundary_condition] dirichlet = [[flag6, val6, [[x0,x1],[y0,y1],[z0,z1]]], ...]
//                        neumann   = [[val3, [[x0,x1],[y0,y1],[z0,z1]]], ...]
//   [this_problem] filterR (optional)
// The VTK file follows the reference's writer (io/export_vtk.hpp:47-190 and the
// drivers' output(), opt_macro_oc.cpp:214-236): ASCII unstructured grid, the
// nodes as float points, hexa8 cells (type 12) in the reference's local node
// order, POINT_DATA "displacement" and CELL_DATA "density" / "mises".
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/icarat_hip.h"
#include "context.hpp"
#include "toml_lite.hpp"

namespace icarat_hip {

// addPlot (io/export_graph.hpp:56-70): "x y" with the stream's default
// formatting; the file is created afresh on the first line
void add_plot(const char *path, double x, double y, bool is_first) {
  std::ofstream f(path, is_first ? std::ios::out : std::ios::app);
  if (!f) throw Error(kInvalidArgument, std::string("cannot write ") + path);
  f << x << " " << y << std::endl;
  if (!f) throw Error(kInternal, std::string("write failed: ") + path);
}

// defined in capi.cpp: wraps a context into the C handle
icarat_hip_ctx *wrap_context(Context *c);
Context *unwrap_context(icarat_hip_ctx *h);
void set_context_error(icarat_hip_ctx *h, const std::string &m);
}  // namespace icarat_hip

namespace {

using icarat_hip::toml::Value;

double num(const Value &v, const char *what) {
  if (!v.is_number()) throw std::runtime_error(std::string(what) + " must be a number");
  return v.number();
}

const Value &req(const icarat_hip::toml::Document &d, const char *table, const char *key) {
  const Value *v = icarat_hip::toml::find(d, table, key);
  if (!v) throw std::runtime_error(std::string("missing [") + table + "] " + key);
  return *v;
}

// [[x0,x1],[y0,y1],[z0,z1]] -> {x0,x1,y0,y1,z0,z1}
void range6(const Value &r, double out[6]) {
  if (r.kind != Value::kArray || r.size() != 3) throw std::runtime_error("a range needs 3 pairs");
  for (int d = 0; d < 3; ++d) {
    const Value &p = r.at(d);
    if (p.kind != Value::kArray || p.size() != 2) throw std::runtime_error("a range pair needs 2 numbers");
    out[2 * d] = num(p.at(0), "range");
    out[2 * d + 1] = num(p.at(1), "range");
  }
}

void copy_err(char *err, int32_t len, const std::string &m) {
  if (!err || len <= 0) return;
  std::strncpy(err, m.c_str(), (size_t)len - 1);
  err[len - 1] = '\0';
}

}  // namespace

extern "C" {

int icarat_hip_create_from_toml(const char *path, icarat_hip_ctx **out, icarat_toml_info *info,
                                char *err, int32_t err_len) {
  if (!path || !out) {
    copy_err(err, err_len, "null argument");
    return ICARAT_HIP_INVALID_ARGUMENT;
  }
  *out = nullptr;
  icarat_hip::Context *c = nullptr;
  try {
    const auto doc = icarat_hip::toml::parse_file(path);
    const Value &dim = req(doc, "mesh", "dimension");
    if (num(dim, "dimension") != 3.0)
      throw std::runtime_error("only 3-D meshes are supported ([mesh] dimension = 3)");
    const Value &type = req(doc, "mesh", "type");
    if (type.kind != Value::kString || type.s != "hexa8")
      throw std::runtime_error("only hexa8 meshes are supported ([mesh] type = \"hexa8\")");
    const Value &div = req(doc, "mesh", "division");
    const Value &len = req(doc, "mesh", "length");
    if (div.size() != 3 || len.size() != 3)
      throw std::runtime_error("[mesh] division and length need 3 entries");
    icarat_hip::GridSpec g{};
    g.nx = (int)num(div.at(0), "division");
    g.ny = (int)num(div.at(1), "division");
    g.nz = (int)num(div.at(2), "division");
    g.lx = num(len.at(0), "length");
    g.ly = num(len.at(1), "length");
    g.lz = num(len.at(2), "length");
    if (g.nx < 1 || g.ny < 1 || g.nz < 1 || !(g.lx > 0) || !(g.ly > 0) || !(g.lz > 0))
      throw std::runtime_error("[mesh] division and length must be positive");
    const Value &young = req(doc, "material", "young");
    if (young.size() != 2) throw std::runtime_error("[material] young needs [E1, E2]");
    g.young1 = num(young.at(0), "young");
    g.young2 = num(young.at(1), "young");
    g.poisson = num(req(doc, "material", "poisson"), "poisson");
    g.penalty = num(req(doc, "material", "penalty"), "penalty");
    const Value *d0 = icarat_hip::toml::find(doc, "material", "designs_0");
    const Value *fr = icarat_hip::toml::find(doc, "this_problem", "filterR");
    c = new icarat_hip::Context(g);
    int nd = 0, nn = 0;
    if (const Value *dir = icarat_hip::toml::find(doc, "boundary_condition", "dirichlet")) {
      for (const Value &e : dir->a) {
        if (e.size() != 3 || e.at(0).size() != 6 || e.at(1).size() != 6)
          throw std::runtime_error("a dirichlet entry is [flag6, val6, range]");
        int flag[6];
        double val[6], rng[6];
        for (int k = 0; k < 6; ++k) {
          flag[k] = (int)num(e.at(0).at(k), "dirichlet flag");
          val[k] = num(e.at(1).at(k), "dirichlet value");
        }
        range6(e.at(2), rng);
        c->problem().add_dirichlet(flag, val, rng);
        ++nd;
      }
    }
    c->make_dof();
    if (const Value *neu = icarat_hip::toml::find(doc, "boundary_condition", "neumann")) {
      for (const Value &e : neu->a) {
        if (e.size() != 2 || e.at(0).size() != 3)
          throw std::runtime_error("a neumann entry is [val3, range]");
        double val[3], rng[6];
        for (int k = 0; k < 3; ++k) val[k] = num(e.at(0).at(k), "neumann value");
        range6(e.at(1), rng);
        c->problem().add_neumann(val, rng);
        ++nn;
      }
    }
    c->mark_fext_dirty();
    if (info) {
      info->nx = g.nx;
      info->ny = g.ny;
      info->nz = g.nz;
      info->lx = g.lx;
      info->ly = g.ly;
      info->lz = g.lz;
      info->designs_0 = d0 ? num(*d0, "designs_0") : 1.0;
      info->filter_radius = fr ? num(*fr, "filterR") : 0.0;
      info->n_dirichlet = nd;
      info->n_neumann = nn;
    }
  } catch (const icarat_hip::Error &e) {
    delete c;
    copy_err(err, err_len, e.what());
    return e.code;
  } catch (const std::exception &e) {
    delete c;
    copy_err(err, err_len, e.what());
    return ICARAT_HIP_INVALID_ARGUMENT;
  }
  *out = icarat_hip::wrap_context(c);
  copy_err(err, err_len, "");
  return ICARAT_HIP_OK;
}

int icarat_hip_add_plot(const char *filename, double x, double y, int32_t is_first) {
  if (!filename) return ICARAT_HIP_INVALID_ARGUMENT;
  try {
    icarat_hip::add_plot(filename, x, y, is_first == 1);
  } catch (const icarat_hip::Error &e) {
    return e.code;
  } catch (const std::exception &) {
    return ICARAT_HIP_INTERNAL;
  }
  return ICARAT_HIP_OK;
}

int icarat_hip_export_graph(const char *filename, const double *xx, const double *yy, int64_t n) {
  if (!filename || !yy || n < 0) return ICARAT_HIP_INVALID_ARGUMENT;
  // exportGraph(yy) / exportGraph(xx, yy) (io/export_graph.hpp:15-40)
  std::ofstream f(filename, std::ios::out);
  if (!f) return ICARAT_HIP_INVALID_ARGUMENT;
  for (int64_t i = 0; i < n; ++i) {
    if (xx)
      f << xx[i] << " " << yy[i] << std::endl;
    else
      f << (int)i << " " << yy[i] << std::endl;
  }
  return f ? ICARAT_HIP_OK : ICARAT_HIP_INTERNAL;
}

int icarat_hip_write_vtk(icarat_hip_ctx *h, const char *path, const double *density) {
  icarat_hip::Context *c = icarat_hip::unwrap_context(h);
  if (!c || !path) return ICARAT_HIP_INVALID_ARGUMENT;
  try {
    const auto &L = c->lattice();
    const auto &P = c->problem();
    // results of the last solve (device): nodal displacement and Mises stress
    std::vector<double> u, mises;
    if (c->solved_) {
      u.resize(3 * L.numnp);
      c->nodal_values_host(u.data());
      mises.resize(L.nelx);
      c->mises_host(mises.data(), nullptr);
    }
    std::ofstream f(path);
    if (!f) throw icarat_hip::Error(ICARAT_HIP_INVALID_ARGUMENT, std::string("cannot write ") + path);
    f << "# vtk DataFile Version 4.1\n";
    f << "vtk output\n";
    f << "ASCII\n";
    f << "DATASET UNSTRUCTURED_GRID\n";
    f << "\nPOINTS\t" << L.numnp << "\tfloat\n";
    for (long long n = 0; n < L.numnp; ++n) {
      for (int d = 0; d < 3; ++d) f << P.coord(n, d) << "\t";
      f << "\n";
    }
    f << "\nCELLS " << L.nelx << "\t" << 9 * L.nelx << "\n";
    for (long long e = 0; e < L.nelx; ++e) {
      const long long i = e % L.nx, j = (e / L.nx) % L.ny, k = e / ((long long)L.nx * L.ny);
      const long long b = i + (long long)L.NX * (j + (long long)L.NY * k);
      f << 8 << "\t";
      for (int a = 0; a < 8; ++a)
        f << b + icarat_hip::lnode_x(a) + (long long)L.NX * (icarat_hip::lnode_y(a) +
                                                               (long long)L.NY * icarat_hip::lnode_z(a))
          << "\t";
      f << "\n";
    }
    f << "\nCELL_TYPES\t" << L.nelx << "\n";
    for (long long e = 0; e < L.nelx; ++e) f << 12 << "\n";
    if (!u.empty()) {
      f << "\nPOINT_DATA\t" << L.numnp << "\n";
      f << "VECTORS displacement float\n";
      for (long long n = 0; n < L.numnp; ++n) {
        for (int d = 0; d < 3; ++d) f << u[3 * n + d] << "\t";
        f << "\n";
      }
    }
    bool header = true;
    auto cell_scalar = [&](const char *name, const double *v) {
      if (header) f << "\nCELL_DATA\t" << L.nelx << "\n";
      header = false;
      f << "SCALARS " << name << " float\n";
      f << "LOOKUP_TABLE default\n";
      for (long long e = 0; e < L.nelx; ++e) f << v[e] << "\n";
    };
    if (density) cell_scalar("density", density);
    if (!mises.empty()) cell_scalar("mises", mises.data());
    if (!f) throw icarat_hip::Error(ICARAT_HIP_INTERNAL, std::string("write failed: ") + path);
  } catch (const icarat_hip::Error &e) {
    icarat_hip::set_context_error(h, e.what());
    return e.code;
  } catch (const std::exception &e) {
    icarat_hip::set_context_error(h, e.what());
    return ICARAT_HIP_INTERNAL;
  }
  icarat_hip::set_context_error(h, "");
  return ICARAT_HIP_OK;
}

}  // extern "C"
