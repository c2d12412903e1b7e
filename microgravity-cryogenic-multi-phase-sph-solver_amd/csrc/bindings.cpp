// Python module `CryoSPH`: the reference's pybind surface (bindings.cpp:133-499)
// over the gfx950 C ABI of include/cryosph.h.  Arrays cross as NumPy copies,
// long calls drop the GIL, and C ABI status codes become Python exceptions
// (invalid argument -> ValueError, out of range -> IndexError, else RuntimeError).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "cryosph.h"

namespace py = pybind11;

namespace {

using Tuple3 = std::array<double, 3>;
using DArr = py::array_t<double, py::array::c_style | py::array::forcecast>;
using U8Arr = py::array_t<uint8_t, py::array::c_style | py::array::forcecast>;

void check(int rc) {
  if (rc >= 0) return;
  const std::string msg = cryo_last_error();
  if (rc == CRYO_ERR_INVALID_ARGUMENT) throw py::value_error(msg);
  if (rc == CRYO_ERR_OUT_OF_RANGE) throw py::index_error(msg);
  throw std::runtime_error(msg);
}

cryo_vec3 to_cv(const Tuple3& t) {
  return cryo_vec3{static_cast<float>(t[0]), static_cast<float>(t[1]), static_cast<float>(t[2])};
}
py::tuple to_tuple(cryo_vec3 v) { return py::make_tuple(double(v.x), double(v.y), double(v.z)); }

template <class T>
py::array_t<T> vec_array(const std::vector<T>& v) {
  py::array_t<T> out(static_cast<py::ssize_t>(v.size()));
  if (!v.empty()) std::memcpy(out.mutable_data(), v.data(), v.size() * sizeof(T));
  return out;
}
py::array_t<float> vec3_array(const std::vector<float>& v) {
  const py::ssize_t n = static_cast<py::ssize_t>(v.size() / 3);
  py::array_t<float> out({n, py::ssize_t{3}});
  if (n) std::memcpy(out.mutable_data(), v.data(), v.size() * sizeof(float));
  return out;
}

// -------------------------------------------------------------- fluids
struct Fluid {
  cryo_fluid f{};
  std::string name = "custom";
  Fluid() { check(cryo_fluid_preset("custom", &f)); }
  static Fluid preset(const char* n) {
    Fluid o;
    check(cryo_fluid_preset(n, &o.f));
    o.name = o.f.name;
    return o;
  }
  cryo_fluid c() const {
    cryo_fluid r = f;
    std::memset(r.name, 0, sizeof(r.name));
    std::strncpy(r.name, name.c_str(), sizeof(r.name) - 1);
    return r;
  }
};

// ------------------------------------------------------------ geometry
struct Geo {
  std::vector<cryo_geo_node> nodes;
  std::vector<std::vector<double>> mesh_v;
  std::vector<std::vector<int32_t>> mesh_t;
  std::string label;
  virtual ~Geo() = default;
  std::vector<cryo_mesh> meshes() const {
    std::vector<cryo_mesh> m;
    for (std::size_t k = 0; k < mesh_v.size(); ++k)
      m.push_back(cryo_mesh{static_cast<int32_t>(mesh_v[k].size() / 3), static_cast<int32_t>(mesh_t[k].size() / 3),
                            mesh_v[k].data(), mesh_t[k].data()});
    return m;
  }
  double sdf(double x, double y, double z) const {
    const auto m = meshes();
    double out = 0;
    check(cryo_geo_sdf(static_cast<int32_t>(nodes.size()), nodes.data(), static_cast<int32_t>(m.size()), m.data(), x, y,
                       z, &out));
    return out;
  }
  double volume() const {
    const auto m = meshes();
    double out = 0;
    check(cryo_geo_volume(static_cast<int32_t>(nodes.size()), nodes.data(), static_cast<int32_t>(m.size()), m.data(),
                          &out));
    return out;
  }
  py::tuple bounds() const {
    const auto m = meshes();
    float lo[3], hi[3];
    check(cryo_geo_bounds(static_cast<int32_t>(nodes.size()), nodes.data(), static_cast<int32_t>(m.size()), m.data(),
                          lo, hi));
    return py::make_tuple(py::make_tuple(double(lo[0]), double(lo[1]), double(lo[2])),
                          py::make_tuple(double(hi[0]), double(hi[1]), double(hi[2])));
  }
  // Appends another tree; returns the index of its root in this node list.
  int absorb(const Geo& o) {
    const int base = static_cast<int>(nodes.size());
    const int mbase = static_cast<int>(mesh_v.size());
    for (cryo_geo_node nd : o.nodes) {
      if (nd.a >= 0) nd.a += base;
      if (nd.b >= 0) nd.b += base;
      if (nd.mesh_id >= 0) nd.mesh_id += mbase;
      nodes.push_back(nd);
    }
    mesh_v.insert(mesh_v.end(), o.mesh_v.begin(), o.mesh_v.end());
    mesh_t.insert(mesh_t.end(), o.mesh_t.begin(), o.mesh_t.end());
    return static_cast<int>(nodes.size()) - 1;
  }
  void validate() const { (void)sdf(0, 0, 0); }
};
cryo_geo_node leaf(int type, double a, double b = 0, double c = 0) {
  return cryo_geo_node{type, -1, -1, -1, {a, b, c}};
}
struct Cylinder : Geo {
  Cylinder(double r, double h) {
    nodes.push_back(leaf(CRYO_GEO_CYLINDER, r, h));
    label = "cylinder";
    validate();
  }
};
struct Capsule : Geo {
  Capsule(double r, double l) {
    nodes.push_back(leaf(CRYO_GEO_CAPSULE, r, l));
    label = "capsule";
    validate();
  }
};
struct Sphere : Geo {
  explicit Sphere(double r) {
    nodes.push_back(leaf(CRYO_GEO_SPHERE, r));
    label = "sphere";
    validate();
  }
};
struct Box : Geo {
  Box(double lx, double ly, double lz) {
    nodes.push_back(leaf(CRYO_GEO_BOX, lx, ly, lz));
    label = "box";
    validate();
  }
};
struct Translated : Geo {
  Translated(std::shared_ptr<Geo> g, const Tuple3& o) {
    if (!g) throw py::value_error("Translated: null geometry");
    const int c = absorb(*g);
    cryo_geo_node nd{CRYO_GEO_TRANSLATED, c, -1, -1,
                     {double(float(o[0])), double(float(o[1])), double(float(o[2]))}};
    nodes.push_back(nd);
    label = "translated(" + g->label + ")";
  }
};
struct Difference : Geo {
  Difference(std::shared_ptr<Geo> a, std::shared_ptr<Geo> b) {
    if (!a || !b) throw py::value_error("Difference: null geometry");
    const int ia = absorb(*a), ib = absorb(*b);
    nodes.push_back(cryo_geo_node{CRYO_GEO_DIFFERENCE, ia, ib, -1, {0, 0, 0}});
    label = "difference(" + a->label + ", " + b->label + ")";
  }
};
struct Union : Geo {
  Union(std::shared_ptr<Geo> a, std::shared_ptr<Geo> b) {
    if (!a || !b) throw py::value_error("Union: null geometry");
    const int ia = absorb(*a), ib = absorb(*b);
    nodes.push_back(cryo_geo_node{CRYO_GEO_UNION, ia, ib, -1, {0, 0, 0}});
    label = "union(" + a->label + ", " + b->label + ")";
  }
};
// Wavefront OBJ (v / f records, fan-triangulated; geometry.cpp:146-176).
struct Mesh : Geo {
  Mesh(const std::string& path, double scale, const Tuple3& offset) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("load_obj: cannot open '" + path + "'");
    std::vector<double> v;
    std::vector<int32_t> t;
    std::string line;
    while (std::getline(in, line)) {
      std::size_t p = line.find_first_not_of(" \t");
      if (p == std::string::npos) continue;
      if (line.compare(p, 2, "v ") == 0 || line.compare(p, 2, "v\t") == 0) {
        double a, b, c;
        if (std::sscanf(line.c_str() + p + 1, "%lf %lf %lf", &a, &b, &c) != 3)
          throw std::runtime_error("load_obj: malformed vertex in '" + path + "'");
        v.push_back(a * scale + float(offset[0]));
        v.push_back(b * scale + float(offset[1]));
        v.push_back(c * scale + float(offset[2]));
      } else if (line.compare(p, 2, "f ") == 0 || line.compare(p, 2, "f\t") == 0) {
        std::vector<int> idx;
        std::size_t q = p + 1;
        const int nv = static_cast<int>(v.size() / 3);
        while (q < line.size()) {
          while (q < line.size() && (line[q] == ' ' || line[q] == '\t')) ++q;
          if (q >= line.size()) break;
          std::size_t e = line.find_first_of(" \t", q);
          const std::string tok = line.substr(q, e == std::string::npos ? std::string::npos : e - q);
          q = e == std::string::npos ? line.size() : e;
          const int k = std::stoi(tok.substr(0, tok.find('/')));
          const int r = k > 0 ? k - 1 : nv + k;
          if (r < 0 || r >= nv) throw std::runtime_error("load_obj: face index out of range in '" + path + "'");
          idx.push_back(r);
        }
        for (std::size_t s = 1; s + 1 < idx.size(); ++s) {
          t.push_back(idx[0]);
          t.push_back(idx[s]);
          t.push_back(idx[s + 1]);
        }
      }
    }
    if (t.empty()) throw std::runtime_error("load_obj: no faces found in '" + path + "'");
    mesh_v.push_back(std::move(v));
    mesh_t.push_back(std::move(t));
    nodes.push_back(cryo_geo_node{CRYO_GEO_MESH, -1, -1, 0, {0, 0, 0}});
    label = "mesh";
  }
  std::size_t num_triangles() const { return mesh_t[0].size() / 3; }
  double winding(double x, double y, double z) const {
    const auto m = meshes();
    double out = 0;
    check(cryo_mesh_winding_number(&m[0], x, y, z, &out));
    return out;
  }
};

// ----------------------------------------------------------- maneuver
struct Maneuver {
  std::vector<double> time;
  std::vector<float> acc, omega;
  Tuple3 origin{0, 0, 0};
  py::tuple evaluate(double t) const {
    float a[3], w[3], wd[3];
    check(cryo_maneuver_evaluate(static_cast<int32_t>(time.size()), time.data(), acc.data(),
                                 omega.empty() ? nullptr : omega.data(), t, a, w, wd));
    auto tp = [](const float* v) { return py::make_tuple(double(v[0]), double(v[1]), double(v[2])); };
    return py::make_tuple(tp(a), tp(w), tp(wd));
  }
};

// ----------------------------------------------------------------- solver
struct Solver {
  cryo_solver* s = nullptr;
  Fluid fluid;
  Solver(const cryo_config& c, const Fluid& f) : fluid(f) {
    const cryo_fluid cf = f.c();
    check(cryo_solver_create(&c, &cf, &s));
  }
  ~Solver() { cryo_solver_destroy(s); }
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;
  int n() const {
    int32_t k = 0;
    check(cryo_solver_num_particles(s, &k));
    return k;
  }
  std::vector<float> f3(int field) const {
    const int k = n();
    std::vector<float> out(3 * static_cast<std::size_t>(k));
    check(cryo_solver_download(s, field, out.data(), k));
    return out;
  }
  template <class T>
  std::vector<T> f1(int field) const {
    const int k = n();
    std::vector<T> out(static_cast<std::size_t>(k));
    check(cryo_solver_download(s, field, out.data(), k));
    return out;
  }
  double time() const {
    double t = 0;
    check(cryo_solver_time(s, &t));
    return t;
  }
  long long steps() const {
    int64_t k = 0;
    check(cryo_solver_step_count(s, &k));
    return k;
  }
  cryo_params params() const {
    cryo_params p{};
    check(cryo_solver_params(s, &p));
    return p;
  }
};

py::dict particles_dict(const Solver& S) {
  py::dict d;
  d["position"] = vec3_array(S.f3(CRYO_F_X));
  d["velocity"] = vec3_array(S.f3(CRYO_F_V));
  d["normal"] = vec3_array(S.f3(CRYO_F_NRM));
  d["mass"] = vec_array(S.f1<float>(CRYO_F_M));
  d["density"] = vec_array(S.f1<float>(CRYO_F_RHO));
  d["temperature"] = vec_array(S.f1<float>(CRYO_F_T));
  d["phase"] = vec_array(S.f1<uint8_t>(CRYO_F_PHASE));
  d["flags"] = vec_array(S.f1<uint8_t>(CRYO_F_FLAGS));
  d["id"] = vec_array(S.f1<uint32_t>(CRYO_F_ID));
  return d;
}

// Solver.write_vtk / write_json (bindings.cpp:491-497 of the reference): the
// current state downloaded, then the C ABI's file writers (io.cpp:52-142 formats).
void write_vtk(const Solver& S, const std::string& path) {
  const auto x = S.f3(CRYO_F_X), v = S.f3(CRYO_F_V);
  const auto T = S.f1<float>(CRYO_F_T), m = S.f1<float>(CRYO_F_M), rho = S.f1<float>(CRYO_F_RHO),
             p = S.f1<float>(CRYO_F_P);
  const auto ph = S.f1<uint8_t>(CRYO_F_PHASE);
  check(cryo_write_vtk(path.c_str(), static_cast<int32_t>(T.size()), x.data(), v.data(), T.data(), m.data(),
                       rho.data(), ph.data(), p.data()));
}

void write_json(const Solver& S, const std::string& path, bool include_walls) {
  const auto x = S.f3(CRYO_F_X), v = S.f3(CRYO_F_V);
  const auto T = S.f1<float>(CRYO_F_T), p = S.f1<float>(CRYO_F_P);
  const auto ph = S.f1<uint8_t>(CRYO_F_PHASE);
  const int32_t n = static_cast<int32_t>(T.size());
  std::vector<float> speed(T.size());
  cryo_frame_speed(n, v.data(), speed.data());
  check(cryo_write_json_frame(path.c_str(), S.time(), S.steps(), n, x.data(), T.data(), p.data(), speed.data(),
                              ph.data(), include_walls ? 1 : 0));
}

cryo_fill default_fill() {
  cryo_fill f;
  cryo_fill_default(&f);
  return f;
}

enum class Phase : int { Liquid = 0, Vapor = 1, Wall = 2 };
enum class KernelType : int { CubicSpline = 0, WendlandC2 = 1 };
enum class DensityMode : int { Summation = 0, Continuity = 1 };
enum class SurfaceTensionModel : int { None = 0, CSF = 1, Akinci = 2 };
enum class WallThermalBC : int { Adiabatic = 0, Isothermal = 1 };
enum class TankThermoMode : int { Vented = 0, Closed = 1 };
enum class FillMode : int { Settled = 0, CentralBubble = 1, Full = 2 };
enum class WettingModel : int { NormalRotation = 0, ContactLineForce = 1, Both = 2 };
enum class TelemetryMode : int { AsyncCopy = 0, ZeroCopy = 1 };

}  // namespace

PYBIND11_MODULE(CryoSPH, m) {
  m.doc() = "CryoSPH on AMD Instinct MI355X (gfx950): multiphase SPH for cryogenic propellant tanks";
  m.attr("__version__") = cryo_version();
  m.attr("cuda_backend") = false;
  m.attr("hip_backend") = true;
  m.attr("double_precision") = false;
  m.def("device_name", []() {
    char buf[256] = {0};
    check(cryo_device_name(buf, sizeof(buf)));
    return std::string(buf);
  });
  m.def("device_count", &cryo_device_count);

  py::enum_<Phase>(m, "Phase").value("Liquid", Phase::Liquid).value("Vapor", Phase::Vapor).value("Wall", Phase::Wall);
  py::enum_<KernelType>(m, "KernelType")
      .value("CubicSpline", KernelType::CubicSpline)
      .value("WendlandC2", KernelType::WendlandC2);
  py::enum_<DensityMode>(m, "DensityMode")
      .value("Summation", DensityMode::Summation)
      .value("Continuity", DensityMode::Continuity);
  py::enum_<SurfaceTensionModel>(m, "SurfaceTensionModel")
      .value("NoTension", SurfaceTensionModel::None)
      .value("CSF", SurfaceTensionModel::CSF)
      .value("Akinci", SurfaceTensionModel::Akinci);
  py::enum_<WallThermalBC>(m, "WallThermalBC")
      .value("Adiabatic", WallThermalBC::Adiabatic)
      .value("Isothermal", WallThermalBC::Isothermal);
  py::enum_<TankThermoMode>(m, "TankThermoMode")
      .value("Vented", TankThermoMode::Vented)
      .value("Closed", TankThermoMode::Closed);
  py::enum_<FillMode>(m, "FillMode")
      .value("Settled", FillMode::Settled)
      .value("CentralBubble", FillMode::CentralBubble)
      .value("Full", FillMode::Full);
  py::enum_<WettingModel>(m, "WettingModel")
      .value("NormalRotation", WettingModel::NormalRotation)
      .value("ContactLineForce", WettingModel::ContactLineForce)
      .value("Both", WettingModel::Both);
  py::enum_<TelemetryMode>(m, "TelemetryMode")
      .value("AsyncCopy", TelemetryMode::AsyncCopy)
      .value("ZeroCopy", TelemetryMode::ZeroCopy);

  // fluids
  py::class_<Fluid>(m, "FluidProperties", "Saturated liquid/vapour property set (SI units).")
      .def(py::init<>())
      .def_readwrite("name", &Fluid::name)
#define FL(field) .def_property(#field, [](const Fluid& f) { return double(f.f.field); }, \
                                [](Fluid& f, double v) { f.f.field = static_cast<float>(v); })
      FL(rho_liquid) FL(rho_vapor) FL(mu_liquid) FL(mu_vapor) FL(k_liquid) FL(k_vapor) FL(cp_liquid) FL(cp_vapor)
      FL(sigma) FL(h_fg) FL(T_sat) FL(p_sat) FL(R_gas)
#undef FL
      .def("saturation_temperature",
           [](const Fluid& f, double p) {
             const cryo_fluid c = f.c();
             return double(cryo_fluid_saturation_temperature(&c, static_cast<float>(p)));
           },
           py::arg("p"))
      .def("saturation_pressure",
           [](const Fluid& f, double T) {
             const cryo_fluid c = f.c();
             return double(cryo_fluid_saturation_pressure(&c, static_cast<float>(T)));
           },
           py::arg("T"))
      .def_static("liquid_hydrogen", []() { return Fluid::preset("LH2"); })
      .def_static("liquid_nitrogen", []() { return Fluid::preset("LN2"); })
      .def_static("liquid_oxygen", []() { return Fluid::preset("LOX"); })
      .def_static("liquid_methane", []() { return Fluid::preset("LCH4"); })
      .def_static("water", []() { return Fluid::preset("H2O"); })
      .def("__repr__", [](const Fluid& f) {
        return "<FluidProperties " + f.name + " rho_l=" + std::to_string(f.f.rho_liquid) +
               " rho_v=" + std::to_string(f.f.rho_vapor) + " T_sat=" + std::to_string(f.f.T_sat) + ">";
      });

  // geometry
  py::class_<Geo, std::shared_ptr<Geo>>(m, "Geometry", "Tank geometry given by a signed distance function.")
      .def("sdf", &Geo::sdf, py::arg("x"), py::arg("y"), py::arg("z"))
      .def("volume", &Geo::volume)
      .def_property_readonly("name", [](const Geo& g) { return g.label; })
      .def_property_readonly("bounds", &Geo::bounds);
  py::class_<Cylinder, Geo, std::shared_ptr<Cylinder>>(m, "CylinderTank")
      .def(py::init<double, double>(), py::arg("radius"), py::arg("height"));
  py::class_<Capsule, Geo, std::shared_ptr<Capsule>>(m, "CapsuleTank")
      .def(py::init<double, double>(), py::arg("radius"), py::arg("length"));
  py::class_<Sphere, Geo, std::shared_ptr<Sphere>>(m, "SphereTank").def(py::init<double>(), py::arg("radius"));
  py::class_<Box, Geo, std::shared_ptr<Box>>(m, "BoxTank")
      .def(py::init<double, double, double>(), py::arg("lx"), py::arg("ly"), py::arg("lz"));
  py::class_<Translated, Geo, std::shared_ptr<Translated>>(m, "Translated", "Geometry shifted by a constant offset.")
      .def(py::init<std::shared_ptr<Geo>, const Tuple3&>(), py::arg("geometry"), py::arg("offset"));
  py::class_<Difference, Geo, std::shared_ptr<Difference>>(m, "Difference",
                                                           "Fluid domain of `a` minus the solid `b`.")
      .def(py::init<std::shared_ptr<Geo>, std::shared_ptr<Geo>>(), py::arg("a"), py::arg("b"));
  py::class_<Union, Geo, std::shared_ptr<Union>>(m, "Union", "Fluid domain of `a` united with `b`.")
      .def(py::init<std::shared_ptr<Geo>, std::shared_ptr<Geo>>(), py::arg("a"), py::arg("b"));
  py::class_<Mesh, Geo, std::shared_ptr<Mesh>>(m, "MeshTank")
      .def(py::init<const std::string&, double, const Tuple3&>(), py::arg("obj_path"), py::arg("scale") = 1.0,
           py::arg("offset") = Tuple3{0, 0, 0})
      .def_property_readonly("num_triangles", &Mesh::num_triangles)
      .def("winding_number", &Mesh::winding, py::arg("x"), py::arg("y"), py::arg("z"));

  py::class_<cryo_fill>(m, "FillSpec", "How the generator fills a tank with liquid and vapour.")
      .def(py::init([]() {
        cryo_fill f;
        cryo_fill_default(&f);
        return f;
      }))
      .def_property("mode", [](const cryo_fill& f) { return static_cast<FillMode>(f.mode); },
                    [](cryo_fill& f, FillMode v) { f.mode = static_cast<int>(v); })
      .def_readwrite("fill_fraction", &cryo_fill::fill_fraction)
      .def_property("settle_direction", [](const cryo_fill& f) { return to_tuple(f.settle_direction); },
                    [](cryo_fill& f, const Tuple3& t) { f.settle_direction = to_cv(t); })
      .def_property("bubble_center", [](const cryo_fill& f) { return to_tuple(f.bubble_center); },
                    [](cryo_fill& f, const Tuple3& t) { f.bubble_center = to_cv(t); })
      .def_property("include_vapor", [](const cryo_fill& f) { return f.include_vapor != 0; },
                    [](cryo_fill& f, bool v) { f.include_vapor = v; })
      .def_readwrite("wall_layers", &cryo_fill::wall_layers)
      .def_property("T_liquid", [](const cryo_fill& f) { return double(f.T_liquid); },
                    [](cryo_fill& f, double v) { f.T_liquid = static_cast<float>(v); })
      .def_property("T_vapor", [](const cryo_fill& f) { return double(f.T_vapor); },
                    [](cryo_fill& f, double v) { f.T_vapor = static_cast<float>(v); })
      .def_property("T_wall", [](const cryo_fill& f) { return double(f.T_wall); },
                    [](cryo_fill& f, double v) { f.T_wall = static_cast<float>(v); })
      .def_readwrite("heater_zmin", &cryo_fill::heater_zmin)
      .def_readwrite("heater_zmax", &cryo_fill::heater_zmax);

  // config
  py::class_<cryo_config> cfg(m, "SolverConfig", "Solver parameters (SolverConfig of the reference).");
  cfg.def(py::init([]() {
       cryo_config c;
       cryo_config_default(&c);
       return c;
     }))
      .def_property("gravity", [](const cryo_config& c) { return to_tuple(c.gravity); },
                    [](cryo_config& c, const Tuple3& t) { c.gravity = to_cv(t); });
#define RWF(field) cfg.def_property(#field, [](const cryo_config& c) { return double(c.field); }, \
                                    [](cryo_config& c, double v) { c.field = static_cast<float>(v); })
#define RWB(field) cfg.def_property(#field, [](const cryo_config& c) { return c.field != 0; }, \
                                    [](cryo_config& c, bool v) { c.field = v ? 1 : 0; })
#define RWE(field, E) cfg.def_property(#field, [](const cryo_config& c) { return static_cast<E>(c.field); }, \
                                       [](cryo_config& c, E v) { c.field = static_cast<int32_t>(v); })
  RWF(dx); RWF(h_factor); RWE(kernel, KernelType); RWE(density_mode, DensityMode); RWB(normalize_initial_masses);
  RWF(velocity_scale); RWF(sound_speed); RWF(tait_gamma); RWF(background_pressure); RWF(vapor_stiffness);
  RWB(clamp_negative_pressure); RWB(transport_velocity); RWF(transport_background_factor); RWB(transport_stress);
  RWB(transport_interface_only); RWB(particle_shifting); RWF(shifting_strength); RWF(artificial_viscosity);
  RWF(artificial_viscosity_vapor); RWF(delta_sph); RWB(no_slip); RWB(sps_turbulence); RWF(smagorinsky_cs);
  RWE(surface_tension, SurfaceTensionModel); RWF(contact_angle_deg); RWF(contact_angle_wall_weight);
  RWB(wall_colour); RWE(wetting_model, WettingModel); RWB(csf_smooth_normals); RWB(csf_balanced);
  RWF(akinci_gamma); RWF(akinci_beta); RWB(thermal); RWB(phase_change); RWE(wall_thermal_bc, WallThermalBC);
  RWF(heater_power); RWF(conductivity_scale); RWF(lee_evaporation); RWF(lee_condensation);
  RWE(thermo_mode, TankThermoMode); RWF(vent_relaxation); RWF(merge_ratio); RWF(split_ratio);
  RWF(max_transfer_fraction); RWF(phase_change_density_limit); RWF(vapor_mass_rebalance); RWF(cfl);
  RWF(viscous_number); RWF(force_number); RWF(capillary_number); RWF(thermal_number); RWF(dt_max); RWF(dt_min);
  RWF(fixed_dt);
  RWB(compensated_sums);
#undef RWF
#undef RWB
#undef RWE

  // diagnostics
  py::class_<cryo_diagnostics> diag(m, "Diagnostics", "Global state summary returned by Solver.diagnostics().");
#define RO(field) diag.def_readonly(#field, &cryo_diagnostics::field)
  RO(time); RO(step); RO(dt); RO(n_total); RO(n_liquid); RO(n_vapor); RO(n_wall); RO(mass_liquid); RO(mass_vapor);
  RO(mass_total); RO(kinetic_energy); RO(enthalpy); RO(heat_input); RO(evaporated_mass); RO(condensed_mass);
  RO(vented_mass); RO(vented_enthalpy); RO(T_sat); RO(p_ullage); RO(max_speed); RO(T_liquid_mean); RO(T_liquid_max);
  RO(T_vapor_mean); RO(splits); RO(merges);
#undef RO
#define ROV(field) diag.def_property_readonly(#field, [](const cryo_diagnostics& d) { return to_tuple(d.field); })
  ROV(liquid_com); ROV(slosh_force); ROV(slosh_moment); ROV(frame_acceleration); ROV(frame_angular_velocity);
  ROV(liquid_momentum);
#undef ROV
  diag.def("as_dict", [](const cryo_diagnostics& d) {
    py::dict r;
    py::object self = py::cast(d);
    for (const char* k :
         {"time", "step", "dt", "n_total", "n_liquid", "n_vapor", "n_wall", "mass_liquid", "mass_vapor", "mass_total",
          "kinetic_energy", "enthalpy", "heat_input", "evaporated_mass", "condensed_mass", "vented_mass",
          "vented_enthalpy", "T_sat", "p_ullage", "max_speed", "T_liquid_mean", "T_liquid_max", "T_vapor_mean",
          "splits", "merges", "liquid_com", "liquid_momentum", "slosh_force", "slosh_moment", "frame_acceleration",
          "frame_angular_velocity"})
      r[k] = self.attr(k);
    return r;
  });

  // maneuver
  py::class_<Maneuver>(m, "ManeuverProfile",
                       "Prescribed tank motion (tank frame): piecewise-linear acceleration and angular velocity.")
      .def(py::init([](std::vector<double> time, DArr acc, py::object omega, const Tuple3& origin) {
             auto rows = [](const DArr& a, std::size_t n, const char* what) {
               if (a.ndim() != 2 || a.shape(1) != 3 || static_cast<std::size_t>(a.shape(0)) != n)
                 throw py::value_error(std::string(what) + " must have shape (len(time), 3)");
               std::vector<float> v(3 * n);
               auto r = a.unchecked<2>();
               for (std::size_t i = 0; i < n; ++i)
                 for (int c = 0; c < 3; ++c) v[3 * i + c] = static_cast<float>(r(i, c));
               return v;
             };
             Maneuver p;
             p.time = std::move(time);
             p.acc = rows(acc, p.time.size(), "acceleration");
             if (!omega.is_none()) p.omega = rows(omega.cast<DArr>(), p.time.size(), "angular_velocity");
             p.origin = origin;
             (void)p.evaluate(p.time.empty() ? 0.0 : p.time.front());  // validates
             return p;
           }),
           py::arg("time"), py::arg("acceleration"), py::arg("angular_velocity") = py::none(),
           py::arg("origin") = Tuple3{0, 0, 0})
      .def("evaluate", &Maneuver::evaluate, py::arg("t"),
           "Returns (acceleration, angular_velocity, angular_acceleration) at time t.");

  // solver
  py::class_<Solver>(m, "Solver", "Weakly-compressible multiphase SPH solver (gfx950).")
      .def(py::init<const cryo_config&, const Fluid&>(), py::arg("config"), py::arg("fluid"))
      .def("build_tank",
           [](Solver& S, const Geo& g, const cryo_fill& f) {
             const auto ms = g.meshes();
             py::gil_scoped_release rel;
             check(cryo_solver_build_tank(S.s, static_cast<int32_t>(g.nodes.size()), g.nodes.data(),
                                          static_cast<int32_t>(ms.size()), ms.data(), &f));
           },
           py::arg("geometry"), py::arg("fill") = default_fill(),
           "Generates wall/liquid/vapour particles for a tank.")
      .def("set_particles",
           [](Solver& S, DArr position, DArr mass, DArr density, DArr temperature, U8Arr phase, const Tuple3& lo,
              const Tuple3& hi, py::object velocity, py::object normal, py::object flags) {
             if (position.ndim() != 2 || position.shape(1) != 3)
               throw py::value_error("position must have shape (N, 3)");
             const py::ssize_t n = position.shape(0);
             auto chk = [n](const py::array& a, const char* name) {
               if (a.size() != n) throw py::value_error(std::string(name) + " must have length N");
             };
             chk(mass, "mass");
             chk(density, "density");
             chk(temperature, "temperature");
             chk(phase, "phase");
             std::vector<float> x(3 * n), v, nr, mm(n), rr(n), tt(n);
             auto P = position.unchecked<2>();
             for (py::ssize_t i = 0; i < n; ++i)
               for (int c = 0; c < 3; ++c) x[3 * i + c] = static_cast<float>(P(i, c));
             auto cp3 = [n](py::object o, std::vector<float>& dst, const char* what) {
               if (o.is_none()) return;
               DArr a = o.cast<DArr>();
               if (a.ndim() != 2 || a.shape(0) != n || a.shape(1) != 3)
                 throw py::value_error(std::string(what) + " must have shape (N, 3)");
               dst.resize(3 * n);
               for (py::ssize_t i = 0; i < 3 * n; ++i) dst[i] = static_cast<float>(a.data()[i]);
             };
             cp3(velocity, v, "velocity");
             cp3(normal, nr, "normal");
             for (py::ssize_t i = 0; i < n; ++i) {
               mm[i] = static_cast<float>(mass.data()[i]);
               rr[i] = static_cast<float>(density.data()[i]);
               tt[i] = static_cast<float>(temperature.data()[i]);
               if (phase.data()[i] > CRYO_WALL)
                 throw py::value_error("phase values must be 0 (liquid), 1 (vapour) or 2 (wall)");
             }
             std::vector<uint8_t> fl;
             if (!flags.is_none()) {
               U8Arr f = flags.cast<U8Arr>();
               chk(f, "flags");
               fl.assign(f.data(), f.data() + n);
             }
             const float l3[3] = {float(lo[0]), float(lo[1]), float(lo[2])};
             const float h3[3] = {float(hi[0]), float(hi[1]), float(hi[2])};
             check(cryo_solver_set_particles(S.s, static_cast<int32_t>(n), x.data(), v.empty() ? nullptr : v.data(),
                                             nr.empty() ? nullptr : nr.data(), mm.data(), rr.data(), tt.data(),
                                             phase.data(), fl.empty() ? nullptr : fl.data(), nullptr, l3, h3));
           },
           py::arg("position"), py::arg("mass"), py::arg("density"), py::arg("temperature"), py::arg("phase"),
           py::arg("domain_lo"), py::arg("domain_hi"), py::arg("velocity") = py::none(),
           py::arg("normal") = py::none(), py::arg("flags") = py::none(),
           "Loads a custom particle set (arrays of length N; position (N,3)).")
      .def("step", [](Solver& S) { check(cryo_solver_step(S.s)); }, py::call_guard<py::gil_scoped_release>())
      .def("run_steps", [](Solver& S, long long n) { check(cryo_solver_run_steps(S.s, n)); }, py::arg("n"),
           py::call_guard<py::gil_scoped_release>())
      .def("run_until", [](Solver& S, double t) { check(cryo_solver_run_until(S.s, t)); }, py::arg("t_end"),
           py::call_guard<py::gil_scoped_release>())
      .def("relax", [](Solver& S, double d, double damping) { check(cryo_solver_relax(S.s, d, float(damping))); },
           py::arg("duration"), py::arg("damping") = 0.98, py::call_guard<py::gil_scoped_release>())
      .def("next_dt", [](Solver& S) {
        double dt = 0;
        check(cryo_solver_next_dt(S.s, &dt));
        return dt;
      })
      .def("diagnostics", [](Solver& S) {
        cryo_diagnostics d{};
        check(cryo_solver_diagnostics(S.s, &d));
        return d;
      })
      .def("particles", &particles_dict, "Copies the particle state into a dict of NumPy arrays.")
      .def("pressure", [](const Solver& S) { return vec_array(S.f1<float>(CRYO_F_P)); })
      .def("curvature", [](const Solver& S) { return vec_array(S.f1<float>(CRYO_F_KAPPA)); })
      .def("field", [](const Solver& S, int field) -> py::object {
        if (field == CRYO_F_X || field == CRYO_F_V || field == CRYO_F_VT || field == CRYO_F_NRM || field == CRYO_F_A ||
            field == CRYO_F_AB || field == CRYO_F_NS || field == CRYO_F_VW || field == CRYO_F_FW || field == CRYO_F_MW)
          return vec3_array(S.f3(field));
        if (field == CRYO_F_PHASE || field == CRYO_F_FLAGS) return vec_array(S.f1<uint8_t>(field));
        if (field == CRYO_F_ID) return vec_array(S.f1<uint32_t>(field));
        return vec_array(S.f1<float>(field));
      }, py::arg("field"), "Any per-particle field by cryo_field id (debugging and parity checks).")
      .def_property_readonly("time", &Solver::time)
      .def_property_readonly("step_count", &Solver::steps)
      .def_property_readonly("last_dt", [](const Solver& S) {
        double dt = 0;
        check(cryo_solver_last_dt(S.s, &dt));
        return dt;
      })
      .def_property_readonly("num_particles", &Solver::n)
      .def_property_readonly("device", [](const Solver&) {
        char buf[256] = {0};
        check(cryo_device_name(buf, sizeof(buf)));
        return std::string(buf);
      })
      .def_property_readonly("config", [](const Solver& S) {
        cryo_config c{};
        check(cryo_solver_config(S.s, &c));
        return c;
      })
      .def_property_readonly("fluid", [](const Solver& S) { return S.fluid; })
      .def_property_readonly("sound_speed", [](const Solver& S) { return double(S.params().c0); })
      .def_property_readonly("smoothing_length", [](const Solver& S) { return double(S.params().h); })
      .def("set_gravity", [](Solver& S, const Tuple3& g) {
        const float v[3] = {float(g[0]), float(g[1]), float(g[2])};
        check(cryo_solver_set_gravity(S.s, v));
      }, py::arg("g"))
      .def("set_heater_power", [](Solver& S, double w) { check(cryo_solver_set_heater_power(S.s, float(w))); },
           py::arg("watts"))
      .def("set_frame_motion",
           [](Solver& S, const Tuple3& a, const Tuple3& w, const Tuple3& wd, const Tuple3& o) {
             const float A[3] = {float(a[0]), float(a[1]), float(a[2])}, W[3] = {float(w[0]), float(w[1]), float(w[2])},
                         D[3] = {float(wd[0]), float(wd[1]), float(wd[2])}, O[3] = {float(o[0]), float(o[1]), float(o[2])};
             check(cryo_solver_set_frame_motion(S.s, A, W, D, O));
           },
           py::arg("acceleration") = Tuple3{0, 0, 0}, py::arg("angular_velocity") = Tuple3{0, 0, 0},
           py::arg("angular_acceleration") = Tuple3{0, 0, 0}, py::arg("origin") = Tuple3{0, 0, 0})
      .def("set_maneuver", [](Solver& S, const Maneuver& p) {
        const float o[3] = {float(p.origin[0]), float(p.origin[1]), float(p.origin[2])};
        check(cryo_solver_set_maneuver(S.s, static_cast<int32_t>(p.time.size()), p.time.data(), p.acc.data(),
                                       p.omega.empty() ? nullptr : p.omega.data(), o));
      }, py::arg("profile"))
      .def("clear_frame_motion", [](Solver& S) { check(cryo_solver_clear_frame_motion(S.s)); })
      .def("set_contact_angle", [](Solver& S, double d) { check(cryo_solver_set_contact_angle(S.s, float(d))); },
           py::arg("degrees"))
      .def("damp_velocities", [](Solver& S, double f) { check(cryo_solver_damp_velocities(S.s, float(f))); },
           py::arg("factor"))
      .def("configure_telemetry",
           [](Solver& S, TelemetryMode mode) { check(cryo_solver_configure_telemetry(S.s, static_cast<int>(mode))); },
           py::arg("mode"))
      .def_property_readonly("telemetry_mode", [](Solver& S) {
        int32_t md = 0;
        check(cryo_solver_telemetry_mode(S.s, &md));
        return static_cast<TelemetryMode>(md);
      })
      .def("probe_pressure", [](Solver& S, DArr pts) {
        if (pts.ndim() != 2 || pts.shape(1) != 3) throw py::value_error("points must have shape (N, 3)");
        const py::ssize_t np = pts.shape(0);
        std::vector<float> p(3 * np), out(np);
        for (py::ssize_t i = 0; i < 3 * np; ++i) p[i] = static_cast<float>(pts.data()[i]);
        check(cryo_solver_probe_pressure(S.s, static_cast<int32_t>(np), p.data(), out.data()));
        return vec_array(out);
      }, py::arg("points"))
      // Extension (cryo_solver_autotune / set_launch_variant): no counterpart in the reference.
      .def("autotune", [](Solver& S, int trial_steps) {
        int32_t v = 0;
        double ms[4] = {0, 0, 0, 0};
        int r;
        {
          py::gil_scoped_release rel;
          r = cryo_solver_autotune(S.s, trial_steps, &v, ms);
        }
        check(r);
        return py::make_tuple(v, py::make_tuple(ms[0], ms[1], ms[2], ms[3]));
      }, py::arg("trial_steps") = 5,
         "Times trial_steps steps of each launch variant of the neighbour passes (LDS cell tiles or list "
         "gathers, for forces and interface), keeps the fastest, restores the state; returns (variant, "
         "ms per step of variants 0..3).  Results do not depend on the variant.")
      .def_property("launch_variant", [](Solver& S) {
        int32_t v = 0;
        check(cryo_solver_launch_variant(S.s, &v));
        return v;
      }, [](Solver& S, int v) { check(cryo_solver_set_launch_variant(S.s, v)); })
      // Extension (cryo_solver_set_tile_shape / autotune_tiles): no counterpart in the reference.
      .def_property("tile_shape", [](Solver& S) {
        int32_t xyz[3] = {0, 0, 0};
        check(cryo_solver_tile_shape(S.s, xyz));
        return py::make_tuple(xyz[0], xyz[1], xyz[2]);
      }, [](Solver& S, std::array<int, 3> xyz) { check(cryo_solver_set_tile_shape(S.s, xyz[0], xyz[1], xyz[2])); })
      .def("autotune_tiles", [](Solver& S, int trial_steps, std::optional<std::vector<std::array<int, 3>>> shapes) {
        std::vector<int32_t> flat;
        if (shapes)
          for (const auto& sh : *shapes) flat.insert(flat.end(), sh.begin(), sh.end());
        // no shapes: the library's six built-in candidates
        const int32_t count = shapes ? static_cast<int32_t>(flat.size() / 3) : 6;
        std::vector<double> ms(count > 0 ? count : 1, 0.0);
        int32_t k = 0;
        int r;
        {
          py::gil_scoped_release rel;
          r = cryo_solver_autotune_tiles(S.s, trial_steps, shapes ? flat.data() : nullptr, count, &k, ms.data());
        }
        check(r);
        int32_t xyz[3] = {0, 0, 0};
        check(cryo_solver_tile_shape(S.s, xyz));
        py::list times;
        for (int32_t i = 0; i < count; ++i) times.append(ms[i]);
        return py::make_tuple(py::make_tuple(xyz[0], xyz[1], xyz[2]), times);
      }, py::arg("trial_steps") = 5, py::arg("shapes") = py::none(),
         "Times trial_steps steps of each cell-tile shape (cells along x, y, z; default: six built-in "
         "candidates) at the current launch variant, keeps the fastest, restores the state; returns "
         "(shape, ms per step of each shape).  Results do not depend on the shape.")
      // Extension (cryo_solver_set_step_graphs / set_device_counts): no counterpart in the reference.
      .def("set_step_graphs", [](Solver& S, int on) { check(cryo_solver_set_step_graphs(S.s, on)); },
           py::arg("on"),
           "1 (or True): run_steps replays the step as a captured HIP graph for the rest of each call; 2: also "
           "lets the ensemble runner capture all members as one graph; 0: off (same results either way).")
      .def("set_device_counts", [](Solver& S, bool on, int room) {
        check(cryo_solver_set_device_counts(S.s, on ? 1 : 0, room));
      }, py::arg("on"), py::arg("room") = 0,
         "run_steps queues a call's phase-change steps with the particle count kept on the device; launches "
         "are sized for the count plus `room` particles (0: max(1024, n/8)), a call whose splits pass it is "
         "restored and run step by step (same results).")
      .def_property_readonly("device_count_restarts", [](Solver& S) {
        int64_t r = 0;
        check(cryo_solver_device_count_restarts(S.s, &r));
        return r;
      })
      // Extension (cryo_solver_adjoint_slosh): no counterpart in the reference.
      .def("adjoint_slosh", [](Solver& S, DArr acc, int checkpoint_every, bool gradient) -> py::tuple {
        if (acc.ndim() != 2 || acc.shape(1) != 3) throw py::value_error("frame_acc must have shape (steps, 3)");
        const py::ssize_t k = acc.shape(0);
        std::vector<float> a(3 * static_cast<std::size_t>(k));
        for (py::ssize_t i = 0; i < 3 * k; ++i) a[i] = static_cast<float>(acc.data()[i]);
        py::array_t<double> load({k, py::ssize_t{3}}), grad({k, py::ssize_t{3}});
        double loss = 0.0, dtheta = 0.0, dlee[2] = {0.0, 0.0};
        int r;
        {
          py::gil_scoped_release rel;
          r = cryo_solver_adjoint_slosh(S.s, k, a.data(), checkpoint_every, &loss,
                                        gradient ? grad.mutable_data() : nullptr, load.mutable_data(),
                                        gradient ? &dtheta : nullptr, gradient ? dlee : nullptr);
        }
        check(r);
        if (!gradient) return py::make_tuple(loss, py::none(), load, py::none(), py::none());
        return py::make_tuple(loss, grad, load, dtheta, py::make_tuple(dlee[0], dlee[1]));
      }, py::arg("frame_acc"), py::arg("checkpoint_every") = 20, py::arg("gradient") = true,
         "Runs len(frame_acc) steps (step k with frame acceleration frame_acc[k]); returns (loss, d loss / "
         "d frame_acc, wall load per step, d loss / d contact angle in degrees, (d loss / d "
         "lee_evaporation, d loss / d lee_condensation)), loss = sum_k dt |load_k,xy|^2; without gradient "
         "the derivatives are None.")
      // Extension (cryo_solver_adjoint_slosh_frame): no counterpart in the reference.
      .def("adjoint_slosh_frame", [](Solver& S, DArr acc, DArr omega, DArr omega_dot, int checkpoint_every,
                                     bool gradient) -> py::tuple {
        for (const DArr* a : {&acc, &omega, &omega_dot})
          if (a->ndim() != 2 || a->shape(1) != 3 || a->shape(0) != acc.shape(0))
            throw py::value_error("frame_acc, frame_omega and frame_omega_dot must have shape (steps, 3)");
        const py::ssize_t k = acc.shape(0);
        std::vector<float> a(3 * static_cast<std::size_t>(k)), w(a.size()), wd(a.size());
        for (py::ssize_t i = 0; i < 3 * k; ++i) {
          a[i] = static_cast<float>(acc.data()[i]);
          w[i] = static_cast<float>(omega.data()[i]);
          wd[i] = static_cast<float>(omega_dot.data()[i]);
        }
        py::array_t<double> load({k, py::ssize_t{3}}), grad({k, py::ssize_t{3}}), gw({k, py::ssize_t{3}}),
            gwd({k, py::ssize_t{3}});
        double loss = 0.0, dtheta = 0.0, dlee[2] = {0.0, 0.0};
        int r;
        {
          py::gil_scoped_release rel;
          r = cryo_solver_adjoint_slosh_frame(S.s, k, a.data(), w.data(), wd.data(), checkpoint_every, &loss,
                                              gradient ? grad.mutable_data() : nullptr,
                                              gradient ? gw.mutable_data() : nullptr,
                                              gradient ? gwd.mutable_data() : nullptr, load.mutable_data(),
                                              gradient ? &dtheta : nullptr, gradient ? dlee : nullptr);
        }
        check(r);
        if (!gradient) return py::make_tuple(loss, py::none(), py::none(), py::none(), load, py::none(), py::none());
        return py::make_tuple(loss, grad, gw, gwd, load, dtheta, py::make_tuple(dlee[0], dlee[1]));
      }, py::arg("frame_acc"), py::arg("frame_omega"), py::arg("frame_omega_dot"), py::arg("checkpoint_every") = 20,
         py::arg("gradient") = true,
         "adjoint_slosh in a rotating frame: step k also sets the frame's angular velocity and rate; returns "
         "(loss, d/d frame_acc, d/d frame_omega, d/d frame_omega_dot, wall loads, d/d contact angle, "
         "(d/d lee_evaporation, d/d lee_condensation)).")
      .def("request_telemetry", [](Solver& S) {
        const int r = cryo_solver_request_telemetry(S.s);
        check(r);
        return r == 1;
      })
      .def("telemetry_ready", [](Solver& S) {
        const int r = cryo_solver_telemetry_ready(S.s);
        check(r);
        return r == 1;
      })
      .def("fetch_telemetry", [](Solver& S) -> py::object {
        int32_t n = 0;
        check(cryo_solver_telemetry_pending_size(S.s, &n));
        std::vector<float> xyz(3 * static_cast<std::size_t>(n)), T(n), p(n), sp(n);
        std::vector<uint8_t> ph(n);
        double t = 0;
        int64_t step = 0;
        int r;
        {
          py::gil_scoped_release rel;
          r = cryo_solver_fetch_telemetry(S.s, &t, &step, &n, xyz.data(), T.data(), p.data(), sp.data(), ph.data());
        }
        check(r);
        if (r == 0) return py::none();
        py::dict d;
        d["time"] = t;
        d["step"] = step;
        d["position"] = vec3_array(xyz);
        d["temperature"] = vec_array(T);
        d["pressure"] = vec_array(p);
        d["speed"] = vec_array(sp);
        d["phase"] = vec_array(ph);
        return d;
      })
      .def("write_vtk", &write_vtk, py::arg("path"))
      .def("write_json", &write_json, py::arg("path"), py::arg("include_walls") = false);
}
