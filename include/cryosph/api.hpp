// cryosph/api.hpp - the C++ API of the reference (namespace cryosph: Solver,
// SolverConfig, FluidProperties, Diagnostics, ManeuverProfile, geometry,
// HostParticles, ParticleSystem, SpatialHash, TelemetryStream, the file
// writers and mem::) as a header-only layer over the C ABI of cryosph.h, so
// that C++ callers of the reference build unchanged against libcryosph.so.
//
// The reference's public headers (include/cryosph/{solvers,physics,core}/*.hpp)
// are mirrored by same-named headers next to this one that include it.
// Names, fields, defaults and exception types follow:
//   SolverConfig / ManeuverProfile / Diagnostics / Solver   sph_solver.hpp:20-232
//   FluidProperties                                         fluid_properties.hpp:11-38
//   AABB, Geometry and its tanks, FillSpec, generators      geometry.hpp:17-182
//   HostParticles, ParticleSystem                           particle_system.hpp:13-142
//   SpatialHash                                             spatial_hash.hpp:58-92
//   TelemetryFrame / TelemetryStream                        telemetry.hpp:12-70
//   write_vtk / write_json_frame / make_frame               io.hpp:12-22
//   Vec3, real, stream_t, mem::                             cuda_utils.cuh
//
// Differences a caller can see (the device internals stay behind the C ABI):
// * ParticleSystem has no view() or DeviceArray members.  A ParticleSystem of
//   one's own keeps its particles in host memory; SpatialHash::build sorts them
//   on the device (the same stable cell sort) and permutes them in place.
// * SimParams is the derived-constant block of the C ABI (cryo_params).
// * stream_t is an opaque handle (a HIP stream).
// * real is float (the reference's default build).
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "cryosph.h"

namespace cryosph {

using real = float;
using stream_t = void*;
using SimParams = cryo_params;

// ------------------------------------------------------------------ errors
namespace detail {
inline void check(int rc) {
  if (rc >= 0) return;
  const std::string msg = cryo_last_error();
  switch (rc) {
    case CRYO_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
    case CRYO_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
    default: throw std::runtime_error(msg);
  }
}
}  // namespace detail

// ------------------------------------------------------------------ vectors
struct Vec3 {
  real x = 0, y = 0, z = 0;
};
inline Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 operator-(Vec3 a) { return {-a.x, -a.y, -a.z}; }
inline Vec3 operator*(Vec3 a, real s) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec3 operator*(real s, Vec3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline Vec3 operator/(Vec3 a, real s) { return {a.x / s, a.y / s, a.z / s}; }
inline Vec3& operator+=(Vec3& a, Vec3 b) { return a = a + b; }
inline Vec3& operator-=(Vec3& a, Vec3 b) { return a = a - b; }
inline Vec3& operator*=(Vec3& a, real s) { return a = a * s; }
inline real dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline real norm2(Vec3 a) { return dot(a, a); }
inline real norm(Vec3 a) { return std::sqrt(dot(a, a)); }
inline Vec3 normalized(Vec3 a, real eps = real(1e-20)) {
  const real n = norm(a);
  return n > eps ? a / n : Vec3{0, 0, 0};
}
inline cryo_vec3 to_c(Vec3 a) { return cryo_vec3{a.x, a.y, a.z}; }
inline Vec3 from_c(cryo_vec3 a) { return Vec3{a.x, a.y, a.z}; }

struct AABB {
  Vec3 lo{0, 0, 0};
  Vec3 hi{0, 0, 0};
};

// ------------------------------------------------------------------ enums
enum class Phase : std::uint8_t { Liquid = 0, Vapor = 1, Wall = 2, Dead = 3 };
inline constexpr std::uint8_t kLiquid = 0, kVapor = 1, kWall = 2, kDead = 3;
inline bool is_fluid(std::uint8_t ph) { return ph <= kVapor; }
namespace pflag {
inline constexpr std::uint8_t Heated = 1u << 0, Source = 1u << 1, TargetVapor = 1u << 2, Dying = 1u << 3;
inline constexpr std::uint8_t ScratchMask = Source | TargetVapor | Dying;
}  // namespace pflag
enum class KernelType : int { CubicSpline = 0, WendlandC2 = 1 };
enum class DensityMode : int { Summation = 0, Continuity = 1 };
enum class SurfaceTensionModel : int { None = 0, CSF = 1, Akinci = 2 };
enum class WallThermalBC : int { Adiabatic = 0, Isothermal = 1 };
enum class TankThermoMode : int { Vented = 0, Closed = 1 };
enum class WettingModel : int { NormalRotation = 0, ContactLineForce = 1, Both = 2 };
enum class FillMode : int { Settled = 0, CentralBubble = 1, Full = 2 };
enum class TelemetryMode : int { AsyncCopy = 0, ZeroCopy = 1 };

// ------------------------------------------------------------------ mem::
namespace mem {
inline stream_t create_stream() {
  void* s = nullptr;
  detail::check(cryo_stream_create(&s));
  return s;
}
inline void destroy_stream(stream_t s) { detail::check(cryo_stream_destroy(s)); }
inline void synchronize(stream_t s) { detail::check(cryo_stream_synchronize(s)); }
inline std::string device_name() {
  char buf[256] = {0};
  detail::check(cryo_device_name(buf, sizeof buf));
  return buf;
}
inline int device_count() { return cryo_device_count(); }
}  // namespace mem

// ------------------------------------------------------------------ fluid
struct FluidProperties {
  std::string name = "custom";
  real rho_liquid = 1000;
  real rho_vapor = 1;
  real mu_liquid = 1e-3;
  real mu_vapor = 1e-5;
  real k_liquid = 0.6;
  real k_vapor = 0.025;
  real cp_liquid = 4180;
  real cp_vapor = 2000;
  real sigma = 0.072;
  real h_fg = 2.26e6;
  real T_sat = 373.15;
  real p_sat = 101325;
  real R_gas = 461.5;

  real saturation_temperature(real p) const {
    const cryo_fluid c = to_c();
    return cryo_fluid_saturation_temperature(&c, p);
  }
  real saturation_pressure(real T) const {
    const cryo_fluid c = to_c();
    return cryo_fluid_saturation_pressure(&c, T);
  }
  static FluidProperties liquid_hydrogen() { return preset("LH2"); }
  static FluidProperties liquid_nitrogen() { return preset("LN2"); }
  static FluidProperties liquid_oxygen() { return preset("LOX"); }
  static FluidProperties liquid_methane() { return preset("LCH4"); }
  static FluidProperties water() { return preset("H2O"); }

  cryo_fluid to_c() const {
    cryo_fluid c{};
    std::snprintf(c.name, sizeof c.name, "%s", name.c_str());
    c.rho_liquid = rho_liquid, c.rho_vapor = rho_vapor, c.mu_liquid = mu_liquid, c.mu_vapor = mu_vapor;
    c.k_liquid = k_liquid, c.k_vapor = k_vapor, c.cp_liquid = cp_liquid, c.cp_vapor = cp_vapor;
    c.sigma = sigma, c.h_fg = h_fg, c.T_sat = T_sat, c.p_sat = p_sat, c.R_gas = R_gas;
    return c;
  }
  static FluidProperties from_c(const cryo_fluid& c) {
    FluidProperties f;
    f.name = std::string(c.name, strnlen(c.name, sizeof c.name));
    f.rho_liquid = c.rho_liquid, f.rho_vapor = c.rho_vapor, f.mu_liquid = c.mu_liquid, f.mu_vapor = c.mu_vapor;
    f.k_liquid = c.k_liquid, f.k_vapor = c.k_vapor, f.cp_liquid = c.cp_liquid, f.cp_vapor = c.cp_vapor;
    f.sigma = c.sigma, f.h_fg = c.h_fg, f.T_sat = c.T_sat, f.p_sat = c.p_sat, f.R_gas = c.R_gas;
    return f;
  }

 private:
  static FluidProperties preset(const char* key) {
    cryo_fluid c{};
    detail::check(cryo_fluid_preset(key, &c));
    return from_c(c);
  }
};

// ------------------------------------------------------------------ particles
struct HostParticles {
  std::vector<Vec3> x, v, nrm;
  std::vector<real> m, rho, T;
  std::vector<std::uint8_t> phase, flags;
  std::vector<std::uint32_t> id;

  std::size_t size() const { return x.size(); }
  void reserve(std::size_t n) {
    x.reserve(n), v.reserve(n), nrm.reserve(n), m.reserve(n), rho.reserve(n), T.reserve(n);
    phase.reserve(n), flags.reserve(n), id.reserve(n);
  }
  void clear() {
    x.clear(), v.clear(), nrm.clear(), m.clear(), rho.clear(), T.clear();
    phase.clear(), flags.clear(), id.clear();
  }
  std::size_t push(Vec3 pos, Vec3 vel, real mass, real density, real temperature, Phase ph, Vec3 normal = {0, 0, 0},
                   std::uint8_t flag = 0) {
    x.push_back(pos), v.push_back(vel), nrm.push_back(normal);
    m.push_back(mass), rho.push_back(density), T.push_back(temperature);
    phase.push_back(static_cast<std::uint8_t>(ph)), flags.push_back(flag);
    id.push_back(static_cast<std::uint32_t>(id.size()));
    return x.size() - 1;
  }
  void append(const HostParticles& o) {
    const auto base = static_cast<std::uint32_t>(size());
    x.insert(x.end(), o.x.begin(), o.x.end()), v.insert(v.end(), o.v.begin(), o.v.end());
    nrm.insert(nrm.end(), o.nrm.begin(), o.nrm.end()), m.insert(m.end(), o.m.begin(), o.m.end());
    rho.insert(rho.end(), o.rho.begin(), o.rho.end()), T.insert(T.end(), o.T.begin(), o.T.end());
    phase.insert(phase.end(), o.phase.begin(), o.phase.end());
    flags.insert(flags.end(), o.flags.begin(), o.flags.end());
    for (std::uint32_t k : o.id) id.push_back(base + k);
  }
};

namespace detail {
// Interleaved float copies of Vec3 arrays, as the C ABI takes them.
inline std::vector<float> flat(const std::vector<Vec3>& a) {
  std::vector<float> out(3 * a.size());
  for (std::size_t i = 0; i < a.size(); ++i) out[3 * i] = a[i].x, out[3 * i + 1] = a[i].y, out[3 * i + 2] = a[i].z;
  return out;
}
inline std::vector<Vec3> unflat(const std::vector<float>& f) {
  std::vector<Vec3> out(f.size() / 3);
  for (std::size_t i = 0; i < out.size(); ++i) out[i] = Vec3{f[3 * i], f[3 * i + 1], f[3 * i + 2]};
  return out;
}
}  // namespace detail

class Solver;

// Particle storage.  A Solver's own (Solver::particle_system()) reads and
// writes the solver's device arrays; one made by the caller holds its
// particles in host memory (see SpatialHash::build).
class ParticleSystem {
 public:
  ParticleSystem() = default;
  ParticleSystem(const ParticleSystem&) = delete;
  ParticleSystem& operator=(const ParticleSystem&) = delete;

  void allocate(std::size_t capacity) {
    if (capacity < host_.size()) throw std::out_of_range("ParticleSystem::allocate: capacity below size");
    capacity_ = capacity;
  }
  void grow(std::size_t new_capacity) {
    if (new_capacity > capacity_) capacity_ = new_capacity;
  }
  inline void upload(const HostParticles& host, double growth_factor = 2.0);
  inline HostParticles download() const;
  inline std::size_t size() const;
  inline std::size_t capacity() const;
  std::vector<real> download_pressure() const;
  std::vector<real> download_curvature() const;
  std::uint32_t next_id = 0;

 private:
  friend class Solver;
  friend class SpatialHash;
  cryo_solver* owner_ = nullptr;  // set for a Solver's own particle system
  AABB* owner_domain_ = nullptr;
  HostParticles host_;
  std::size_t capacity_ = 0;
};

// ------------------------------------------------------------------ spatial hash
class SpatialHash {
 public:
  void configure(Vec3 lo, Vec3 hi, real cell_size) {
    if (!(cell_size > 0)) throw std::invalid_argument("SpatialHash: cell size must be positive");
    lo_ = lo, hi_ = hi, cell_ = cell_size;
    const Vec3 origin = lo - Vec3{cell_size, cell_size, cell_size};
    const real inv = real(1) / cell_size;
    const Vec3 ext = (hi + Vec3{cell_size, cell_size, cell_size}) - origin;
    nx_ = std::max(1, static_cast<int>(std::ceil(ext.x * inv)));
    ny_ = std::max(1, static_cast<int>(std::ceil(ext.y * inv)));
    nz_ = std::max(1, static_cast<int>(std::ceil(ext.z * inv)));
    origin_ = origin;
    configured_ = true;
  }
  // Stable sort of `ps` by cell (dead particles dropped); returns the number
  // removed.  For a caller's own ParticleSystem the sort runs on the device
  // and permutes the host-resident particles.
  int build(ParticleSystem& ps, stream_t /*stream*/) {
    if (ps.owner_) {
      int32_t removed = 0;
      detail::check(cryo_solver_rebuild_grid(ps.owner_, &removed));
      return removed;
    }
    require_configured();
    HostParticles& h = ps.host_;
    const std::vector<std::uint32_t> perm = sort(h, 0, nullptr, nullptr);
    const std::size_t before = h.size();
    HostParticles out;
    out.reserve(perm.size());
    for (std::uint32_t q : perm) {
      out.x.push_back(h.x[q]), out.v.push_back(h.v[q]), out.nrm.push_back(h.nrm[q]);
      out.m.push_back(h.m[q]), out.rho.push_back(h.rho[q]), out.T.push_back(h.T[q]);
      out.phase.push_back(h.phase[q]), out.flags.push_back(h.flags[q]), out.id.push_back(h.id[q]);
    }
    h = std::move(out);
    return static_cast<int>(before - h.size());
  }
  // For every particle (in the current order) the number of particles within
  // `radius` (itself included) and the sum of their ids.
  void neighbor_stats(ParticleSystem& ps, real radius, std::vector<int>& counts, std::vector<std::uint64_t>& id_sums,
                      stream_t /*stream*/) {
    if (ps.owner_) {
      int32_t n = 0;
      detail::check(cryo_solver_num_particles(ps.owner_, &n));
      counts.assign(static_cast<std::size_t>(n), 0);
      id_sums.assign(static_cast<std::size_t>(n), 0);
      detail::check(cryo_solver_neighbor_stats(ps.owner_, radius, counts.data(), id_sums.data()));
      return;
    }
    require_configured();
    counts.assign(ps.host_.size(), 0);
    id_sums.assign(ps.host_.size(), 0);
    sort(ps.host_, radius, counts.data(), id_sums.data());
  }
  int num_cells() const {
    sync_from_owner();
    return nx_ * ny_ * nz_;
  }
  real cell_size() const {
    sync_from_owner();
    return cell_;
  }
  Vec3 lower() const {
    sync_from_owner();
    return origin_;
  }
  Vec3 upper() const {
    sync_from_owner();
    return origin_ + Vec3{cell_ * static_cast<real>(nx_), cell_ * static_cast<real>(ny_), cell_ * static_cast<real>(nz_)};
  }

 private:
  friend class Solver;
  // a Solver's own grid reports the solver's cell table
  void sync_from_owner() const {
    if (!owner_) return;
    int32_t dims[3], cells = 0;
    float origin[3], cell = 0;
    detail::check(cryo_solver_grid_info(owner_, dims, origin, &cell, &cells));
    nx_ = dims[0], ny_ = dims[1], nz_ = dims[2];
    origin_ = Vec3{origin[0], origin[1], origin[2]};
    cell_ = cell;
  }
  cryo_solver* owner_ = nullptr;
  void require_configured() const {
    if (!configured_) throw std::runtime_error("SpatialHash: configure() first");
  }
  std::vector<std::uint32_t> sort(const HostParticles& h, real radius, int* counts, std::uint64_t* sums) const {
    const int32_t n = static_cast<int32_t>(h.size());
    const std::vector<float> xs = detail::flat(h.x);
    const float lo[3] = {lo_.x, lo_.y, lo_.z}, hi[3] = {hi_.x, hi_.y, hi_.z};
    std::vector<std::uint32_t> perm(h.size());
    int32_t alive = 0;
    std::vector<uint64_t> s64(sums ? h.size() : 0);
    detail::check(cryo_hash_sort_ids(n, xs.data(), h.phase.data(), h.id.data(), lo, hi, cell_, perm.data(), &alive,
                                     radius, counts, sums ? s64.data() : nullptr));
    if (sums) std::copy(s64.begin(), s64.end(), sums);
    perm.resize(static_cast<std::size_t>(alive));
    return perm;
  }
  Vec3 lo_{0, 0, 0}, hi_{1, 1, 1};
  mutable Vec3 origin_{0, 0, 0};
  mutable real cell_ = 1;
  mutable int nx_ = 1, ny_ = 1, nz_ = 1;
  bool configured_ = false;
};

// ------------------------------------------------------------------ telemetry
struct TelemetryFrame {
  double time = 0;
  long long step = 0;
  int n = 0;
  std::vector<float> position, temperature, pressure, speed;
  std::vector<std::uint8_t> phase;
};

struct TelemetryRecord {
  float x, y, z, T, p, speed;
  std::uint32_t phase;
};

// A Solver's frame capture (telemetry.hpp:33-70): one frame in flight at a time.
class TelemetryStream {
 public:
  TelemetryStream(const TelemetryStream&) = delete;
  TelemetryStream& operator=(const TelemetryStream&) = delete;
  void configure(std::size_t /*capacity: the solver sizes its buffers itself*/, TelemetryMode mode) {
    detail::check(cryo_solver_configure_telemetry(s_, static_cast<int32_t>(mode)));
  }
  TelemetryMode mode() const {
    int32_t m = 0;
    detail::check(cryo_solver_telemetry_mode(s_, &m));
    return static_cast<TelemetryMode>(m);
  }
  bool request(ParticleSystem&, stream_t, double, long long) { return request(); }
  bool ready() const {
    const int r = cryo_solver_telemetry_ready(s_);
    detail::check(r);
    return r != 0;
  }
  bool pending() const {
    int32_t n = 0;
    detail::check(cryo_solver_telemetry_pending_size(s_, &n));
    return n > 0;
  }
  bool fetch(TelemetryFrame& out) {
    int32_t n = 0;
    detail::check(cryo_solver_telemetry_pending_size(s_, &n));
    if (n <= 0) return false;
    out.position.resize(3 * static_cast<std::size_t>(n));
    out.temperature.resize(n), out.pressure.resize(n), out.speed.resize(n), out.phase.resize(n);
    int64_t step = 0;
    const int r = cryo_solver_fetch_telemetry(s_, &out.time, &step, &n, out.position.data(), out.temperature.data(),
                                              out.pressure.data(), out.speed.data(), out.phase.data());
    detail::check(r);
    out.step = step;
    out.n = n;
    return r == 1;
  }

 private:
  friend class Solver;
  explicit TelemetryStream(cryo_solver* s) : s_(s) {}
  bool request() {
    const int r = cryo_solver_request_telemetry(s_);
    detail::check(r);
    return r == 1;
  }
  cryo_solver* s_;
};

// ------------------------------------------------------------------ files
inline void write_vtk(const std::string& path, const HostParticles& hp, const std::vector<real>& pressure = {}) {
  const std::vector<float> x = detail::flat(hp.x), v = detail::flat(hp.v);
  const int32_t n = static_cast<int32_t>(hp.size());
  detail::check(cryo_write_vtk(path.c_str(), n, x.data(), v.data(), hp.T.data(), hp.m.data(), hp.rho.data(),
                               hp.phase.data(), pressure.size() == hp.size() ? pressure.data() : nullptr));
}
inline TelemetryFrame make_frame(const HostParticles& hp, const std::vector<real>& pressure, double time,
                                 long long step) {
  TelemetryFrame f;
  const std::size_t n = hp.size();
  f.time = time, f.step = step, f.n = static_cast<int>(n);
  f.position = detail::flat(hp.x);
  f.temperature = hp.T;
  f.pressure = pressure.size() == n ? pressure : std::vector<real>(n, 0.0f);
  f.speed.resize(n);
  const std::vector<float> v = detail::flat(hp.v);
  cryo_frame_speed(static_cast<int32_t>(n), v.data(), f.speed.data());
  f.phase = hp.phase;
  return f;
}
inline void write_json_frame(const std::string& path, const TelemetryFrame& fr, bool include_walls = true) {
  detail::check(cryo_write_json_frame(path.c_str(), fr.time, fr.step, fr.n, fr.position.data(),
                                      fr.temperature.data(), fr.pressure.data(), fr.speed.data(), fr.phase.data(),
                                      include_walls ? 1 : 0));
}

// ------------------------------------------------------------------ geometry
// A tank as a CSG tree.  The built-in shapes serialise to the C ABI's node
// list; a caller's own subclass enters it as a distance callback.
struct GeoList {
  std::vector<cryo_geo_node> nodes;
  std::vector<cryo_mesh> meshes;
  std::vector<cryo_geo_callback> callbacks;
};

class Geometry {
 public:
  virtual ~Geometry() = default;
  virtual double sdf(double x, double y, double z) const = 0;
  virtual AABB bounds() const = 0;
  virtual std::string name() const = 0;
  virtual double volume() const { return -1.0; }
  // central differences, as the lattice generator computes wall normals
  Vec3 outward_normal(double x, double y, double z, double eps) const {
    const double gx = sdf(x + eps, y, z) - sdf(x - eps, y, z);
    const double gy = sdf(x, y + eps, z) - sdf(x, y - eps, z);
    const double gz = sdf(x, y, z + eps) - sdf(x, y, z - eps);
    const double n = std::sqrt(gx * gx + gy * gy + gz * gz);
    if (n < 1e-300) return Vec3{0, 0, 0};
    return Vec3{static_cast<real>(gx / n), static_cast<real>(gy / n), static_cast<real>(gz / n)};
  }
  // Appends this tree's nodes; returns the index of its root.
  virtual int emit(GeoList& out) const {
    cryo_geo_callback cb{};
    cb.sdf = [](void* ctx, double x, double y, double z) { return static_cast<const Geometry*>(ctx)->sdf(x, y, z); };
    cb.ctx = const_cast<Geometry*>(this);
    const AABB b = bounds();
    cb.lo[0] = b.lo.x, cb.lo[1] = b.lo.y, cb.lo[2] = b.lo.z, cb.hi[0] = b.hi.x, cb.hi[1] = b.hi.y, cb.hi[2] = b.hi.z;
    cb.volume = volume();
    out.callbacks.push_back(cb);
    return leaf(out, CRYO_GEO_CALLBACK, {0, 0, 0}, static_cast<int>(out.callbacks.size()) - 1);
  }

 protected:
  static int leaf(GeoList& out, int type, std::array<double, 3> p, int mesh_id = -1, int a = -1, int b = -1) {
    cryo_geo_node nd{};
    nd.type = type, nd.a = a, nd.b = b, nd.mesh_id = mesh_id;
    nd.p[0] = p[0], nd.p[1] = p[1], nd.p[2] = p[2];
    out.nodes.push_back(nd);
    return static_cast<int>(out.nodes.size()) - 1;
  }
  // Distance / bounds / volume of a built-in tree through the C ABI.
  double tree_sdf(double x, double y, double z) const {
    GeoList g;
    emit(g);
    double d = 0;
    detail::check(cryo_geo_sdf(static_cast<int32_t>(g.nodes.size()), g.nodes.data(),
                               static_cast<int32_t>(g.meshes.size()), g.meshes.data(), x, y, z, &d));
    return d;
  }
  AABB tree_bounds() const {
    GeoList g;
    emit(g);
    float lo[3], hi[3];
    detail::check(cryo_geo_bounds(static_cast<int32_t>(g.nodes.size()), g.nodes.data(),
                                  static_cast<int32_t>(g.meshes.size()), g.meshes.data(), lo, hi));
    return AABB{Vec3{lo[0], lo[1], lo[2]}, Vec3{hi[0], hi[1], hi[2]}};
  }
  double tree_volume() const {
    GeoList g;
    emit(g);
    double v = -1;
    detail::check(cryo_geo_volume(static_cast<int32_t>(g.nodes.size()), g.nodes.data(),
                                  static_cast<int32_t>(g.meshes.size()), g.meshes.data(), &v));
    return v;
  }
};

namespace detail {
// The analytic shapes: one leaf node each, evaluated by the library.
template <int Type>
class Primitive : public Geometry {
 public:
  double sdf(double x, double y, double z) const override { return tree_sdf(x, y, z); }
  AABB bounds() const override { return tree_bounds(); }
  double volume() const override { return tree_volume(); }
  int emit(GeoList& out) const override { return leaf(out, Type, p_); }

 protected:
  explicit Primitive(std::array<double, 3> p) : p_(p) {}
  std::array<double, 3> p_;
};
}  // namespace detail

class CylinderTank final : public detail::Primitive<CRYO_GEO_CYLINDER> {
 public:
  CylinderTank(double radius, double height) : Primitive({radius, height, 0}) {
    if (!(radius > 0 && height > 0)) throw std::invalid_argument("CylinderTank: radius and height must be > 0");
  }
  std::string name() const override { return "cylinder"; }
  double radius() const { return p_[0]; }
  double height() const { return p_[1]; }
};

class CapsuleTank final : public detail::Primitive<CRYO_GEO_CAPSULE> {
 public:
  CapsuleTank(double radius, double length) : Primitive({radius, length, 0}) {
    if (!(radius > 0 && length >= 0)) throw std::invalid_argument("CapsuleTank: invalid dimensions");
  }
  std::string name() const override { return "capsule"; }
};

class SphereTank final : public detail::Primitive<CRYO_GEO_SPHERE> {
 public:
  explicit SphereTank(double radius) : Primitive({radius, 0, 0}) {
    if (!(radius > 0)) throw std::invalid_argument("SphereTank: radius must be > 0");
  }
  std::string name() const override { return "sphere"; }
};

class BoxTank final : public detail::Primitive<CRYO_GEO_BOX> {
 public:
  BoxTank(double lx, double ly, double lz) : Primitive({lx, ly, lz}) {
    if (!(lx > 0 && ly > 0 && lz > 0)) throw std::invalid_argument("BoxTank: side lengths must be > 0");
  }
  std::string name() const override { return "box"; }
};

class Translated final : public Geometry {
 public:
  Translated(std::shared_ptr<const Geometry> g, Vec3 offset) : g_(std::move(g)), o_(offset) {
    if (!g_) throw std::invalid_argument("Translated: null geometry");
  }
  double sdf(double x, double y, double z) const override {
    return g_->sdf(x - static_cast<double>(o_.x), y - static_cast<double>(o_.y), z - static_cast<double>(o_.z));
  }
  AABB bounds() const override {
    const AABB b = g_->bounds();
    return AABB{b.lo + o_, b.hi + o_};
  }
  std::string name() const override { return "translated(" + g_->name() + ")"; }
  double volume() const override { return g_->volume(); }
  int emit(GeoList& out) const override {
    const int child = g_->emit(out);
    return leaf(out, CRYO_GEO_TRANSLATED, {o_.x, o_.y, o_.z}, -1, child);
  }

 private:
  std::shared_ptr<const Geometry> g_;
  Vec3 o_;
};

class Difference final : public Geometry {
 public:
  Difference(std::shared_ptr<const Geometry> a, std::shared_ptr<const Geometry> b) : a_(std::move(a)), b_(std::move(b)) {
    if (!a_ || !b_) throw std::invalid_argument("Difference: null geometry");
  }
  double sdf(double x, double y, double z) const override {
    const double da = a_->sdf(x, y, z), db = -b_->sdf(x, y, z);
    return (da < db) ? db : da;
  }
  AABB bounds() const override { return a_->bounds(); }
  std::string name() const override { return "difference(" + a_->name() + ", " + b_->name() + ")"; }
  int emit(GeoList& out) const override {
    const int ia = a_->emit(out), ib = b_->emit(out);
    return leaf(out, CRYO_GEO_DIFFERENCE, {0, 0, 0}, -1, ia, ib);
  }

 private:
  std::shared_ptr<const Geometry> a_, b_;
};

class Union final : public Geometry {
 public:
  Union(std::shared_ptr<const Geometry> a, std::shared_ptr<const Geometry> b) : a_(std::move(a)), b_(std::move(b)) {
    if (!a_ || !b_) throw std::invalid_argument("Union: null geometry");
  }
  double sdf(double x, double y, double z) const override {
    const double da = a_->sdf(x, y, z), db = b_->sdf(x, y, z);
    return (db < da) ? db : da;
  }
  AABB bounds() const override {
    const AABB a = a_->bounds(), b = b_->bounds();
    return AABB{Vec3{std::min(a.lo.x, b.lo.x), std::min(a.lo.y, b.lo.y), std::min(a.lo.z, b.lo.z)},
                Vec3{std::max(a.hi.x, b.hi.x), std::max(a.hi.y, b.hi.y), std::max(a.hi.z, b.hi.z)}};
  }
  std::string name() const override { return "union(" + a_->name() + ", " + b_->name() + ")"; }
  int emit(GeoList& out) const override {
    const int ia = a_->emit(out), ib = b_->emit(out);
    return leaf(out, CRYO_GEO_UNION, {0, 0, 0}, -1, ia, ib);
  }

 private:
  std::shared_ptr<const Geometry> a_, b_;
};

// Wavefront OBJ (v and f records; polygons fan-triangulated; 1-based or
// negative relative indices; "i/t/n" forms take the vertex index).
inline void load_obj(const std::string& path, std::vector<std::array<double, 3>>& vertices,
                     std::vector<std::array<int, 3>>& triangles) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("load_obj: cannot open '" + path + "'");
  vertices.clear();
  triangles.clear();
  for (std::string line; std::getline(in, line);) {
    std::istringstream fields(line);
    std::string kind;
    if (!(fields >> kind)) continue;
    if (kind == "v") {
      std::array<double, 3> p{};
      if (!(fields >> p[0] >> p[1] >> p[2])) throw std::runtime_error("load_obj: malformed vertex in '" + path + "'");
      vertices.push_back(p);
    } else if (kind == "f") {
      std::vector<int> corner;
      for (std::string tok; fields >> tok;) {
        const int k = std::stoi(tok.substr(0, tok.find('/')));
        const int at = k > 0 ? k - 1 : static_cast<int>(vertices.size()) + k;
        if (at < 0 || at >= static_cast<int>(vertices.size()))
          throw std::runtime_error("load_obj: face index out of range in '" + path + "'");
        corner.push_back(at);
      }
      for (std::size_t c = 1; c + 1 < corner.size(); ++c) triangles.push_back({corner[0], corner[c], corner[c + 1]});
    }
  }
  if (triangles.empty()) throw std::runtime_error("load_obj: no faces found in '" + path + "'");
}

class MeshTank final : public Geometry {
 public:
  MeshTank(const std::string& obj_path, double scale = 1.0, Vec3 offset = {0, 0, 0}) {
    std::vector<std::array<double, 3>> v;
    std::vector<std::array<int, 3>> t;
    load_obj(obj_path, v, t);
    for (auto& p : v) p = {p[0] * scale + offset.x, p[1] * scale + offset.y, p[2] * scale + offset.z};
    set(std::move(v), std::move(t));
  }
  MeshTank(std::vector<std::array<double, 3>> vertices, std::vector<std::array<int, 3>> triangles) {
    set(std::move(vertices), std::move(triangles));
  }
  double sdf(double x, double y, double z) const override { return tree_sdf(x, y, z); }
  AABB bounds() const override { return tree_bounds(); }
  std::string name() const override { return "mesh"; }
  double volume() const override { return tree_volume(); }
  std::size_t num_triangles() const { return tris_.size() / 3; }
  double winding_number(double x, double y, double z) const {
    const cryo_mesh m = mesh();
    double w = 0;
    detail::check(cryo_mesh_winding_number(&m, x, y, z, &w));
    return w;
  }
  int emit(GeoList& out) const override {
    out.meshes.push_back(mesh());
    return leaf(out, CRYO_GEO_MESH, {0, 0, 0}, static_cast<int>(out.meshes.size()) - 1);
  }

 private:
  void set(std::vector<std::array<double, 3>> v, std::vector<std::array<int, 3>> t) {
    if (v.empty() || t.empty()) throw std::invalid_argument("MeshTank: empty mesh");
    for (const auto& p : v) verts_.insert(verts_.end(), p.begin(), p.end());
    for (const auto& tri : t)
      for (int k : tri) {
        if (k < 0 || k >= static_cast<int>(v.size())) throw std::invalid_argument("MeshTank: bad triangle index");
        tris_.push_back(k);
      }
  }
  cryo_mesh mesh() const {
    return cryo_mesh{static_cast<int32_t>(verts_.size() / 3), static_cast<int32_t>(tris_.size() / 3), verts_.data(),
                     tris_.data()};
  }
  std::vector<double> verts_;
  std::vector<int32_t> tris_;
};

struct FillSpec {
  FillMode mode = FillMode::Settled;
  double fill_fraction = 0.5;
  Vec3 settle_direction{0, 0, -1};
  Vec3 bubble_center{0, 0, 0};
  bool include_vapor = true;
  int wall_layers = 3;
  real T_liquid = std::numeric_limits<real>::quiet_NaN();
  real T_vapor = std::numeric_limits<real>::quiet_NaN();
  real T_wall = std::numeric_limits<real>::quiet_NaN();
  double heater_zmin = -std::numeric_limits<double>::infinity();
  double heater_zmax = std::numeric_limits<double>::infinity();

  cryo_fill to_c() const {
    cryo_fill f{};
    f.mode = static_cast<int32_t>(mode), f.fill_fraction = fill_fraction;
    f.settle_direction = cryosph::to_c(settle_direction), f.bubble_center = cryosph::to_c(bubble_center);
    f.include_vapor = include_vapor ? 1 : 0, f.wall_layers = wall_layers;
    f.T_liquid = T_liquid, f.T_vapor = T_vapor, f.T_wall = T_wall;
    f.heater_zmin = heater_zmin, f.heater_zmax = heater_zmax;
    return f;
  }
};

inline HostParticles build_tank_particles(const Geometry& geo, const FluidProperties& fluid, real dx,
                                          const FillSpec& fill) {
  GeoList g;
  geo.emit(g);
  const cryo_fluid cf = fluid.to_c();
  const cryo_fill cfill = fill.to_c();
  const auto call = [&](int32_t* n, float* x, float* v, float* nr, float* m, float* rho, float* T, uint8_t* ph,
                        uint8_t* fl, uint32_t* id) {
    detail::check(cryo_build_tank_particles_cb(static_cast<int32_t>(g.nodes.size()), g.nodes.data(),
                                               static_cast<int32_t>(g.meshes.size()), g.meshes.data(),
                                               static_cast<int32_t>(g.callbacks.size()), g.callbacks.data(), &cf, dx,
                                               &cfill, n, x, v, nr, m, rho, T, ph, fl, id));
  };
  int32_t n = 0;
  call(&n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  std::vector<float> x(3 * static_cast<std::size_t>(n)), v(x.size()), nr(x.size());
  HostParticles hp;
  hp.m.resize(n), hp.rho.resize(n), hp.T.resize(n), hp.phase.resize(n), hp.flags.resize(n), hp.id.resize(n);
  call(&n, x.data(), v.data(), nr.data(), hp.m.data(), hp.rho.data(), hp.T.data(), hp.phase.data(), hp.flags.data(),
       hp.id.data());
  hp.x = detail::unflat(x), hp.v = detail::unflat(v), hp.nrm = detail::unflat(nr);
  return hp;
}

inline HostParticles make_block(Vec3 lo, Vec3 hi, real dx, Phase phase, real density, real temperature) {
  HostParticles out;
  const int nx = static_cast<int>(std::floor((hi.x - lo.x) / dx + real(1e-4)));
  const int ny = static_cast<int>(std::floor((hi.y - lo.y) / dx + real(1e-4)));
  const int nz = static_cast<int>(std::floor((hi.z - lo.z) / dx + real(1e-4)));
  const real mass = density * dx * dx * dx;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        out.push(Vec3{lo.x + (i + real(0.5)) * dx, lo.y + (j + real(0.5)) * dx, lo.z + (k + real(0.5)) * dx},
                 Vec3{0, 0, 0}, mass, density, temperature, phase);
  return out;
}

// ------------------------------------------------------------------ solver
struct SolverConfig {
  real dx = real(0.005);
  real h_factor = real(1.3);
  KernelType kernel = KernelType::WendlandC2;
  DensityMode density_mode = DensityMode::Summation;
  bool normalize_initial_masses = true;
  Vec3 gravity{0, 0, real(-9.81)};
  real velocity_scale = 0;
  real sound_speed = 0;
  real tait_gamma = 7;
  real background_pressure = -1;
  real vapor_stiffness = -1;
  bool clamp_negative_pressure = false;
  bool transport_velocity = true;
  real transport_background_factor = real(2.0);
  bool transport_stress = false;
  bool transport_interface_only = true;
  bool particle_shifting = false;
  real shifting_strength = real(0.25);
  real artificial_viscosity = real(0.02);
  real artificial_viscosity_vapor = real(0.2);
  real delta_sph = real(0.1);
  bool no_slip = true;
  bool sps_turbulence = false;
  real smagorinsky_cs = real(0.12);
  SurfaceTensionModel surface_tension = SurfaceTensionModel::CSF;
  real contact_angle_deg = 90;
  real contact_angle_wall_weight = real(0.01);
  bool wall_colour = true;
  WettingModel wetting_model = WettingModel::NormalRotation;
  bool csf_smooth_normals = false;
  bool csf_balanced = false;
  real akinci_gamma = -1;
  real akinci_beta = -1;
  bool thermal = true;
  bool phase_change = true;
  WallThermalBC wall_thermal_bc = WallThermalBC::Adiabatic;
  real heater_power = 0;
  real conductivity_scale = 1;
  real lee_evaporation = 10;
  real lee_condensation = 10;
  TankThermoMode thermo_mode = TankThermoMode::Vented;
  real vent_relaxation = real(0.02);
  real merge_ratio = real(0.3);
  real split_ratio = real(2.0);
  real max_transfer_fraction = real(0.25);
  real phase_change_density_limit = real(0.01);
  real vapor_mass_rebalance = real(1.0);
  real cfl = real(0.25);
  real viscous_number = real(0.125);
  real force_number = real(0.25);
  real capillary_number = real(0.25);
  real thermal_number = real(0.1);
  real dt_max = real(1e-2);
  real dt_min = real(1e-9);
  real fixed_dt = 0;
  // Extension without a reference counterpart: the precision form
  // (double-float positions, compensated FP32 neighbour sums, the summation
  // density's EOS in double), compared with the FP64 build.
  bool compensated_sums = false;

  cryo_config to_c() const {
    cryo_config c{};
    c.dx = dx, c.h_factor = h_factor, c.kernel = static_cast<int32_t>(kernel);
    c.density_mode = static_cast<int32_t>(density_mode), c.normalize_initial_masses = normalize_initial_masses;
    c.gravity = cryosph::to_c(gravity), c.velocity_scale = velocity_scale, c.sound_speed = sound_speed;
    c.tait_gamma = tait_gamma, c.background_pressure = background_pressure, c.vapor_stiffness = vapor_stiffness;
    c.clamp_negative_pressure = clamp_negative_pressure, c.transport_velocity = transport_velocity;
    c.transport_background_factor = transport_background_factor, c.transport_stress = transport_stress;
    c.transport_interface_only = transport_interface_only, c.particle_shifting = particle_shifting;
    c.shifting_strength = shifting_strength, c.artificial_viscosity = artificial_viscosity;
    c.artificial_viscosity_vapor = artificial_viscosity_vapor, c.delta_sph = delta_sph, c.no_slip = no_slip;
    c.sps_turbulence = sps_turbulence, c.smagorinsky_cs = smagorinsky_cs;
    c.surface_tension = static_cast<int32_t>(surface_tension), c.contact_angle_deg = contact_angle_deg;
    c.contact_angle_wall_weight = contact_angle_wall_weight, c.wall_colour = wall_colour;
    c.wetting_model = static_cast<int32_t>(wetting_model), c.csf_smooth_normals = csf_smooth_normals;
    c.csf_balanced = csf_balanced, c.akinci_gamma = akinci_gamma, c.akinci_beta = akinci_beta;
    c.thermal = thermal, c.phase_change = phase_change, c.wall_thermal_bc = static_cast<int32_t>(wall_thermal_bc);
    c.heater_power = heater_power, c.conductivity_scale = conductivity_scale;
    c.lee_evaporation = lee_evaporation, c.lee_condensation = lee_condensation;
    c.thermo_mode = static_cast<int32_t>(thermo_mode), c.vent_relaxation = vent_relaxation;
    c.merge_ratio = merge_ratio, c.split_ratio = split_ratio, c.max_transfer_fraction = max_transfer_fraction;
    c.phase_change_density_limit = phase_change_density_limit, c.vapor_mass_rebalance = vapor_mass_rebalance;
    c.cfl = cfl, c.viscous_number = viscous_number, c.force_number = force_number;
    c.capillary_number = capillary_number, c.thermal_number = thermal_number;
    c.dt_max = dt_max, c.dt_min = dt_min, c.fixed_dt = fixed_dt, c.compensated_sums = compensated_sums;
    return c;
  }
};

struct ManeuverProfile {
  std::vector<double> time;
  std::vector<Vec3> acceleration;
  std::vector<Vec3> angular_velocity;
  Vec3 origin{0, 0, 0};

  void evaluate(double t, Vec3& a, Vec3& omega, Vec3& omega_dot) const {
    validate();
    const std::vector<float> acc = detail::flat(acceleration), om = detail::flat(angular_velocity);
    float ra[3], rw[3], rwd[3];
    detail::check(cryo_maneuver_evaluate(static_cast<int32_t>(time.size()), time.data(), acc.data(),
                                         om.empty() ? nullptr : om.data(), t, ra, rw, rwd));
    a = Vec3{ra[0], ra[1], ra[2]}, omega = Vec3{rw[0], rw[1], rw[2]}, omega_dot = Vec3{rwd[0], rwd[1], rwd[2]};
  }
  void validate() const {
    if (time.empty()) throw std::invalid_argument("ManeuverProfile: empty time table");
    if (acceleration.size() != time.size())
      throw std::invalid_argument("ManeuverProfile: acceleration must have one entry per time");
    if (!angular_velocity.empty() && angular_velocity.size() != time.size())
      throw std::invalid_argument("ManeuverProfile: angular_velocity must be empty or have one entry per time");
    for (std::size_t k = 1; k < time.size(); ++k)
      if (!(time[k] > time[k - 1])) throw std::invalid_argument("ManeuverProfile: time must be strictly increasing");
  }
};

struct Diagnostics {
  double time = 0;
  long long step = 0;
  double dt = 0;
  int n_total = 0, n_liquid = 0, n_vapor = 0, n_wall = 0;
  double mass_liquid = 0, mass_vapor = 0, mass_total = 0;
  double kinetic_energy = 0;
  double enthalpy = 0;
  double heat_input = 0;
  double evaporated_mass = 0;
  double condensed_mass = 0;
  double vented_mass = 0;
  double vented_enthalpy = 0;
  double T_sat = 0;
  double p_ullage = 0;
  double max_speed = 0;
  double T_liquid_mean = 0, T_liquid_max = 0, T_vapor_mean = 0;
  Vec3 liquid_com{0, 0, 0};
  Vec3 liquid_momentum{0, 0, 0};
  int splits = 0, merges = 0;
  Vec3 slosh_force{0, 0, 0};
  Vec3 slosh_moment{0, 0, 0};
  Vec3 frame_acceleration{0, 0, 0};
  Vec3 frame_angular_velocity{0, 0, 0};

  static Diagnostics from_c(const cryo_diagnostics& c) {
    Diagnostics d;
    d.time = c.time, d.step = c.step, d.dt = c.dt;
    d.n_total = c.n_total, d.n_liquid = c.n_liquid, d.n_vapor = c.n_vapor, d.n_wall = c.n_wall;
    d.mass_liquid = c.mass_liquid, d.mass_vapor = c.mass_vapor, d.mass_total = c.mass_total;
    d.kinetic_energy = c.kinetic_energy, d.enthalpy = c.enthalpy, d.heat_input = c.heat_input;
    d.evaporated_mass = c.evaporated_mass, d.condensed_mass = c.condensed_mass;
    d.vented_mass = c.vented_mass, d.vented_enthalpy = c.vented_enthalpy;
    d.T_sat = c.T_sat, d.p_ullage = c.p_ullage, d.max_speed = c.max_speed;
    d.T_liquid_mean = c.T_liquid_mean, d.T_liquid_max = c.T_liquid_max, d.T_vapor_mean = c.T_vapor_mean;
    d.liquid_com = from_c(c.liquid_com), d.liquid_momentum = from_c(c.liquid_momentum);
    d.splits = c.splits, d.merges = c.merges;
    d.slosh_force = from_c(c.slosh_force), d.slosh_moment = from_c(c.slosh_moment);
    d.frame_acceleration = from_c(c.frame_acceleration);
    d.frame_angular_velocity = from_c(c.frame_angular_velocity);
    return d;
  }

 private:
  static Vec3 from_c(cryo_vec3 a) { return cryosph::from_c(a); }
};

class Solver {
 public:
  Solver(const SolverConfig& config, const FluidProperties& fluid) : cfg_(config), fluid_(fluid) {
    const cryo_config c = config.to_c();
    const cryo_fluid f = fluid.to_c();
    detail::check(cryo_solver_create(&c, &f, &s_));
    ps_.owner_ = s_;
    ps_.owner_domain_ = &domain_;
    telemetry_.reset(new TelemetryStream(s_));
    grid_.owner_ = s_;
  }
  ~Solver() {
    if (s_) cryo_solver_destroy(s_);
  }
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;

  void set_particles(const HostParticles& p, const AABB& domain) {
    const std::vector<float> x = detail::flat(p.x), v = detail::flat(p.v), nr = detail::flat(p.nrm);
    const float lo[3] = {domain.lo.x, domain.lo.y, domain.lo.z}, hi[3] = {domain.hi.x, domain.hi.y, domain.hi.z};
    detail::check(cryo_solver_set_particles(s_, static_cast<int32_t>(p.size()), x.data(), v.data(), nr.data(),
                                            p.m.data(), p.rho.data(), p.T.data(), p.phase.data(), p.flags.data(),
                                            p.id.data(), lo, hi));
    domain_ = domain;
  }
  void build_tank(const Geometry& geometry, const FillSpec& fill) {
    GeoList g;
    geometry.emit(g);
    const cryo_fill f = fill.to_c();
    detail::check(cryo_solver_build_tank_cb(s_, static_cast<int32_t>(g.nodes.size()), g.nodes.data(),
                                            static_cast<int32_t>(g.meshes.size()), g.meshes.data(),
                                            static_cast<int32_t>(g.callbacks.size()), g.callbacks.data(), &f));
    const AABB b = geometry.bounds();
    const real pad = static_cast<real>(fill.wall_layers + 1) * cfg_.dx;
    domain_ = AABB{b.lo - Vec3{pad, pad, pad}, b.hi + Vec3{pad, pad, pad}};
  }
  void step() { detail::check(cryo_solver_step(s_)); }
  void run_steps(long long n) { detail::check(cryo_solver_run_steps(s_, n)); }
  void run_until(double t_end) { detail::check(cryo_solver_run_until(s_, t_end)); }
  double time() const {
    double t = 0;
    detail::check(cryo_solver_time(s_, &t));
    return t;
  }
  long long step_count() const {
    int64_t k = 0;
    detail::check(cryo_solver_step_count(s_, &k));
    return k;
  }
  double last_dt() const {
    double dt = 0;
    detail::check(cryo_solver_last_dt(s_, &dt));
    return dt;
  }
  double next_dt() {
    double dt = 0;
    detail::check(cryo_solver_next_dt(s_, &dt));
    return dt;
  }
  Diagnostics diagnostics() {
    cryo_diagnostics d{};
    detail::check(cryo_solver_diagnostics(s_, &d));
    return Diagnostics::from_c(d);
  }
  HostParticles particles() const { return ps_.download(); }
  std::vector<real> pressure() const { return field(CRYO_F_P); }
  std::vector<real> curvature() const { return field(CRYO_F_KAPPA); }
  std::size_t num_particles() const { return ps_.size(); }

  void set_gravity(Vec3 g) {
    const float a[3] = {g.x, g.y, g.z};
    detail::check(cryo_solver_set_gravity(s_, a));
    cfg_.gravity = g;
  }
  void set_frame_motion(Vec3 acceleration, Vec3 angular_velocity, Vec3 angular_acceleration, Vec3 origin) {
    const float a[3] = {acceleration.x, acceleration.y, acceleration.z};
    const float w[3] = {angular_velocity.x, angular_velocity.y, angular_velocity.z};
    const float wd[3] = {angular_acceleration.x, angular_acceleration.y, angular_acceleration.z};
    const float o[3] = {origin.x, origin.y, origin.z};
    detail::check(cryo_solver_set_frame_motion(s_, a, w, wd, o));
  }
  void set_maneuver(const ManeuverProfile& profile) {
    profile.validate();
    const std::vector<float> acc = detail::flat(profile.acceleration), om = detail::flat(profile.angular_velocity);
    const float o[3] = {profile.origin.x, profile.origin.y, profile.origin.z};
    detail::check(cryo_solver_set_maneuver(s_, static_cast<int32_t>(profile.time.size()), profile.time.data(),
                                           acc.data(), om.empty() ? nullptr : om.data(), o));
  }
  void clear_frame_motion() { detail::check(cryo_solver_clear_frame_motion(s_)); }
  void set_heater_power(real watts) { detail::check(cryo_solver_set_heater_power(s_, watts)); }
  void set_contact_angle(real degrees) { detail::check(cryo_solver_set_contact_angle(s_, degrees)); }
  void damp_velocities(real factor) { detail::check(cryo_solver_damp_velocities(s_, factor)); }
  void relax(double duration, real damping = real(0.98)) { detail::check(cryo_solver_relax(s_, duration, damping)); }

  const SolverConfig& config() const noexcept { return cfg_; }
  const FluidProperties& fluid() const noexcept { return fluid_; }
  SimParams params() const {
    SimParams p{};
    detail::check(cryo_solver_params(s_, &p));
    return p;
  }
  ParticleSystem& particle_system() noexcept { return ps_; }
  SpatialHash& grid() noexcept { return grid_; }
  TelemetryStream& telemetry() noexcept { return *telemetry_; }
  stream_t stream() const {
    void* h = nullptr;
    detail::check(cryo_solver_stream(s_, &h));
    return h;
  }
  std::string device() const { return mem::device_name(); }
  bool request_telemetry() { return telemetry_->request(); }
  std::vector<real> probe_pressure(const std::vector<Vec3>& points) {
    const std::vector<float> pts = detail::flat(points);
    std::vector<real> out(points.size());
    detail::check(cryo_solver_probe_pressure(s_, static_cast<int32_t>(points.size()), pts.data(), out.data()));
    return out;
  }
  // The C ABI handle, for the entry points with no reference counterpart
  // (ensembles, the adjoint, launch variants).
  cryo_solver* handle() const noexcept { return s_; }

 private:
  std::vector<real> field(int f) const {
    std::vector<real> out(num_particles());
    detail::check(cryo_solver_download(s_, f, out.data(), static_cast<int32_t>(out.size())));
    return out;
  }
  SolverConfig cfg_;
  FluidProperties fluid_;
  cryo_solver* s_ = nullptr;
  AABB domain_{};
  ParticleSystem ps_;
  SpatialHash grid_;
  std::unique_ptr<TelemetryStream> telemetry_;
};

// ------------------------------------------------------------------ particle system members
inline std::size_t ParticleSystem::size() const {
  if (!owner_) return host_.size();
  int32_t n = 0;
  detail::check(cryo_solver_num_particles(owner_, &n));
  return static_cast<std::size_t>(n);
}
inline std::size_t ParticleSystem::capacity() const {
  if (!owner_) return std::max(capacity_, host_.size());
  int32_t c = 0;
  detail::check(cryo_solver_capacity(owner_, &c));
  return static_cast<std::size_t>(c);
}
inline void ParticleSystem::upload(const HostParticles& host, double growth_factor) {
  if (owner_) {
    const std::vector<float> x = detail::flat(host.x), v = detail::flat(host.v), nr = detail::flat(host.nrm);
    const AABB d = *owner_domain_;
    const float lo[3] = {d.lo.x, d.lo.y, d.lo.z}, hi[3] = {d.hi.x, d.hi.y, d.hi.z};
    detail::check(cryo_solver_set_particles(owner_, static_cast<int32_t>(host.size()), x.data(), v.data(), nr.data(),
                                            host.m.data(), host.rho.data(), host.T.data(), host.phase.data(),
                                            host.flags.data(), host.id.data(), lo, hi));
    return;
  }
  host_ = host;
  // set_particles semantics (particle_system.cpp:74-87): capacity grows to
  // max(growth * n, n + 1024), ids continue after the largest one
  const std::size_t n = host.size();
  capacity_ = std::max(capacity_, std::max(static_cast<std::size_t>(growth_factor * static_cast<double>(n)), n + 1024));
  next_id = 0;
  for (std::uint32_t k : host.id) next_id = std::max(next_id, k + 1);
}
inline HostParticles ParticleSystem::download() const {
  if (!owner_) return host_;
  const std::size_t n = size();
  HostParticles hp;
  std::vector<float> x(3 * n), v(3 * n), nr(3 * n);
  hp.m.resize(n), hp.rho.resize(n), hp.T.resize(n), hp.phase.resize(n), hp.flags.resize(n), hp.id.resize(n);
  const int32_t cnt = static_cast<int32_t>(n);
  detail::check(cryo_solver_download(owner_, CRYO_F_X, x.data(), cnt));
  detail::check(cryo_solver_download(owner_, CRYO_F_V, v.data(), cnt));
  detail::check(cryo_solver_download(owner_, CRYO_F_NRM, nr.data(), cnt));
  detail::check(cryo_solver_download(owner_, CRYO_F_M, hp.m.data(), cnt));
  detail::check(cryo_solver_download(owner_, CRYO_F_RHO, hp.rho.data(), cnt));
  detail::check(cryo_solver_download(owner_, CRYO_F_T, hp.T.data(), cnt));
  detail::check(cryo_solver_download(owner_, CRYO_F_PHASE, hp.phase.data(), cnt));
  detail::check(cryo_solver_download(owner_, CRYO_F_FLAGS, hp.flags.data(), cnt));
  detail::check(cryo_solver_download(owner_, CRYO_F_ID, hp.id.data(), cnt));
  hp.x = detail::unflat(x), hp.v = detail::unflat(v), hp.nrm = detail::unflat(nr);
  return hp;
}
inline std::vector<real> ParticleSystem::download_pressure() const {
  if (!owner_) return std::vector<real>(host_.size(), 0.0f);
  std::vector<real> out(size());
  detail::check(cryo_solver_download(owner_, CRYO_F_P, out.data(), static_cast<int32_t>(out.size())));
  return out;
}
inline std::vector<real> ParticleSystem::download_curvature() const {
  if (!owner_) return std::vector<real>(host_.size(), 0.0f);
  std::vector<real> out(size());
  detail::check(cryo_solver_download(owner_, CRYO_F_KAPPA, out.data(), static_cast<int32_t>(out.size())));
  return out;
}

}  // namespace cryosph
