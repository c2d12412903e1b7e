// C-ABI shim over the reference CryoSPH CPU/OpenMP build (test infrastructure).
//
// Compiled by oracle/Makefile against the reference headers where they lie
// (include/cryosph/...) and linked with the reference's own sources built as
// C++ (CMakeLists.txt:68-88 CPU path), once with real=float and once with
// real=double (CRYOSPH_DOUBLE_PRECISION, cuda_utils.cuh:50-55).  Output goes
// to oracle/_ref/ only.  Nothing in the product links or loads this; tests use
// it to pin the oracle restatement and the HIP path against the reference.
#include <omp.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cryosph.h"
#include "cryosph/core/geometry.hpp"
#include "cryosph/core/io.hpp"
#include "cryosph/core/particle_system.hpp"
#include "cryosph/core/spatial_hash.hpp"
#include "cryosph/physics/fluid_properties.hpp"
#include "cryosph/solvers/sph_solver.hpp"

using namespace cryosph;

namespace {

thread_local std::string g_err;

int fail(const std::exception& e, int code) {
  g_err = e.what();
  return code;
}

#define GUARD(...)                                               \
  try {                                                           \
    __VA_ARGS__;                                            \
    return 0;                                                     \
  } catch (const std::invalid_argument& e) {                      \
    return fail(e, CRYO_ERR_INVALID_ARGUMENT);                    \
  } catch (const std::out_of_range& e) {                          \
    return fail(e, CRYO_ERR_OUT_OF_RANGE);                        \
  } catch (const std::exception& e) {                             \
    return fail(e, CRYO_ERR_RUNTIME);                             \
  }

Vec3 v3(cryo_vec3 a) { return Vec3{real(a.x), real(a.y), real(a.z)}; }
Vec3 v3p(const float* a) { return Vec3{real(a[0]), real(a[1]), real(a[2])}; }
cryo_vec3 cv(Vec3 a) { return cryo_vec3{float(a.x), float(a.y), float(a.z)}; }

SolverConfig to_cfg(const cryo_config& c) {
  SolverConfig o;
  o.dx = c.dx;
  o.h_factor = c.h_factor;
  o.kernel = static_cast<KernelType>(c.kernel);
  o.density_mode = static_cast<DensityMode>(c.density_mode);
  o.normalize_initial_masses = c.normalize_initial_masses != 0;
  o.gravity = v3(c.gravity);
  o.velocity_scale = c.velocity_scale;
  o.sound_speed = c.sound_speed;
  o.tait_gamma = c.tait_gamma;
  o.background_pressure = c.background_pressure;
  o.vapor_stiffness = c.vapor_stiffness;
  o.clamp_negative_pressure = c.clamp_negative_pressure != 0;
  o.transport_velocity = c.transport_velocity != 0;
  o.transport_background_factor = c.transport_background_factor;
  o.transport_stress = c.transport_stress != 0;
  o.transport_interface_only = c.transport_interface_only != 0;
  o.particle_shifting = c.particle_shifting != 0;
  o.shifting_strength = c.shifting_strength;
  o.artificial_viscosity = c.artificial_viscosity;
  o.artificial_viscosity_vapor = c.artificial_viscosity_vapor;
  o.delta_sph = c.delta_sph;
  o.no_slip = c.no_slip != 0;
  o.sps_turbulence = c.sps_turbulence != 0;
  o.smagorinsky_cs = c.smagorinsky_cs;
  o.surface_tension = static_cast<SurfaceTensionModel>(c.surface_tension);
  o.contact_angle_deg = c.contact_angle_deg;
  o.contact_angle_wall_weight = c.contact_angle_wall_weight;
  o.wall_colour = c.wall_colour != 0;
  o.wetting_model = static_cast<WettingModel>(c.wetting_model);
  o.csf_smooth_normals = c.csf_smooth_normals != 0;
  o.csf_balanced = c.csf_balanced != 0;
  o.akinci_gamma = c.akinci_gamma;
  o.akinci_beta = c.akinci_beta;
  o.thermal = c.thermal != 0;
  o.phase_change = c.phase_change != 0;
  o.wall_thermal_bc = static_cast<WallThermalBC>(c.wall_thermal_bc);
  o.heater_power = c.heater_power;
  o.conductivity_scale = c.conductivity_scale;
  o.lee_evaporation = c.lee_evaporation;
  o.lee_condensation = c.lee_condensation;
  o.thermo_mode = static_cast<TankThermoMode>(c.thermo_mode);
  o.vent_relaxation = c.vent_relaxation;
  o.merge_ratio = c.merge_ratio;
  o.split_ratio = c.split_ratio;
  o.max_transfer_fraction = c.max_transfer_fraction;
  o.phase_change_density_limit = c.phase_change_density_limit;
  o.vapor_mass_rebalance = c.vapor_mass_rebalance;
  o.cfl = c.cfl;
  o.viscous_number = c.viscous_number;
  o.force_number = c.force_number;
  o.capillary_number = c.capillary_number;
  o.thermal_number = c.thermal_number;
  o.dt_max = c.dt_max;
  o.dt_min = c.dt_min;
  o.fixed_dt = c.fixed_dt;
  return o;
}

FluidProperties to_fluid(const cryo_fluid& f) {
  FluidProperties o;
  o.name = f.name;
  o.rho_liquid = f.rho_liquid;
  o.rho_vapor = f.rho_vapor;
  o.mu_liquid = f.mu_liquid;
  o.mu_vapor = f.mu_vapor;
  o.k_liquid = f.k_liquid;
  o.k_vapor = f.k_vapor;
  o.cp_liquid = f.cp_liquid;
  o.cp_vapor = f.cp_vapor;
  o.sigma = f.sigma;
  o.h_fg = f.h_fg;
  o.T_sat = f.T_sat;
  o.p_sat = f.p_sat;
  o.R_gas = f.R_gas;
  return o;
}

void from_fluid(const FluidProperties& f, cryo_fluid* o) {
  std::memset(o, 0, sizeof(*o));
  std::strncpy(o->name, f.name.c_str(), sizeof(o->name) - 1);
  o->rho_liquid = float(f.rho_liquid);
  o->rho_vapor = float(f.rho_vapor);
  o->mu_liquid = float(f.mu_liquid);
  o->mu_vapor = float(f.mu_vapor);
  o->k_liquid = float(f.k_liquid);
  o->k_vapor = float(f.k_vapor);
  o->cp_liquid = float(f.cp_liquid);
  o->cp_vapor = float(f.cp_vapor);
  o->sigma = float(f.sigma);
  o->h_fg = float(f.h_fg);
  o->T_sat = float(f.T_sat);
  o->p_sat = float(f.p_sat);
  o->R_gas = float(f.R_gas);
}

std::shared_ptr<const Geometry> build_geo(int n_nodes, const cryo_geo_node* nodes, int n_meshes,
                                          const cryo_mesh* meshes) {
  if (n_nodes <= 0) throw std::invalid_argument("geometry: no nodes");
  std::vector<std::shared_ptr<const Geometry>> g(static_cast<std::size_t>(n_nodes));
  for (int k = 0; k < n_nodes; ++k) {
    const cryo_geo_node& nd = nodes[k];
    auto child = [&](int c) {
      if (c < 0 || c >= k) throw std::invalid_argument("geometry: bad child index");
      return g[static_cast<std::size_t>(c)];
    };
    switch (nd.type) {
      case CRYO_GEO_CYLINDER: g[k] = std::make_shared<CylinderTank>(nd.p[0], nd.p[1]); break;
      case CRYO_GEO_CAPSULE: g[k] = std::make_shared<CapsuleTank>(nd.p[0], nd.p[1]); break;
      case CRYO_GEO_SPHERE: g[k] = std::make_shared<SphereTank>(nd.p[0]); break;
      case CRYO_GEO_BOX: g[k] = std::make_shared<BoxTank>(nd.p[0], nd.p[1], nd.p[2]); break;
      case CRYO_GEO_TRANSLATED:
        g[k] = std::make_shared<Translated>(child(nd.a), Vec3{real(nd.p[0]), real(nd.p[1]), real(nd.p[2])});
        break;
      case CRYO_GEO_DIFFERENCE: g[k] = std::make_shared<Difference>(child(nd.a), child(nd.b)); break;
      case CRYO_GEO_UNION: g[k] = std::make_shared<Union>(child(nd.a), child(nd.b)); break;
      case CRYO_GEO_MESH: {
        if (nd.mesh_id < 0 || nd.mesh_id >= n_meshes) throw std::invalid_argument("geometry: bad mesh id");
        const cryo_mesh& m = meshes[nd.mesh_id];
        std::vector<std::array<double, 3>> v(static_cast<std::size_t>(m.n_vertices));
        std::vector<std::array<int, 3>> t(static_cast<std::size_t>(m.n_triangles));
        for (int q = 0; q < m.n_vertices; ++q) v[q] = {m.vertices[3 * q], m.vertices[3 * q + 1], m.vertices[3 * q + 2]};
        for (int q = 0; q < m.n_triangles; ++q)
          t[q] = {m.triangles[3 * q], m.triangles[3 * q + 1], m.triangles[3 * q + 2]};
        g[k] = std::make_shared<MeshTank>(std::move(v), std::move(t));
        break;
      }
      default: throw std::invalid_argument("geometry: unknown node type");
    }
  }
  return g.back();
}

FillSpec to_fill(const cryo_fill& f) {
  FillSpec o;
  o.mode = static_cast<FillMode>(f.mode);
  o.fill_fraction = f.fill_fraction;
  o.settle_direction = v3(f.settle_direction);
  o.bubble_center = v3(f.bubble_center);
  o.include_vapor = f.include_vapor != 0;
  o.wall_layers = f.wall_layers;
  o.T_liquid = f.T_liquid;
  o.T_vapor = f.T_vapor;
  o.T_wall = f.T_wall;
  o.heater_zmin = f.heater_zmin;
  o.heater_zmax = f.heater_zmax;
  return o;
}

template <class T>
void put(void* dst, const T* src, int n) {
  std::memcpy(dst, src, sizeof(T) * static_cast<std::size_t>(n));
}

}  // namespace

extern "C" {

const char* ref_last_error(void) { return g_err.c_str(); }
int ref_real_size(void) { return static_cast<int>(sizeof(real)); }

int ref_create(const cryo_config* cfg, const cryo_fluid* fluid, void** out) {
  GUARD(*out = new Solver(to_cfg(*cfg), to_fluid(*fluid)))
}
void ref_destroy(void* s) { delete static_cast<Solver*>(s); }

int ref_set_particles(void* sp, int n, const float* x, const float* v, const float* nrm, const float* m,
                      const float* rho, const float* T, const uint8_t* phase, const uint8_t* flags,
                      const uint32_t* id, const float* lo, const float* hi) {
  GUARD({
    HostParticles hp;
    hp.reserve(static_cast<std::size_t>(n));
    for (int i = 0; i < n; ++i) {
      hp.push(v3p(x + 3 * i), v ? v3p(v + 3 * i) : Vec3{0, 0, 0}, real(m[i]), real(rho[i]), real(T[i]),
              static_cast<Phase>(phase[i]), nrm ? v3p(nrm + 3 * i) : Vec3{0, 0, 0}, flags ? flags[i] : 0);
      if (id) hp.id.back() = id[i];
    }
    static_cast<Solver*>(sp)->set_particles(hp, AABB{v3p(lo), v3p(hi)});
  })
}

int ref_build_tank(void* sp, int n_nodes, const cryo_geo_node* nodes, int n_meshes, const cryo_mesh* meshes,
                   const cryo_fill* fill) {
  GUARD(static_cast<Solver*>(sp)->build_tank(*build_geo(n_nodes, nodes, n_meshes, meshes), to_fill(*fill)))
}

int ref_step(void* s) { GUARD(static_cast<Solver*>(s)->step()) }
int ref_run_steps(void* s, long long n) { GUARD(static_cast<Solver*>(s)->run_steps(n)) }
int ref_run_until(void* s, double t) { GUARD(static_cast<Solver*>(s)->run_until(t)) }
int ref_relax(void* s, double d, double damping) { GUARD(static_cast<Solver*>(s)->relax(d, real(damping))) }
int ref_next_dt(void* s, double* dt) { GUARD(*dt = static_cast<Solver*>(s)->next_dt()) }
int ref_damp_velocities(void* s, double f) { GUARD(static_cast<Solver*>(s)->damp_velocities(real(f))) }
double ref_time(void* s) { return static_cast<Solver*>(s)->time(); }
long long ref_step_count(void* s) { return static_cast<Solver*>(s)->step_count(); }
double ref_last_dt(void* s) { return static_cast<Solver*>(s)->last_dt(); }
int ref_num_particles(void* s) { return static_cast<int>(static_cast<Solver*>(s)->num_particles()); }

int ref_set_gravity(void* s, const float* g) { GUARD(static_cast<Solver*>(s)->set_gravity(v3p(g))) }
int ref_set_frame_motion(void* s, const float* a, const float* w, const float* wd, const float* o) {
  GUARD(static_cast<Solver*>(s)->set_frame_motion(v3p(a), v3p(w), v3p(wd), v3p(o)))
}
int ref_set_maneuver(void* s, int n, const double* t, const float* a, const float* w, const float* o) {
  GUARD({
    ManeuverProfile p;
    p.time.assign(t, t + n);
    for (int k = 0; k < n; ++k) p.acceleration.push_back(v3p(a + 3 * k));
    if (w)
      for (int k = 0; k < n; ++k) p.angular_velocity.push_back(v3p(w + 3 * k));
    p.origin = v3p(o);
    static_cast<Solver*>(s)->set_maneuver(p);
  })
}
int ref_clear_frame_motion(void* s) { GUARD(static_cast<Solver*>(s)->clear_frame_motion()) }
int ref_set_heater_power(void* s, double w) { GUARD(static_cast<Solver*>(s)->set_heater_power(real(w))) }
int ref_set_contact_angle(void* s, double d) { GUARD(static_cast<Solver*>(s)->set_contact_angle(real(d))) }

int ref_params(void* sp, cryo_params* o) {
  GUARD({
    const SimParams& P = static_cast<Solver*>(sp)->params();
    std::memset(o, 0, sizeof(*o));
    o->h = float(P.kernel.h);
    o->inv_h = float(P.kernel.inv_h);
    o->support = float(P.kernel.support);
    o->kernel_norm = float(P.kernel.norm);
    o->support2 = float(P.support2);
    o->dx = float(P.dx);
    o->vol0 = float(P.vol0);
    for (int k = 0; k < 2; ++k) {
      o->rho0[k] = float(P.rho0[k]);
      o->mu[k] = float(P.mu[k]);
      o->k[k] = float(P.k[k]);
      o->cp[k] = float(P.cp[k]);
      o->mref[k] = float(P.mref[k]);
    }
    o->c0 = float(P.c0);
    o->tait_B = float(P.tait_B);
    o->tait_gamma = float(P.tait_gamma);
    o->p_bg = float(P.p_bg);
    o->gas_R = float(P.gas_R);
    o->vapor_K = float(P.vapor_K);
    o->gravity = cv(P.gravity);
    o->omega = cv(P.omega);
    o->omega_dot = cv(P.omega_dot);
    o->frame_origin = cv(P.frame_origin);
    o->frame_rot = P.frame_rot;
    o->sigma = float(P.sigma);
    o->cos_theta = float(P.cos_theta);
    o->sin_theta = float(P.sin_theta);
    o->normal_eps = float(P.normal_eps);
    o->akinci_gamma = float(P.akinci_gamma);
    o->akinci_beta = float(P.akinci_beta);
    o->tvf_sigma_min = float(P.tvf_sigma_min);
    o->shift_strength = float(P.shift_strength);
    o->T_ref = float(P.T_ref);
    o->T_sat = float(P.T_sat);
    o->h_fg = float(P.h_fg);
    o->heat_inv_norm = float(P.heat_inv_norm);
    o->dt = float(P.dt);
  })
}

int ref_diagnostics(void* sp, cryo_diagnostics* o) {
  GUARD({
    const Diagnostics d = static_cast<Solver*>(sp)->diagnostics();
    std::memset(o, 0, sizeof(*o));
    o->time = d.time;
    o->step = d.step;
    o->dt = d.dt;
    o->n_total = d.n_total;
    o->n_liquid = d.n_liquid;
    o->n_vapor = d.n_vapor;
    o->n_wall = d.n_wall;
    o->mass_liquid = d.mass_liquid;
    o->mass_vapor = d.mass_vapor;
    o->mass_total = d.mass_total;
    o->kinetic_energy = d.kinetic_energy;
    o->enthalpy = d.enthalpy;
    o->heat_input = d.heat_input;
    o->evaporated_mass = d.evaporated_mass;
    o->condensed_mass = d.condensed_mass;
    o->vented_mass = d.vented_mass;
    o->vented_enthalpy = d.vented_enthalpy;
    o->T_sat = d.T_sat;
    o->p_ullage = d.p_ullage;
    o->max_speed = d.max_speed;
    o->T_liquid_mean = d.T_liquid_mean;
    o->T_liquid_max = d.T_liquid_max;
    o->T_vapor_mean = d.T_vapor_mean;
    o->liquid_com = cv(d.liquid_com);
    o->liquid_momentum = cv(d.liquid_momentum);
    o->splits = d.splits;
    o->merges = d.merges;
    o->slosh_force = cv(d.slosh_force);
    o->slosh_moment = cv(d.slosh_moment);
    o->frame_acceleration = cv(d.frame_acceleration);
    o->frame_angular_velocity = cv(d.frame_angular_velocity);
  })
}

// Field download in the build's native `real` (float or double); vec3 fields
// as 3 reals per particle, phase/flags as uint8, id as uint32.
int ref_download(void* sp, int field, void* dst, int count) {
  GUARD({
    Solver& s = *static_cast<Solver*>(sp);
    ParticleSystem& ps = s.particle_system();
    if (count < 0 || static_cast<std::size_t>(count) > ps.size()) throw std::out_of_range("ref_download: count");
    const ParticleView pv = ps.view();
    const Vec3* vf = nullptr;
    const real* rf = nullptr;
    switch (field) {
      case CRYO_F_X: vf = pv.x; break;
      case CRYO_F_V: vf = pv.v; break;
      case CRYO_F_VT: vf = pv.vt; break;
      case CRYO_F_NRM: vf = pv.nrm; break;
      case CRYO_F_A: vf = pv.a; break;
      case CRYO_F_AB: vf = pv.ab; break;
      case CRYO_F_NS: vf = pv.ns; break;
      case CRYO_F_VW: vf = pv.vw; break;
      case CRYO_F_FW: vf = pv.fw; break;
      case CRYO_F_MW: vf = pv.mw; break;
      case CRYO_F_M: rf = pv.m; break;
      case CRYO_F_M0: rf = pv.m0; break;
      case CRYO_F_RHO: rf = pv.rho; break;
      case CRYO_F_T: rf = pv.T; break;
      case CRYO_F_P: rf = pv.p; break;
      case CRYO_F_DRHO: rf = pv.drho; break;
      case CRYO_F_DTDT: rf = pv.dTdt; break;
      case CRYO_F_SIGMA: rf = pv.sigma; break;
      case CRYO_F_KAPPA: rf = pv.kappa; break;
      case CRYO_F_NU_T: rf = pv.nu_t; break;
      case CRYO_F_ALPHA: rf = pv.alpha; break;
      case CRYO_F_WALL_W: rf = pv.wall_w; break;
      case CRYO_F_HEAT_W: rf = pv.heat_w; break;
      case CRYO_F_OUT_M: rf = pv.out_m; break;
      case CRYO_F_OUT_H: rf = pv.out_h; break;
      case CRYO_F_W_NORM: rf = pv.w_norm; break;
      case CRYO_F_P_WV: rf = pv.p_wv; break;
      case CRYO_F_RMIN: rf = pv.rmin; break;
      case CRYO_F_PHASE: put(dst, pv.phase, count); return 0;
      case CRYO_F_FLAGS: put(dst, pv.flags, count); return 0;
      case CRYO_F_ID: put(dst, pv.id, count); return 0;
      default: throw std::invalid_argument("ref_download: unknown field");
    }
    if (vf) put(dst, vf, count);
    else put(dst, rf, count);
  })
}

int ref_grid_info(void* sp, int* dims, float* origin, float* cell, int* num_cells) {
  GUARD({
    SpatialHash& g = static_cast<Solver*>(sp)->grid();
    const GridView gv = g.view();
    dims[0] = gv.nx;
    dims[1] = gv.ny;
    dims[2] = gv.nz;
    origin[0] = float(gv.origin.x);
    origin[1] = float(gv.origin.y);
    origin[2] = float(gv.origin.z);
    *cell = float(g.cell_size());
    *num_cells = g.num_cells();
  })
}

int ref_grid_tables(void* sp, int* cs, int* ce) {
  GUARD({
    SpatialHash& g = static_cast<Solver*>(sp)->grid();
    const GridView gv = g.view();
    std::memcpy(cs, gv.cell_start, sizeof(int) * static_cast<std::size_t>(g.num_cells()));
    std::memcpy(ce, gv.cell_end, sizeof(int) * static_cast<std::size_t>(g.num_cells()));
  })
}

int ref_neighbor_stats(void* sp, double radius, int* counts, unsigned long long* sums) {
  GUARD({
    Solver& s = *static_cast<Solver*>(sp);
    std::vector<int> c;
    std::vector<std::uint64_t> q;
    s.grid().neighbor_stats(s.particle_system(), real(radius), c, q, s.stream());
    std::memcpy(counts, c.data(), sizeof(int) * c.size());
    std::memcpy(sums, q.data(), sizeof(std::uint64_t) * q.size());
  })
}

// Neighbour sets via the reference's own enumerator (spatial_hash.hpp:37-56)
// on the solver's current cell lists, r^2 < radius^2.
int ref_neighbor_csr(void* sp, double radius, long long* offsets, int* indices) {
  GUARD({
    Solver& s = *static_cast<Solver*>(sp);
    ParticleSystem& ps = s.particle_system();
    const ParticleView pv = ps.view();
    const GridView g = s.grid().view();
    const real r2max = real(radius) * real(radius);
    long long k = 0;
    for (int i = 0; i < pv.n; ++i) {
      offsets[i] = k;
      const Vec3 xi = pv.x[i];
      for_each_neighbor(g, xi, [&](int j) {
        if (norm2(xi - pv.x[j]) < r2max) {
          if (indices) indices[k] = j;
          ++k;
        }
      });
    }
    offsets[pv.n] = k;
  })
}

int ref_hash_sort(int n, const float* x, const uint8_t* phase, const float* lo, const float* hi, double cell,
                  uint32_t* perm, int* n_alive, double radius, int* counts, unsigned long long* sums) {
  GUARD({
    HostParticles hp;
    for (int i = 0; i < n; ++i) hp.push(v3p(x + 3 * i), Vec3{0, 0, 0}, 1, 1, 1, static_cast<Phase>(phase[i]));
    ParticleSystem ps;
    ps.upload(hp);
    SpatialHash grid;
    grid.configure(v3p(lo), v3p(hi), real(cell));
    stream_t st = mem::create_stream();
    grid.build(ps, st);
    *n_alive = static_cast<int>(ps.size());
    const HostParticles out = ps.download();
    for (std::size_t i = 0; i < out.size(); ++i) perm[i] = out.id[i];
    if (counts) {
      std::vector<int> c;
      std::vector<std::uint64_t> q;
      grid.neighbor_stats(ps, real(radius), c, q, st);
      std::memcpy(counts, c.data(), sizeof(int) * c.size());
      std::memcpy(sums, q.data(), sizeof(std::uint64_t) * q.size());
    }
    mem::destroy_stream(st);
  })
}

int ref_fluid_preset(const char* name, cryo_fluid* out) {
  GUARD({
    const std::string n = name;
    FluidProperties f;
    if (n == "LH2") f = FluidProperties::liquid_hydrogen();
    else if (n == "LN2") f = FluidProperties::liquid_nitrogen();
    else if (n == "LOX") f = FluidProperties::liquid_oxygen();
    else if (n == "LCH4") f = FluidProperties::liquid_methane();
    else if (n == "H2O") f = FluidProperties::water();
    else if (n != "custom") throw std::invalid_argument("unknown preset");
    from_fluid(f, out);
  })
}

double ref_saturation_temperature(const cryo_fluid* f, double p) {
  return double(to_fluid(*f).saturation_temperature(real(p)));
}
double ref_saturation_pressure(const cryo_fluid* f, double T) {
  return double(to_fluid(*f).saturation_pressure(real(T)));
}

int ref_geo_sdf(int nn, const cryo_geo_node* nodes, int nm, const cryo_mesh* meshes, double x, double y, double z,
                double* out) {
  GUARD(*out = build_geo(nn, nodes, nm, meshes)->sdf(x, y, z))
}

// build_tank_particles (geometry.cpp:242-330) into caller arrays (native real
// for m/rho/T and x/v/nrm), count-only when x == NULL.
int ref_build_tank_particles(int nn, const cryo_geo_node* nodes, int nm, const cryo_mesh* meshes,
                             const cryo_fluid* fluid, double dx, const cryo_fill* fill, int* n, void* x, void* v,
                             void* nrm, void* m, void* rho, void* T, uint8_t* phase, uint8_t* flags,
                             uint32_t* id) {
  GUARD({
    const HostParticles hp =
        build_tank_particles(*build_geo(nn, nodes, nm, meshes), to_fluid(*fluid), real(dx), to_fill(*fill));
    const int cnt = static_cast<int>(hp.size());
    if (x) {
      if (*n < cnt) throw std::out_of_range("ref_build_tank_particles: arrays too small");
      put(x, hp.x.data(), cnt);
      put(v, hp.v.data(), cnt);
      put(nrm, hp.nrm.data(), cnt);
      put(m, hp.m.data(), cnt);
      put(rho, hp.rho.data(), cnt);
      put(T, hp.T.data(), cnt);
      put(phase, hp.phase.data(), cnt);
      put(flags, hp.flags.data(), cnt);
      put(id, hp.id.data(), cnt);
    }
    *n = cnt;
  })
}

// write_vtk (io.cpp:52-87) of float host arrays; pressure NULL = omitted.
int ref_write_vtk(const char* path, int n, const float* x, const float* v, const float* T, const float* m,
                  const float* rho, const uint8_t* phase, const float* pressure) {
  GUARD({
    HostParticles hp;
    for (int i = 0; i < n; ++i)
      hp.push(v3p(x + 3 * i), v3p(v + 3 * i), real(m[i]), real(rho[i]), real(T[i]), static_cast<Phase>(phase[i]));
    std::vector<real> p;
    if (pressure) p.assign(pressure, pressure + n);
    write_vtk(path, hp, p);
  })
}

// make_frame + write_json_frame (io.cpp:89-142) of float host arrays.
int ref_write_json_frame(const char* path, double time, long long step, int n, const float* x, const float* v,
                         const float* T, const float* pressure, const uint8_t* phase, int include_walls) {
  GUARD({
    HostParticles hp;
    for (int i = 0; i < n; ++i)
      hp.push(v3p(x + 3 * i), v3p(v + 3 * i), real(1), real(1), real(T[i]), static_cast<Phase>(phase[i]));
    std::vector<real> p;
    if (pressure) p.assign(pressure, pressure + n);
    write_json_frame(path, make_frame(hp, p, time, step), include_walls != 0);
  })
}

}  // extern "C"

extern "C" {
void ref_set_threads(int n) { omp_set_num_threads(n); }
void ref_kernel_eval(int type, double h, double r, double* W, double* dWdr) {
  const KernelParams k = KernelParams::make(static_cast<KernelType>(type), real(h));
  *W = double(kernel_W(k, real(r)));
  *dWdr = double(kernel_dWdr(k, real(r)));
}
}
