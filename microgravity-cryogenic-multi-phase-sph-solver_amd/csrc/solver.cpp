// Solver orchestration on the host and the C ABI of include/cryosph.h.
//
// Mirrors Solver (sph_solver.hpp:134-232, sph_solver.cu:273-756) with the
// step kept on the device: dt, the manoeuvre frame, bookkeeping and the
// ullage state live in DevState, so a step is a fixed sequence of launches on
// one blocking HIP stream.  The host synchronises only where it must know the
// particle count (phase change can split or merge particles) or the clock
// (run_until, relax), and when a caller reads results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "host.hpp"

namespace cryo {
namespace {

thread_local std::string g_err;

// Host pow used where the reference evaluates std::pow in float
// (akinci_adhesion, SPH_kernels.cuh:81-86).
float akinci_adh_host(float r, float c) {
  if (r >= c || 2.0f * r <= c) return 0.0f;
  const float arg = -4.0f * r * r / c + 6.0f * r - 2.0f * c;
  if (arg <= 0.0f) return 0.0f;
  return 0.007f / std::pow(c, 3.25f) * std::pow(arg, 0.25f);
}

// Akinci Fowler integral (sph_solver.cu:181-193).
template <class Fn>
double radial_moment4(Fn f, double c) {
  const int n = 2000;
  const double hstep = c / n;
  double acc = 0;
  for (int k = 0; k <= n; ++k) {
    const double r = k * hstep;
    const double w = (k == 0 || k == n) ? 1.0 : (k % 2 ? 4.0 : 2.0);
    acc += w * r * r * r * r * f(r);
  }
  return acc * hstep / 3.0;
}

float saturation_temperature(const cryo_fluid& f, float p) {  // fluid_properties.cpp:7-10
  const double inv_T = 1.0 / f.T_sat - (double(f.R_gas) / f.h_fg) * std::log(double(p) / f.p_sat);
  return static_cast<float>(1.0 / inv_T);
}
float saturation_pressure(const cryo_fluid& f, float T) {  // fluid_properties.cpp:12-14
  return static_cast<float>(f.p_sat * std::exp(double(f.h_fg) / f.R_gas * (1.0 / f.T_sat - 1.0 / double(T))));
}

bool preset(const std::string& n, cryo_fluid* o) {  // fluid_properties.cpp:16-79
  std::memset(o, 0, sizeof(*o));
  auto set = [&](const char* name, float rl, float rv, float ml, float mv, float kl, float kv, float cl, float cvv,
                 float sg, float hfg, float ts, float ps, float rg) {
    std::strncpy(o->name, name, sizeof(o->name) - 1);
    o->rho_liquid = rl;
    o->rho_vapor = rv;
    o->mu_liquid = ml;
    o->mu_vapor = mv;
    o->k_liquid = kl;
    o->k_vapor = kv;
    o->cp_liquid = cl;
    o->cp_vapor = cvv;
    o->sigma = sg;
    o->h_fg = hfg;
    o->T_sat = ts;
    o->p_sat = ps;
    o->R_gas = rg;
  };
  if (n == "custom") set("custom", 1000.f, 1.f, 1e-3f, 1e-5f, 0.6f, 0.025f, 4180.f, 2000.f, 0.072f, 2.26e6f, 373.15f, 101325.f, 461.5f);
  else if (n == "LH2") set("LH2", 70.85f, 1.338f, 1.33e-5f, 1.11e-6f, 0.1034f, 0.0166f, 9770.f, 12200.f, 1.93e-3f, 445.6e3f, 20.37f, 101325.f, 4124.2f);
  else if (n == "LN2") set("LN2", 806.1f, 4.61f, 1.61e-4f, 5.4e-6f, 0.1446f, 0.0074f, 2042.f, 1123.f, 8.85e-3f, 199.2e3f, 77.36f, 101325.f, 296.8f);
  else if (n == "LOX") set("LOX", 1141.f, 4.47f, 1.96e-4f, 6.9e-6f, 0.152f, 0.0084f, 1699.f, 972.f, 13.2e-3f, 213.1e3f, 90.19f, 101325.f, 259.8f);
  else if (n == "LCH4") set("LCH4", 422.4f, 1.816f, 1.17e-4f, 4.4e-6f, 0.184f, 0.0119f, 3481.f, 2218.f, 13.4e-3f, 510.8e3f, 111.67f, 101325.f, 518.3f);
  else if (n == "H2O") set("H2O", 998.2f, 0.598f, 1.0e-3f, 1.23e-5f, 0.598f, 0.0248f, 4182.f, 2080.f, 0.0728f, 2.257e6f, 373.15f, 101325.f, 461.5f);
  else return false;
  return true;
}

void config_default(cryo_config* c) {  // sph_solver.hpp:20-91
  std::memset(c, 0, sizeof(*c));
  c->dx = 0.005f;
  c->h_factor = 1.3f;
  c->kernel = CRYO_KERNEL_WENDLAND_C2;
  c->density_mode = CRYO_DENSITY_SUMMATION;
  c->normalize_initial_masses = 1;
  c->gravity = cryo_vec3{0.0f, 0.0f, -9.81f};
  c->velocity_scale = 0.0f;
  c->sound_speed = 0.0f;
  c->tait_gamma = 7.0f;
  c->background_pressure = -1.0f;
  c->vapor_stiffness = -1.0f;
  c->clamp_negative_pressure = 0;
  c->transport_velocity = 1;
  c->transport_background_factor = 2.0f;
  c->transport_stress = 0;
  c->transport_interface_only = 1;
  c->particle_shifting = 0;
  c->shifting_strength = 0.25f;
  c->artificial_viscosity = 0.02f;
  c->artificial_viscosity_vapor = 0.2f;
  c->delta_sph = 0.1f;
  c->no_slip = 1;
  c->sps_turbulence = 0;
  c->smagorinsky_cs = 0.12f;
  c->surface_tension = CRYO_ST_CSF;
  c->contact_angle_deg = 90.0f;
  c->contact_angle_wall_weight = 0.01f;
  c->wall_colour = 1;
  c->wetting_model = CRYO_WET_NORMAL_ROTATION;
  c->csf_smooth_normals = 0;
  c->csf_balanced = 0;
  c->akinci_gamma = -1.0f;
  c->akinci_beta = -1.0f;
  c->thermal = 1;
  c->phase_change = 1;
  c->wall_thermal_bc = CRYO_WALL_ADIABATIC;
  c->heater_power = 0.0f;
  c->conductivity_scale = 1.0f;
  c->lee_evaporation = 10.0f;
  c->lee_condensation = 10.0f;
  c->thermo_mode = CRYO_VENTED;
  c->vent_relaxation = 0.02f;
  c->merge_ratio = 0.3f;
  c->split_ratio = 2.0f;
  c->max_transfer_fraction = 0.25f;
  c->phase_change_density_limit = 0.01f;
  c->vapor_mass_rebalance = 1.0f;
  c->cfl = 0.25f;
  c->viscous_number = 0.125f;
  c->force_number = 0.25f;
  c->capillary_number = 0.25f;
  c->thermal_number = 0.1f;
  c->dt_max = 1e-2f;
  c->dt_min = 1e-9f;
  c->fixed_dt = 0.0f;
  c->compensated_sums = 0;
}

V3 v3(cryo_vec3 a) { return V3{a.x, a.y, a.z}; }
V3 v3p(const float* a) { return V3{a[0], a[1], a[2]}; }
cryo_vec3 cv(V3 a) { return cryo_vec3{a.x, a.y, a.z}; }

// ManeuverProfile host copy (sph_solver.hpp:96-105, sph_solver.cu:234-268).
struct HostManeuver {
  std::vector<double> time;
  std::vector<V3> acc, omega;
  V3 origin{0, 0, 0};
  void validate() const {
    if (time.empty()) throw std::invalid_argument("ManeuverProfile: empty time table");
    if (acc.size() != time.size())
      throw std::invalid_argument("ManeuverProfile: acceleration must have one entry per time");
    if (!omega.empty() && omega.size() != time.size())
      throw std::invalid_argument("ManeuverProfile: angular_velocity must be empty or have one entry per time");
    for (std::size_t k = 1; k < time.size(); ++k)
      if (!(time[k] > time[k - 1])) throw std::invalid_argument("ManeuverProfile: time must be strictly increasing");
  }
  void evaluate(double t, V3& a, V3& w, V3& wd) const {
    const std::size_t n = time.size();
    const bool rot = !omega.empty();
    wd = V3{0, 0, 0};
    if (t <= time.front() || n == 1) {
      a = acc.front();
      w = rot ? omega.front() : V3{0, 0, 0};
      return;
    }
    if (t >= time.back()) {
      a = acc.back();
      w = rot ? omega.back() : V3{0, 0, 0};
      return;
    }
    const std::size_t k = static_cast<std::size_t>(std::upper_bound(time.begin(), time.end(), t) - time.begin()) - 1;
    const double span = time[k + 1] - time[k];
    const float f = static_cast<float>((t - time[k]) / span);
    a = acc[k] * (1.0f - f) + acc[k + 1] * f;
    if (rot) {
      w = omega[k] * (1.0f - f) + omega[k + 1] * f;
      wd = (omega[k + 1] - omega[k]) / static_cast<float>(span);
    } else {
      w = V3{0, 0, 0};
    }
  }
};

void check_device() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    throw std::runtime_error("CryoSPH: no HIP device visible (this library runs on AMD Instinct MI355X, gfx950)");
  int dev = 0;
  CRYO_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop{};
  CRYO_HIP(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    throw std::runtime_error(std::string("CryoSPH: device arch '") + prop.gcnArchName +
                             "' is not gfx950; this library carries gfx950 code only");
}

}  // namespace
}  // namespace cryo

using namespace cryo;

// ------------------------------------------------------------------ solver
struct cryo_solver {
  cryo_config cfg{};
  cryo_fluid fluid{};
  SimParams P{};
  DtLimits L{};
  hipStream_t stream = nullptr;

  // particle arrays: two persistent sets (gather ping-pong) + scratch
  DevBuf<float4> x_[2], v_[2], vt_[2], nrm_[2];
  DevBuf<float4> xlo_[2];  // precision form only: low-order parts of the positions
  DevBuf<float> m_[2], m0_[2], rho_[2], T_[2];
  DevBuf<uint8_t> phase_[2], flags_[2];
  DevBuf<uint32_t> id_[2];
  int set = 0;
  DevBuf<float4> a, ab, ns, vw, fw, mw, nrec;
  DevBuf<float> p, drho, dTdt, sigma, kappa, nu_t, alpha, wall_w, heat_w, out_m, out_h, w_norm, p_wv, rmin, vol;
  // grid and scans
  DevBuf<uint32_t> keys_[2], vals_[2], tile_hist, row_total, split_flag, split_off, scan_sums, scan_total;
  DevBuf<int> cell_off;
  DevBuf<uint8_t> cell_flags;
  // neighbour lists of the current step (NbrList in cryo_device.hpp)
  static constexpr int kListK = 96;
  DevBuf<int> nbr_idx, nbr_count;
  bool lists_valid = false;
  // Every particle [0, n) sits in the cell it was sorted into (no move or
  // append since the last sort): the cell-tile passes may be used.
  bool tiles_ok = false;
  // Launch variants of the neighbour passes (results are the same bits either
  // way): LDS-staged cell tiles or per-particle list gathers, for the
  // interface and the forces pass.  autotune() times them.
  bool tile_iface = true, tile_forces = true;
  // cell-tile shape of the tiled passes (cells along x, y, z); any accepted
  // shape gives the same bits.  autotune_tiles() times candidates.
  int tile_shape[3] = {3, 3, 3};
  const uint32_t* last_perm = nullptr;
  int n_sorted = 0;
  // reductions
  DevBuf<double> partials, diag_out;
  DevBuf<unsigned> diag_max;
  int stride = 0;
  DevBuf<DevState> state;
  DevState* hs = nullptr;  // pinned host mirror
  // manoeuvre table on the device
  DevBuf<double> man_time;
  DevBuf<V3> man_acc, man_omega;
  Maneuver man{};
  HostManeuver hman;
  bool has_maneuver = false;
  V3 frame_acc{0, 0, 0};
  V3 frame_omega{0, 0, 0}, frame_omega_dot{0, 0, 0}, frame_origin{0, 0, 0};
  int frame_rot = 0;
  // grid configuration
  GridDev g{};
  float cell = 1.0f;
  int key_bits = 1;
  // host knowledge
  int n = 0, capacity = 0, pending_dead = 0, merges_seen = 0;
  // Device counts: phase-change steps queued back to back without the count
  // coming back in between.  Launches are sized for count_ub (the count at
  // the start plus count_room, default max(1024, n / 8)); kernels read the
  // count from the device state.  A step whose splits would pass count_ub
  // flags it, and the run is restored and re-run one step at a time.
  bool device_counts = false;  // run_steps and the ensemble runner use them
  bool dev_count = false;      // set while such steps are queued
  int count_ub = 0, count_room = 0;
  int64_t count_restarts = 0;  // queued runs restored and re-run step by step
  const int* state_field(std::size_t off) const {
    return reinterpret_cast<const int*>(reinterpret_cast<const char*>(state.get()) + off);
  }
  // the count of the step's passes, and of its cell sort (dead included)
  Count cn() const { return dev_count ? Count(count_ub, state_field(offsetof(DevState, n))) : Count(n); }
  Count cn_sort() const {
    return dev_count ? Count(count_ub, state_field(offsetof(DevState, n_sort))) : Count(n);
  }
  int n_bound() const { return dev_count ? count_ub : capacity; }  // bounds the count after a step's splits
  uint32_t next_id = 0;
  bool initialized = false, has_vapor = false;
  float char_length = 0.0f;
  // telemetry (telemetry.hpp:33-70)
  int tele_mode = CRYO_TELEMETRY_ASYNC_COPY;
  int tele_cap = 0;
  float* tele_host = nullptr;
  float* tele_host_dev = nullptr;
  DevBuf<float> tele_staging;
  hipStream_t copy_stream = nullptr;
  hipEvent_t tele_computed = nullptr, tele_landed = nullptr;
  bool tele_pending = false;
  double tele_time = 0;
  long long tele_step = 0;
  int tele_n = 0;

  // Per-section HIP-event timing on the solver stream (cryo_solver_profile_*).
  static constexpr int kSections = 10;
  static constexpr const char* kSectionNames[kSections] = {
      "begin_kick_keys", "cell_sort", "gather", "density_eos", "wall_bc", "interface", "forces", "kick_sources",
      "phase_change", "end_step"};
  unsigned prof_mask = 0;
  struct Mark {
    int sec;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> event_pool;
  hipEvent_t open_event[kSections] = {};
  double prof_ms[kSections] = {};
  long long prof_count[kSections] = {};

  hipEvent_t take_event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    CRYO_HIP(hipEventCreate(&e));
    return e;
  }
  void prof_begin(int sec) {
    if (!(prof_mask & (1u << sec))) return;
    open_event[sec] = take_event();
    CRYO_HIP(hipEventRecord(open_event[sec], stream));
  }
  void prof_end(int sec) {
    if (!(prof_mask & (1u << sec)) || !open_event[sec]) return;
    hipEvent_t b = take_event();
    CRYO_HIP(hipEventRecord(b, stream));
    marks.push_back(Mark{sec, open_event[sec], b});
    open_event[sec] = nullptr;
  }
  void prof_collect() {
    if (marks.empty()) return;
    CRYO_HIP(hipStreamSynchronize(stream));
    for (const Mark& m : marks) {
      float ms = 0;
      CRYO_HIP(hipEventElapsedTime(&ms, m.a, m.b));
      prof_ms[m.sec] += ms;
      prof_count[m.sec] += 1;
      event_pool.push_back(m.a);
      event_pool.push_back(m.b);
    }
    marks.clear();
  }

  cryo_solver(const cryo_config& c, const cryo_fluid& f) : cfg(c), fluid(f) {
    if (!(cfg.dx > 0)) throw std::invalid_argument("SolverConfig.dx must be > 0");
    if (cfg.contact_angle_deg < 0 || cfg.contact_angle_deg > 180)
      throw std::invalid_argument("contact angle must be within [0, 180] degrees");
    check_device();
    CRYO_HIP(hipStreamCreate(&stream));  // blocking, default priority
    CRYO_HIP(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
    CRYO_HIP(hipEventCreateWithFlags(&tele_computed, hipEventDisableTiming));
    CRYO_HIP(hipEventCreateWithFlags(&tele_landed, hipEventDisableTiming));
    CRYO_HIP(hipHostMalloc(reinterpret_cast<void**>(&hs), sizeof(DevState), hipHostMallocDefault));
    std::memset(hs, 0, sizeof(DevState));
    state.resize(1);
    hs->p_ullage = fluid.p_sat;
    hs->T_sat = fluid.T_sat;
    hs->spacing_ratio = 1.0f;
    hs->cfg_gravity = v3(cfg.gravity);
    hs->gravity = v3(cfg.gravity);
    push_state();
  }

  ~cryo_solver() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    graph_reset();
    release_telemetry();
    for (const Mark& m : marks) {
      (void)hipEventDestroy(m.a);
      (void)hipEventDestroy(m.b);
    }
    for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
    if (hs) (void)hipHostFree(hs);
    if (tele_computed) (void)hipEventDestroy(tele_computed);
    if (tele_landed) (void)hipEventDestroy(tele_landed);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // ------------------------------------------------------- state mirror
  void sync_state() {
    CRYO_HIP(hipMemcpyAsync(hs, state.get(), sizeof(DevState), hipMemcpyDeviceToHost, stream));
    CRYO_HIP(hipStreamSynchronize(stream));
  }
  void push_state() {
    CRYO_HIP(hipMemcpyAsync(state.get(), hs, sizeof(DevState), hipMemcpyHostToDevice, stream));
    CRYO_HIP(hipStreamSynchronize(stream));
  }
  void check_launch() { CRYO_HIP(hipGetLastError()); }

  // -------------------------------------------------------------- views
  PV view() {
    PV pv{};
    const int k = set;
    pv.s = Persist{x_[k].get(), v_[k].get(), vt_[k].get(), nrm_[k].get(), m_[k].get(), m0_[k].get(),
                   rho_[k].get(), T_[k].get(), phase_[k].get(), flags_[k].get(), id_[k].get(), xlo_[k].get()};
    pv.t = Scratch{a.get(), ab.get(), ns.get(), vw.get(), fw.get(), mw.get(), p.get(), drho.get(), dTdt.get(),
                   sigma.get(), kappa.get(), nu_t.get(), alpha.get(), wall_w.get(), heat_w.get(), out_m.get(),
                   out_h.get(), w_norm.get(), p_wv.get(), rmin.get(), nrec.get(), vol.get()};
    pv.capacity = capacity;
    return pv;
  }
  Persist persist(int k) {
    return Persist{x_[k].get(), v_[k].get(), vt_[k].get(), nrm_[k].get(), m_[k].get(), m0_[k].get(),
                   rho_[k].get(), T_[k].get(), phase_[k].get(), flags_[k].get(), id_[k].get(), xlo_[k].get()};
  }
  GridDev grid() {
    GridDev gd = g;
    gd.off = cell_off.get();
    gd.flags = cell_flags.get();
    gd.tx = tile_shape[0];
    gd.ty = tile_shape[1];
    gd.tz = tile_shape[2];
    return gd;
  }
  NbrList lists() {
    return NbrList{nbr_idx.get(), nbr_count.get(), capacity, kListK, lists_valid ? 1 : 0};
  }

  template <class F>
  void each_particle_array(F&& f) {
    for (int k = 0; k < 2; ++k) {
      f(x_[k], k == set);
      f(v_[k], k == set);
      f(vt_[k], k == set);
      f(nrm_[k], k == set);
      f(m_[k], k == set);
      f(m0_[k], k == set);
      f(rho_[k], k == set);
      f(T_[k], k == set);
      f(phase_[k], k == set);
      f(flags_[k], k == set);
      f(id_[k], k == set);
      if (cfg.compensated_sums) f(xlo_[k], k == set);
    }
    f(a, true); f(ab, true); f(ns, true); f(vw, true); f(fw, true); f(mw, true); f(nrec, false); f(vol, false);
    f(p, true); f(drho, true); f(dTdt, true); f(sigma, true); f(kappa, true); f(nu_t, true); f(alpha, true);
    f(wall_w, true); f(heat_w, true); f(out_m, true); f(out_h, true); f(w_norm, true); f(p_wv, true); f(rmin, true);
  }

  void size_aux(int cap) {
    for (int k = 0; k < 2; ++k) {
      keys_[k].resize(cap);
      vals_[k].resize(cap);
    }
    tile_hist.resize(static_cast<std::size_t>(256) * radix_tiles(cap));
    row_total.resize(256);
    split_flag.resize(cap);
    split_off.resize(cap);
    scan_sums.resize(static_cast<std::size_t>((cap + 1023) / 1024) + 1);
    scan_total.resize(1);
    // list chunks interleave over groups of 64 particles: whole groups
    nbr_idx.resize(static_cast<std::size_t>(kListK) * ((static_cast<std::size_t>(cap) + 63) / 64 * 64));
    nbr_count.resize(cap);
    lists_valid = false;
    stride = (cap + 255) / 256 + 1;
    partials.resize(static_cast<std::size_t>(std::max(diag_quantities(), 8)) * stride);
    diag_out.resize(diag_quantities());
    diag_max.resize(2);
  }

  void allocate(int cap) {  // ParticleSystem::allocate (particle_system.cpp:50-54)
    each_particle_array([cap](auto& b, bool) { b.resize(cap); });
    size_aux(cap);
    capacity = cap;
  }

  void grow(int cap) {  // ParticleSystem::grow (particle_system.cpp:62-67)
    if (cap <= capacity) return;
    const std::size_t keep = static_cast<std::size_t>(capacity);
    each_particle_array([&](auto& b, bool live) {
      if (live) b.grow(cap, keep, stream);
      else b.resize(cap);
    });
    size_aux(cap);
    capacity = cap;
    configure_telemetry(tele_mode);
  }

  // ---------------------------------------------------------- init_params
  void init_params() {  // sph_solver.cu:292-405
    const cryo_config& c = cfg;
    SimParams& Q = P;
    Q = SimParams{};
    const float dx = c.dx;
    const float h = c.h_factor * dx;
    Q.kernel.type = c.kernel;
    Q.kernel.h = h;
    Q.kernel.inv_h = 1.0f / h;
    Q.kernel.support = 2.0f * h;
    const float h3 = h * h * h;
    Q.kernel.norm = c.kernel == CRYO_KERNEL_CUBIC ? 1.0f / (kPi * h3) : 21.0f / (16.0f * kPi * h3);
    Q.dx = dx;
    Q.support2 = Q.kernel.support * Q.kernel.support;
    {
      double sigma0 = 0;
      const int r = static_cast<int>(std::ceil(Q.kernel.support / dx));
      for (int k = -r; k <= r; ++k)
        for (int j = -r; j <= r; ++j)
          for (int i = -r; i <= r; ++i)
            sigma0 += kW(Q.kernel, dx * std::sqrt(static_cast<float>(i * i + j * j + k * k)));
      Q.vol0 = static_cast<float>(1.0 / sigma0);
    }
    Q.rho0[0] = fluid.rho_liquid;
    Q.rho0[1] = fluid.rho_vapor;
    Q.mu[0] = fluid.mu_liquid;
    Q.mu[1] = fluid.mu_vapor;
    Q.k[0] = fluid.k_liquid * c.conductivity_scale;
    Q.k[1] = fluid.k_vapor * c.conductivity_scale;
    Q.cp[0] = fluid.cp_liquid;
    Q.cp[1] = fluid.cp_vapor;
    Q.mref[0] = Q.rho0[0] * Q.vol0;
    Q.mref[1] = Q.rho0[1] * Q.vol0;
    const bool st_on = c.surface_tension != CRYO_ST_NONE;
    const float gn = norm(v3(c.gravity));
    float vref = c.velocity_scale;
    if (!(vref > 0)) {
      vref = std::sqrt(gn * char_length);
      if (st_on) vref = std::max(vref, std::sqrt(fluid.sigma / (fluid.rho_liquid * h)));
      vref = std::max(vref, 1e-3f);
    }
    Q.c0 = c.sound_speed > 0 ? c.sound_speed : 10.0f * vref;
    Q.tait_gamma = c.tait_gamma;
    Q.tait_B = Q.rho0[0] * Q.c0 * Q.c0 / Q.tait_gamma;
    Q.p_bg = c.background_pressure >= 0 ? c.background_pressure
                                        : (has_vapor ? 0.1f * Q.rho0[1] * Q.c0 * Q.c0 : 0.0f);
    Q.vapor_K = c.vapor_stiffness > 0 ? c.vapor_stiffness : (has_vapor ? Q.rho0[1] * Q.c0 * Q.c0 : Q.p_bg);
    Q.gas_R = Q.vapor_K > 0 ? Q.vapor_K / (Q.rho0[1] * fluid.T_sat) : fluid.R_gas;
    Q.clamp_neg = c.clamp_negative_pressure ? 1 : 0;
    Q.alpha_av = c.artificial_viscosity;
    Q.alpha_av_vap = c.artificial_viscosity_vapor;
    auto harmonic = [](float a, float b) {
      const float sum = a + b;
      return sum > 0.0f ? 2.0f * a * b / sum : 0.0f;
    };
    Q.eta_ll = harmonic(Q.mu[0], Q.mu[0]);
    Q.eta_lv = harmonic(Q.mu[0], Q.mu[1]);
    Q.eta_vv = harmonic(Q.mu[1], Q.mu[1]);
    Q.delta_sph = c.delta_sph;
    Q.density_mode = c.density_mode;
    Q.no_slip = c.no_slip ? 1 : 0;
    Q.sps = c.sps_turbulence ? 1 : 0;
    Q.smag_cs = c.smagorinsky_cs;
    Q.tvf = c.transport_velocity ? 1 : 0;
    Q.tvf_stress = c.transport_stress ? 1 : 0;
    Q.tvf_factor = c.transport_background_factor;
    Q.tvf_interface_only = c.transport_interface_only ? 1 : 0;
    Q.shift_strength = c.particle_shifting ? c.shifting_strength : 0.0f;
    Q.tvf_sigma_min = has_vapor ? 0.0f : 0.95f / Q.vol0;
    Q.st_model = c.surface_tension;
    Q.sigma = st_on ? fluid.sigma : 0.0f;
    const float theta = c.contact_angle_deg * kPi / 180.0f;
    Q.cos_theta = std::cos(theta);
    Q.sin_theta = std::sin(theta);
    Q.wall_adh_w = c.contact_angle_wall_weight;
    Q.wall_colour = c.wall_colour ? 1 : 0;
    Q.wetting = c.wetting_model;
    Q.csf_smooth = c.csf_smooth_normals ? 1 : 0;
    Q.csf_balanced = c.csf_balanced ? 1 : 0;
    Q.normal_eps = 0.1f / h;
    {
      const double sup = Q.kernel.support;
      const double i_coh =
          radial_moment4([&](double r) { return double(akinci_coh(static_cast<float>(r), static_cast<float>(sup))); }, sup);
      const double i_adh = radial_moment4(
          [&](double r) { return double(akinci_adh_host(static_cast<float>(r), static_cast<float>(sup))); }, sup);
      const double rho = fluid.rho_liquid;
      const double gamma =
          c.akinci_gamma >= 0 ? double(c.akinci_gamma) : (i_coh > 0 ? 8.0 * fluid.sigma / (kPi * rho * rho * i_coh) : 0.0);
      Q.ak_gamma = static_cast<float>(gamma);
      Q.ak_beta = c.akinci_beta >= 0
                      ? c.akinci_beta
                      : static_cast<float>(i_adh > 0 ? gamma * (i_coh / i_adh) * 0.5 * (1.0 + Q.cos_theta) : 0.0);
      Q.ak_adh_scale = 0.007f / std::pow(Q.kernel.support, 3.25f);
    }
    Q.thermal = c.thermal ? 1 : 0;
    Q.phase_change = c.phase_change ? 1 : 0;
    Q.wall_bc = c.wall_thermal_bc;
    Q.T_wall = fluid.T_sat;
    Q.T_ref = fluid.T_sat;
    Q.h_fg = fluid.h_fg;
    Q.lee_evap = c.lee_evaporation;
    Q.lee_cond = c.lee_condensation;
    Q.max_transfer = c.max_transfer_fraction;
    Q.pc_density_limit = c.phase_change_density_limit;
    Q.vapor_rebalance = c.vapor_mass_rebalance;
    Q.merge_ratio = c.merge_ratio;
    Q.split_ratio = c.split_ratio;
    Q.heat_power = c.heater_power;
    Q.has_vapor = has_vapor ? 1 : 0;
    Q.compensated = c.compensated_sums ? 1 : 0;
    L = DtLimits{c.cfl, c.viscous_number, c.force_number, c.capillary_number, c.thermal_number,
                 c.dt_max, c.dt_min, c.fixed_dt};
  }

  FluidConst fluid_const() const {
    return FluidConst{fluid.rho_liquid, fluid.rho_vapor, fluid.T_sat, fluid.p_sat, fluid.R_gas, fluid.h_fg};
  }

  // ---------------------------------------------------------- particles
  void set_particles(int n_in, const float* x, const float* v, const float* nrm, const float* m, const float* rho,
                     const float* T, const uint8_t* phase, const uint8_t* flags, const uint32_t* id,
                     const float* lo, const float* hi) {
    if (n_in <= 0) throw std::invalid_argument("Solver::set_particles: no particles");
    if (!x || !m || !rho || !T || !phase || !lo || !hi) throw std::invalid_argument("set_particles: null array");
    bool vap = false;
    int dead = 0;
    for (int i = 0; i < n_in; ++i) {
      if (phase[i] > kDead) throw std::invalid_argument("phase values must be 0 (liquid), 1 (vapour) or 2 (wall)");
      vap = vap || phase[i] == kVapor;
      dead += phase[i] == kDead;
    }
    if (vap && cfg.surface_tension == CRYO_ST_AKINCI)
      throw std::invalid_argument(
          "Solver: the Akinci surface-tension model is a free-surface (liquid-only) model; use CSF when vapour "
          "particles are present (FillSpec.include_vapor = false for Akinci runs)");
    sync_state();
    has_vapor = vap;
    const V3 lo3 = v3p(lo), hi3 = v3p(hi);
    const V3 ext = hi3 - lo3;
    char_length = std::max({ext.x, ext.y, ext.z});
    init_params();
    const int cap = static_cast<int>(std::max<std::size_t>(static_cast<std::size_t>(static_cast<double>(n_in) * 2.0),
                                                           static_cast<std::size_t>(n_in) + 1024));
    if (cap > capacity) allocate(cap);
    // persistent arrays
    std::vector<float4> h4(static_cast<std::size_t>(n_in));
    auto up4 = [&](DevBuf<float4>& b, const float* src) {
      for (int i = 0; i < n_in; ++i)
        h4[i] = src ? make_float4(src[3 * i], src[3 * i + 1], src[3 * i + 2], 0.0f) : make_float4(0, 0, 0, 0);
      CRYO_HIP(hipMemcpyAsync(b.get(), h4.data(), sizeof(float4) * n_in, hipMemcpyHostToDevice, stream));
      CRYO_HIP(hipStreamSynchronize(stream));
    };
    auto up = [&](auto& b, const auto* src) {
      CRYO_HIP(hipMemcpyAsync(b.get(), src, sizeof(*src) * n_in, hipMemcpyHostToDevice, stream));
    };
    up4(x_[set], x);
    up4(v_[set], v);
    up4(vt_[set], v);
    up4(nrm_[set], nrm);
    up(m_[set], m);
    up(m0_[set], m);
    up(rho_[set], rho);
    up(T_[set], T);
    up(phase_[set], phase);
    std::vector<uint8_t> fl(static_cast<std::size_t>(n_in), 0);
    if (flags) std::copy(flags, flags + n_in, fl.begin());
    up(flags_[set], fl.data());
    std::vector<uint32_t> ids(static_cast<std::size_t>(n_in));
    if (id) std::copy(id, id + n_in, ids.begin());
    else std::iota(ids.begin(), ids.end(), 0u);
    up(id_[set], ids.data());
    if (xlo_[set].get())  // the given positions are exact: no low-order parts yet
      CRYO_HIP(hipMemsetAsync(xlo_[set].get(), 0, sizeof(float4) * n_in, stream));
    CRYO_HIP(hipStreamSynchronize(stream));
    next_id = 0;
    for (uint32_t k : ids) next_id = std::max(next_id, k + 1);
    // scratch fields start from a clean state (particle_system.cpp:81-83)
    auto zero = [&](auto& b) { CRYO_HIP(hipMemsetAsync(b.get(), 0, sizeof(*b.get()) * b.size(), stream)); };
    zero(a); zero(ab); zero(ns); zero(vw); zero(fw); zero(mw); zero(p); zero(drho); zero(dTdt); zero(sigma);
    zero(kappa); zero(nu_t); zero(alpha); zero(wall_w); zero(heat_w); zero(out_m); zero(out_h); zero(w_norm);
    zero(p_wv);
    n = n_in;
    pending_dead = dead;
    configure_grid(lo3, hi3, P.kernel.support);
    configure_telemetry(CRYO_TELEMETRY_ASYNC_COPY);
    initialized = false;
    merges_seen = 0;
    // clock and counters (sph_solver.cu:420-427); the frame is kept
    DevState& s = *hs;
    s.time = 0;
    s.step = 0;
    s.last_dt = 0;
    s.heat_input = s.evaporated = s.condensed = s.vented = s.vented_h = 0;
    s.splits = s.merges = 0;
    s.merges_sorted = s.count_overflow = 0;
    s.p_ullage = fluid.p_sat;
    s.fluid_volume0 = 0;
    s.dt = 0;
    s.T_sat = fluid.T_sat;
    s.heat_inv_norm = 0;
    s.cfg_gravity = v3(cfg.gravity);
    s.gravity = v3(cfg.gravity) - frame_acc;
    s.frame_acc = frame_acc;
    s.omega = frame_omega;
    s.omega_dot = frame_omega_dot;
    s.origin = frame_origin;
    s.frame_rot = frame_rot;
    s.n = n;
    s.n_alive = n;
    s.next_id = next_id;
    push_state();
  }

  void configure_grid(V3 lo, V3 hi, float cell_size) {  // spatial_hash.cu:15-30
    if (!(cell_size > 0)) throw std::invalid_argument("SpatialHash: cell size must be positive");
    cell = cell_size;
    const V3 origin = lo - V3{cell, cell, cell};
    const V3 ext = hi - lo;
    const int nx = std::max(1, static_cast<int>(std::ceil(ext.x / cell)) + 2);
    const int ny = std::max(1, static_cast<int>(std::ceil(ext.y / cell)) + 2);
    const int nz = std::max(1, static_cast<int>(std::ceil(ext.z / cell)) + 2);
    const long long cells = static_cast<long long>(nx) * ny * nz;
    if (cells > 400000000LL) throw std::runtime_error("SpatialHash: grid too large; increase the cell size");
    g.ox = origin.x;
    g.oy = origin.y;
    g.oz = origin.z;
    g.inv_cell = 1.0f / cell;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.num_cells = static_cast<int>(cells);
    // |error| of (p - origin) * inv_cell is a few ulp of the largest index
    g.cull_pad = 1e-3f + 5e-7f * static_cast<float>(std::max(nx, std::max(ny, nz)));
    cell_off.resize(static_cast<std::size_t>(cells) + 2);
    cell_flags.resize(static_cast<std::size_t>(cells));
    key_bits = 1;
    while ((1ll << key_bits) <= cells) ++key_bits;
    last_perm = nullptr;
    n_sorted = 0;
    tiles_ok = false;
  }

  // ----------------------------------------------------------- the step
  // SpatialHash::build (spatial_hash.cu:48-114), optionally after kick_drift.
  void grid_build(bool do_kick) {
    PV pv = view();
    const GridDev gd = grid();
    prof_begin(0);
    const Count ns = cn_sort();
    launch_kick_drift_keys(pv, gd, state.get(), ns, do_kick, cfg.density_mode == CRYO_DENSITY_CONTINUITY,
                           P.tvf != 0, keys_[0].get(), vals_[0].get(), stream);
    prof_end(0);
    prof_begin(1);
    bool alt = false;
    launch_radix_sort(keys_[0].get(), vals_[0].get(), keys_[1].get(), vals_[1].get(), ns, key_bits, tile_hist.get(),
                      row_total.get(), stream, &alt);
    const uint32_t* keys = alt ? keys_[1].get() : keys_[0].get();
    const uint32_t* perm = alt ? vals_[1].get() : vals_[0].get();
    launch_cell_offsets(keys, ns, g.num_cells, cell_off.get(), state.get(), stream);
    prof_end(1);
    prof_begin(2);
    launch_gather(persist(1 - set), persist(set), perm, state.get(), ns.ub, stream);
    launch_cell_flags(cell_off.get(), persist(1 - set).phase, g.num_cells, cell_flags.get(), stream);
    prof_end(2);
    check_launch();
    set = 1 - set;
    last_perm = perm;
    lists_valid = false;
    tiles_ok = true;
    if (dev_count) return;  // the device drops them and holds the count
    n_sorted = n;
    n -= pending_dead;  // dead particles sort last and are dropped
    pending_dead = 0;
  }

  // forces = false: everything but the forces pass (what the adjoint's
  // reverse sweep reads: densities, pressures, wall values, normals, lists).
  void evaluate_rates(bool forces = true) {  // sph_solver.cu:449-480
    PV pv = view();
    const GridDev gd = grid();
    const bool csf_mode = cfg.surface_tension == CRYO_ST_CSF && has_vapor;
    const bool heat_sum = P.thermal && P.heat_power != 0.0f;
    prof_begin(3);
    lists_valid = false;
    const Count nc = cn();
    launch_density(pv, gd, P, lists(), nc, stream);
    lists_valid = true;
    const NbrList L = lists();
    prof_end(3);
    prof_begin(4);
    launch_wall_bc(pv, gd, P, L, state.get(), nc, stream);
    prof_end(4);
    prof_begin(5);
    const int nb =
        launch_interface(pv, gd, P, L, state.get(), nc, csf_mode, heat_sum ? partials.get() : nullptr,
                         tiles_ok && tile_iface, stream);
    if (csf_mode && P.csf_smooth) launch_smooth_normals(pv, gd, P, L, nc, stream);
    if (P.thermal) launch_heat_norm(state.get(), partials.get(), nb, heat_sum, stream);
    prof_end(5);
    if (!forces) {
      check_launch();
      return;
    }
    prof_begin(6);
    launch_forces(pv, gd, P, L, state.get(), nc, csf_mode, cfg.surface_tension == CRYO_ST_AKINCI,
                  cfg.density_mode == CRYO_DENSITY_CONTINUITY, P.thermal != 0, tiles_ok && tile_forces, stream);
    prof_end(6);
    check_launch();
  }

  // Phase inventory + ullage update + clock (sph_solver.cu:586-607, 530-532).
  void end_step(int nb_pc, int nb_vent, bool advance) {
    const bool closed = cfg.thermo_mode == CRYO_CLOSED && has_vapor;
    int nb_inv = 0;
    if (closed) nb_inv = launch_inventory(view(), state.get(), n_bound(), partials.get() + 5 * stride, stride, stream);
    launch_end_step(state.get(), partials.get(), stride, nb_pc, nb_vent, nb_inv, fluid_const(), closed, advance,
                    stream);
    check_launch();
  }

  void ensure_initialized() {  // sph_solver.cu:439-447
    if (initialized) return;
    if (n == 0) throw std::runtime_error("Solver: no particles loaded");
    grid_build(false);
    if (cfg.normalize_initial_masses) launch_normalize_masses(view(), grid(), P, n, stream);
    evaluate_rates();
    end_step(0, 0, false);
    sync_state();
    n = hs->n;
    initialized = true;
  }

  void phase_change_step(int& nb_pc, int& nb_vent) {  // sph_solver.cu:545-584
    const GridDev gd = grid();
    PV pv = view();
    const NbrList L = lists();
    const Count nc = cn();
    nb_pc = launch_pc_normalize(pv, gd, P, L, nc, partials.get(), stride, stream);
    launch_pc_apply(pv, gd, P, L, nc, stream);
    const bool rebalance = P.vapor_rebalance > 0.0f && has_vapor;
    if (rebalance) launch_rebalance(pv, gd, P, L, nc, stream);
    // (with device counts the run grew the arrays before it was queued)
    if (!dev_count && static_cast<std::size_t>(n) * 4 > static_cast<std::size_t>(capacity) * 3) {
      launch_pc_fold(state.get(), partials.get(), stride, nb_pc, stream);  // growing drops the partials
      nb_pc = 0;
      grow(capacity * 3 / 2 + 1024);
      pv = view();
    }
    launch_rebalance_apply_mark(pv, P, nc, rebalance, split_flag.get(), stream);
    launch_exclusive_scan(split_flag.get(), split_off.get(), nc, scan_sums.get(), scan_total.get(), stream);
    launch_split(pv, P, state.get(), nc, split_flag.get(), split_off.get(), stream);
    lists_valid = false;  // split moves parents and appends particles
    tiles_ok = false;
    launch_split_commit(state.get(), nc, n_bound(), scan_total.get(), stream);
    nb_vent = 0;
    if (cfg.thermo_mode == CRYO_VENTED && has_vapor)
      nb_vent = launch_venting(pv, P, state.get(), n_bound(), cfg.vent_relaxation, partials.get() + 3 * stride, stride,
                               stream);
    check_launch();
  }

  void step() {  // sph_solver.cu:511-533
    step_issue();
    step_finish();
  }
  // Everything of one step that the device can do without the host: queued
  // on the solver stream, returns at once.
  // ------------------------------------------------------ captured steps
  // One step's launches as a HIP graph, replayed while the host state it was
  // captured from holds: the persistent buffer set, the particle count, the
  // merged particles to drop and the capacity (a step's kernels, grids and
  // pointers depend on nothing else).  Replay sets the host state the
  // captured issue left.  With use_graphs set, cryo_solver_run_steps and the
  // ensemble runner enable it for the duration of their call, during which
  // SimParams cannot change.
  struct StepGraph {
    hipGraphExec_t exec = nullptr;
    bool dev = false;  // captured with device counts (keyed by count_ub, not n)
    int n = -1, pending_dead = -1, capacity = -1, ub = -1;
    int set_after = 0, n_after = 0, n_sorted = 0, pd_after = 0;
    const uint32_t* last_perm = nullptr;
    bool lists_valid = false, tiles_ok = false;
  };
  StepGraph sg[2];  // by persistent buffer set
  bool use_graphs = false, graphs_on = false;
  bool ensemble_graph_on = false;  // the ensemble runner may capture all members as one graph
  void graph_reset() {
    for (StepGraph& x : sg) {
      if (x.exec) (void)hipGraphExecDestroy(x.exec);
      x = StepGraph{};
    }
  }
  void step_issue_graph() {
    ensure_initialized();
    const bool grows = !dev_count && cfg.phase_change != 0 &&
                       static_cast<std::size_t>(n) * 4 > static_cast<std::size_t>(capacity) * 3;  // allocates
    if (!graphs_on || prof_mask != 0 || tape_slot || grows || n == 0) {
      step_issue();
      return;
    }
    StepGraph& G = sg[set];
    const bool same = G.dev == dev_count && G.capacity == capacity &&
                      (dev_count ? G.ub == count_ub : G.n == n && G.pending_dead == pending_dead);
    if (!G.exec || !same) {
      if (G.exec) (void)hipGraphExecDestroy(G.exec);
      G = StepGraph{};
      const int n0 = n, pd0 = pending_dead, cap0 = capacity;
      G.dev = dev_count;
      G.ub = count_ub;
      CRYO_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      try {
        step_issue();
      } catch (...) {
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(stream, &partial);
        if (partial) (void)hipGraphDestroy(partial);
        throw;
      }
      hipGraph_t graph = nullptr;
      CRYO_HIP(hipStreamEndCapture(stream, &graph));
      const hipError_t e = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      CRYO_HIP(e);
      G.n = n0, G.pending_dead = pd0, G.capacity = cap0;
      G.set_after = set, G.n_after = n, G.n_sorted = n_sorted, G.pd_after = pending_dead;
      G.last_perm = last_perm, G.lists_valid = lists_valid, G.tiles_ok = tiles_ok;
      CRYO_HIP(hipGraphLaunch(G.exec, stream));
      return;
    }
    CRYO_HIP(hipGraphLaunch(G.exec, stream));
    set = G.set_after, n = G.n_after, n_sorted = G.n_sorted, pending_dead = G.pd_after;
    last_perm = G.last_perm, lists_valid = G.lists_valid, tiles_ok = G.tiles_ok;
  }

  void step_issue() {
    ensure_initialized();
    launch_begin_step(state.get(), P, L, man, stream);
    grid_build(true);
    if (tape_slot) snap_save(*tape_slot, false, last_perm);
    if (marks.size() > 4096) prof_collect();
    evaluate_rates();
    if (tape_slot) snap_save_rates(*tape_slot);
    const bool shifting = P.shift_strength > 0.0f;
    const bool pc = cfg.phase_change != 0;
    PV pv = view();
    prof_begin(7);
    const Count nc = cn();
    launch_kick_sources(pv, P, state.get(), nc, true, P.thermal != 0, pc && !shifting, stream);
    if (shifting) {
      launch_shifting(pv, grid(), P, lists(), state.get(), nc, stream);
      lists_valid = false;  // positions moved
      tiles_ok = false;
      if (pc) launch_kick_sources(pv, P, state.get(), nc, false, false, true, stream);
    }
    prof_end(7);
    int nb_pc = 0, nb_vent = 0;
    prof_begin(8);
    if (pc) phase_change_step(nb_pc, nb_vent);
    prof_end(8);
    prof_begin(9);
    end_step(nb_pc, nb_vent, true);
    prof_end(9);
  }
  // Host bookkeeping the next step needs: with phase change the particle
  // count (splits append, merges drop at the next sort) comes back from the
  // device, which waits for this solver's stream.
  void step_finish() {
    if (cfg.phase_change != 0) {
      sync_state();
      take_counts();
    }
  }
  // The same in two halves for callers that poll (the ensemble runner): the
  // state copy is queued behind the step, and once the stream has drained the
  // host takes the counts from it.
  bool needs_counts() const { return cfg.phase_change != 0; }
  void queue_state_copy() {
    CRYO_HIP(hipMemcpyAsync(hs, state.get(), sizeof(DevState), hipMemcpyDeviceToHost, stream));
  }
  bool stream_idle() const {
    const hipError_t e = hipStreamQuery(stream);
    if (e == hipErrorNotReady) return false;
    CRYO_HIP(e);
    return true;
  }
  void take_counts() {
    n = hs->n;
    pending_dead = hs->merges - merges_seen;
    merges_seen = hs->merges;
  }

  // A run of steps queued with device counts (device_counts set, phase change
  // on): count_begin() applies the host's growth rule once for the run, saves
  // the state and switches the launches to the device count; count_end(),
  // once they are queued, takes the counts back, or, when a step's splits
  // passed the launch bound, restores the saved state and returns false (the
  // caller then runs the same steps one at a time).
  bool uses_device_counts() const { return device_counts && cfg.phase_change != 0; }
  struct CountRun;  // the saved state (after Snap)
  std::unique_ptr<CountRun> count_run;
  void count_begin() {
    ensure_initialized();
    sync_state();
    const int room = count_room > 0 ? count_room : std::max(1024, n / 8);
    if (static_cast<std::size_t>(n) * 4 > static_cast<std::size_t>(capacity) * 3 ||
        static_cast<long long>(n) + room > capacity)
      grow(static_cast<int>(std::max<long long>(static_cast<long long>(capacity) * 3 / 2 + 1024,
                                                static_cast<long long>(n) + room)));
    if (hs->count_overflow) {
      hs->count_overflow = 0;
      push_state();
    }
    if (!count_run) count_run = std::make_unique<CountRun>();
    CountRun& r = *count_run;
    snap_save(r.z, true, nullptr);
    snap_save_scratch(r.z);
    r.pending_dead = pending_dead, r.merges_seen = merges_seen, r.n_sorted = n_sorted, r.last_perm = last_perm;
    count_ub = n + room;
    dev_count = true;
  }
  bool count_end() {
    dev_count = false;
    sync_state();
    CountRun& r = *count_run;
    if (hs->count_overflow) {
      snap_load(r.z, true);
      snap_load_scratch(r.z);
      pending_dead = r.pending_dead, merges_seen = r.merges_seen, n_sorted = r.n_sorted, last_perm = r.last_perm;
      sync_state();
      ++count_restarts;
      return false;
    }
    n = hs->n;
    n_sorted = hs->n_sort;
    pending_dead = hs->merges - hs->merges_sorted;
    merges_seen = hs->merges;
    return true;
  }
  // The whole-ensemble graph of cryo_ensemble_run_steps (this solver is member
  // 0; every member initialised, none needing its count back between steps).
  void ensemble_graph(cryo_solver* const* v, int32_t count, int64_t steps) {
    std::vector<hipEvent_t> ev(static_cast<std::size_t>(count) + 1, nullptr);
    struct Events {
      std::vector<hipEvent_t>& e;
      ~Events() {
        for (hipEvent_t x : e)
          if (x) (void)hipEventDestroy(x);
      }
    } owned_events{ev};
    for (hipEvent_t& e : ev) CRYO_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    hipEvent_t fork = ev[static_cast<std::size_t>(count)];
    for (int32_t t = 0; t < count; ++t) CRYO_HIP(hipStreamSynchronize(v[t]->stream));  // saved states copied
    CRYO_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      CRYO_HIP(hipEventRecord(fork, stream));
      for (int32_t t = 0; t < count; ++t) {
        cryo_solver& z = *v[t];
        if (z.stream != stream) CRYO_HIP(hipStreamWaitEvent(z.stream, fork, 0));
        z.step_issue();
        z.step_issue();
        if (z.stream != stream) {
          CRYO_HIP(hipEventRecord(ev[static_cast<std::size_t>(t)], z.stream));
          CRYO_HIP(hipStreamWaitEvent(stream, ev[static_cast<std::size_t>(t)], 0));
        }
      }
    } catch (...) {
      hipGraph_t partial = nullptr;
      (void)hipStreamEndCapture(stream, &partial);
      if (partial) (void)hipGraphDestroy(partial);
      throw;
    }
    hipGraph_t graph = nullptr;
    CRYO_HIP(hipStreamEndCapture(stream, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    CRYO_HIP(e);
    struct Exec {
      hipGraphExec_t x;
      hipStream_t s;
      ~Exec() {
        (void)hipStreamSynchronize(s);
        (void)hipGraphExecDestroy(x);
      }
    } owned_exec{exec, stream};
    // the capture queued nothing; the host state of every member already
    // reads two steps on, which is where each launch leaves the device
    for (int64_t q = 0; q + 1 < steps; q += 2) CRYO_HIP(hipGraphLaunch(exec, stream));
    CRYO_HIP(hipEventRecord(fork, stream));  // the members' streams go on behind the graph
    for (int32_t t = 0; t < count; ++t) {
      cryo_solver& z = *v[t];
      if (z.stream != stream) CRYO_HIP(hipStreamWaitEvent(z.stream, fork, 0));
      if (steps % 2) z.step_issue();
    }
  }
  void count_abort() {  // an error while the run was queued: take what the device holds
    if (!dev_count) return;
    dev_count = false;
    if (hipMemcpyAsync(hs, state.get(), sizeof(DevState), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
      return;
    n = hs->n;
    n_sorted = hs->n_sort;
    pending_dead = hs->merges - hs->merges_sorted;
    merges_seen = hs->merges;
  }

  void scale_velocities(float f) { launch_scale_velocities(view(), n, f, stream); }

  void relax(double duration, float damping) {  // sph_solver.cu:707-736
    ensure_initialized();
    const int thermal = P.thermal;
    const int pc = cfg.phase_change;
    P.thermal = 0;
    cfg.phase_change = 0;
    sync_state();
    const double t_end = hs->time + duration;
    try {
      while (hs->time < t_end) {
        step();
        scale_velocities(damping);
        sync_state();
      }
    } catch (...) {
      P.thermal = thermal;
      cfg.phase_change = pc;
      throw;
    }
    P.thermal = thermal;
    cfg.phase_change = pc;
    scale_velocities(0.0f);
    sync_state();
    hs->time = 0;
    hs->step = 0;
    hs->last_dt = 0;
    hs->heat_input = hs->evaporated = hs->condensed = hs->vented = hs->vented_h = 0;
    hs->splits = hs->merges = 0;
    hs->merges_sorted = 0;
    merges_seen = 0;
    push_state();
    evaluate_rates();
    end_step(0, 0, false);
    sync_state();
  }

  // ------------------------------------------------------------ adjoint
  // A device copy of the persistent particle state, the device state block,
  // optionally `a` (at step boundaries) and the step's sort permutation (taped
  // post-sort states).
  struct Snap {
    DevBuf<float4> x, v, vt, nrm, a, ab, nrec, xlo;
    DevBuf<float> m, m0, rho, T, vol;
    DevBuf<int> lidx, lcnt, coff;  // neighbour lists and cell table of a taped rates evaluation
    DevBuf<uint8_t> cflags;
    DevBuf<uint8_t> phase, flags;
    DevBuf<uint32_t> id, perm;
    DevBuf<DevState> st;
    DevBuf<float4> sv[6];  // every scratch field (the end state: what the user reads after the call)
    DevBuf<float> sf[14];
    int n = 0;
  };
  // The scratch fields in a fixed order (view().t).
  void scratch_fields(float4* (&v)[6], float* (&f)[14]) {
    const Scratch t = view().t;
    float4* vv[6] = {t.a, t.ab, t.ns, t.vw, t.fw, t.mw};
    float* ff[14] = {t.p,      t.drho,   t.dTdt,  t.sigma, t.kappa,  t.nu_t, t.alpha,
                     t.wall_w, t.heat_w, t.out_m, t.out_h, t.w_norm, t.p_wv, t.rmin};
    for (int k = 0; k < 6; ++k) v[k] = vv[k];
    for (int k = 0; k < 14; ++k) f[k] = ff[k];
  }
  void snap_save_scratch(Snap& z) {
    float4* v[6];
    float* f[14];
    scratch_fields(v, f);
    {
      CopyBatch cb(*this);
      for (int k = 0; k < 6; ++k) fit(z.sv[k], n), d2d(z.sv[k].get(), v[k], n);
      for (int k = 0; k < 14; ++k) fit(z.sf[k], n), d2d(z.sf[k].get(), f[k], n);
    }
  }
  // A taped step's rates evaluation, after evaluate_rates(): the densities and
  // normals it wrote, every scratch field, the neighbour lists, the cell table
  // and the device state.  The reverse step loads it instead of re-sorting and
  // re-evaluating.
  void snap_save_rates(Snap& z) {
    const Persist q = persist(set);
    const Scratch t = view().t;
    const std::size_t nl = static_cast<std::size_t>(kListK) * ((static_cast<std::size_t>(n) + 63) / 64 * 64);
    {
      CopyBatch cb(*this);
      d2d(z.rho.get(), q.rho, n), d2d(z.nrm.get(), q.nrm, n);
      snap_save_scratch(z);
      fit(z.nrec, n), d2d(z.nrec.get(), t.nrec, n), fit(z.vol, n), d2d(z.vol.get(), t.vol, n);
      fit(z.lidx, nl), d2d(z.lidx.get(), nbr_idx.get(), static_cast<int>(nl));
      fit(z.lcnt, n), d2d(z.lcnt.get(), nbr_count.get(), n);
      fit(z.coff, g.num_cells + 2), d2d(z.coff.get(), cell_off.get(), g.num_cells + 2);
      fit(z.cflags, g.num_cells), d2d(z.cflags.get(), cell_flags.get(), g.num_cells);
      d2d(z.st.get(), state.get(), 1);
    }
  }
  void snap_load_rates(const Snap& z) {
    const Scratch t = view().t;
    const std::size_t nl = static_cast<std::size_t>(kListK) * ((static_cast<std::size_t>(z.n) + 63) / 64 * 64);
    {
      CopyBatch cb(*this);
      snap_load(z, false);
      snap_load_scratch(z);
      d2d(t.nrec, z.nrec.get(), z.n), d2d(t.vol, z.vol.get(), z.n);
      d2d(nbr_idx.get(), z.lidx.get(), static_cast<int>(nl)), d2d(nbr_count.get(), z.lcnt.get(), z.n);
      d2d(cell_off.get(), z.coff.get(), g.num_cells + 2), d2d(cell_flags.get(), z.cflags.get(), g.num_cells);
    }
    lists_valid = true;
    tiles_ok = true;
  }
  void snap_load_scratch(const Snap& z) {
    float4* v[6];
    float* f[14];
    scratch_fields(v, f);
    {
      CopyBatch cb(*this);
      for (int k = 0; k < 6; ++k) d2d(v[k], z.sv[k].get(), z.n);
      for (int k = 0; k < 14; ++k) d2d(f[k], z.sf[k].get(), z.n);
    }
  }
  struct CountRun {
    Snap z;
    int pending_dead = 0, merges_seen = 0, n_sorted = 0;
    const uint32_t* last_perm = nullptr;
  };
  std::vector<std::unique_ptr<Snap>> adj_ckpt, adj_tape;
  std::unique_ptr<Snap> adj_end;
  DevBuf<float4> adj_c[2][4];  // two cotangent sets of (x, v, a, ab)
  DevBuf<float4> adj_at, adj_wcp, adj_wcv, adj_xr, adj_vr, adj_cs, adj_gcb, adj_wnb, adj_xc, adj_tvs, adj_gbal, adj_sacc, adj_gpb,
      adj_vtr;
  DevBuf<float> adj_sb, adj_acc, adj_csv;
  DevBuf<double> adj_S, adj_G, adj_T, adj_part;
  // heat and phase change: m and T cotangents (two sets), the phase-change re-run, its cotangents
  DevBuf<float> adj_cm[2], adj_cT[2], adj_mr, adj_Tr, adj_qb, adj_T1, adj_om, adj_oh, adj_m2, adj_T2, adj_pV, adj_pr;
  DevBuf<float4> adj_v1, adj_pa, adj_po, adj_pb4, adj_pdp, adj_v1b, adj_pn, adj_xp;
  DevBuf<float2> adj_pm2;
  DevBuf<uint8_t> adj_fl;
  DevBuf<int> adj_amax;
  DevBuf<double> adj_hsum, adj_lee, adj_satv, adj_W;
  DevBuf<float> adj_w, adj_wd;
  Snap* tape_slot = nullptr;  // set: step_issue() records its post-sort state and its rates here

  // d2d between copy_begin() and copy_end() queues the copy; copy_end() issues
  // every queued copy in one launch (they must not depend on each other).
  std::vector<CopyItem> copy_q;
  int copy_depth = 0;
  void copy_begin() { ++copy_depth; }
  void copy_end() {  // launch errors surface at the next check_launch(); nothing here throws
    if (--copy_depth > 0) return;
    if (!copy_q.empty()) launch_copy_batch(copy_q.data(), static_cast<int>(copy_q.size()), stream);
    copy_q.clear();
  }
  struct CopyBatch {  // copy_begin() ... copy_end(), also when the scope unwinds
    cryo_solver& z;
    explicit CopyBatch(cryo_solver& s) : z(s) { z.copy_begin(); }
    ~CopyBatch() { z.copy_end(); }
    CopyBatch(const CopyBatch&) = delete;
    CopyBatch& operator=(const CopyBatch&) = delete;
  };
  template <class T>
  void d2d(T* dst, const T* src, int count) {
    if (count <= 0) return;
    const std::size_t bytes = sizeof(T) * static_cast<std::size_t>(count);
    if (copy_depth > 0 && ((reinterpret_cast<std::uintptr_t>(dst) | reinterpret_cast<std::uintptr_t>(src)) & 15) == 0) {
      copy_q.push_back(CopyItem{dst, src, bytes});
      return;
    }
    CRYO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream));
  }
  template <class T>
  static void fit(DevBuf<T>& b, std::size_t count) {
    if (b.size() < count) b.resize(count);
  }
  void snap_save(Snap& z, bool with_a, const uint32_t* perm) {
    const int m = n;
    const Persist q = persist(set);
    fit(z.x, m), fit(z.v, m), fit(z.vt, m), fit(z.nrm, m), fit(z.m, m), fit(z.m0, m), fit(z.rho, m), fit(z.T, m);
    fit(z.phase, m), fit(z.flags, m), fit(z.id, m), fit(z.st, 1);
    if (q.xlo) fit(z.xlo, m);
    if (with_a) fit(z.a, m), fit(z.ab, m);
    if (perm) fit(z.perm, m);
    {
      CopyBatch cb(*this);
      d2d(z.x.get(), q.x, m), d2d(z.v.get(), q.v, m), d2d(z.vt.get(), q.vt, m), d2d(z.nrm.get(), q.nrm, m);
      d2d(z.m.get(), q.m, m), d2d(z.m0.get(), q.m0, m), d2d(z.rho.get(), q.rho, m), d2d(z.T.get(), q.T, m);
      d2d(z.phase.get(), q.phase, m), d2d(z.flags.get(), q.flags, m), d2d(z.id.get(), q.id, m);
      if (q.xlo) d2d(z.xlo.get(), q.xlo, m);
      d2d(z.st.get(), state.get(), 1);
      if (with_a) d2d(z.a.get(), a.get(), m), d2d(z.ab.get(), ab.get(), m);
      if (perm) d2d(z.perm.get(), perm, m);
    }
    z.n = m;
  }
  void snap_load(const Snap& z, bool with_a) {
    const int m = z.n;
    const Persist q = persist(set);
    {
      CopyBatch cb(*this);
      d2d(q.x, z.x.get(), m), d2d(q.v, z.v.get(), m), d2d(q.vt, z.vt.get(), m), d2d(q.nrm, z.nrm.get(), m);
      d2d(q.m, z.m.get(), m), d2d(q.m0, z.m0.get(), m), d2d(q.rho, z.rho.get(), m), d2d(q.T, z.T.get(), m);
      d2d(q.phase, z.phase.get(), m), d2d(q.flags, z.flags.get(), m), d2d(q.id, z.id.get(), m);
      if (q.xlo) d2d(q.xlo, z.xlo.get(), m);
      d2d(state.get(), z.st.get(), 1);
      if (with_a) d2d(a.get(), z.a.get(), m), d2d(ab.get(), z.ab.get(), m);
    }
    n = m;
    lists_valid = false;
    tiles_ok = false;
  }

  bool adj_heat() const { return P.thermal != 0; }
  bool adj_pc() const { return cfg.phase_change != 0; }
  AdjArgs adj_args(int cur, const uint32_t* perm, const double* S, bool csf, double* lee_out = nullptr,
                   bool sat_prev = false) {
    AdjArgs A{adj_c[cur][0].get(), adj_c[cur][1].get(), adj_c[cur][2].get(),
              adj_c[1 - cur][0].get(), adj_c[1 - cur][1].get(), adj_c[1 - cur][2].get(),
              adj_at.get(), adj_wcp.get(), adj_wcv.get(), adj_xr.get(), adj_vr.get(), adj_sb.get(),
              perm, S, adj_part.get(), stride, csf ? 1 : 0, adj_cs.get(), adj_csv.get(), adj_gcb.get(),
              adj_wnb.get(), adj_xc.get(), P.tvf ? 1 : 0, adj_c[cur][3].get(), adj_c[1 - cur][3].get(),
              adj_tvs.get()};
    A.gbal = csf && P.csf_balanced ? adj_gbal.get() : nullptr;
    A.sacc = csf && P.csf_smooth ? adj_sacc.get() : nullptr;
    A.gpb = csf && P.wetting != 0 ? adj_gpb.get() : nullptr;
    A.vtr = P.tvf && P.tvf_stress ? adj_vtr.get() : nullptr;
    A.heat = adj_heat() ? 1 : 0;
    A.pc = adj_pc() ? 1 : 0;
    A.mT = (A.heat || A.pc) ? 1 : 0;
    A.reb = A.pc && P.vapor_rebalance > 0.0f && has_vapor;
    A.vent = A.pc && cfg.thermo_mode == CRYO_VENTED && has_vapor;
    A.relax = cfg.vent_relaxation;
    if (A.mT) {
      A.cm = adj_cm[cur].get(), A.cT = adj_cT[cur].get(), A.nm = adj_cm[1 - cur].get(), A.nT = adj_cT[1 - cur].get();
      A.mr = adj_mr.get(), A.Tr = adj_Tr.get(), A.qb = adj_qb.get(), A.hsum = adj_hsum.get();
    }
    if (A.pc) {
      A.v1 = adj_v1.get(), A.T1 = adj_T1.get(), A.om = adj_om.get(), A.oh = adj_oh.get(), A.m2 = adj_m2.get();
      A.T2 = adj_T2.get(), A.fl = adj_fl.get(), A.pa = adj_pa.get(), A.po = adj_po.get(), A.pb4 = adj_pb4.get();
      A.pdp = adj_pdp.get(), A.v1b = adj_v1b.get(), A.pm2 = adj_pm2.get(), A.pn = adj_pn.get();
      A.amax = adj_amax.get(), A.pV = adj_pV.get(), A.pr = adj_pr.get(), A.xp = adj_xp.get();
      A.lee_out = lee_out;
      A.sat = sat_prev && cfg.thermo_mode == CRYO_CLOSED && has_vapor ? 1 : 0;
      A.fluid = fluid_const();
      A.satv = adj_satv.get();
    }
    return A;
  }
  // The reverse step's forward re-run of the kick and of phase change up to
  // apply on the taped state t (rates evaluated): keeps the kicked v and T,
  // the sources' flags, the normalised out_m / out_h and the applied m and T,
  // then puts the taped v, T, m, flags and phase back.
  void adj_pc_rerun(const Snap& t) {
    const Persist q = persist(set);
    if (!adj_pc()) return;
    launch_kick_sources(view(), P, state.get(), n, true, P.thermal != 0, true, stream);
    {
      CopyBatch cb(*this);
      d2d(adj_v1.get(), q.v, n), d2d(adj_T1.get(), q.T, n), d2d(adj_fl.get(), q.flags, n);
    }
    launch_pc_normalize(view(), grid(), P, lists(), n, partials.get(), stride, stream);
    {
      CopyBatch cb(*this);
      d2d(adj_om.get(), view().t.out_m, n), d2d(adj_oh.get(), view().t.out_h, n);
    }
    launch_pc_apply(view(), grid(), P, lists(), n, stream);
    {
      CopyBatch cb(*this);
      d2d(adj_m2.get(), q.m, n), d2d(adj_T2.get(), q.T, n);
    }
    {
      CopyBatch cb(*this);
      d2d(q.v, t.v.get(), n), d2d(q.T, t.T.get(), n), d2d(q.m, t.m.get(), n);
      d2d(q.flags, t.flags.get(), n), d2d(q.phase, t.phase.get(), n);
    }
  }

  void check_adjoint_scope(bool explicit_rotation = false) const {
    const cryo_config& c = cfg;
    if (!(c.fixed_dt > 0.0f)) throw std::invalid_argument("adjoint: needs fixed_dt > 0");
    if (c.density_mode != CRYO_DENSITY_SUMMATION) throw std::invalid_argument("adjoint: density summation only");
    if (c.surface_tension == CRYO_ST_AKINCI) throw std::invalid_argument("adjoint: Akinci surface tension not differentiated");
    if (c.sps_turbulence || c.particle_shifting)
      throw std::invalid_argument("adjoint: SPS and shifting not differentiated");
    if (has_maneuver) throw std::invalid_argument("adjoint: manoeuvre table not allowed (the frame is per-step input)");
    if (frame_rot && !explicit_rotation)
      throw std::invalid_argument("adjoint: a rotating frame is differentiated only with per-step rotation inputs");
  }

  // loss = sum_k dt (S_x^2 + S_y^2), S_k = sum of fw over fluid particles
  // after step k's rates; step k runs with frame acceleration acc[3k..3k+2].
  // grad (optional) = d loss / d acc (and grad_theta = d loss / d contact
  // angle) by the reverse sweep: checkpoints every `every` steps, each segment
  // re-run forward with its post-sort states taped, then walked back one step
  // at a time (the rates but the forces pass re-evaluated on the taped
  // state).  The solver ends in the forward end state.
  // omega / omega_dot (optional, 3 per step): the frame's angular velocity and
  // its rate per step (a -= wd x r + w x (w x r) + 2 w x v); grad_omega /
  // grad_omega_dot (optional) receive their derivatives.
  void adjoint_slosh(long long steps, const float* acc, int every, double* loss, double* grad, double* force,
                     double* grad_theta, double* grad_lee, const float* omega = nullptr,
                     const float* omega_dot = nullptr, double* grad_omega = nullptr, double* grad_omega_dot = nullptr) {
    const bool rot = omega || omega_dot;
    check_adjoint_scope(rot);
    if (steps < 0) throw std::invalid_argument("adjoint: steps must be >= 0");
    if (every <= 0) throw std::invalid_argument("adjoint: checkpoint_every must be > 0");
    if (!loss) throw std::invalid_argument("adjoint: loss output is required");
    if (grad_theta && !grad) throw std::invalid_argument("adjoint: the contact-angle gradient needs grad_frame_acc");
    if (grad_lee && !grad) throw std::invalid_argument("adjoint: the Lee-coefficient gradient needs grad_frame_acc");
    if ((grad_omega || grad_omega_dot) && !grad)
      throw std::invalid_argument("adjoint: the rotation gradients need grad_frame_acc");
    const bool fresh = !initialized;  // the start-up rates (and their a) belong to this run
    ensure_initialized();
    *loss = 0.0;
    if (grad_theta) *grad_theta = 0.0;
    if (grad_lee) grad_lee[0] = grad_lee[1] = 0.0;
    if (steps == 0) return;
    const bool mT = adj_heat() || adj_pc();
    sync_state();
    const long long events0 = static_cast<long long>(hs->merges) + hs->splits;
    const bool csf = cfg.surface_tension == CRYO_ST_CSF && has_vapor;
    const std::size_t ns = static_cast<std::size_t>(steps);
    fit(adj_acc, 3 * ns), fit(adj_S, 3 * ns), fit(adj_G, 3 * ns), fit(adj_T, 2 * ns + 2);
    fit(adj_part, 21 * static_cast<std::size_t>(stride));
    fit(adj_satv, 2);
    if (adj_pc()) fit(adj_lee, 2 * ns);
    CRYO_HIP(hipMemcpyAsync(adj_acc.get(), acc, sizeof(float) * 3 * ns, hipMemcpyHostToDevice, stream));
    const float* dw = nullptr;
    const float* dwd = nullptr;
    if (omega) {
      fit(adj_w, 3 * ns);
      CRYO_HIP(hipMemcpyAsync(adj_w.get(), omega, sizeof(float) * 3 * ns, hipMemcpyHostToDevice, stream));
      dw = adj_w.get();
    }
    if (omega_dot) {
      fit(adj_wd, 3 * ns);
      CRYO_HIP(hipMemcpyAsync(adj_wd.get(), omega_dot, sizeof(float) * 3 * ns, hipMemcpyHostToDevice, stream));
      dwd = adj_wd.get();
    }
    const bool want_w = grad && (grad_omega || grad_omega_dot);
    if (want_w) fit(adj_W, 6 * ns);
    auto set_frame = [&](long long k) {
      launch_adj_set_frame(state.get(), adj_acc.get() + 3 * k, dw ? dw + 3 * k : nullptr, dwd ? dwd + 3 * k : nullptr,
                           stream);
    };
    const long long K = every;
    const long long nseg = (steps + K - 1) / K;
    const long long k_last = (nseg - 1) * K;  // the last segment is taped by the forward pass itself
    if (grad) {
      adj_ckpt.resize(static_cast<std::size_t>(nseg));
      for (auto& c : adj_ckpt)
        if (!c) c = std::make_unique<Snap>();
      adj_tape.resize(static_cast<std::size_t>(std::min(K, steps)));
      for (auto& t : adj_tape)
        if (!t) t = std::make_unique<Snap>();
    }
    for (long long k = 0; k < steps; ++k) {
      if (grad && k % K == 0) snap_save(*adj_ckpt[static_cast<std::size_t>(k / K)], true, nullptr);
      set_frame(k);
      tape_slot = (grad && k >= k_last) ? adj_tape[static_cast<std::size_t>(k - k_last)].get() : nullptr;
      step_issue();
      tape_slot = nullptr;
      step_finish();
      launch_adj_load_sum(view(), n, adj_part.get(), stride, adj_S.get() + 3 * k, stream);
    }
    check_launch();
    if (grad && mT) {
      sync_state();
      if (static_cast<long long>(hs->merges) + hs->splits != events0)
        throw std::invalid_argument("adjoint: a phase-change merge or split inside the window is not differentiable");
    }
    if (grad) {
      if (!adj_end) adj_end = std::make_unique<Snap>();
      snap_save(*adj_end, true, nullptr);
      snap_save_scratch(*adj_end);
      const std::size_t m = static_cast<std::size_t>(n);
      for (auto& set3 : adj_c)
        for (auto& b : set3) fit(b, m);
      fit(adj_at, m), fit(adj_wcp, m), fit(adj_wcv, m), fit(adj_xr, m), fit(adj_vr, m), fit(adj_sb, m);
      if (csf) fit(adj_cs, m), fit(adj_csv, m), fit(adj_gcb, m), fit(adj_wnb, m), fit(adj_xc, m);
      if (csf && P.csf_balanced) fit(adj_gbal, m);
      if (csf && P.csf_smooth) fit(adj_sacc, m);
      if (csf && P.wetting != 0) fit(adj_gpb, m);
      if (P.tvf && P.tvf_stress) fit(adj_vtr, m);
      fit(adj_tvs, m);
      launch_adj_fill(adj_c[0][0].get(), adj_c[0][1].get(), adj_c[0][2].get(), n, stream);
      launch_adj_fill(adj_c[0][3].get(), adj_c[1][3].get(), adj_tvs.get(), n, stream);
      if (mT) {
        for (auto& b : adj_cm) fit(b, m);
        for (auto& b : adj_cT) fit(b, m);
        fit(adj_mr, m), fit(adj_Tr, m), fit(adj_qb, m), fit(adj_hsum, 1);
        CRYO_HIP(hipMemsetAsync(adj_cm[0].get(), 0, sizeof(float) * m, stream));
        CRYO_HIP(hipMemsetAsync(adj_cT[0].get(), 0, sizeof(float) * m, stream));
        CRYO_HIP(hipMemsetAsync(adj_hsum.get(), 0, sizeof(double), stream));
      }
      if (adj_pc()) {
        fit(adj_v1, m), fit(adj_T1, m), fit(adj_om, m), fit(adj_oh, m), fit(adj_m2, m), fit(adj_T2, m);
        fit(adj_fl, m), fit(adj_pa, m), fit(adj_po, m), fit(adj_pb4, m), fit(adj_pdp, m), fit(adj_v1b, m);
        fit(adj_pm2, m), fit(adj_pn, m), fit(adj_amax, m), fit(adj_pV, m), fit(adj_pr, m), fit(adj_xp, m);
      }
      int cur = 0;
      for (long long seg = nseg - 1; seg >= 0; --seg) {
        const long long k0 = seg * K, k1 = std::min(steps, k0 + K);
        if (seg < nseg - 1) snap_load(*adj_ckpt[static_cast<std::size_t>(seg)], true);
        for (long long k = k0; k < k1 && seg < nseg - 1; ++k) {
          set_frame(k);
          tape_slot = adj_tape[static_cast<std::size_t>(k - k0)].get();
          step_issue();
          tape_slot = nullptr;
          step_finish();
        }
        for (long long k = k1 - 1; k >= k0; --k) {
          const Snap& t = *adj_tape[static_cast<std::size_t>(k - k0)];
          snap_load_rates(t);  // the step's sorted state and its rates
          if (mT) adj_pc_rerun(t);
          AdjArgs A = adj_args(cur, t.perm.get(), adj_S.get() + 3 * k, csf,
                               adj_pc() ? adj_lee.get() + 2 * k : nullptr, k > 0);
          A.wout = want_w ? adj_W.get() + 6 * k : nullptr;
          launch_adj_step(view(), grid(), P, lists(), state.get(), A, n, adj_G.get() + 3 * k, adj_T.get() + 2 * k,
                          stream);
          check_launch();
          cur = 1 - cur;
        }
      }
      if (fresh && csf && grad_theta) {
        // the start-up a (checkpoint 0) depends on the contact angle too:
        // adj_c[cur][2] holds its cotangent
        snap_load(*adj_ckpt[0], true);
        grid_build(false);
        evaluate_rates(false);
        d2d(adj_at.get(), adj_c[cur][2].get(), n);
        const AdjArgs A = adj_args(cur, nullptr, nullptr, true);
        launch_adj_csf_only(view(), grid(), P, lists(), state.get(), A, n, adj_T.get() + 2 * ns, stream);
        check_launch();
      }
      snap_load(*adj_end, true);
      grid_build(false);  // the end state's own cell table (identity permutation)
      snap_load_scratch(*adj_end);
    }
    std::vector<double> S(3 * ns), T(2 * ns + 2, 0.0);
    CRYO_HIP(hipMemcpyAsync(S.data(), adj_S.get(), sizeof(double) * 3 * ns, hipMemcpyDeviceToHost, stream));
    if (grad) CRYO_HIP(hipMemcpyAsync(grad, adj_G.get(), sizeof(double) * 3 * ns, hipMemcpyDeviceToHost, stream));
    const bool start_term = fresh && csf && grad_theta;
    if (grad && csf) CRYO_HIP(hipMemcpyAsync(T.data(), adj_T.get(), sizeof(double) * (2 * ns + (start_term ? 2 : 0)),
                                             hipMemcpyDeviceToHost, stream));
    std::vector<double> lee(2 * ns, 0.0), W(want_w ? 6 * ns : 0);
    if (grad && adj_pc())
      CRYO_HIP(hipMemcpyAsync(lee.data(), adj_lee.get(), sizeof(double) * 2 * ns, hipMemcpyDeviceToHost, stream));
    if (want_w) CRYO_HIP(hipMemcpyAsync(W.data(), adj_W.get(), sizeof(double) * 6 * ns, hipMemcpyDeviceToHost, stream));
    CRYO_HIP(hipStreamSynchronize(stream));
    for (std::size_t q = 0; q < (want_w ? ns : 0); ++q)
      for (int c = 0; c < 3; ++c) {
        if (grad_omega) grad_omega[3 * q + c] = W[6 * q + c];
        if (grad_omega_dot) grad_omega_dot[3 * q + c] = W[6 * q + 3 + c];
      }
    if (grad_lee)
      for (std::size_t q = 0; q < ns; ++q) grad_lee[0] += lee[2 * q], grad_lee[1] += lee[2 * q + 1];
    if (grad_theta) {  // d loss / d theta = (cos_bar (-sin theta) + sin_bar cos theta) pi / 180
      double cb = 0.0, sb = 0.0;
      for (std::size_t q = 0; q < ns + (start_term ? 1 : 0); ++q) cb += T[2 * q], sb += T[2 * q + 1];
      *grad_theta = (cb * -static_cast<double>(P.sin_theta) + sb * static_cast<double>(P.cos_theta)) *
                    (3.14159265358979323846 / 180.0);
    }
    const double dt = static_cast<double>(cfg.fixed_dt);
    double Lsum = 0.0;
    for (std::size_t k = 0; k < ns; ++k) Lsum += dt * (S[3 * k] * S[3 * k] + S[3 * k + 1] * S[3 * k + 1]);
    *loss = Lsum;
    if (force) std::memcpy(force, S.data(), sizeof(double) * 3 * ns);
    sync_state();
    // the host's frame is the last step's, as on the device
    frame_acc = hs->frame_acc;
    frame_omega = hs->omega;
    frame_omega_dot = hs->omega_dot;
    frame_rot = hs->frame_rot;
  }

  // ----------------------------------------------------------- autotune
  int launch_variant() const { return (tile_forces ? 0 : 1) | (tile_iface ? 0 : 2); }
  void set_launch_variant(int v) {
    if (v < 0 || v > 3) throw std::invalid_argument("launch variant must be 0..3");
    tile_forces = (v & 1) == 0;
    tile_iface = (v & 2) == 0;
  }
  void set_tile_shape(int x, int y, int z) {
    if (!tile_shape_ok(x, y, z))
      throw std::invalid_argument("tile shape: need 1 <= x <= 16, y, z >= 1, (y + 2)(z + 2) <= 25 and y z <= 9");
    tile_shape[0] = x, tile_shape[1] = y, tile_shape[2] = z;
  }
  // Times `trials` steps of each of `count` configurations from the current
  // state (after one untimed step each; apply(k) selects configuration k),
  // selects the fastest and puts the state back.  ms[k] (optional) receives
  // configuration k's milliseconds per step.  Returns the chosen k; on an
  // error the configuration `current` is selected again.
  template <class Apply>
  int time_configurations(int trials, int count, int current, Apply&& apply, double* ms) {
    if (trials <= 0) throw std::invalid_argument("autotune: trial_steps must be > 0");
    ensure_initialized();
    sync_state();
    Snap z;
    snap_save(z, true, nullptr);
    snap_save_scratch(z);
    const int pd = pending_dead;
    const int ms_seen = merges_seen;
    hipEvent_t e0, e1;
    CRYO_HIP(hipEventCreate(&e0));
    CRYO_HIP(hipEventCreate(&e1));
    double best = 0.0;
    int pick = current;
    auto restore = [&]() {
      snap_load(z, true);
      snap_load_scratch(z);
      pending_dead = pd;
      merges_seen = ms_seen;
      sync_state();
    };
    try {
      for (int k = 0; k < count; ++k) {
        restore();
        apply(k);
        step();
        CRYO_HIP(hipEventRecord(e0, stream));
        for (int r = 0; r < trials; ++r) step();
        CRYO_HIP(hipEventRecord(e1, stream));
        CRYO_HIP(hipEventSynchronize(e1));
        float elapsed = 0.0f;
        CRYO_HIP(hipEventElapsedTime(&elapsed, e0, e1));
        const double per = static_cast<double>(elapsed) / trials;
        if (ms) ms[k] = per;
        if (k == 0 || per < best) best = per, pick = k;
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      apply(current);
      restore();
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    apply(pick);
    restore();
    return pick;
  }
  // Every launch variant (ms: 4 entries).
  int autotune(int trials, double* ms) {
    return time_configurations(trials, 4, launch_variant(), [this](int v) { set_launch_variant(v); }, ms);
  }
  // Tile shapes shapes[3k..3k+2], k < count, at the current launch variant;
  // shapes == nullptr: the built-in candidates below (count ignored; ms then
  // needs kTileCandidates entries).  Returns the index of the chosen shape.
  static constexpr int kTileCandidates = 6;
  int autotune_tiles(int trials, const int32_t* shapes, int count, double* ms) {
    static constexpr int32_t builtin[kTileCandidates * 3] = {3, 3, 3, 4, 3, 2, 4, 2, 3, 2, 3, 3, 3, 3, 2, 3, 2, 3};
    if (!shapes) shapes = builtin, count = kTileCandidates;
    if (count <= 0) throw std::invalid_argument("autotune_tiles: need at least one tile shape");
    for (int k = 0; k < count; ++k)
      if (!tile_shape_ok(shapes[3 * k], shapes[3 * k + 1], shapes[3 * k + 2]))
        throw std::invalid_argument("autotune_tiles: tile shape " + std::to_string(k) + " out of range");
    const int keep[3] = {tile_shape[0], tile_shape[1], tile_shape[2]};
    // the current shape is put back on an error: apply(count) means "keep"
    return time_configurations(
        trials, count, count,
        [&](int k) {
          if (k >= count) set_tile_shape(keep[0], keep[1], keep[2]);
          else set_tile_shape(shapes[3 * k], shapes[3 * k + 1], shapes[3 * k + 2]);
        },
        ms);
  }

  // -------------------------------------------------------------- frame
  void apply_frame(V3 acc, V3 w, V3 wd, V3 origin) {  // sph_solver.cu:660-667
    sync_state();
    frame_acc = acc;
    frame_omega = w;
    frame_omega_dot = wd;
    frame_origin = origin;
    frame_rot = (norm2(w) > 0.0f || norm2(wd) > 0.0f) ? 1 : 0;
    hs->frame_acc = acc;
    hs->gravity = v3(cfg.gravity) - acc;
    hs->omega = w;
    hs->omega_dot = wd;
    hs->origin = origin;
    hs->frame_rot = frame_rot;
    push_state();
  }

  void set_maneuver(const HostManeuver& mv) {  // sph_solver.cu:674-681
    mv.validate();
    hman = mv;
    const int cnt = static_cast<int>(mv.time.size());
    man_time.resize(cnt);
    man_acc.resize(cnt);
    CRYO_HIP(hipMemcpyAsync(man_time.get(), mv.time.data(), sizeof(double) * cnt, hipMemcpyHostToDevice, stream));
    CRYO_HIP(hipMemcpyAsync(man_acc.get(), mv.acc.data(), sizeof(V3) * cnt, hipMemcpyHostToDevice, stream));
    if (!mv.omega.empty()) {
      man_omega.resize(cnt);
      CRYO_HIP(hipMemcpyAsync(man_omega.get(), mv.omega.data(), sizeof(V3) * cnt, hipMemcpyHostToDevice, stream));
    }
    man = Maneuver{cnt, man_time.get(), man_acc.get(), mv.omega.empty() ? nullptr : man_omega.get(), mv.origin};
    has_maneuver = true;
    sync_state();
    V3 a3, w3, wd3;
    hman.evaluate(hs->time, a3, w3, wd3);
    apply_frame(a3, w3, wd3, mv.origin);
  }

  void clear_maneuver() {
    has_maneuver = false;
    man = Maneuver{};
  }

  // ----------------------------------------------------------- queries
  void diagnostics(cryo_diagnostics* d) {  // sph_solver.cu:609-653
    ensure_initialized();
    launch_diagnostics(view(), P, n, partials.get(), stride, diag_max.get(), diag_out.get(), stream);
    check_launch();
    std::vector<double> q(static_cast<std::size_t>(diag_quantities()));
    unsigned mx[2];
    CRYO_HIP(hipMemcpyAsync(q.data(), diag_out.get(), sizeof(double) * q.size(), hipMemcpyDeviceToHost, stream));
    CRYO_HIP(hipMemcpyAsync(mx, diag_max.get(), sizeof(mx), hipMemcpyDeviceToHost, stream));
    sync_state();
    float fmx[2];
    std::memcpy(fmx, mx, sizeof(fmx));
    std::memset(d, 0, sizeof(*d));
    const DevState& s = *hs;
    d->time = s.time;
    d->step = s.step;
    d->dt = s.last_dt;
    d->n_total = n;
    d->n_liquid = static_cast<int>(q[0]);
    d->n_vapor = static_cast<int>(q[1]);
    d->n_wall = static_cast<int>(q[2]);
    d->mass_liquid = q[3];
    d->mass_vapor = q[4];
    d->mass_total = d->mass_liquid + d->mass_vapor;
    d->kinetic_energy = q[5] + q[6];
    d->enthalpy = q[7] + q[8];
    d->heat_input = s.heat_input;
    d->evaporated_mass = s.evaporated;
    d->condensed_mass = s.condensed;
    d->vented_mass = s.vented;
    d->vented_enthalpy = s.vented_h;
    d->T_sat = s.T_sat;
    d->p_ullage = s.p_ullage;
    d->max_speed = fmx[0];
    if (d->mass_liquid > 0) {
      d->T_liquid_mean = q[9] / d->mass_liquid;
      d->T_liquid_max = fmx[1];
      d->liquid_com = cryo_vec3{static_cast<float>(q[11] / d->mass_liquid), static_cast<float>(q[12] / d->mass_liquid),
                                static_cast<float>(q[13] / d->mass_liquid)};
      d->liquid_momentum = cryo_vec3{static_cast<float>(q[14]), static_cast<float>(q[15]), static_cast<float>(q[16])};
    }
    if (d->mass_vapor > 0) d->T_vapor_mean = q[10] / d->mass_vapor;
    d->splits = s.splits;
    d->merges = s.merges;
    d->slosh_force = cryo_vec3{-static_cast<float>(q[17]), -static_cast<float>(q[18]), -static_cast<float>(q[19])};
    d->slosh_moment = cryo_vec3{-static_cast<float>(q[20]), -static_cast<float>(q[21]), -static_cast<float>(q[22])};
    d->frame_acceleration = cv(s.frame_acc);
    d->frame_angular_velocity = cv(s.omega);
  }

  void params(cryo_params* o) {
    sync_state();
    std::memset(o, 0, sizeof(*o));
    o->h = P.kernel.h;
    o->inv_h = P.kernel.inv_h;
    o->support = P.kernel.support;
    o->kernel_norm = P.kernel.norm;
    o->support2 = P.support2;
    o->dx = P.dx;
    o->vol0 = P.vol0;
    for (int k = 0; k < 2; ++k) {
      o->rho0[k] = P.rho0[k];
      o->mu[k] = P.mu[k];
      o->k[k] = P.k[k];
      o->cp[k] = P.cp[k];
      o->mref[k] = P.mref[k];
    }
    o->c0 = P.c0;
    o->tait_B = P.tait_B;
    o->tait_gamma = P.tait_gamma;
    o->p_bg = P.p_bg;
    o->gas_R = P.gas_R;
    o->vapor_K = P.vapor_K;
    o->gravity = cv(hs->gravity);
    o->omega = cv(hs->omega);
    o->omega_dot = cv(hs->omega_dot);
    o->frame_origin = cv(hs->origin);
    o->frame_rot = hs->frame_rot;
    o->sigma = P.sigma;
    o->cos_theta = P.cos_theta;
    o->sin_theta = P.sin_theta;
    o->normal_eps = P.normal_eps;
    o->akinci_gamma = P.ak_gamma;
    o->akinci_beta = P.ak_beta;
    o->tvf_sigma_min = P.tvf_sigma_min;
    o->shift_strength = P.shift_strength;
    o->T_ref = P.T_ref;
    o->T_sat = hs->T_sat;
    o->h_fg = P.h_fg;
    o->heat_inv_norm = hs->heat_inv_norm;
    o->dt = hs->dt;
  }

  void download(int field, void* dst, int count) {
    if (count < 0 || count > n) throw std::out_of_range("cryo_solver_download: count exceeds num_particles");
    if (count == 0) return;
    const int k = set;
    auto d4 = [&](const DevBuf<float4>& b) {
      std::vector<float4> h(static_cast<std::size_t>(count));
      CRYO_HIP(hipMemcpyAsync(h.data(), b.get(), sizeof(float4) * count, hipMemcpyDeviceToHost, stream));
      CRYO_HIP(hipStreamSynchronize(stream));
      float* o = static_cast<float*>(dst);
      for (int i = 0; i < count; ++i) {
        o[3 * i] = h[i].x;
        o[3 * i + 1] = h[i].y;
        o[3 * i + 2] = h[i].z;
      }
    };
    auto d1 = [&](const auto& b) {
      CRYO_HIP(hipMemcpyAsync(dst, b.get(), sizeof(*b.get()) * count, hipMemcpyDeviceToHost, stream));
      CRYO_HIP(hipStreamSynchronize(stream));
    };
    switch (field) {
      case CRYO_F_X: d4(x_[k]); break;
      case CRYO_F_V: d4(v_[k]); break;
      case CRYO_F_VT: d4(vt_[k]); break;
      case CRYO_F_NRM: d4(nrm_[k]); break;
      case CRYO_F_M: d1(m_[k]); break;
      case CRYO_F_M0: d1(m0_[k]); break;
      case CRYO_F_RHO: d1(rho_[k]); break;
      case CRYO_F_T: d1(T_[k]); break;
      case CRYO_F_PHASE: d1(phase_[k]); break;
      case CRYO_F_FLAGS: d1(flags_[k]); break;
      case CRYO_F_ID: d1(id_[k]); break;
      case CRYO_F_A: d4(a); break;
      case CRYO_F_AB: d4(ab); break;
      case CRYO_F_NS: d4(ns); break;
      case CRYO_F_VW: d4(vw); break;
      case CRYO_F_FW: d4(fw); break;
      case CRYO_F_MW: d4(mw); break;
      case CRYO_F_P: d1(p); break;
      case CRYO_F_DRHO: d1(drho); break;
      case CRYO_F_DTDT: d1(dTdt); break;
      case CRYO_F_SIGMA: d1(sigma); break;
      case CRYO_F_KAPPA: d1(kappa); break;
      case CRYO_F_NU_T: d1(nu_t); break;
      case CRYO_F_ALPHA: d1(alpha); break;
      case CRYO_F_WALL_W: d1(wall_w); break;
      case CRYO_F_HEAT_W: d1(heat_w); break;
      case CRYO_F_OUT_M: d1(out_m); break;
      case CRYO_F_OUT_H: d1(out_h); break;
      case CRYO_F_W_NORM: d1(w_norm); break;
      case CRYO_F_P_WV: d1(p_wv); break;
      case CRYO_F_RMIN: d1(rmin); break;
      default: throw std::invalid_argument("cryo_solver_download: unknown field");
    }
  }

  void neighbor_counts(float radius, int32_t* counts, uint64_t* sums) {
    DevBuf<int> dc;
    DevBuf<unsigned long long> ds;
    dc.resize(std::max(n, 1));
    ds.resize(std::max(n, 1));
    const PV pv = view();
    launch_neighbor_stats(pv.s.x, pv.s.id, grid(), n, radius * radius, dc.get(), ds.get(), stream);
    check_launch();
    if (n > 0) {
      CRYO_HIP(hipMemcpyAsync(counts, dc.get(), sizeof(int) * n, hipMemcpyDeviceToHost, stream));
      if (sums) CRYO_HIP(hipMemcpyAsync(sums, ds.get(), sizeof(uint64_t) * n, hipMemcpyDeviceToHost, stream));
    }
    CRYO_HIP(hipStreamSynchronize(stream));
  }

  void neighbor_csr(float radius, int64_t* offsets, int32_t* indices) {
    std::vector<int32_t> c(static_cast<std::size_t>(n));
    neighbor_counts(radius, c.data(), nullptr);
    offsets[0] = 0;
    for (int i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + c[i];
    if (!indices || n == 0) return;
    DevBuf<long long> doff;
    DevBuf<int> didx;
    doff.resize(n + 1);
    didx.resize(std::max<int64_t>(offsets[n], 1));
    CRYO_HIP(hipMemcpyAsync(doff.get(), offsets, sizeof(long long) * (n + 1), hipMemcpyHostToDevice, stream));
    launch_neighbor_fill(view().s.x, grid(), n, radius * radius, doff.get(), didx.get(), stream);
    check_launch();
    CRYO_HIP(hipMemcpyAsync(indices, didx.get(), sizeof(int) * offsets[n], hipMemcpyDeviceToHost, stream));
    CRYO_HIP(hipStreamSynchronize(stream));
  }

  // ---------------------------------------------------------- telemetry
  void release_telemetry() {
    if (tele_pending) {
      (void)hipEventSynchronize(tele_landed);
      tele_pending = false;
    }
    if (tele_host) (void)hipHostFree(tele_host);
    tele_host = nullptr;
    tele_host_dev = nullptr;
    tele_staging.release();
    tele_cap = 0;
  }
  void configure_telemetry(int mode) {  // telemetry.cu:27-54
    if (mode != CRYO_TELEMETRY_ASYNC_COPY && mode != CRYO_TELEMETRY_ZERO_COPY)
      throw std::invalid_argument("telemetry mode must be AsyncCopy or ZeroCopy");
    release_telemetry();
    tele_mode = mode;
    if (capacity <= 0) return;
    const std::size_t bytes = sizeof(float) * 7 * static_cast<std::size_t>(capacity);
    CRYO_HIP(hipHostMalloc(reinterpret_cast<void**>(&tele_host), bytes,
                           mode == CRYO_TELEMETRY_ZERO_COPY ? hipHostMallocMapped : hipHostMallocDefault));
    if (mode == CRYO_TELEMETRY_ZERO_COPY)
      CRYO_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&tele_host_dev), tele_host, 0));
    else
      tele_staging.resize(7 * static_cast<std::size_t>(capacity));
    tele_cap = capacity;
  }
  bool request_telemetry() {  // telemetry.cu:56-96
    ensure_initialized();
    if (tele_pending || n > tele_cap) return false;
    sync_state();
    tele_time = hs->time;
    tele_step = hs->step;
    tele_n = n;
    if (tele_mode == CRYO_TELEMETRY_ZERO_COPY) {
      launch_telemetry_pack(view(), n, tele_host_dev, stream);
      CRYO_HIP(hipEventRecord(tele_landed, stream));
    } else {
      launch_telemetry_pack(view(), n, tele_staging.get(), stream);
      CRYO_HIP(hipEventRecord(tele_computed, stream));
      CRYO_HIP(hipStreamWaitEvent(copy_stream, tele_computed, 0));
      CRYO_HIP(hipMemcpyAsync(tele_host, tele_staging.get(), sizeof(float) * 7 * n, hipMemcpyDeviceToHost,
                              copy_stream));
      CRYO_HIP(hipEventRecord(tele_landed, copy_stream));
    }
    check_launch();
    tele_pending = true;
    return true;
  }
  bool telemetry_ready() {
    if (!tele_pending) return true;
    const hipError_t e = hipEventQuery(tele_landed);
    if (e == hipSuccess) return true;
    if (e == hipErrorNotReady) return false;
    CRYO_HIP(e);
    return false;
  }
};

// --------------------------------------- probe scratch (allocation-free entry)
// cryo_solver_probe_pressure_noalloc keeps its probe points and results in
// this fixed device array instead of buffers allocated per call: points in
// [0, 3 kProbeScratchPoints), results after them (1 MiB).  Every device the
// library runs on holds its own copy until the process ends; its address on
// each device is looked up once (hipGetSymbolAddress) and used from then on,
// never the array's name in host code.  Calls that overlap on one device
// share it and race.
constexpr int kProbeScratchPoints = 65536;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) float scratch[4 * kProbeScratchPoints];
std::mutex g_scratch_lock;
float* g_scratch_on[kScratchDevices] = {};

float* probe_scratch_on_current_device() {
  int dev = 0;
  CRYO_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices) throw std::runtime_error("probe scratch: device index out of range");
  std::lock_guard<std::mutex> g(g_scratch_lock);
  if (!g_scratch_on[dev]) {
    void* p = nullptr;
    CRYO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw std::runtime_error("probe scratch: no device address");
    g_scratch_on[dev] = static_cast<float*>(p);
  }
  return g_scratch_on[dev];
}

// ------------------------------------------------------------ C ABI glue
int cryo::record_error(const char* what, int code) {
  g_err = what;
  return code;
}

namespace {

int fail(const char* what, int code) { return cryo::record_error(what, code); }


cryo_solver& S(cryo_solver* s) {
  if (!s) throw std::invalid_argument("null solver");
  return *s;
}

}  // namespace

namespace {
// The byte image of every input of one lattice build: two calls with equal
// images build the same particles.
std::string lattice_inputs(bool on_device, int32_t nn, const cryo_geo_node* nodes, int32_t nm, const cryo_mesh* meshes,
                           const cryo_fluid* fluid, float dx, const cryo_fill* fill) {
  std::string key;
  const auto put = [&key](const void* p, std::size_t bytes) {
    if (bytes) key.append(static_cast<const char*>(p), bytes);
  };
  put(&on_device, sizeof on_device);
  put(&nn, sizeof nn);
  put(&nm, sizeof nm);
  put(&dx, sizeof dx);
  put(fluid, sizeof *fluid);
  put(fill, sizeof *fill);
  if (nodes && nn > 0) put(nodes, sizeof(cryo_geo_node) * static_cast<std::size_t>(nn));
  for (int32_t k = 0; meshes && k < nm; ++k) {
    put(&meshes[k].n_vertices, sizeof(int32_t));
    put(&meshes[k].n_triangles, sizeof(int32_t));
    if (meshes[k].vertices && meshes[k].n_vertices > 0)
      put(meshes[k].vertices, 3 * sizeof(double) * static_cast<std::size_t>(meshes[k].n_vertices));
    if (meshes[k].triangles && meshes[k].n_triangles > 0)
      put(meshes[k].triangles, 3 * sizeof(int32_t) * static_cast<std::size_t>(meshes[k].n_triangles));
  }
  return key;
}

// The particles of the last counting call (x == NULL), kept for the filling
// call with the same inputs that normally follows it, so that the two-call
// protocol builds the lattice once.
struct LatticeMemo {
  std::mutex lock;
  std::string inputs;
  std::shared_ptr<const HostParticles> particles;
};
LatticeMemo& lattice_memo() {
  static LatticeMemo memo;
  return memo;
}

int build_tank_export(bool on_device, int32_t nn, const cryo_geo_node* nodes, int32_t nm, const cryo_mesh* meshes,
                      int32_t ncb, const cryo_geo_callback* callbacks, const cryo_fluid* fluid, float dx,
                      const cryo_fill* fill, int32_t* n, float* x, float* v, float* nrm, float* m, float* rho,
                      float* T, uint8_t* phase, uint8_t* flags, uint32_t* id) {
  CRYO_GUARD({
    if (!fluid || !fill || !n) throw std::invalid_argument("cryo_build_tank_particles: null argument");
    // a caller's distance function may change between the two calls, so such
    // trees are built afresh every time
    const bool memoise = ncb <= 0;
    const std::string inputs = memoise ? lattice_inputs(on_device, nn, nodes, nm, meshes, fluid, dx, fill) : "";
    LatticeMemo& memo = lattice_memo();
    std::shared_ptr<const HostParticles> built;
    if (memoise) {
      std::lock_guard<std::mutex> g(memo.lock);
      if (memo.particles && memo.inputs == inputs) built = memo.particles;
      if (x) {  // a filling call ends the pair
        memo.particles.reset();
        memo.inputs.clear();
      }
    }
    if (!built) {
      built = std::make_shared<const HostParticles>(
          build_tank_particles(Geometry(nn, nodes, nm, meshes, ncb, callbacks), *fluid, dx, *fill, on_device));
      if (!x && memoise) {
        std::lock_guard<std::mutex> g(memo.lock);
        memo.inputs = inputs;
        memo.particles = built;
      }
    }
    const HostParticles& hp = *built;
    const int cnt = static_cast<int>(hp.size());
    if (x) {
      if (*n < cnt) throw std::out_of_range("cryo_build_tank_particles: arrays too small");
      for (int i = 0; i < cnt; ++i) {
        x[3 * i] = hp.x[i].x; x[3 * i + 1] = hp.x[i].y; x[3 * i + 2] = hp.x[i].z;
        if (v) { v[3 * i] = hp.v[i].x; v[3 * i + 1] = hp.v[i].y; v[3 * i + 2] = hp.v[i].z; }
        if (nrm) { nrm[3 * i] = hp.nrm[i].x; nrm[3 * i + 1] = hp.nrm[i].y; nrm[3 * i + 2] = hp.nrm[i].z; }
      }
      if (m) std::copy(hp.m.begin(), hp.m.end(), m);
      if (rho) std::copy(hp.rho.begin(), hp.rho.end(), rho);
      if (T) std::copy(hp.T.begin(), hp.T.end(), T);
      if (phase) std::copy(hp.phase.begin(), hp.phase.end(), phase);
      if (flags) std::copy(hp.flags.begin(), hp.flags.end(), flags);
      if (id) std::copy(hp.id.begin(), hp.id.end(), id);
    }
    *n = cnt;
  })
}
}  // namespace

extern "C" {

const char* cryo_last_error(void) { return g_err.c_str(); }
const char* cryo_version(void) { return "0.2.0-gfx950"; }

int cryo_device_name(char* buf, size_t len) {
  CRYO_GUARD({
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) throw std::runtime_error("no HIP device");
    int dev = 0;
    CRYO_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop{};
    CRYO_HIP(hipGetDeviceProperties(&prop, dev));
    const std::string name = std::string(prop.name) + " (" + prop.gcnArchName + ")";
    if (buf && len) {
      std::strncpy(buf, name.c_str(), len - 1);
      buf[len - 1] = 0;
    }
  })
}
int cryo_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void cryo_config_default(cryo_config* out) {
  if (out) config_default(out);
}
int cryo_fluid_preset(const char* name, cryo_fluid* out) {
  CRYO_GUARD({
    if (!name || !out) throw std::invalid_argument("null argument");
    if (!preset(name, out)) throw std::invalid_argument(std::string("unknown fluid preset '") + name + "'");
  })
}
float cryo_fluid_saturation_temperature(const cryo_fluid* f, float p) { return saturation_temperature(*f, p); }
float cryo_fluid_saturation_pressure(const cryo_fluid* f, float T) { return saturation_pressure(*f, T); }

int cryo_solver_create(const cryo_config* cfg, const cryo_fluid* fluid, cryo_solver** out) {
  try {
    if (!cfg || !fluid || !out) throw std::invalid_argument("null argument");
    *out = new cryo_solver(*cfg, *fluid);
    return CRYO_OK;
  } catch (const std::invalid_argument& e) {
    return fail(e.what(), CRYO_ERR_INVALID_ARGUMENT);
  } catch (const std::exception& e) {
    const std::string w = e.what();
    return fail(e.what(), w.find("CryoSPH:") == 0 ? CRYO_ERR_NO_DEVICE : CRYO_ERR_RUNTIME);
  }
}
void cryo_solver_destroy(cryo_solver* s) {
  try {
    delete s;
  } catch (...) {
  }
}

int cryo_solver_set_particles(cryo_solver* s, int32_t n, const float* x, const float* v, const float* nrm,
                              const float* m, const float* rho, const float* T, const uint8_t* phase,
                              const uint8_t* flags, const uint32_t* id, const float domain_lo[3],
                              const float domain_hi[3]) {
  CRYO_GUARD(S(s).set_particles(n, x, v, nrm, m, rho, T, phase, flags, id, domain_lo, domain_hi))
}
int cryo_solver_step(cryo_solver* s) { CRYO_GUARD(S(s).step()) }
int cryo_solver_run_steps(cryo_solver* s, int64_t k) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    struct Off {
      cryo_solver& z;
      ~Off() {
        z.count_abort();
        if (!z.graphs_on) return;
        (void)hipStreamSynchronize(z.stream);
        z.graph_reset();
        z.graphs_on = false;
      }
    } off{z};
    z.ensure_initialized();
    z.graphs_on = z.use_graphs;
    auto one_by_one = [&]() {
      for (int64_t q = 0; q < k; ++q) {
        z.step_issue_graph();
        z.step_finish();
      }
    };
    if (z.uses_device_counts() && k > 1) {
      z.count_begin();
      for (int64_t q = 0; q < k; ++q) z.step_issue_graph();
      if (!z.count_end()) one_by_one();
    } else {
      one_by_one();
    }
    z.sync_state();
  })
}
int cryo_solver_set_launch_variant(cryo_solver* s, int32_t variant) { CRYO_GUARD(S(s).set_launch_variant(variant)) }
int cryo_solver_set_step_graphs(cryo_solver* s, int32_t on) {
  if (on < 0 || on > 2) return fail("step graphs: on must be 0, 1 or 2", CRYO_ERR_INVALID_ARGUMENT);
  CRYO_GUARD({
    S(s).use_graphs = on != 0;
    S(s).ensemble_graph_on = on == 2;
  })
}
int cryo_solver_set_device_counts(cryo_solver* s, int32_t on, int32_t room) {
  if (room < 0) return fail("device counts: room must be >= 0", CRYO_ERR_INVALID_ARGUMENT);
  CRYO_GUARD({
    S(s).device_counts = on != 0;
    S(s).count_room = room;
  })
}
int cryo_solver_device_count_restarts(cryo_solver* s, int64_t* restarts) {
  if (!restarts) return fail("restarts is required", CRYO_ERR_INVALID_ARGUMENT);
  CRYO_GUARD(*restarts = S(s).count_restarts)
}
int cryo_solver_launch_variant(cryo_solver* s, int32_t* variant) {
  if (!variant) return fail("variant is required", CRYO_ERR_INVALID_ARGUMENT);
  CRYO_GUARD(*variant = S(s).launch_variant())
}
int cryo_solver_autotune(cryo_solver* s, int32_t trial_steps, int32_t* variant, double* ms_per_step) {
  CRYO_GUARD({
    const int v = S(s).autotune(trial_steps, ms_per_step);
    if (variant) *variant = v;
  })
}
int cryo_solver_set_tile_shape(cryo_solver* s, int32_t cells_x, int32_t cells_y, int32_t cells_z) {
  CRYO_GUARD(S(s).set_tile_shape(cells_x, cells_y, cells_z))
}
int cryo_solver_tile_shape(cryo_solver* s, int32_t* cells_xyz) {
  if (!cells_xyz) return fail("cells_xyz is required", CRYO_ERR_INVALID_ARGUMENT);
  CRYO_GUARD({
    for (int k = 0; k < 3; ++k) cells_xyz[k] = S(s).tile_shape[k];
  })
}
int cryo_solver_autotune_tiles(cryo_solver* s, int32_t trial_steps, const int32_t* shapes, int32_t count,
                               int32_t* pick, double* ms_per_step) {
  CRYO_GUARD({
    const int k = S(s).autotune_tiles(trial_steps, shapes, count, ms_per_step);
    if (pick) *pick = k;
  })
}
int cryo_ensemble_run_steps(cryo_solver* const* solvers, int32_t count, int64_t steps) {
  CRYO_GUARD({
    if (count < 0 || steps < 0) throw std::invalid_argument("cryo_ensemble_run_steps: negative count or steps");
    if (count > 0 && !solvers) throw std::invalid_argument("cryo_ensemble_run_steps: null solver array");
    for (int32_t t = 0; t < count; ++t) {
      if (!solvers[t]) throw std::invalid_argument("cryo_ensemble_run_steps: null solver");
      for (int32_t u = 0; u < t; ++u)
        if (solvers[u] == solvers[t]) throw std::invalid_argument("cryo_ensemble_run_steps: a solver appears twice");
    }
    // Every member's step is queued on its own stream before any member
    // waits, so the device runs the tanks' kernels side by side and a
    // member's host round trip overlaps the others' device work.  A member
    // whose next step needs its particle count back (phase change) queues
    // the state copy behind its step; the host polls the members' streams and
    // queues a member's next step as soon as its own copy has landed,
    // whichever member that is, so no member waits behind a slower one.
    // Members without phase change queue all their steps at once.
    // Each member's step is a captured graph, replayed step after step.
    struct Graphs {
      cryo_solver* const* v;
      int32_t c;
      ~Graphs() {
        for (int32_t t = 0; t < c; ++t) {
          (void)hipStreamSynchronize(v[t]->stream);
          v[t]->graph_reset();
          v[t]->graphs_on = false;
        }
      }
    } graphs{solvers, count};
    for (int32_t t = 0; t < count; ++t) {
      cryo_solver& z = S(solvers[t]);
      z.ensure_initialized();
      z.graphs_on = z.use_graphs;
    }
    if (steps > 0) {
      std::vector<int64_t> issued(static_cast<std::size_t>(count), 0);
      std::vector<char> dc(static_cast<std::size_t>(count), 0);  // queued with device counts
      struct Abort {
        cryo_solver* const* v;
        int32_t c;
        ~Abort() {
          for (int32_t t = 0; t < c; ++t) v[t]->count_abort();
        }
      } abort_runs{solvers, count};
      // One graph for the whole ensemble: when every member asks for it (step
      // graphs mode 2) and none needs its particle count back between steps
      // (no phase change, or device counts), two steps of every member (one per persistent buffer
      // set, so the graph ends where it began) are captured into one graph,
      // each member on its own stream, forked from and joined to member 0's
      // stream, and that graph is launched once per two ensemble steps; an odd
      // last step is queued eagerly behind it.
      bool one_graph = steps >= 2;
      for (int32_t t = 0; t < count && one_graph; ++t) {
        const cryo_solver& z = S(solvers[t]);
        one_graph = z.ensemble_graph_on && z.prof_mask == 0 && z.n > 0 && (!z.needs_counts() || z.uses_device_counts());
      }
      if (one_graph) {
        for (int32_t t = 0; t < count; ++t)
          if (S(solvers[t]).needs_counts()) S(solvers[t]).count_begin(), dc[t] = 1;
        S(solvers[0]).ensemble_graph(solvers, count, steps);
      }
      int32_t waiting = 0;  // members with a step whose counts are not back yet
      for (int32_t t = 0; t < count && !one_graph; ++t) {
        cryo_solver& z = S(solvers[t]);
        if (!z.needs_counts() || (z.uses_device_counts() && steps > 1)) {
          if (z.needs_counts()) z.count_begin(), dc[t] = 1;
          for (int64_t q = 0; q < steps; ++q) z.step_issue_graph();
          issued[t] = steps;
          continue;
        }
        z.step_issue_graph();
        z.queue_state_copy();
        issued[t] = 1;
        ++waiting;
      }
      std::vector<char> pending(static_cast<std::size_t>(count), 0);
      for (int32_t t = 0; t < count; ++t) pending[t] = S(solvers[t]).needs_counts() && !dc[t] ? 1 : 0;
      while (waiting > 0) {
        bool progressed = false;
        for (int32_t t = 0; t < count; ++t) {
          if (!pending[t]) continue;
          cryo_solver& z = S(solvers[t]);
          if (!z.stream_idle()) continue;
          z.take_counts();
          progressed = true;
          if (issued[t] < steps) {
            z.step_issue_graph();
            z.queue_state_copy();
            ++issued[t];
          } else {
            pending[t] = 0;
            --waiting;
          }
        }
        if (!progressed) std::this_thread::yield();
      }
      for (int32_t t = 0; t < count; ++t) {
        if (!dc[t]) continue;
        cryo_solver& z = S(solvers[t]);
        if (z.count_end()) continue;
        for (int64_t q = 0; q < steps; ++q) {  // its splits passed the bound: one step at a time
          z.step_issue_graph();
          z.step_finish();
        }
      }
    }
    for (int32_t t = 0; t < count; ++t) S(solvers[t]).sync_state();
  })
}
int cryo_solver_adjoint_slosh(cryo_solver* s, int64_t steps, const float* frame_acc, int32_t checkpoint_every,
                              double* loss, double* grad_frame_acc, double* wall_load, double* grad_contact_angle,
                              double* grad_lee) {
  if (steps > 0 && !frame_acc) return fail("frame_acc is required", CRYO_ERR_INVALID_ARGUMENT);
  CRYO_GUARD(S(s).adjoint_slosh(steps, frame_acc, checkpoint_every, loss, grad_frame_acc, wall_load,
                                grad_contact_angle, grad_lee))
}
int cryo_solver_adjoint_slosh_frame(cryo_solver* s, int64_t steps, const float* frame_acc, const float* frame_omega,
                                    const float* frame_omega_dot, int32_t checkpoint_every, double* loss,
                                    double* grad_frame_acc, double* grad_frame_omega, double* grad_frame_omega_dot,
                                    double* wall_load, double* grad_contact_angle, double* grad_lee) {
  if (steps > 0 && !frame_acc) return fail("frame_acc is required", CRYO_ERR_INVALID_ARGUMENT);
  CRYO_GUARD(S(s).adjoint_slosh(steps, frame_acc, checkpoint_every, loss, grad_frame_acc, wall_load,
                                grad_contact_angle, grad_lee, frame_omega, frame_omega_dot, grad_frame_omega,
                                grad_frame_omega_dot))
}
int cryo_solver_run_until(cryo_solver* s, double t_end) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    z.ensure_initialized();
    z.sync_state();
    while (z.hs->time < t_end) {
      z.step();
      z.sync_state();
    }
  })
}
int cryo_solver_relax(cryo_solver* s, double duration, float damping) { CRYO_GUARD(S(s).relax(duration, damping)) }
int cryo_solver_next_dt(cryo_solver* s, double* dt) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    z.ensure_initialized();
    z.sync_state();
    *dt = compute_dt(z.P, z.L, z.hs->v_max, z.hs->a_max, z.hs->Tv_max, z.hs->nu_t_max, z.hs->spacing_ratio);
  })
}
int cryo_solver_damp_velocities(cryo_solver* s, float factor) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    z.scale_velocities(factor);
    z.sync_state();
    z.hs->v_max *= factor;
    z.push_state();
  })
}
int cryo_solver_time(cryo_solver* s, double* t) { CRYO_GUARD(S(s).sync_state(); *t = S(s).hs->time) }
int cryo_solver_step_count(cryo_solver* s, int64_t* k) { CRYO_GUARD(S(s).sync_state(); *k = S(s).hs->step) }
int cryo_solver_last_dt(cryo_solver* s, double* dt) { CRYO_GUARD(S(s).sync_state(); *dt = S(s).hs->last_dt) }
int cryo_solver_num_particles(cryo_solver* s, int32_t* n) { CRYO_GUARD(*n = S(s).n) }
int cryo_solver_capacity(cryo_solver* s, int32_t* c) { CRYO_GUARD(*c = S(s).capacity) }
int cryo_solver_config(cryo_solver* s, cryo_config* o) { CRYO_GUARD(*o = S(s).cfg) }
int cryo_solver_fluid(cryo_solver* s, cryo_fluid* o) { CRYO_GUARD(*o = S(s).fluid) }
int cryo_solver_params(cryo_solver* s, cryo_params* o) { CRYO_GUARD(S(s).params(o)) }
int cryo_solver_diagnostics(cryo_solver* s, cryo_diagnostics* o) { CRYO_GUARD(S(s).diagnostics(o)) }
int cryo_solver_download(cryo_solver* s, int32_t field, void* dst, int32_t count) {
  CRYO_GUARD(S(s).download(field, dst, count))
}

int cryo_solver_set_gravity(cryo_solver* s, const float g[3]) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    z.cfg.gravity = cryo_vec3{g[0], g[1], g[2]};
    z.sync_state();
    z.hs->cfg_gravity = v3p(g);
    z.hs->gravity = v3p(g) - z.hs->frame_acc;
    z.push_state();
  })
}
int cryo_solver_set_frame_motion(cryo_solver* s, const float acc[3], const float omega[3], const float omega_dot[3],
                                 const float origin[3]) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    z.clear_maneuver();
    z.apply_frame(v3p(acc), v3p(omega), v3p(omega_dot), v3p(origin));
  })
}
int cryo_solver_set_maneuver(cryo_solver* s, int32_t n, const double* time, const float* acc, const float* omega,
                             const float origin[3]) {
  CRYO_GUARD({
    if (n < 0 || (n > 0 && (!time || !acc))) throw std::invalid_argument("ManeuverProfile: bad table");
    HostManeuver m;
    m.time.assign(time, time + n);
    for (int k = 0; k < n; ++k) m.acc.push_back(v3p(acc + 3 * k));
    if (omega)
      for (int k = 0; k < n; ++k) m.omega.push_back(v3p(omega + 3 * k));
    m.origin = origin ? v3p(origin) : V3{0, 0, 0};
    S(s).set_maneuver(m);
  })
}
int cryo_solver_clear_frame_motion(cryo_solver* s) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    z.clear_maneuver();
    z.apply_frame(V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0});
  })
}
int cryo_solver_set_heater_power(cryo_solver* s, float watts) {
  CRYO_GUARD({
    S(s).cfg.heater_power = watts;
    S(s).P.heat_power = watts;
  })
}
int cryo_solver_set_contact_angle(cryo_solver* s, float degrees) {
  CRYO_GUARD({
    if (degrees < 0 || degrees > 180) throw std::invalid_argument("contact angle must be within [0, 180] degrees");
    cryo_solver& z = S(s);
    z.cfg.contact_angle_deg = degrees;
    const float theta = degrees * kPi / 180.0f;
    z.P.cos_theta = std::cos(theta);
    z.P.sin_theta = std::sin(theta);
  })
}

int cryo_maneuver_evaluate(int32_t n, const double* time, const float* acc, const float* omega, double t,
                           float a_out[3], float w_out[3], float wdot_out[3]) {
  CRYO_GUARD({
    if (n < 0 || (n > 0 && (!time || !acc))) throw std::invalid_argument("ManeuverProfile: bad table");
    HostManeuver m;
    m.time.assign(time, time + n);
    for (int k = 0; k < n; ++k) m.acc.push_back(v3p(acc + 3 * k));
    if (omega)
      for (int k = 0; k < n; ++k) m.omega.push_back(v3p(omega + 3 * k));
    m.validate();
    V3 a3, w3, wd3;
    m.evaluate(t, a3, w3, wd3);
    a_out[0] = a3.x; a_out[1] = a3.y; a_out[2] = a3.z;
    w_out[0] = w3.x; w_out[1] = w3.y; w_out[2] = w3.z;
    wdot_out[0] = wd3.x; wdot_out[1] = wd3.y; wdot_out[2] = wd3.z;
  })
}

int cryo_solver_probe_pressure(cryo_solver* s, int32_t np, const float* points, float* out) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    z.ensure_initialized();
    if (np <= 0) return CRYO_OK;
    DevBuf<float> dp, dout;
    dp.resize(3 * static_cast<std::size_t>(np));
    dout.resize(np);
    CRYO_HIP(hipMemcpyAsync(dp.get(), points, sizeof(float) * 3 * np, hipMemcpyHostToDevice, z.stream));
    launch_probe(z.view(), z.grid(), z.P, dp.get(), dout.get(), np, z.stream);
    z.check_launch();
    CRYO_HIP(hipMemcpyAsync(out, dout.get(), sizeof(float) * np, hipMemcpyDeviceToHost, z.stream));
    CRYO_HIP(hipStreamSynchronize(z.stream));
  })
}

int cryo_solver_probe_pressure_noalloc(cryo_solver* s, int32_t np, const float* points, float* out) {
  CRYO_GUARD({
    if (np < 0 || np > kProbeScratchPoints)
      throw std::out_of_range("cryo_solver_probe_pressure_noalloc: np must be within [0, " +
                              std::to_string(kProbeScratchPoints) + "]");
    cryo_solver& z = S(s);
    // the start-up evaluation would allocate (cell tables, lists)
    if (!z.initialized)
      throw std::runtime_error(
          "cryo_solver_probe_pressure_noalloc: the solver is not initialised yet (call next_dt or step first)");
    if (np == 0) return CRYO_OK;
    if (!points || !out) throw std::invalid_argument("cryo_solver_probe_pressure_noalloc: null argument");
    float* const pts = probe_scratch_on_current_device();
    float* const res = pts + 3 * kProbeScratchPoints;
    CRYO_HIP(hipMemcpyAsync(pts, points, sizeof(float) * 3 * np, hipMemcpyHostToDevice, z.stream));
    launch_probe(z.view(), z.grid(), z.P, pts, res, np, z.stream);
    z.check_launch();
    CRYO_HIP(hipMemcpyAsync(out, res, sizeof(float) * np, hipMemcpyDeviceToHost, z.stream));
    CRYO_HIP(hipStreamSynchronize(z.stream));
  })
}

int cryo_solver_neighbor_stats(cryo_solver* s, float radius, int32_t* counts, uint64_t* id_sums) {
  CRYO_GUARD(S(s).neighbor_counts(radius, counts, id_sums))
}
int cryo_solver_neighbor_csr(cryo_solver* s, float radius, int64_t* offsets, int32_t* indices) {
  CRYO_GUARD(S(s).neighbor_csr(radius, offsets, indices))
}
int cryo_solver_grid_info(cryo_solver* s, int32_t dims[3], float origin[3], float* cell_size, int32_t* num_cells) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    dims[0] = z.g.nx;
    dims[1] = z.g.ny;
    dims[2] = z.g.nz;
    origin[0] = z.g.ox;
    origin[1] = z.g.oy;
    origin[2] = z.g.oz;
    *cell_size = z.cell;
    *num_cells = z.g.num_cells;
  })
}
int cryo_solver_grid_tables(cryo_solver* s, int32_t* cell_start, int32_t* cell_end, uint32_t* perm,
                            int32_t* n_sorted) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    const int nc = z.g.num_cells;
    std::vector<int> off(static_cast<std::size_t>(nc) + 1);
    CRYO_HIP(hipMemcpyAsync(off.data(), z.cell_off.get(), sizeof(int) * (nc + 1), hipMemcpyDeviceToHost, z.stream));
    if (perm && z.last_perm && z.n_sorted > 0)
      CRYO_HIP(hipMemcpyAsync(perm, z.last_perm, sizeof(uint32_t) * z.n_sorted, hipMemcpyDeviceToHost, z.stream));
    CRYO_HIP(hipStreamSynchronize(z.stream));
    for (int c = 0; c < nc; ++c) {
      const bool nonempty = off[c + 1] > off[c];
      if (cell_start) cell_start[c] = nonempty ? off[c] : -1;
      if (cell_end) cell_end[c] = nonempty ? off[c + 1] : -1;
    }
    if (n_sorted) *n_sorted = z.n_sorted;
  })
}

int cryo_solver_initialize(cryo_solver* s) { CRYO_GUARD(S(s).ensure_initialized()) }
int cryo_solver_rebuild_grid(cryo_solver* s, int32_t* removed) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    const int before = z.n;
    z.grid_build(false);
    z.sync_state();
    z.n = z.hs->n_alive;
    if (removed) *removed = before - z.n;
  })
}
int cryo_solver_evaluate_rates(cryo_solver* s) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    z.evaluate_rates();
    CRYO_HIP(hipStreamSynchronize(z.stream));
  })
}

int cryo_hash_sort(int32_t n, const float* x, const uint8_t* phase, const float lo[3], const float hi[3], float cell,
                   uint32_t* perm, int32_t* n_alive, float radius, int32_t* counts, uint64_t* id_sums) {
  return cryo_hash_sort_ids(n, x, phase, nullptr, lo, hi, cell, perm, n_alive, radius, counts, id_sums);
}

int cryo_hash_sort_ids(int32_t n, const float* x, const uint8_t* phase, const uint32_t* id, const float lo[3],
                       const float hi[3], float cell, uint32_t* perm, int32_t* n_alive, float radius, int32_t* counts,
                       uint64_t* id_sums) {
  CRYO_GUARD({
    if (n < 0) throw std::invalid_argument("cryo_hash_sort: n < 0");
    // the particles' ids ride through the sort; perm maps them back to input
    // positions (ids must then be unique)
    std::vector<uint32_t> where;
    if (id) {
      uint32_t top = 0;
      for (int32_t k = 0; k < n; ++k) top = std::max(top, id[k]);
      where.assign(static_cast<std::size_t>(top) + 1, UINT32_MAX);
      for (int32_t k = 0; k < n; ++k) {
        if (where[id[k]] != UINT32_MAX) throw std::invalid_argument("cryo_hash_sort: particle ids are not unique");
        where[id[k]] = static_cast<uint32_t>(k);
      }
    }
    check_device();
    cryo_config c;
    config_default(&c);
    cryo_fluid f;
    preset("custom", &f);
    cryo_solver z(c, f);
    std::vector<float> ones(static_cast<std::size_t>(std::max(n, 1)), 1.0f);
    if (n == 0) {
      *n_alive = 0;
      return CRYO_OK;
    }
    z.set_particles(n, x, nullptr, nullptr, ones.data(), ones.data(), ones.data(), phase, nullptr, id, lo, hi);
    z.configure_grid(v3p(lo), v3p(hi), cell);
    const int before = z.n;
    z.grid_build(false);
    z.sync_state();
    z.n = z.hs->n_alive;
    *n_alive = z.n;
    (void)before;
    if (z.n > 0)
      CRYO_HIP(hipMemcpyAsync(perm, z.view().s.id, sizeof(uint32_t) * z.n, hipMemcpyDeviceToHost, z.stream));
    CRYO_HIP(hipStreamSynchronize(z.stream));
    if (id)
      for (int32_t k = 0; k < z.n; ++k) perm[k] = where[perm[k]];
    if (counts) z.neighbor_counts(radius, counts, id_sums);
  })
}

int cryo_solver_profile_enable(cryo_solver* s, uint32_t mask) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    z.prof_collect();
    z.prof_mask = mask;
  })
}
int cryo_solver_profile_reset(cryo_solver* s) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    z.prof_collect();
    for (int k = 0; k < cryo_solver::kSections; ++k) {
      z.prof_ms[k] = 0;
      z.prof_count[k] = 0;
    }
  })
}
int cryo_solver_profile_section(cryo_solver* s, int32_t index, char* name, size_t len, double* total_ms,
                                int64_t* count) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    if (index < 0 || index >= cryo_solver::kSections) throw std::out_of_range("profile section index");
    z.prof_collect();
    if (name && len) {
      std::strncpy(name, cryo_solver::kSectionNames[index], len - 1);
      name[len - 1] = 0;
    }
    if (total_ms) *total_ms = z.prof_ms[index];
    if (count) *count = z.prof_count[index];
  })
}

int cryo_solver_configure_telemetry(cryo_solver* s, int32_t mode) { CRYO_GUARD(S(s).configure_telemetry(mode)) }
int cryo_solver_telemetry_mode(cryo_solver* s, int32_t* mode) { CRYO_GUARD(*mode = S(s).tele_mode) }
int cryo_solver_request_telemetry(cryo_solver* s) {
  try {
    return S(s).request_telemetry() ? 1 : 0;
  } catch (const std::exception& e) {
    return fail(e.what(), CRYO_ERR_RUNTIME);
  }
}
int cryo_solver_telemetry_ready(cryo_solver* s) {
  try {
    return S(s).telemetry_ready() ? 1 : 0;
  } catch (const std::exception& e) {
    return fail(e.what(), CRYO_ERR_RUNTIME);
  }
}
int cryo_solver_telemetry_pending_size(cryo_solver* s, int32_t* n) {
  CRYO_GUARD(*n = S(s).tele_pending ? S(s).tele_n : 0)
}
int cryo_solver_fetch_telemetry(cryo_solver* s, double* time, int64_t* step, int32_t* n, float* xyz, float* T,
                                float* p, float* speed, uint8_t* phase) {
  try {
    cryo_solver& z = S(s);
    if (!z.tele_pending) return 0;
    CRYO_HIP(hipEventSynchronize(z.tele_landed));
    z.tele_pending = false;
    *time = z.tele_time;
    *step = z.tele_step;
    *n = z.tele_n;
    for (int i = 0; i < z.tele_n; ++i) {
      const float* r = z.tele_host + 7 * static_cast<std::size_t>(i);
      if (xyz) {
        xyz[3 * i] = r[0];
        xyz[3 * i + 1] = r[1];
        xyz[3 * i + 2] = r[2];
      }
      if (T) T[i] = r[3];
      if (p) p[i] = r[4];
      if (speed) speed[i] = r[5];
      uint32_t ph;
      std::memcpy(&ph, &r[6], sizeof(ph));
      if (phase) phase[i] = static_cast<uint8_t>(ph);
    }
    return 1;
  } catch (const std::exception& e) {
    return fail(e.what(), CRYO_ERR_RUNTIME);
  }
}

void cryo_fill_default(cryo_fill* f) {
  if (!f) return;
  std::memset(f, 0, sizeof(*f));
  f->mode = CRYO_FILL_SETTLED;
  f->fill_fraction = 0.5;
  f->settle_direction = cryo_vec3{0, 0, -1};
  f->bubble_center = cryo_vec3{0, 0, 0};
  f->include_vapor = 1;
  f->wall_layers = 3;
  f->T_liquid = f->T_vapor = f->T_wall = std::numeric_limits<float>::quiet_NaN();
  f->heater_zmin = -std::numeric_limits<double>::infinity();
  f->heater_zmax = std::numeric_limits<double>::infinity();
}

int cryo_geo_sdf(int32_t nn, const cryo_geo_node* nodes, int32_t nm, const cryo_mesh* meshes, double x, double y,
                 double z, double* out) {
  CRYO_GUARD(*out = Geometry(nn, nodes, nm, meshes).sdf(x, y, z))
}
int cryo_geo_bounds(int32_t nn, const cryo_geo_node* nodes, int32_t nm, const cryo_mesh* meshes, float lo[3],
                    float hi[3]) {
  CRYO_GUARD({
    V3 l, h;
    Geometry(nn, nodes, nm, meshes).bounds(l, h);
    lo[0] = l.x; lo[1] = l.y; lo[2] = l.z;
    hi[0] = h.x; hi[1] = h.y; hi[2] = h.z;
  })
}
int cryo_geo_volume(int32_t nn, const cryo_geo_node* nodes, int32_t nm, const cryo_mesh* meshes, double* out) {
  CRYO_GUARD(*out = Geometry(nn, nodes, nm, meshes).volume())
}
int cryo_mesh_winding_number(const cryo_mesh* mesh, double x, double y, double z, double* out) {
  CRYO_GUARD({
    cryo_geo_node nd{CRYO_GEO_MESH, -1, -1, 0, {0, 0, 0}};
    *out = Geometry(1, &nd, 1, mesh).mesh_winding(0, x, y, z);
  })
}


int cryo_build_tank_particles(int32_t nn, const cryo_geo_node* nodes, int32_t nm, const cryo_mesh* meshes,
                              const cryo_fluid* fluid, float dx, const cryo_fill* fill, int32_t* n, float* x,
                              float* v, float* nrm, float* m, float* rho, float* T, uint8_t* phase, uint8_t* flags,
                              uint32_t* id) {
  return build_tank_export(false, nn, nodes, nm, meshes, 0, nullptr, fluid, dx, fill, n, x, v, nrm, m, rho, T, phase,
                           flags, id);
}
int cryo_build_tank_particles_cb(int32_t nn, const cryo_geo_node* nodes, int32_t nm, const cryo_mesh* meshes,
                                 int32_t ncb, const cryo_geo_callback* callbacks, const cryo_fluid* fluid, float dx,
                                 const cryo_fill* fill, int32_t* n, float* x, float* v, float* nrm, float* m,
                                 float* rho, float* T, uint8_t* phase, uint8_t* flags, uint32_t* id) {
  return build_tank_export(false, nn, nodes, nm, meshes, ncb, callbacks, fluid, dx, fill, n, x, v, nrm, m, rho, T,
                           phase, flags, id);
}
int cryo_build_tank_particles_device(int32_t nn, const cryo_geo_node* nodes, int32_t nm, const cryo_mesh* meshes,
                                     const cryo_fluid* fluid, float dx, const cryo_fill* fill, int32_t* n, float* x,
                                     float* v, float* nrm, float* m, float* rho, float* T, uint8_t* phase,
                                     uint8_t* flags, uint32_t* id) {
  return build_tank_export(true, nn, nodes, nm, meshes, 0, nullptr, fluid, dx, fill, n, x, v, nrm, m, rho, T, phase,
                           flags, id);
}

int cryo_solver_build_tank(cryo_solver* s, int32_t nn, const cryo_geo_node* nodes, int32_t nm,
                           const cryo_mesh* meshes, const cryo_fill* fill) {
  return cryo_solver_build_tank_cb(s, nn, nodes, nm, meshes, 0, nullptr, fill);
}

int cryo_solver_build_tank_cb(cryo_solver* s, int32_t nn, const cryo_geo_node* nodes, int32_t nm,
                              const cryo_mesh* meshes, int32_t ncb, const cryo_geo_callback* callbacks,
                              const cryo_fill* fill) {
  CRYO_GUARD({
    cryo_solver& z = S(s);
    if (!fill) throw std::invalid_argument("cryo_solver_build_tank: null fill");
    const Geometry geo(nn, nodes, nm, meshes, ncb, callbacks);
    const HostParticles hp = build_tank_particles(geo, z.fluid, z.cfg.dx, *fill);
    V3 lo, hi;
    geo.bounds(lo, hi);
    const float pad = (fill->wall_layers + 1) * z.cfg.dx;
    lo = lo - V3{pad, pad, pad};
    hi = hi + V3{pad, pad, pad};
    const int cnt = static_cast<int>(hp.size());
    std::vector<float> xs(3 * static_cast<std::size_t>(cnt)), vs(xs.size()), ns(xs.size());
    for (int i = 0; i < cnt; ++i) {
      xs[3 * i] = hp.x[i].x; xs[3 * i + 1] = hp.x[i].y; xs[3 * i + 2] = hp.x[i].z;
      vs[3 * i] = hp.v[i].x; vs[3 * i + 1] = hp.v[i].y; vs[3 * i + 2] = hp.v[i].z;
      ns[3 * i] = hp.nrm[i].x; ns[3 * i + 1] = hp.nrm[i].y; ns[3 * i + 2] = hp.nrm[i].z;
    }
    const float l3[3] = {lo.x, lo.y, lo.z}, h3[3] = {hi.x, hi.y, hi.z};
    z.set_particles(cnt, xs.data(), vs.data(), ns.data(), hp.m.data(), hp.rho.data(), hp.T.data(), hp.phase.data(),
                    hp.flags.data(), hp.id.data(), l3, h3);
  })
}

int cryo_stream_create(void** stream) {
  CRYO_GUARD({
    if (!stream) throw std::invalid_argument("cryo_stream_create: null argument");
    check_device();
    hipStream_t h = nullptr;
    CRYO_HIP(hipStreamCreate(&h));  // blocking, default priority
    *stream = h;
  })
}
int cryo_stream_destroy(void* stream) {
  CRYO_GUARD(if (stream) CRYO_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream))))
}
int cryo_stream_synchronize(void* stream) {
  CRYO_GUARD(CRYO_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream))))
}
int cryo_solver_stream(cryo_solver* s, void** stream) {
  CRYO_GUARD({
    if (!stream) throw std::invalid_argument("cryo_solver_stream: null argument");
    *stream = S(s).stream;
  })
}

}  // extern "C"
