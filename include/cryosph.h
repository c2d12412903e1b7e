/*
 * cryosph.h - C ABI of the MI355X (gfx950) CryoSPH per-step particle pipeline.
 *
 * Every entry point returns 0 on success or a negative cryo_status; the
 * message of the last failure on the calling thread is cryo_last_error().
 * Nothing here throws, allocates caller-visible memory or exposes a device
 * type: buffers are plain host pointers and sizes.  The opaque cryo_solver
 * owns its HIP stream and every device array.
 *
 * Each declaration names the reference interface it replaces
 * (reference = Semedo2002/Microgravity-Cryogenic-Multi-Phase-SPH-Solver).
 */
#ifndef CRYOSPH_H
#define CRYOSPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------- status ---- */
typedef enum cryo_status {
  CRYO_OK = 0,
  CRYO_ERR_INVALID_ARGUMENT = -1, /* std::invalid_argument in the reference (Python ValueError) */
  CRYO_ERR_RUNTIME = -2,          /* std::runtime_error (device failures, bad state)            */
  CRYO_ERR_OUT_OF_RANGE = -3,     /* std::out_of_range (buffer ranges, capacity)                */
  CRYO_ERR_NO_DEVICE = -4         /* no gfx950 device / HIP runtime failure at start-up          */
} cryo_status;

/* Message of the last failing call on this thread ("" if none). */
const char* cryo_last_error(void);
/* Library version string. */
const char* cryo_version(void);
/* Name of the device a new solver binds to, and the number of visible devices
 * (mem::device_name / mem::device_count, src/core/cuda_utils.cpp:341-352). */
int cryo_device_name(char* buf, size_t len);
int cryo_device_count(void);

/* ------------------------------------------------------------ enums ---- */
/* Values match the reference enums (sim_params.cuh:12-18, SPH_kernels.cuh:6,
 * particle_system.hpp:13-34, geometry.hpp:156-160, telemetry.hpp:12). */
enum { CRYO_LIQUID = 0, CRYO_VAPOR = 1, CRYO_WALL = 2, CRYO_DEAD = 3 };
enum { CRYO_FLAG_HEATED = 1, CRYO_FLAG_SOURCE = 2, CRYO_FLAG_TARGET_VAPOR = 4, CRYO_FLAG_DYING = 8 };
enum { CRYO_KERNEL_CUBIC = 0, CRYO_KERNEL_WENDLAND_C2 = 1 };
enum { CRYO_DENSITY_SUMMATION = 0, CRYO_DENSITY_CONTINUITY = 1 };
enum { CRYO_ST_NONE = 0, CRYO_ST_CSF = 1, CRYO_ST_AKINCI = 2 };
enum { CRYO_WALL_ADIABATIC = 0, CRYO_WALL_ISOTHERMAL = 1 };
enum { CRYO_VENTED = 0, CRYO_CLOSED = 1 };
enum { CRYO_WET_NORMAL_ROTATION = 0, CRYO_WET_CONTACT_LINE_FORCE = 1, CRYO_WET_BOTH = 2 };
enum { CRYO_FILL_SETTLED = 0, CRYO_FILL_CENTRAL_BUBBLE = 1, CRYO_FILL_FULL = 2 };
enum { CRYO_TELEMETRY_ASYNC_COPY = 0, CRYO_TELEMETRY_ZERO_COPY = 1 };

/* -------------------------------------------------------- aggregates ---- */
typedef struct cryo_vec3 { float x, y, z; } cryo_vec3;

/* SolverConfig, sph_solver.hpp:20-91: same fields, names, defaults. */
typedef struct cryo_config {
  float dx, h_factor;
  int32_t kernel, density_mode, normalize_initial_masses;
  cryo_vec3 gravity;
  float velocity_scale, sound_speed, tait_gamma, background_pressure, vapor_stiffness;
  int32_t clamp_negative_pressure;
  int32_t transport_velocity;
  float transport_background_factor;
  int32_t transport_stress, transport_interface_only;
  int32_t particle_shifting;
  float shifting_strength;
  float artificial_viscosity, artificial_viscosity_vapor, delta_sph;
  int32_t no_slip, sps_turbulence;
  float smagorinsky_cs;
  int32_t surface_tension;
  float contact_angle_deg, contact_angle_wall_weight;
  int32_t wall_colour, wetting_model, csf_smooth_normals, csf_balanced;
  float akinci_gamma, akinci_beta;
  int32_t thermal, phase_change, wall_thermal_bc;
  float heater_power, conductivity_scale, lee_evaporation, lee_condensation;
  int32_t thermo_mode;
  float vent_relaxation, merge_ratio, split_ratio, max_transfer_fraction, phase_change_density_limit,
      vapor_mass_rebalance;
  float cfl, viscous_number, force_number, capillary_number, thermal_number, dt_max, dt_min, fixed_dt;
  /* Extension (no SolverConfig counterpart): 1 = precision form, for small
   * loads out of large, nearly cancelling forces (microgravity slosh).
   * Positions are carried as double-float pairs (a float and its low-order
   * part), pair distances taken from the pairs; neighbour sums use
   * compensated (Kahan-Babuska) FP32 accumulation; the summation density and
   * its EOS are evaluated in double.  Neighbour sets still come from the float
   * positions.  Results differ from the reference in the last bits and are
   * compared with the FP64 build instead.  The adjoint runs on this forward
   * (its reverse sweep stays FP32).  Default 0. */
  int32_t compensated_sums;
} cryo_config;

/* FluidProperties, fluid_properties.hpp:11-38. */
typedef struct cryo_fluid {
  char name[32];
  float rho_liquid, rho_vapor, mu_liquid, mu_vapor, k_liquid, k_vapor, cp_liquid, cp_vapor, sigma, h_fg, T_sat,
      p_sat, R_gas;
} cryo_fluid;

/* Diagnostics, sph_solver.hpp:108-132 (all sums in double precision). */
typedef struct cryo_diagnostics {
  double time;
  int64_t step;
  double dt;
  int32_t n_total, n_liquid, n_vapor, n_wall;
  double mass_liquid, mass_vapor, mass_total, kinetic_energy, enthalpy, heat_input, evaporated_mass,
      condensed_mass, vented_mass, vented_enthalpy, T_sat, p_ullage, max_speed, T_liquid_mean, T_liquid_max,
      T_vapor_mean;
  cryo_vec3 liquid_com, liquid_momentum;
  int32_t splits, merges;
  cryo_vec3 slosh_force, slosh_moment, frame_acceleration, frame_angular_velocity;
} cryo_diagnostics;

/* Derived constants of SimParams (sim_params.cuh:20-95) that init_params
 * (sph_solver.cu:292-405) computes from the config; read-only. */
typedef struct cryo_params {
  float h, inv_h, support, kernel_norm, support2, dx, vol0;
  float rho0[2], mu[2], k[2], cp[2], mref[2];
  float c0, tait_B, tait_gamma, p_bg, gas_R, vapor_K;
  cryo_vec3 gravity, omega, omega_dot, frame_origin;
  int32_t frame_rot;
  float sigma, cos_theta, sin_theta, normal_eps, akinci_gamma, akinci_beta, tvf_sigma_min, shift_strength;
  float T_ref, T_sat, h_fg, heat_inv_norm, dt;
} cryo_params;

/* Per-particle fields of ParticleSystem (particle_system.hpp:119-131).
 * Vec3 fields cross as 3 floats per particle, scalars as float, phase and
 * flags as uint8, id as uint32. */
typedef enum cryo_field {
  CRYO_F_X = 0, CRYO_F_V, CRYO_F_VT, CRYO_F_NRM, CRYO_F_M, CRYO_F_M0, CRYO_F_RHO, CRYO_F_T, CRYO_F_PHASE,
  CRYO_F_FLAGS, CRYO_F_ID,
  CRYO_F_A, CRYO_F_AB, CRYO_F_NS, CRYO_F_VW, CRYO_F_FW, CRYO_F_MW, CRYO_F_P, CRYO_F_DRHO, CRYO_F_DTDT,
  CRYO_F_SIGMA, CRYO_F_KAPPA, CRYO_F_NU_T, CRYO_F_ALPHA, CRYO_F_WALL_W, CRYO_F_HEAT_W, CRYO_F_OUT_M,
  CRYO_F_OUT_H, CRYO_F_W_NORM, CRYO_F_P_WV, CRYO_F_RMIN,
  CRYO_F_COUNT
} cryo_field;

/* ------------------------------------------------------ config/fluid ---- */
/* SolverConfig defaults (sph_solver.hpp:20-91). */
void cryo_config_default(cryo_config* out);
/* FluidProperties{} and the presets "LH2" "LN2" "LOX" "LCH4" "H2O" / "custom"
 * (fluid_properties.cpp:16-79). */
int cryo_fluid_preset(const char* name, cryo_fluid* out);
/* FluidProperties::saturation_temperature / saturation_pressure
 * (fluid_properties.cpp:7-14). */
float cryo_fluid_saturation_temperature(const cryo_fluid* f, float p);
float cryo_fluid_saturation_pressure(const cryo_fluid* f, float T);

/* ------------------------------------------------------------ solver ---- */
typedef struct cryo_solver cryo_solver;

/* Solver::Solver (sph_solver.cu:273-279): validates dx and the contact angle. */
int cryo_solver_create(const cryo_config* cfg, const cryo_fluid* fluid, cryo_solver** out);
/* Solver::~Solver (sph_solver.cu:281-288). */
void cryo_solver_destroy(cryo_solver* s);

/* Solver::set_particles (sph_solver.cu:407-428) with HostParticles
 * (particle_system.hpp:76-90) as SoA host arrays of length n.  v, nrm, flags
 * and id may be NULL (zero / zero / 0 / 0..n-1).  domain = AABB lo, hi. */
int cryo_solver_set_particles(cryo_solver* s, int32_t n, const float* x, const float* v, const float* nrm,
                              const float* m, const float* rho, const float* T, const uint8_t* phase,
                              const uint8_t* flags, const uint32_t* id, const float domain_lo[3],
                              const float domain_hi[3]);

/* Solver::step / run_steps / run_until / relax / next_dt
 * (sph_solver.cu:511-543,707-736,506-509). */
int cryo_solver_step(cryo_solver* s);
int cryo_solver_run_steps(cryo_solver* s, int64_t n);
int cryo_solver_run_until(cryo_solver* s, double t_end);
int cryo_solver_relax(cryo_solver* s, double duration, float damping);
int cryo_solver_next_dt(cryo_solver* s, double* dt);
/* Solver::damp_velocities (sph_solver.cu:701-705). */
int cryo_solver_damp_velocities(cryo_solver* s, float factor);

/* Extension (no counterpart in the reference, whose ensembles are a loop of
 * independent Solvers): advances every solver of an ensemble `steps` steps.
 * Each member's step is queued on its own stream before any member waits for
 * its host bookkeeping, so independent tanks run side by side on the GPU.
 * Results are those of calling cryo_solver_run_steps on each member alone.
 * Members must be distinct; a solver is not thread-safe, so no other thread
 * may use a member during the call. */
int cryo_ensemble_run_steps(cryo_solver* const* solvers, int32_t count, int64_t steps);

/* Extension (no counterpart in the reference, whose CUDA kernels have one launch
 * shape): the launch variant of the interface and forces passes, bit 0 set =
 * forces by per-particle neighbour-list gathers instead of LDS-staged cell
 * tiles, bit 1 set = the same for the interface pass.  Every variant gives the
 * same results bit for bit; only the speed differs with tank size and fill.
 * cryo_solver_autotune times `trial_steps` steps of each variant from the
 * solver's current state (after one untimed step each), keeps the fastest,
 * returns it in *variant (may be NULL) and each variant's milliseconds per step
 * in ms_per_step[0..3] (may be NULL), and puts the solver's state back as it was. */
int cryo_solver_set_launch_variant(cryo_solver* s, int32_t variant);
/* Extension (no counterpart in the reference): on != 0 makes cryo_solver_run_steps
 * and cryo_ensemble_run_steps capture this solver's step as a HIP graph (one per
 * persistent buffer set, captured again when the particle count or capacity
 * changes) and replay it for the rest of the call.  Results are the same bits.
 * Default off: on MI355X the replay has not been faster than plain launches for
 * these steps (the device, not the launch path, bounds them).  on = 2 also lets
 * cryo_ensemble_run_steps, when every member is set so and none reads its
 * particle count back between steps (no phase change, or device counts),
 * capture two steps of every member as one graph and launch it once per two
 * ensemble steps (same bits; slower here than separate member streams).  Other
 * values fail with CRYO_ERR_INVALID_ARGUMENT. */
int cryo_solver_set_step_graphs(cryo_solver* s, int32_t on);
/* Extension (no counterpart in the reference, which reads the particle count
 * back after every phase-change step): on != 0 makes cryo_solver_run_steps and
 * cryo_ensemble_run_steps queue a call's phase-change steps back to back with
 * the particle count kept on the device.  The launches are sized for the count
 * at the start of the call plus `room` particles (0: max(1024, n / 8)) and the
 * kernels read the count when they run; splits append, a cell sort drops
 * merged particles, all on the device, and the host takes the counts once at
 * the end.  If a step's splits would pass that bound, the call restores the
 * state it started from and runs the steps one at a time instead (the solver
 * keeps a device copy of its particle state for that, about the size of the
 * state itself).  Results are the same bits either way.  Default off.  Fails
 * with CRYO_ERR_INVALID_ARGUMENT for room < 0. */
int cryo_solver_set_device_counts(cryo_solver* s, int32_t on, int32_t room);
/* How many such calls passed their bound and ran step by step (since the
 * solver was created): a count that grows says `room` is too small. */
int cryo_solver_device_count_restarts(cryo_solver* s, int64_t* restarts);
int cryo_solver_launch_variant(cryo_solver* s, int32_t* variant);
int cryo_solver_autotune(cryo_solver* s, int32_t trial_steps, int32_t* variant, double* ms_per_step);

/* Extension (no counterpart in the reference): the cell-tile shape of the
 * LDS-staged interface and forces passes, in grid cells along x, y and z
 * (default 3 x 3 x 3).  Accepted: 1 <= x <= 16, y, z >= 1, (y + 2)(z + 2) <= 25
 * and y z <= 9; other shapes fail with CRYO_ERR_INVALID_ARGUMENT.  Every shape
 * gives the same results bit for bit; the shape sets how many particles each
 * workgroup stages in LDS and serves.  cryo_solver_tile_shape writes the
 * current shape to cells_xyz[0..2].
 * cryo_solver_autotune_tiles times `trial_steps` steps of each of `count`
 * shapes shapes[3k..3k+2] (shapes NULL: six built-in candidates, 3x3x3, 4x3x2,
 * 4x2x3, 2x3x3, 3x3x2 and 3x2x3, and count is ignored) at the current launch
 * variant, from the solver's current state (after one untimed step each),
 * keeps the fastest, returns its index in *pick (may be NULL) and each shape's
 * milliseconds per step in ms_per_step[k] (may be NULL; count entries, six
 * for the built-in list), and puts the solver's state back as it was. */
int cryo_solver_set_tile_shape(cryo_solver* s, int32_t cells_x, int32_t cells_y, int32_t cells_z);
int cryo_solver_tile_shape(cryo_solver* s, int32_t* cells_xyz);
int cryo_solver_autotune_tiles(cryo_solver* s, int32_t trial_steps, const int32_t* shapes, int32_t count,
                               int32_t* pick, double* ms_per_step);

/* Extension (no counterpart in the reference, which has no adjoint or autodiff):
 * runs `steps` steps, step k with frame acceleration frame_acc[3k..3k+2]
 * (g_eff = gravity - acc, as cryo_solver_set_frame_motion sets it), and returns
 *   loss = sum_k dt (S_k.x^2 + S_k.y^2),  S_k = sum of fw over the fluid
 *   particles after step k's rates (the negated slosh force),
 * S_k in wall_load[3k..3k+2] (NULL: not returned) and, when grad_frame_acc is
 * not NULL, d loss / d frame_acc by a hand-written reverse sweep that
 * checkpoints every `checkpoint_every` steps and re-runs each segment, and
 * (grad_contact_angle not NULL, which needs grad_frame_acc) d loss / d the
 * contact angle in degrees (0 without CSF; on a solver not yet initialised
 * the start-up rates' share is included), and (grad_lee not NULL, two
 * doubles, which needs grad_frame_acc) d loss / d lee_evaporation and
 * d loss / d lee_condensation.  The solver ends in the forward end state
 * either way.  Differentiated: density summation, EOS, Shepard wall pressure,
 * ghost velocity and wall colour, momentum (pressure, viscosity, artificial
 * viscosity, wall terms), CSF surface tension (colour gradient, contact-angle
 * rotation and/or contact-line force, plain or smoothed normals, curvature,
 * classic or balanced force), the transport velocity and its
 * momentum stress, conduction and the
 * heater, Lee phase change (sources, receiver normalisation, apply, vapour
 * rebalance, venting, a closed tank's saturation-temperature feedback;
 * masses and temperatures are state), kick-drift-kick.  Requires
 * fixed_dt > 0 (either kernel: Wendland C2 or cubic spline), density summation, no SPS
 * or shifting, no manoeuvre table and no rotating frame (see
 * cryo_solver_adjoint_slosh_frame); a window in which a particle splits or
 * merges is not differentiable (CRYO_ERR_INVALID_ARGUMENT otherwise). */
int cryo_solver_adjoint_slosh(cryo_solver* s, int64_t steps, const float* frame_acc, int32_t checkpoint_every,
                              double* loss, double* grad_frame_acc, double* wall_load, double* grad_contact_angle,
                              double* grad_lee);
/* Extension (no counterpart in the reference): cryo_solver_adjoint_slosh in a
 * rotating frame.  Step k also sets the frame's angular velocity
 * frame_omega[3k..3k+2] and its rate frame_omega_dot[3k..3k+2] (either may be
 * NULL: zero), as cryo_solver_set_frame_motion would (a -= wd x r + w x (w x r)
 * + 2 w x v about the frame origin), and grad_frame_omega /
 * grad_frame_omega_dot (NULL: not returned; they need grad_frame_acc) receive
 * d loss / d those inputs.  A rotating frame set on the solver beforehand is
 * replaced by these inputs; cryo_solver_adjoint_slosh rejects one. */
int cryo_solver_adjoint_slosh_frame(cryo_solver* s, int64_t steps, const float* frame_acc, const float* frame_omega,
                                    const float* frame_omega_dot, int32_t checkpoint_every, double* loss,
                                    double* grad_frame_acc, double* grad_frame_omega, double* grad_frame_omega_dot,
                                    double* wall_load, double* grad_contact_angle, double* grad_lee);

/* Solver::time / step_count / last_dt / num_particles (sph_solver.hpp:153-163). */
int cryo_solver_time(cryo_solver* s, double* t);
int cryo_solver_step_count(cryo_solver* s, int64_t* step);
int cryo_solver_last_dt(cryo_solver* s, double* dt);
int cryo_solver_num_particles(cryo_solver* s, int32_t* n);
int cryo_solver_capacity(cryo_solver* s, int32_t* cap);
/* Solver::config / fluid / params (sph_solver.hpp:182-184). */
int cryo_solver_config(cryo_solver* s, cryo_config* out);
int cryo_solver_fluid(cryo_solver* s, cryo_fluid* out);
int cryo_solver_params(cryo_solver* s, cryo_params* out);

/* Solver::diagnostics (sph_solver.cu:609-653): one fused FP64 reduction. */
int cryo_solver_diagnostics(cryo_solver* s, cryo_diagnostics* out);

/* ParticleSystem::download / download_pressure / download_curvature
 * (particle_system.cpp:88-95, particle_system.hpp:115-117): copies field
 * `field` of the first `count` live particles into dst (count must be <=
 * num_particles; bytes = count * element size). */
int cryo_solver_download(cryo_solver* s, int32_t field, void* dst, int32_t count);

/* Run-time controls (sph_solver.cu:655-699). */
int cryo_solver_set_gravity(cryo_solver* s, const float g[3]);
int cryo_solver_set_frame_motion(cryo_solver* s, const float acc[3], const float omega[3],
                                 const float omega_dot[3], const float origin[3]);
/* Solver::set_maneuver with ManeuverProfile (sph_solver.hpp:96-105): n rows of
 * time (double), acceleration (3 floats), angular velocity (3 floats or NULL). */
int cryo_solver_set_maneuver(cryo_solver* s, int32_t n, const double* time, const float* acc,
                             const float* omega, const float origin[3]);
int cryo_solver_clear_frame_motion(cryo_solver* s);
int cryo_solver_set_heater_power(cryo_solver* s, float watts);
int cryo_solver_set_contact_angle(cryo_solver* s, float degrees);

/* ManeuverProfile::validate / evaluate (sph_solver.cu:234-268), host only. */
int cryo_maneuver_evaluate(int32_t n, const double* time, const float* acc, const float* omega, double t,
                           float a_out[3], float w_out[3], float wdot_out[3]);

/* Solver::probe_pressure (sph_solver.cu:738-748; kernel :197-216). */
int cryo_solver_probe_pressure(cryo_solver* s, int32_t np, const float* points, float* out);
/* cryo_solver_probe_pressure without allocating: no counterpart in the
 * reference.  For callers that cannot allocate device memory (HIP graph
 * capture, tight serving loops).  At most 65536 points per call
 * (CRYO_ERR_OUT_OF_RANGE beyond); the solver must already be initialised
 * (a step or next_dt has run, CRYO_ERR_RUNTIME otherwise).  Points and
 * results pass through one fixed 1 MiB device buffer per GPU that the
 * library holds until the process ends, so calls that overlap on one GPU
 * race.  Same results as cryo_solver_probe_pressure. */
int cryo_solver_probe_pressure_noalloc(cryo_solver* s, int32_t np, const float* points, float* out);

/* Spatial hash of the solver's current state (spatial_hash.cu:140-171):
 * per-particle count and id-sum of particles closer than `radius`. */
int cryo_solver_neighbor_stats(cryo_solver* s, float radius, int32_t* counts, uint64_t* id_sums);
/* Full neighbour sets (r^2 < radius^2, self included) in enumeration order
 * (27 cells z->y->x, index ascending; spatial_hash.hpp:37-56) as CSR.  Call
 * with indices == NULL to get offsets (n+1 entries) only. */
int cryo_solver_neighbor_csr(cryo_solver* s, float radius, int64_t* offsets, int32_t* indices);
/* Grid dimensions, the sort permutation of the last build (new index ->
 * index before the build) and the cell start/end tables (-1 = empty). */
int cryo_solver_grid_info(cryo_solver* s, int32_t dims[3], float origin[3], float* cell_size, int32_t* num_cells);
int cryo_solver_grid_tables(cryo_solver* s, int32_t* cell_start, int32_t* cell_end, uint32_t* perm,
                            int32_t* n_sorted);

/* Operator-level entry points for parity testing: ensure_initialized
 * (sph_solver.cu:439-447), a grid rebuild (SpatialHash::build,
 * spatial_hash.cu:48-114) and evaluate_rates (sph_solver.cu:449-480) on the
 * current state. */
int cryo_solver_initialize(cryo_solver* s);
int cryo_solver_rebuild_grid(cryo_solver* s, int32_t* removed);
int cryo_solver_evaluate_rates(cryo_solver* s);

/* Standalone spatial hash over host arrays (SpatialHash::configure/build/
 * neighbor_stats as the reference tests use them, test_spatial_hash.cpp:25-93):
 * sorts x/phase by cell (dead last), returns the permutation, n_alive and,
 * if counts != NULL, neighbour counts and id sums (ids = original index)
 * of the sorted particles within `radius`. */
int cryo_hash_sort(int32_t n, const float* x, const uint8_t* phase, const float lo[3], const float hi[3],
                   float cell, uint32_t* perm, int32_t* n_alive, float radius, int32_t* counts,
                   uint64_t* id_sums);
/* cryo_hash_sort for particles that carry their own (unique) ids: the id
 * sums add those ids, perm still gives input positions. */
int cryo_hash_sort_ids(int32_t n, const float* x, const uint8_t* phase, const uint32_t* id, const float lo[3],
                       const float hi[3], float cell, uint32_t* perm, int32_t* n_alive, float radius, int32_t* counts,
                       uint64_t* id_sums);

/* ------------------------------------------------------ profiling ---- */
/* HIP-event timing of the step's sections on the solver stream (no
 * counterpart in the reference, which only times whole steps with
 * steady_clock, benchmarks/perf_scaling.cpp:86-104).  mask bit k enables
 * section k: 0 begin+kick-drift+keys, 1 cell sort, 2 gather, 3 density+EOS,
 * 4 wall BC, 5 interface fields, 6 forces (TVF, momentum, CSF, continuity,
 * conduction), 7 kick+sources, 8 phase change, 9 end of step. */
int cryo_solver_profile_enable(cryo_solver* s, uint32_t mask);
int cryo_solver_profile_reset(cryo_solver* s);
/* Name, accumulated milliseconds and launch count of section `index`;
 * CRYO_ERR_OUT_OF_RANGE past the last section. */
int cryo_solver_profile_section(cryo_solver* s, int32_t index, char* name, size_t len, double* total_ms,
                                int64_t* count);

/* ---------------------------------------------------- telemetry ---- */
/* TelemetryStream (telemetry.hpp:33-70): configure / request / ready /
 * fetch.  A frame holds n records of (x, y, z, T, p, speed) floats + phase. */
int cryo_solver_configure_telemetry(cryo_solver* s, int32_t mode);
int cryo_solver_telemetry_mode(cryo_solver* s, int32_t* mode);
/* Returns 1 if a capture was enqueued, 0 if one is already pending. */
int cryo_solver_request_telemetry(cryo_solver* s);
int cryo_solver_telemetry_ready(cryo_solver* s);
/* Blocks for the pending frame.  Returns 1 and fills the outputs (arrays of
 * capacity >= *n as sized by cryo_solver_telemetry_pending_size) or 0 if
 * nothing was requested. */
int cryo_solver_telemetry_pending_size(cryo_solver* s, int32_t* n);
int cryo_solver_fetch_telemetry(cryo_solver* s, double* time, int64_t* step, int32_t* n, float* xyz,
                                float* T, float* p, float* speed, uint8_t* phase);

/* ------------------------------------------------------------ files ---- */
/* Host-side writers; they need no device.
 * write_vtk (io.cpp:52-87): legacy binary VTK POLYDATA (big-endian) of n
 * particles from host arrays (x, v: 3n floats); pressure may be NULL, which
 * leaves that point array out as the reference does for a size mismatch. */
int cryo_write_vtk(const char* path, int32_t n, const float* x, const float* v, const float* T, const float* m,
                   const float* rho, const uint8_t* phase, const float* pressure);
/* write_json_frame (io.cpp:89-142) of a TelemetryFrame (telemetry.hpp:15-24)
 * given as arrays: xyz 3n, T, p (NULL = zeros), speed, phase; walls are left
 * out unless include_walls. */
int cryo_write_json_frame(const char* path, double time, int64_t step, int32_t n, const float* xyz, const float* T,
                          const float* p, const float* speed, const uint8_t* phase, int32_t include_walls);
/* make_frame's speed column (io.cpp:100-117): |v| of each particle in float. */
void cryo_frame_speed(int32_t n, const float* v, float* speed);

/* ------------------------------------------------------- geometry ---- */
/* Tank geometry as a flat node list (geometry.hpp:22-150).  Node k refers to
 * children by index < k; the last node is the root. */
enum {
  CRYO_GEO_CYLINDER = 0, /* p0 = radius, p1 = height            */
  CRYO_GEO_CAPSULE = 1,  /* p0 = radius, p1 = barrel length     */
  CRYO_GEO_SPHERE = 2,   /* p0 = radius                         */
  CRYO_GEO_BOX = 3,      /* p0..p2 = full side lengths          */
  CRYO_GEO_TRANSLATED = 4, /* a = child, p0..p2 = offset        */
  CRYO_GEO_DIFFERENCE = 5, /* a minus solid b                   */
  CRYO_GEO_UNION = 6,      /* a united with b                   */
  CRYO_GEO_MESH = 7,       /* mesh_id = index into the mesh list */
  CRYO_GEO_CALLBACK = 8    /* mesh_id = index into the callback list (a caller's own distance function) */
};
typedef struct cryo_geo_node {
  int32_t type, a, b, mesh_id;
  double p[3];
} cryo_geo_node;
typedef struct cryo_mesh {
  int32_t n_vertices, n_triangles;
  const double* vertices;   /* 3 * n_vertices */
  const int32_t* triangles; /* 3 * n_triangles */
} cryo_mesh;

/* A caller-defined distance function: a Geometry subclass of the caller's
 * (geometry.hpp:22-34 lets callers override sdf / bounds / volume).  sdf is
 * called from several host threads at once and must be thread-safe; volume < 0
 * means unknown.  Tanks holding one are built on the host only. */
typedef double (*cryo_sdf_fn)(void* ctx, double x, double y, double z);
typedef struct cryo_geo_callback {
  cryo_sdf_fn sdf;
  void* ctx;
  float lo[3], hi[3];
  double volume;
} cryo_geo_callback;

/* FillSpec (geometry.hpp:162-175). NaN temperatures mean T_sat. */
typedef struct cryo_fill {
  int32_t mode;
  double fill_fraction;
  cryo_vec3 settle_direction, bubble_center;
  int32_t include_vapor, wall_layers;
  float T_liquid, T_vapor, T_wall;
  double heater_zmin, heater_zmax;
} cryo_fill;
void cryo_fill_default(cryo_fill* out);

/* Geometry::sdf / bounds / volume (geometry.cpp:51-239). */
int cryo_geo_sdf(int32_t n_nodes, const cryo_geo_node* nodes, int32_t n_meshes, const cryo_mesh* meshes,
                 double x, double y, double z, double* out);
int cryo_geo_bounds(int32_t n_nodes, const cryo_geo_node* nodes, int32_t n_meshes, const cryo_mesh* meshes,
                    float lo[3], float hi[3]);
int cryo_geo_volume(int32_t n_nodes, const cryo_geo_node* nodes, int32_t n_meshes, const cryo_mesh* meshes,
                    double* out);
int cryo_mesh_winding_number(const cryo_mesh* mesh, double x, double y, double z, double* out);

/* build_tank_particles (geometry.cpp:242-330): lattice count first
 * (x == NULL), then fill the caller's arrays (x, v, nrm 3n floats). */
int cryo_build_tank_particles(int32_t n_nodes, const cryo_geo_node* nodes, int32_t n_meshes,
                              const cryo_mesh* meshes, const cryo_fluid* fluid, float dx, const cryo_fill* fill,
                              int32_t* n, float* x, float* v, float* nrm, float* m, float* rho, float* T,
                              uint8_t* phase, uint8_t* flags, uint32_t* id);
/* build_tank_particles run on the current device (next to
 * geometry.cpp:242-330, which runs on the host): the lattice's node classes,
 * their compaction to wall and fluid lists, the wall normals, the fill cut and
 * the particle arrays are all computed there and downloaded; the same
 * particles bit for bit, for analytic, CSG and mesh tanks (a mesh's distances
 * against every triangle and its winding-number sign per lattice node, on the
 * device). Caller-defined distance functions are refused (host only). Same
 * arguments and two-call protocol as cryo_build_tank_particles. */
int cryo_build_tank_particles_device(int32_t n_nodes, const cryo_geo_node* nodes, int32_t n_meshes,
                                     const cryo_mesh* meshes, const cryo_fluid* fluid, float dx,
                                     const cryo_fill* fill, int32_t* n, float* x, float* v, float* nrm, float* m,
                                     float* rho, float* T, uint8_t* phase, uint8_t* flags, uint32_t* id);
/* Solver::build_tank (sph_solver.cu:430-437). */
int cryo_solver_build_tank(cryo_solver* s, int32_t n_nodes, const cryo_geo_node* nodes, int32_t n_meshes,
                           const cryo_mesh* meshes, const cryo_fill* fill);
/* build_tank_particles and Solver::build_tank for trees that hold
 * CRYO_GEO_CALLBACK nodes (a caller's own Geometry subclasses); otherwise as
 * cryo_build_tank_particles / cryo_solver_build_tank. */
int cryo_build_tank_particles_cb(int32_t n_nodes, const cryo_geo_node* nodes, int32_t n_meshes,
                                 const cryo_mesh* meshes, int32_t n_callbacks, const cryo_geo_callback* callbacks,
                                 const cryo_fluid* fluid, float dx, const cryo_fill* fill, int32_t* n, float* x,
                                 float* v, float* nrm, float* m, float* rho, float* T, uint8_t* phase,
                                 uint8_t* flags, uint32_t* id);
int cryo_solver_build_tank_cb(cryo_solver* s, int32_t n_nodes, const cryo_geo_node* nodes, int32_t n_meshes,
                              const cryo_mesh* meshes, int32_t n_callbacks, const cryo_geo_callback* callbacks,
                              const cryo_fill* fill);

/* -------------------------------------------------------- streams ---- */
/* mem::create_stream / destroy_stream / synchronize (cuda_utils.cpp:48-60)
 * as opaque handles (a blocking HIP stream on the current device), and
 * Solver::stream (sph_solver.hpp:181): the solver's own compute stream, for
 * a caller that waits on it. */
int cryo_stream_create(void** stream);
int cryo_stream_destroy(void* stream);
int cryo_stream_synchronize(void* stream);
int cryo_solver_stream(cryo_solver* s, void** stream);

#ifdef __cplusplus
}
#endif
#endif /* CRYOSPH_H */
