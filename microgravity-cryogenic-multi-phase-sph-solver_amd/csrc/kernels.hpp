// Host-callable launchers of the gfx950 kernels (all asynchronous on `s`).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "cryo_device.hpp"

namespace cryo {

// Prescribed tank motion table on the device (ManeuverProfile, sph_solver.hpp:96-105).
struct Maneuver {
  int n;               // 0 = no manoeuvre
  const double* time;  // n
  const V3* acc;       // n
  const V3* omega;     // n or nullptr
  V3 origin;
};

// A particle count as a kernel sees it.  Normally the host's count (`dev` =
// nullptr).  While the host queues several phase-change steps without reading
// the count back (splits append particles, a sort drops merged ones), `dev`
// points at the count in the device state, read by each kernel when it runs,
// and `ub`, an upper bound of it, sizes the launch: threads at or past the
// count exit, block partial sums past it are zero.
struct Count {
  int ub = 0;
  const int* dev = nullptr;
  Count() = default;
  Count(int n) : ub(n) {}  // a host count converts implicitly
  Count(int bound, const int* d) : ub(bound), dev(d) {}
  __device__ int get() const { return dev ? *dev : ub; }
};

// Fluid constants used by the ullage update (FluidProperties, fluid_properties.hpp:11-38).
struct FluidConst {
  float rho_liquid, rho_vapor, T_sat, p_sat, R_gas, h_fg;
};

// grid
void launch_kick_drift_keys(const PV& pv, const GridDev& g, DevState* st, Count n, bool do_kick, bool continuity,
                            bool tvf, uint32_t* keys, uint32_t* vals, hipStream_t s);
int radix_tiles(int n);
void launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, Count n,
                       int key_bits, uint32_t* tile_hist, uint32_t* row_total, hipStream_t s, bool* result_in_alt);
void launch_cell_offsets(const uint32_t* keys, Count n, int num_cells, int* off, DevState* st, hipStream_t s);
void launch_cell_flags(const int* off, const uint8_t* phase, int num_cells, uint8_t* flags, hipStream_t s);
void launch_gather(const Persist& dst, const Persist& src, const uint32_t* perm, const DevState* st, int n_ub,
                   hipStream_t s);
void launch_neighbor_stats(const float4* x, const uint32_t* id, const GridDev& g, int n, float r2max, int* counts,
                           unsigned long long* sums, hipStream_t s);
void launch_neighbor_fill(const float4* x, const GridDev& g, int n, float r2max, const long long* offsets,
                          int* indices, hipStream_t s);
void launch_scan_one_block(uint32_t* data, int total, uint32_t* total_out, hipStream_t s);

// physics
void launch_density(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, Count n, hipStream_t s);
void launch_normalize_masses(const PV& pv, const GridDev& g, const SimParams& P, int n, hipStream_t s);
void launch_wall_bc(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, const DevState* st, Count n,
                    hipStream_t s);
int launch_interface(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, DevState* st, Count n,
                     bool csf_mode, double* heat_partials, bool tiled, hipStream_t s);
void launch_smooth_normals(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, Count n,
                           hipStream_t s);
void launch_heat_norm(DevState* st, const double* partials, int nb, bool active, hipStream_t s);
void launch_forces(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, DevState* st, Count n,
                   bool csf_mode, bool akinci, bool continuity, bool thermal, bool tiled, hipStream_t s);
void launch_probe(const PV& pv, const GridDev& g, const SimParams& P, const float* pts, float* out, int np,
                  hipStream_t s);
void launch_shifting(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, const DevState* st, Count n,
                     hipStream_t s);

// step
int diag_quantities();
void launch_begin_step(DevState* st, const SimParams& P, const DtLimits& L, const Maneuver& M, hipStream_t s);
void launch_kick_sources(const PV& pv, const SimParams& P, DevState* st, Count n, bool do_kick, bool thermal,
                         bool sources, hipStream_t s);
int launch_pc_normalize(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, Count n,
                        double* partials, int stride, hipStream_t s);
void launch_pc_apply(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, Count n, hipStream_t s);
void launch_rebalance(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, Count n, hipStream_t s);
void launch_rebalance_apply_mark(const PV& pv, const SimParams& P, Count n, bool rebalance, uint32_t* flag,
                                 hipStream_t s);
void launch_exclusive_scan(const uint32_t* in, uint32_t* out, Count n, uint32_t* block_sums, uint32_t* total,
                           hipStream_t s);
void launch_split(const PV& pv, const SimParams& P, const DevState* st, Count n, const uint32_t* flag,
                  const uint32_t* offset, hipStream_t s);
void launch_split_commit(DevState* st, Count n, int capacity, const uint32_t* total, hipStream_t s);
int launch_venting(const PV& pv, const SimParams& P, const DevState* st, int n_ub, float relax, double* partials,
                   int stride, hipStream_t s);
int launch_inventory(const PV& pv, const DevState* st, int n_ub, double* partials, int stride, hipStream_t s);
void launch_end_step(DevState* st, const double* partials, int stride, int nb_pc, int nb_vent, int nb_inv,
                     const FluidConst& fl, bool closed, bool advance_clock, hipStream_t s);
void launch_pc_fold(DevState* st, const double* partials, int stride, int nb_pc, hipStream_t s);
void launch_scale_velocities(const PV& pv, int n, float f, hipStream_t s);
void launch_diagnostics(const PV& pv, const SimParams& P, int n, double* partials, int stride, unsigned* maxbits,
                        double* out, hipStream_t s);
void launch_telemetry_pack(const PV& pv, int n, float* rec, hipStream_t s);

// Reverse-mode adjoint of the hydrodynamic step (kernels_adjoint.hip).
struct AdjArgs {
  const float4 *cx, *cv, *ca;  // cotangents of x, v, a at the step's end (this step's sorted order)
  float4 *nx, *nv, *na;        // cotangents at the step's start (previous order), written through perm
  float4* at;                  // a_bar + (dt/2) v_bar of each fluid particle
  float4 *wcp, *wcv;           // walls: {cp0, cp1} and cvw, Shepard-weighted cotangents of pw0, pw1, vw
  float4 *xr, *vr;             // the rates' share of the x and v cotangents
  float* sb;                   // cotangent of sigma
  const uint32_t* perm;        // the step's sort permutation (new i <- old perm[i])
  const double* S;             // the step's sum of fw over fluid particles
  double* partials;            // 8 rows of `stride` doubles
  int stride;
  // surface tension (CSF with the contact-angle rotation); csf = 0: not used
  int csf;
  float4* cs;   // {num_bar, den_bar, kappa} of each fluid particle's curvature sums
  float* csv;   // cotangent of V = m / rho from the curvature and colour sums
  float4* gcb;  // cotangents of the colour gradient (xyz) and the wall weight ww (w)
  float4* wnb;  // cotangent of the wall-normal sum
  float4* xc;   // the curvature sums' share of the x cotangent
  // transport velocity; tvf = 0: not used
  int tvf;
  const float4* ct;  // cotangent of the transport acceleration ab at the step's end
  float4* nt;        // the same before the step (the drift reads ab), written through perm
  float4* tvs;       // {K * ab_bar, 1} of each active transport owner, else 0
  // heat and phase change; mT = 0: masses and temperatures are constants
  int mT, heat, pc, reb, vent;
  float relax;              // venting relaxation
  const float *cm, *cT;     // cotangents of m and T at the step's end
  float *nm, *nT;           // the same at the step's start, written through perm
  float *mr, *Tr;           // the step's m and T cotangents (sorted order)
  float* qb;                // cotangent of each fluid particle's conduction sum
  double* hsum;             // sum of qb heat_w (the heater normalisation's cotangent)
  // the phase-change re-run: kicked v and T, normalised out_m, out_h, applied m and T,
  // the sources' flags
  const float4* v1;
  const float *T1, *om, *oh, *m2, *T2;
  const uint8_t* fl;
  float4* pa;    // {m_bar before phase change, kicked T_bar, in_m_bar, in_H_bar}
  float4* po;    // {out_m_bar, out_h_bar, w_norm_bar (apply share)}
  float4* pb4;   // rebalance {dm_bar, dH_bar, m2_bar (own), T2_bar (own)}
  float4* pdp;   // rebalance {dP_bar, rho_bar (venting)}
  float4* v1b;   // cotangent of the kicked velocity
  float2* pm2;   // {m2_bar, T2_bar}
  float4* pn;    // {w_norm_bar, wr_max_bar, s_all_bar, s_other_bar}
  int* amax;     // the neighbour giving a capped source's wr_max, else -1
  float *pV, *pr;  // phase change's share of the V = m / rho and rho cotangents
  float4* xp;      // phase change's share of the x cotangent
  double* lee_out; // this step's d loss / d (lee_evap, lee_cond)
  // closed tank: this step's T_sat came from the previous step's end state
  int sat;
  FluidConst fluid;
  double* satv;    // {sum_L m bar, sum_V m T bar} of that feedback
  // rotating frame: this step's d loss / d (omega, omega_dot), 6 doubles (null: not wanted)
  double* wout;
  // balanced CSF (csf and P.csf_balanced): {G_bar (xyz), the 1 / m factor's mass cotangent (w)}
  float4* gbal = nullptr;
  // transport stress (P.tvf and P.tvf_stress): the rates' share of the vt cotangent
  float4* vtr = nullptr;
  // smoothed normals: each particle's cotangent of its smoothing sum
  float4* sacc = nullptr;
  // contact-line force: each fluid particle's cotangent of its wall-indicator gradient
  float4* gpb = nullptr;
};
// Several independent device-to-device copies in one launch (16-byte aligned
// pointers): the adjoint's snapshots.
struct CopyItem {
  void* dst;
  const void* src;
  unsigned long long bytes;
};
void launch_copy_batch(const CopyItem* items, int count, hipStream_t s);
void launch_adj_fill(float4* a, float4* b, float4* c, int n, hipStream_t s);
void launch_adj_set_frame(DevState* st, const float* acc, const float* omega, const float* omega_dot, hipStream_t s);
void launch_adj_load_sum(const PV& pv, int n, double* partials, int stride, double* out, hipStream_t s);
void launch_adj_step(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, const DevState* st,
                     const AdjArgs& A, int n, double* grad_out, double* theta_out, hipStream_t s);
// The contact-angle share of the start-up rates: A.at holds a's cotangent.
void launch_adj_csf_only(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, const DevState* st,
                         const AdjArgs& A, int n, double* theta_out, hipStream_t s);

}  // namespace cryo
