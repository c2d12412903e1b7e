// Device-side data layout and per-pair maths of the CryoSPH step on gfx950.
//
// Layout in HBM (structure of arrays, capacity-sized, one allocation each):
//   persistent (permuted by the cell sort, double-buffered for the gather):
//     x, v, vt, nrm : float4 (xyz, w = 0)      m, m0, rho, T : float
//     phase, flags  : uint8                     id            : uint32
//   scratch (indexed like the reference: never permuted, rewritten before use):
//     a, ab, ns, vw, fw, mw : float4            p, drho, dTdt, sigma, kappa, nu_t,
//     alpha, wall_w, heat_w, out_m, out_h, w_norm, p_wv, rmin : float
// float4 keeps every vector load/store a single 16-byte access.
//
// Arithmetic: FP32 in the same operation order as the reference per-pair
// bodies (SPH_kernels.cu, navier_stokes.cu, surface_tension.cu,
// phase_change.cu); everything is compiled with -ffp-contract=off so that no
// multiply-add is fused, and FP32 division / sqrt are IEEE correctly rounded.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cryo {

constexpr float kPi = 3.14159265358979323846f;
constexpr uint8_t kLiquid = 0, kVapor = 1, kWall = 2, kDead = 3;
constexpr uint8_t kHeated = 1, kSource = 2, kTargetVapor = 4, kDying = 8, kScratchMask = 2 | 4 | 8;

// ------------------------------------------------------------------ vectors
struct V3 {
  float x, y, z;
};
__host__ __device__ inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__host__ __device__ inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline V3& operator+=(V3& a, V3 b) {
  a.x += b.x;
  a.y += b.y;
  a.z += b.z;
  return a;
}
__host__ __device__ inline V3& operator-=(V3& a, V3 b) {
  a.x -= b.x;
  a.y -= b.y;
  a.z -= b.z;
  return a;
}
__host__ __device__ inline V3& operator*=(V3& a, float s) {
  a.x *= s;
  a.y *= s;
  a.z *= s;
  return a;
}
// Neighbour-sum accumulators.  KS = false: a plain FP32 running sum, the
// reference's rounding.  KS = true: the compensated (Kahan-Babuska /
// Neumaier) sum, whose rounding error stays O(eps) instead of growing with
// the number of neighbours; the opt-in precision form.
template <bool KS>
struct Acc;
template <>
struct Acc<false> {
  float s = 0.0f;
  __host__ __device__ Acc& operator+=(float x) {
    s += x;
    return *this;
  }
  __host__ __device__ Acc& operator-=(float x) {
    s -= x;
    return *this;
  }
  __host__ __device__ operator float() const { return s; }
};
template <>
struct Acc<true> {
  float s = 0.0f, c = 0.0f;
  __host__ __device__ Acc& operator+=(float x) {
    const float t = s + x;
    const float lost = fabsf(s) >= fabsf(x) ? (s - t) + x : (x - t) + s;
    c += lost;
    s = t;
    return *this;
  }
  __host__ __device__ Acc& operator-=(float x) { return *this += -x; }
  __host__ __device__ operator float() const { return s + c; }
};
template <bool KS>
struct AccV3 {
  Acc<KS> x, y, z;
  __host__ __device__ AccV3& operator+=(V3 a) {
    x += a.x;
    y += a.y;
    z += a.z;
    return *this;
  }
  __host__ __device__ AccV3& operator-=(V3 a) {
    x -= a.x;
    y -= a.y;
    z -= a.z;
    return *this;
  }
  __host__ __device__ operator V3() const { return V3{float(x), float(y), float(z)}; }
};

__host__ __device__ inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline float norm2(V3 a) { return dot(a, a); }
__host__ __device__ inline float norm(V3 a) { return sqrtf(dot(a, a)); }
__host__ __device__ inline V3 normalized(V3 a, float eps = 1e-20f) {
  const float n = norm(a);
  return n > eps ? a / n : V3{0, 0, 0};
}
__host__ __device__ inline V3 ld3(const float4& q) { return V3{q.x, q.y, q.z}; }
__host__ __device__ inline float4 st3(V3 a) { return make_float4(a.x, a.y, a.z, 0.0f); }
template <class T>
__host__ __device__ inline T clampv(T v, T lo, T hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}
__host__ __device__ inline bool is_fluid(uint8_t ph) { return ph <= kVapor; }
__host__ __device__ inline int pidx(uint8_t ph) { return ph == kVapor ? 1 : 0; }

// ------------------------------------------------------------ smoothing kernel
struct KernelP {
  int type;  // 0 cubic spline, 1 Wendland C2
  float h, inv_h, support, norm;
};
__host__ __device__ inline float kW(const KernelP& k, float r) {
  const float q = r * k.inv_h;
  if (q >= 2.0f) return 0.0f;
  if (k.type == 0) {
    if (q < 1.0f) return k.norm * (1.0f - 1.5f * q * q + 0.75f * q * q * q);
    const float t = 2.0f - q;
    return k.norm * 0.25f * t * t * t;
  }
  const float t = 1.0f - 0.5f * q;
  const float t2 = t * t;
  return k.norm * t2 * t2 * (2.0f * q + 1.0f);
}
__host__ __device__ inline float kdW(const KernelP& k, float r) {
  const float q = r * k.inv_h;
  if (q >= 2.0f) return 0.0f;
  if (k.type == 0) {
    if (q < 1.0f) return k.norm * k.inv_h * (-3.0f * q + 2.25f * q * q);
    const float t = 2.0f - q;
    return k.norm * k.inv_h * -0.75f * t * t;
  }
  const float t = 1.0f - 0.5f * q;
  return k.norm * k.inv_h * -5.0f * q * t * t * t;
}
__host__ __device__ inline V3 kgradW(const KernelP& k, V3 rij, float r) {
  if (r <= 1e-12f * k.h) return V3{0, 0, 0};
  return rij * (kdW(k, r) / r);
}
__host__ __device__ inline float akinci_coh(float r, float c) {
  if (r >= c || r <= 0.0f) return 0.0f;
  const float c3 = c * c * c;
  const float c9 = c3 * c3 * c3;
  const float coeff = 32.0f / (kPi * c9);
  const float t = (c - r);
  const float base = t * t * t * r * r * r;
  if (2.0f * r > c) return coeff * base;
  return coeff * (2.0f * base - c3 * c3 / 64.0f);
}

// Element k (0 or 1) of a two-entry table without indexing it: a dynamic
// index into a by-value parameter struct would put the struct in scratch.
template <class T>
__host__ __device__ inline T sel2(const T (&a)[2], int k) {
  return k ? a[1] : a[0];
}

// ---------------------------------------------------------- static params
// Everything init_params derives once (sph_solver.cu:292-405).  Passed by
// value to every kernel; values that change during a run live in DevState.
struct SimParams {
  KernelP kernel;
  float dx, vol0, support2;
  float rho0[2], mu[2], k[2], cp[2], mref[2];
  float c0, tait_B, tait_gamma, p_bg, gas_R, vapor_K;
  int clamp_neg;
  float alpha_av, alpha_av_vap, delta_sph;
  // harmonic viscosity 2 mu_a mu_b / (mu_a + mu_b) of the phase pairs LL, LV,
  // VV (0 when mu_a + mu_b = 0), evaluated once in the pair loops' order
  float eta_ll, eta_lv, eta_vv;
  int density_mode, no_slip, sps, tvf, tvf_stress;
  float tvf_factor, tvf_sigma_min, shift_strength;
  int tvf_interface_only;
  float smag_cs;
  int st_model;
  float sigma, cos_theta, sin_theta, normal_eps, wall_adh_w;
  int wetting, wall_colour, csf_smooth, csf_balanced;
  float ak_gamma, ak_beta, ak_adh_scale;  // ak_adh_scale = 0.007 / c^3.25 (host)
  int thermal, phase_change, wall_bc;
  float T_wall, T_ref, h_fg, lee_evap, lee_cond, max_transfer, vapor_rebalance, pc_density_limit, merge_ratio,
      split_ratio, heat_power;
  int has_vapor;
  int compensated;  // compensated neighbour sums (precision form)
};

// ------------------------------------------------------- device state block
// Device-resident per-step state: the host never needs it to launch a step.
struct DevState {
  double time, last_dt;
  long long step;
  double heat_input, evaporated, condensed, vented, vented_h, p_ullage, fluid_volume0;
  float dt;             // P.dt of the reference (time step of the current step)
  float T_sat;          // current saturation temperature
  float heat_inv_norm;  // 1 / sum heat_w
  float v_max, a_max, Tv_max, nu_t_max, spacing_ratio;  // evaluate_rates maxima
  // frame (sph_solver.cu:660-667)
  V3 cfg_gravity, frame_acc, gravity, omega, omega_dot, origin;
  int frame_rot;
  int n;        // live particles
  int n_alive;  // after the last cell sort
  int splits, merges;
  unsigned next_id;
  unsigned split_requested;
  unsigned max_bits[5];  // atomicMax scratch for the maxima (float bits, all >= 0)
  int n_sort;            // the count the step's cell sort runs over (dead particles included)
  int merges_sorted;     // merges at the last cell sort (later ones are still in the arrays)
  int count_overflow;    // a step's splits passed the launch bound of a queued run of steps
  int pad;
};

// --------------------------------------------------------------- equations
__host__ __device__ inline float eos_pressure(const SimParams& P, uint8_t ph, float rho, float T) {
  float p;
  if (ph == kVapor) {
    p = rho * P.gas_R * T + (P.p_bg - P.vapor_K);
  } else {
    const float ratio = rho / P.rho0[0];
    float rg;
    if (P.tait_gamma == 7.0f) {
      const float r2 = ratio * ratio;
      rg = r2 * r2 * r2 * ratio;
    } else {
      rg = (float)pow((double)ratio, (double)P.tait_gamma);
    }
    p = P.tait_B * (rg - 1.0f) + P.p_bg;
  }
  return (P.clamp_neg && p < 0.0f) ? 0.0f : p;
}
// The same equation of state from a density carried in double (the
// precision form's summation density, m times the compensated kernel sum):
// near rest p is a small difference of large terms, and an FP32 density
// would quantise it to B gamma 2^-24.  Rounded to float once, at the end.
__host__ __device__ inline float eos_pressure_hp(const SimParams& P, uint8_t ph, double rho, float T) {
  double p;
  if (ph == kVapor) {
    p = rho * static_cast<double>(P.gas_R) * static_cast<double>(T) + (static_cast<double>(P.p_bg) - P.vapor_K);
  } else {
    const double ratio = rho / static_cast<double>(P.rho0[0]);
    double rg;
    if (P.tait_gamma == 7.0f) {
      const double r2 = ratio * ratio;
      rg = r2 * r2 * r2 * ratio;
    } else {
      rg = pow(ratio, static_cast<double>(P.tait_gamma));
    }
    p = static_cast<double>(P.tait_B) * (rg - 1.0) + P.p_bg;
  }
  return (P.clamp_neg && p < 0.0) ? 0.0f : static_cast<float>(p);
}
__host__ __device__ inline float sound_speed(const SimParams& P, uint8_t ph, float T) {
  if (ph == kVapor) return sqrtf(P.gas_R * (T > 0.0f ? T : 0.0f));
  return P.c0;
}
__host__ __device__ inline float enthalpy(const SimParams& P, uint8_t ph, float T) {
  return ph == kVapor ? P.h_fg + P.cp[1] * (T - P.T_ref) : P.cp[0] * (T - P.T_ref);
}
__host__ __device__ inline float temp_from_h(const SimParams& P, uint8_t ph, float h) {
  return ph == kVapor ? P.T_ref + (h - P.h_fg) / P.cp[1] : P.T_ref + h / P.cp[0];
}

// Time-step limits (sph_solver.cu:482-504); cfg numbers passed in DtLimits.
struct DtLimits {
  float cfl, viscous_number, force_number, capillary_number, thermal_number, dt_max, dt_min, fixed_dt;
};
__host__ __device__ inline float compute_dt(const SimParams& P, const DtLimits& L, float v_max, float a_max,
                                            float Tv_max, float nu_t_max, float spacing_ratio) {
  if (L.fixed_dt > 0.0f) return L.fixed_dt;
  const float h = P.kernel.h;
  float c_max = P.c0;
  if (P.has_vapor) {
    const float tv = Tv_max > P.T_ref ? Tv_max : P.T_ref;
    const float cv = sqrtf(P.gas_R * tv);
    c_max = c_max < cv ? cv : c_max;
  }
  const float sr = 1.0f < spacing_ratio ? spacing_ratio : 1.0f;
  float dt = L.cfl * h / ((c_max + v_max) * sr);
  const float n0 = P.mu[0] / P.rho0[0], n1 = P.mu[1] / P.rho0[1];
  const float nu = (n0 < n1 ? n1 : n0) + nu_t_max;
  if (nu > 0.0f) {
    const float d = L.viscous_number * h * h / nu;
    dt = d < dt ? d : dt;
  }
  if (a_max > 0.0f) {
    const float d = L.force_number * sqrtf(h / a_max);
    dt = d < dt ? d : dt;
  }
  if (P.sigma > 0.0f) {
    const float rho_avg = 0.5f * (P.rho0[0] + (P.has_vapor ? P.rho0[1] : P.rho0[0]));
    const float d = L.capillary_number * sqrtf(rho_avg * h * h * h / (2.0f * kPi * P.sigma));
    dt = d < dt ? d : dt;
  }
  if (P.thermal) {
    float alpha = P.k[0] / (P.rho0[0] * P.cp[0]);
    if (P.has_vapor) {
      const float a1 = P.k[1] / (P.rho0[1] * P.cp[1]);
      alpha = alpha < a1 ? a1 : alpha;
    }
    if (alpha > 0.0f) {
      const float d = L.thermal_number * h * h / alpha;
      dt = d < dt ? d : dt;
    }
  }
  return dt < L.dt_min ? L.dt_min : (L.dt_max < dt ? L.dt_max : dt);
}

// --------------------------------------------------------------------- grid
// Uniform grid, cell = 2h (spatial_hash.hpp:14-56).  `off` is the exclusive
// prefix of particles per cell (num_cells + 1 entries): the particles of cell c
// are [off[c], off[c+1]).  Cells x0..x1 of one (y, z) row are consecutive keys,
// so a row is one contiguous index range, and the 9 rows in z, y order visit
// neighbours in the reference's order (27 cells z->y->x, index ascending).
struct GridDev {
  float ox, oy, oz, inv_cell;
  int nx, ny, nz, num_cells;
  const int* off;
  float cull_pad;  // cell units: rounding allowance of the cell-culled sweep
  // per cell: which kinds of particle it holds (kCellWall, kCellLiquid,
  // kCellVapor; built with the cell table each sort) and, during the phase
  // change pass, whether one of them gives mass to liquid / vapour receivers
  // (kCellSrcL, kCellSrcV)
  uint8_t* flags;
  // cell-tile shape of the LDS-staged passes: cells along x, y, z
  int tx = 3, ty = 3, tz = 3;
};
// Cell-tile shapes the tiled passes accept: at most 25 staged (z, y) rows,
// (ty + 2)(tz + 2), and at most 9 target rows, ty tz (the sizes of their row
// tables); at most 16 cells along x.
__host__ __device__ inline bool tile_shape_ok(int tx, int ty, int tz) {
  return tx >= 1 && ty >= 1 && tz >= 1 && tx <= 16 && (ty + 2) * (tz + 2) <= 25 && ty * tz <= 9;
}
constexpr int kCellWall = 1, kCellLiquid = 2, kCellVapor = 4, kCellFluid = 6, kCellSrcL = 8, kCellSrcV = 16;
__host__ __device__ inline void cell_coords(const GridDev& g, V3 p, int& cx, int& cy, int& cz) {
  cx = clampv(static_cast<int>(floorf((p.x - g.ox) * g.inv_cell)), 0, g.nx - 1);
  cy = clampv(static_cast<int>(floorf((p.y - g.oy) * g.inv_cell)), 0, g.ny - 1);
  cz = clampv(static_cast<int>(floorf((p.z - g.oz) * g.inv_cell)), 0, g.nz - 1);
}
__host__ __device__ inline int cell_index(const GridDev& g, V3 p) {
  int cx, cy, cz;
  cell_coords(g, p, cx, cy, cz);
  return (cz * g.ny + cy) * g.nx + cx;
}

// Visits every particle of the 27 cells around p, in the reference order.
template <class F>
__host__ __device__ inline void for_each_nbr(const GridDev& g, V3 p, F&& f) {
  int cx, cy, cz;
  cell_coords(g, p, cx, cy, cz);
  const int z0 = cz > 0 ? cz - 1 : 0, z1 = cz < g.nz - 1 ? cz + 1 : cz;
  const int y0 = cy > 0 ? cy - 1 : 0, y1 = cy < g.ny - 1 ? cy + 1 : cy;
  const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx < g.nx - 1 ? cx + 1 : cx;
  for (int z = z0; z <= z1; ++z) {
    for (int y = y0; y <= y1; ++y) {
      const int row = (z * g.ny + y) * g.nx;
      const int s = g.off[row + x0], e = g.off[row + x1 + 1];
      for (int j = s; j < e; ++j) f(j);
    }
  }
}

// Whether any of the 27 cells around p holds a particle of the given kind
// (a neighbour of that kind is possible only then).
__device__ inline bool cells_any(const GridDev& g, V3 p, int mask) {
  int cx, cy, cz;
  cell_coords(g, p, cx, cy, cz);
  const int z0 = cz > 0 ? cz - 1 : 0, z1 = cz < g.nz - 1 ? cz + 1 : cz;
  const int y0 = cy > 0 ? cy - 1 : 0, y1 = cy < g.ny - 1 ? cy + 1 : cy;
  const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx < g.nx - 1 ? cx + 1 : cx;
  int any = 0;
  for (int z = z0; z <= z1; ++z)
    for (int y = y0; y <= y1; ++y) {
      const int row = (z * g.ny + y) * g.nx;
      for (int x = x0; x <= x1; ++x) any |= g.flags[row + x];
    }
  return (any & mask) != 0;
}

// The same sweep without the cells that provably hold nothing within `radius`
// of p: a row (or the first / last cell of a row) is skipped when the gap
// between p and that cell, measured in cell units from the same
// (p - origin) * inv_cell that assigns cells, is at least radius plus a
// rounding allowance.  Particles clamped into the outer cells lie beyond those
// cells, so the gap stays a lower bound of their distance.  The visiting order
// of what remains is unchanged (ascending sorted index).
template <class F>
__host__ __device__ inline void for_each_cand(const GridDev& g, V3 p, float radius, F&& f) {
  const float ux = (p.x - g.ox) * g.inv_cell, uy = (p.y - g.oy) * g.inv_cell, uz = (p.z - g.oz) * g.inv_cell;
  const int cx = clampv(static_cast<int>(floorf(ux)), 0, g.nx - 1);
  const int cy = clampv(static_cast<int>(floorf(uy)), 0, g.ny - 1);
  const int cz = clampv(static_cast<int>(floorf(uz)), 0, g.nz - 1);
  const float fx = ux - static_cast<float>(cx), fy = uy - static_cast<float>(cy), fz = uz - static_cast<float>(cz);
  const float rc = radius * g.inv_cell + g.cull_pad;
  const float rr = rc * rc;
  const float gxl = fx * fx, gxr = (1.0f - fx) * (1.0f - fx);
  const int z0 = cz > 0 ? cz - 1 : 0, z1 = cz < g.nz - 1 ? cz + 1 : cz;
  const int y0 = cy > 0 ? cy - 1 : 0, y1 = cy < g.ny - 1 ? cy + 1 : cy;
  for (int z = z0; z <= z1; ++z) {
    const float gz = z < cz ? fz : (z > cz ? 1.0f - fz : 0.0f);
    for (int y = y0; y <= y1; ++y) {
      const float gy = y < cy ? fy : (y > cy ? 1.0f - fy : 0.0f);
      const float gyz = gy * gy + gz * gz;
      if (gyz >= rr) continue;
      const int x0 = (cx > 0 && gyz + gxl < rr) ? cx - 1 : cx;
      const int x1 = (cx < g.nx - 1 && gyz + gxr < rr) ? cx + 1 : cx;
      const int row = (z * g.ny + y) * g.nx;
      const int s = g.off[row + x0], e = g.off[row + x1 + 1];
      for (int j = s; j < e; ++j) f(j);
    }
  }
}

// for_each_cand with the candidates' positions (and, WP, phases) gathered
// eight at a time (one wait per eight loads instead of one per candidate):
// f(j, x[j]) or f(j, x[j], phase[j]).
template <bool WP, class F>
__device__ inline void for_each_cand_impl(const GridDev& g, V3 p, float radius, const float4* __restrict__ xs,
                                          const uint8_t* __restrict__ phs, F&& f) {
  const float ux = (p.x - g.ox) * g.inv_cell, uy = (p.y - g.oy) * g.inv_cell, uz = (p.z - g.oz) * g.inv_cell;
  const int cx = clampv(static_cast<int>(floorf(ux)), 0, g.nx - 1);
  const int cy = clampv(static_cast<int>(floorf(uy)), 0, g.ny - 1);
  const int cz = clampv(static_cast<int>(floorf(uz)), 0, g.nz - 1);
  const float fx = ux - static_cast<float>(cx), fy = uy - static_cast<float>(cy), fz = uz - static_cast<float>(cz);
  const float rc = radius * g.inv_cell + g.cull_pad;
  const float rr = rc * rc;
  const float gxl = fx * fx, gxr = (1.0f - fx) * (1.0f - fx);
  const int z0 = cz > 0 ? cz - 1 : 0, z1 = cz < g.nz - 1 ? cz + 1 : cz;
  const int y0 = cy > 0 ? cy - 1 : 0, y1 = cy < g.ny - 1 ? cy + 1 : cy;
  for (int z = z0; z <= z1; ++z) {
    const float gz = z < cz ? fz : (z > cz ? 1.0f - fz : 0.0f);
    for (int y = y0; y <= y1; ++y) {
      const float gy = y < cy ? fy : (y > cy ? 1.0f - fy : 0.0f);
      const float gyz = gy * gy + gz * gz;
      if (gyz >= rr) continue;
      const int x0 = (cx > 0 && gyz + gxl < rr) ? cx - 1 : cx;
      const int x1 = (cx < g.nx - 1 && gyz + gxr < rr) ? cx + 1 : cx;
      const int row = (z * g.ny + y) * g.nx;
      const int s = g.off[row + x0], e = g.off[row + x1 + 1];
      int j = s;
      for (; j + 8 <= e; j += 8) {
        float4 b[8];
        uint8_t q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          b[u] = xs[j + u];
          if constexpr (WP) q[u] = phs[j + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if constexpr (WP) f(j + u, b[u], q[u]);
          else f(j + u, b[u]);
        }
      }
      for (; j < e; ++j) {
        if constexpr (WP) f(j, xs[j], phs[j]);
        else f(j, xs[j]);
      }
    }
  }
}
template <class F>
__device__ inline void for_each_cand_x(const GridDev& g, V3 p, float radius, const float4* __restrict__ xs, F&& f) {
  for_each_cand_impl<false>(g, p, radius, xs, nullptr, f);
}
template <class F>
__device__ inline void for_each_cand_xph(const GridDev& g, V3 p, float radius, const float4* __restrict__ xs,
                                         const uint8_t* __restrict__ phs, F&& f) {
  for_each_cand_impl<true>(g, p, radius, xs, phs, f);
}

// ---------------------------------------------------------- neighbour lists
// Built once per step by the density pass from the same 27-cell sweep, in
// the same ascending order, holding every j with r^2 < (2h)^2 (fluid targets:
// all particles, self included; wall targets: fluid particles only).
// Layout: 16-byte chunks of four entries, interleaved over groups of 64
// particles: chunk q of particle i sits at int4 index
// ((i / 64) * (K / 4) + q) * 64 + i % 64, so when the 64 lanes of a wave
// read (or write) chunk q of 64 neighbouring particles they touch one
// contiguous kilobyte, and a pair loop waits for memory once per four
// neighbours.  K is a multiple of 4; the buffer holds capacity rounded up to
// a multiple of 64 particles.  count[i] = -1 marks an overflow (more than K
// neighbours): that particle's passes fall back to the cell sweep, which
// visits the same j in the same order.
struct NbrList {
  int* idx;
  int* count;
  int stride;  // unused (the layout above needs only K)
  int K;
  int valid;  // 0: positions or cells changed since the lists were built
};
constexpr int kListGroup = 64;  // particles whose chunks interleave
// Chunk 0 of particle i; chunk q is at [q * kListGroup].
__host__ __device__ inline int4* list_chunks(const NbrList& L, int i) {
  const long long g = i / kListGroup;
  return reinterpret_cast<int4*>(L.idx) + (g * (L.K / 4)) * kListGroup + (i % kListGroup);
}

// Appends one entry of particle i's list, staging four in registers.
struct ListWriter {
  int4* out;
  int4 buf = make_int4(0, 0, 0, 0);
  int c = 0;
  int K;
  __device__ ListWriter(const NbrList& L, int i) : out(list_chunks(L, i)), K(L.K) {}
  // The chunk is a shift register: the newest entry enters at w, so after
  // four pushes x..w hold them in order (moves, no compares or selects).
  __device__ inline void push(int j) {
    buf.x = buf.y;
    buf.y = buf.z;
    buf.z = buf.w;
    buf.w = j;
    if ((c & 3) == 3 && c < K) out[(c >> 2) * kListGroup] = buf;
    ++c;
  }
  // Writes the last partial chunk (its m entries moved to the front);
  // returns the count to store (-1: overflow).
  __device__ inline int finish() {
    const int m = c & 3;
    if (m != 0 && c < K) {
      int4 o = make_int4(buf.y, buf.z, buf.w, 0);
      if (m == 1) o = make_int4(buf.w, 0, 0, 0);
      if (m == 2) o = make_int4(buf.z, buf.w, 0, 0);
      out[(c >> 2) * kListGroup] = o;
    }
    return c <= K ? c : -1;
  }
};

// Visits particle i's c (>= 0) listed neighbours in order; the next chunk
// of four is in flight while the current one is processed.
template <class F>
__device__ inline void for_each_listed(const NbrList& L, int i, int c, F&& f) {
  const int4* p = list_chunks(L, i);
  const int nq = (c + 3) >> 2;
  int j0 = 0, j1 = 0, j2 = 0, j3 = 0;
  if (nq > 0) {
    const int4 t = p[0];
    j0 = t.x, j1 = t.y, j2 = t.z, j3 = t.w;
  }
  for (int q = 0; q < nq; ++q) {
    int n0 = j0, n1 = j1, n2 = j2, n3 = j3;
    if (q + 1 < nq) {
      const int4 t = p[(q + 1) * kListGroup];
      n0 = t.x, n1 = t.y, n2 = t.z, n3 = t.w;
    }
    const int m = c - (q << 2) < 4 ? c - (q << 2) : 4;
#pragma unroll 1
    for (int u = 0; u < m; ++u) {  // one copy of f; the entry picked by selects
      f(u == 0 ? j0 : (u == 1 ? j1 : (u == 2 ? j2 : j3)));
    }
    j0 = n0, j1 = n1, j2 = n2, j3 = n3;
  }
}

template <class F>
__device__ inline void for_each_pair(const NbrList& L, const GridDev& g, int i, V3 xi, F&& f) {
  const int c = L.valid ? L.count[i] : -1;
  if (c >= 0) {
    for_each_listed(L, i, c, f);
  } else {
    for_each_nbr(g, xi, f);
  }
}

// for_each_pair with the neighbours' data gathered four at a time:
// load(j) returns a struct of what the pair needs (all loads issued before
// any is used), f(j, data) consumes it, in the same order.  The sweep
// fallback (count < 0 or invalid lists) loads one at a time.
template <class Load, class F>
__device__ inline void for_each_pair4(const NbrList& L, const GridDev& g, int i, V3 xi, Load&& load, F&& f) {
  const int c = L.valid ? L.count[i] : -1;
  if (c < 0) {
    for_each_nbr(g, xi, [&](int j) { f(j, load(j)); });
    return;
  }
  const int4* p = list_chunks(L, i);
  const int nq = (c + 3) >> 2;
  for (int q = 0; q < nq; ++q) {
    const int4 t = p[q * kListGroup];
    const int m = c - (q << 2);
    const auto da = load(t.x);
    const auto db = load(m > 1 ? t.y : t.x);
    const auto dc = load(m > 2 ? t.z : t.x);
    const auto dd = load(m > 3 ? t.w : t.x);
    f(t.x, da);
    if (m > 1) f(t.y, db);
    if (m > 2) f(t.z, dc);
    if (m > 3) f(t.w, dd);
  }
}

// Holds the listed values in registers here: every load issued before this
// point completes here, together, instead of each at its first use.
__device__ inline void settle1(float v) { asm volatile("" ::"v"(v)); }
__device__ inline void settle1(int v) { asm volatile("" ::"v"(v)); }
template <class... T>
__device__ inline void settle(T... v) {
  (settle1(v), ...);
}

// ------------------------------------------------------------- particle SoA
struct Persist {
  float4 *x, *v, *vt, *nrm;
  float *m, *m0, *rho, *T;
  uint8_t *phase, *flags;
  uint32_t* id;
  // Precision form only (null otherwise): the low-order part of each
  // position, x + xlo with |xlo| <= ulp(x) / 2, so that positions resolve
  // displacements far below float spacing (see pos_diff).
  float4* xlo;
};

// ------------------------------------------------ double-float positions
// Knuth's TwoSum of a and -b: s = fl(a - b) and the exact error (a - b) - s
// (round to nearest, no fused operations).
__host__ __device__ inline float diff_err(float a, float b, float s) {
  const float nb = -b;
  const float bv = s - a;
  const float av = s - bv;
  return (a - av) + (nb - bv);
}
// One coordinate of (a + al) - (b + bl), rounded once: the float parts'
// difference, its exact error and the low parts' difference.
__host__ __device__ inline float df_diff(float a, float al, float b, float bl) {
  const float s = a - b;
  return s + (diff_err(a, b, s) + (al - bl));
}
// Pair vector x_i - x_j.  HP = false: the float difference of the stored
// positions, the reference's arithmetic.  HP = true (the precision form):
// the double-float positions' difference, to which particle j's low part is
// fetched here; li is particle i's.  Exactly antisymmetric in (i, j), as the
// float difference is.
template <bool HP>
__device__ inline V3 pos_diff(V3 xi, V3 li, V3 xj, const float4* __restrict__ xlo, int j) {
  if constexpr (!HP) {
    return xi - xj;
  } else {
    const float4 lj = xlo[j];
    return V3{df_diff(xi.x, li.x, xj.x, lj.x), df_diff(xi.y, li.y, xj.y, lj.y), df_diff(xi.z, li.z, xj.z, lj.z)};
  }
}
template <bool HP>
__device__ inline V3 pos_lo(const float4* __restrict__ xlo, int i) {
  if constexpr (!HP) return V3{0.0f, 0.0f, 0.0f};
  else return ld3(xlo[i]);
}
// x + d into the double-float (x, lo): the float sum, its exact rounding
// error folded into the low part, then renormalised so |lo| <= ulp(x) / 2.
__host__ __device__ inline void df_add(float& x, float& lo, float d) {
  const float s = x + d;
  const float bv = s - x;
  const float err = (x - (s - bv)) + (d - bv);
  const float t = lo + err;
  const float hi = s + t;
  lo = t - (hi - s);
  x = hi;
}
struct Scratch {
  float4 *a, *ab, *ns, *vw, *fw, *mw;
  float *p, *drho, *dTdt, *sigma, *kappa, *nu_t, *alpha, *wall_w, *heat_w, *out_m, *out_h, *w_norm, *p_wv, *rmin;
  // Neighbour record of one rates evaluation, so a pair loop gathers one
  // float4 instead of four scalars: {p, m, rho, T} for fluid particles
  // (written by the density pass), {p, p_wv, rho_tvf, T} for walls (wall
  // pass; rho_tvf is the wall density liquid targets use for the transport
  // velocity).
  float4* nrec;
  // V = m / rho of each fluid particle for one rates evaluation (density
  // pass): the pair loops read it instead of dividing per pair.
  float* vol;
};
struct PV {  // kernel view
  Persist s;
  Scratch t;
  int capacity;
};

}  // namespace cryo
