// CPU oracle: a restatement of the reference CryoSPH per-step pipeline
// (Semedo2002/Microgravity-Cryogenic-Multi-Phase-SPH-Solver) in plain C++.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the CPU
// baseline leg of bench.py load this library, as the checker.  The product
// (libcryosph.so, the CryoSPH module) never links, loads or calls it.
//
// Built twice by oracle/Makefile: ORC_REAL=float (the FP32 oracle) and
// ORC_REAL=double (the FP64 oracle), with -ffp-contract=off and no -march so
// that every float operation rounds as in the reference's default CPU build.
// Every function cites the reference file:line it restates.  Operation order
// inside each expression follows the reference, because FP32 results are
// compared bit for bit against the reference build (oracle/_ref).
//
// Entry points mirror include/cryosph.h with the prefix orc_ and fields in the
// oracle's native real type.
//
// orc_adjoint_slosh (the reverse-mode adjoint, which the reference does not
// have) takes its per-pair derivative formulas from the library's
// csrc/adjoint_pair.hpp and drives them in its own serial scatter form; the
// tests pin it by central differences of this file's FP64 forward step.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "adjoint_pair.hpp"
#include "cryosph.h"

#ifndef ORC_REAL
#define ORC_REAL float
#endif

namespace orc {

using R = ORC_REAL;
using u8 = std::uint8_t;
using u32 = std::uint32_t;

// orc_round_positions (a diagnostic switch; off = the reference algorithm)
inline bool g_round_positions = false;

constexpr R kPi = R(3.14159265358979323846);  // cuda_utils.cuh:67
constexpr u8 L = 0, V = 1, W = 2, D = 3;       // particle_system.hpp:13-23
constexpr u8 F_HEATED = 1, F_SOURCE = 2, F_TARGET_VAPOR = 4, F_DYING = 8, F_SCRATCH = 2 | 4 | 8;

// ------------------------------------------------ vector maths (cuda_utils.cuh:85-113)
struct V3 {
  R x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, R s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, R s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3& operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline V3& operator-=(V3& a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
inline V3& operator*=(V3& a, R s) { a.x *= s; a.y *= s; a.z *= s; return a; }
inline R dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline R norm2(V3 a) { return dot(a, a); }
inline R norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalized(V3 a, R eps = R(1e-20)) {
  const R n = norm(a);
  return n > eps ? a / n : V3{0, 0, 0};
}
template <class T>
inline T clampv(T v, T lo, T hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline bool is_fluid(u8 ph) { return ph <= V; }
inline int pidx(u8 ph) { return ph == V ? 1 : 0; }

// ---------------------------------------------- reductions (cuda_utils.cuh:303-349)
// Sum: each OpenMP thread owns a fixed contiguous block; block sums are added
// in thread order.  The thread count therefore fixes the rounding.
template <class T, class Fn>
T reduce_sum(int n, Fn f, T init) {
  std::vector<T> part(static_cast<std::size_t>(omp_get_max_threads()), T(0));
#pragma omp parallel
  {
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
    const int lo = static_cast<int>(static_cast<long long>(n) * t / nt);
    const int hi = static_cast<int>(static_cast<long long>(n) * (t + 1) / nt);
    T acc = T(0);
    for (int i = lo; i < hi; ++i) acc += f(i);
    part[static_cast<std::size_t>(t)] = acc;
  }
  T acc = init;
  for (const T& v : part) acc += v;
  return acc;
}
template <class T, class Fn>
T reduce_max(int n, Fn f, T init) {  // order-free
  T acc = init;
  for (int i = 0; i < n; ++i) {
    const T v = f(i);
    acc = v > acc ? v : acc;
  }
  return acc;
}
template <class Fn>
void pfor(int n, Fn f) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) f(i);
}

// ------------------------------------------------- kernels (SPH_kernels.cuh:14-86)
struct Kern {
  int type = 1;  // 0 cubic spline, 1 Wendland C2
  R h = 1, inv_h = 1, support = 2, norm = 1;
  static Kern make(int t, R h) {  // SPH_kernels.cuh:14-31
    Kern k;
    k.type = t;
    k.h = h;
    k.inv_h = R(1) / h;
    k.support = R(2) * h;
    const R h3 = h * h * h;
    k.norm = t == 0 ? R(1) / (kPi * h3) : R(21) / (R(16) * kPi * h3);
    return k;
  }
};
inline R kW(const Kern& k, R r) {  // SPH_kernels.cuh:34-45
  const R q = r * k.inv_h;
  if (q >= R(2)) return R(0);
  if (k.type == 0) {
    if (q < R(1)) return k.norm * (R(1) - R(1.5) * q * q + R(0.75) * q * q * q);
    const R t = R(2) - q;
    return k.norm * R(0.25) * t * t * t;
  }
  const R t = R(1) - R(0.5) * q;
  const R t2 = t * t;
  return k.norm * t2 * t2 * (R(2) * q + R(1));
}
inline R kdW(const Kern& k, R r) {  // SPH_kernels.cuh:48-58
  const R q = r * k.inv_h;
  if (q >= R(2)) return R(0);
  if (k.type == 0) {
    if (q < R(1)) return k.norm * k.inv_h * (R(-3) * q + R(2.25) * q * q);
    const R t = R(2) - q;
    return k.norm * k.inv_h * R(-0.75) * t * t;
  }
  const R t = R(1) - R(0.5) * q;
  return k.norm * k.inv_h * R(-5) * q * t * t * t;
}
inline V3 kgradW(const Kern& k, V3 rij, R r) {  // SPH_kernels.cuh:61-64
  if (r <= R(1e-12) * k.h) return V3{0, 0, 0};
  return rij * (kdW(k, r) / r);
}
inline R akinci_coh(R r, R c) {  // SPH_kernels.cuh:69-78
  if (r >= c || r <= R(0)) return R(0);
  const R c3 = c * c * c;
  const R c9 = c3 * c3 * c3;
  const R coeff = R(32) / (kPi * c9);
  const R t = (c - r);
  const R base = t * t * t * r * r * r;
  if (R(2) * r > c) return coeff * base;
  return coeff * (R(2) * base - c3 * c3 / R(64));
}
inline R akinci_adh(R r, R c) {  // SPH_kernels.cuh:81-86
  if (r >= c || R(2) * r <= c) return R(0);
  const R arg = R(-4) * r * r / c + R(6) * r - R(2) * c;
  if (arg <= R(0)) return R(0);
  return R(0.007) / std::pow(c, R(3.25)) * std::pow(arg, R(0.25));
}

// ------------------------------------------------ parameters (sim_params.cuh:20-95)
struct Params {
  Kern kernel;
  R dx = 0, vol0 = 0, support2 = 0;
  R rho0[2] = {1, 1}, mu[2] = {0, 0}, k[2] = {0, 0}, cp[2] = {1, 1}, mref[2] = {1, 1};
  R c0 = 1, tait_B = 1, tait_gamma = 7, p_bg = 0, gas_R = 1, vapor_K = 1;
  int clamp_neg = 0;
  V3 gravity{0, 0, 0};
  int frame_rot = 0;
  V3 omega{0, 0, 0}, omega_dot{0, 0, 0}, frame_origin{0, 0, 0};
  R alpha_av = 0, alpha_av_vap = 0, delta_sph = R(0.1f);
  int density_mode = 0, no_slip = 1, sps = 0, tvf = 1, tvf_stress = 0;
  R tvf_factor = R(2.0f), tvf_sigma_min = 0, shift_strength = R(0.25f);
  int tvf_interface_only = 1;
  R smag_cs = R(0.12f);
  int st_model = 1;
  R sigma = 0, cos_theta = 0, sin_theta = 1, normal_eps = 0, wall_adh_w = R(0.01f);
  int wetting = 0, wall_colour = 1, csf_smooth = 0, csf_balanced = 0;
  R ak_gamma = 0, ak_beta = 0;
  int thermal = 1, phase_change = 1, wall_bc = 0;
  R T_wall = 0, T_ref = 0, T_sat = 0, h_fg = 0, lee_evap = 0, lee_cond = 0, max_transfer = R(0.25f),
    vapor_rebalance = R(1.0f), pc_density_limit = R(0.01f), merge_ratio = R(0.3f), split_ratio = R(2.0f),
    heat_power = 0, heat_inv_norm = 0, dt = 0;
};

// sim_params.cuh:100-135
inline R eos_pressure(const Params& P, u8 ph, R rho, R T) {
  R p;
  if (ph == V) {
    p = rho * P.gas_R * T + (P.p_bg - P.vapor_K);
  } else {
    const R ratio = rho / P.rho0[0];
    R rg;
    if (P.tait_gamma == R(7)) {
      const R r2 = ratio * ratio;
      rg = r2 * r2 * r2 * ratio;
    } else {
      rg = std::pow(ratio, P.tait_gamma);
    }
    p = P.tait_B * (rg - R(1)) + P.p_bg;
  }
  return (P.clamp_neg && p < R(0)) ? R(0) : p;
}
inline R sound_speed(const Params& P, u8 ph, R T) {
  if (ph == V) return std::sqrt(P.gas_R * (T > R(0) ? T : R(0)));
  return P.c0;
}
inline R enthalpy(const Params& P, u8 ph, R T) {
  return ph == V ? P.h_fg + P.cp[1] * (T - P.T_ref) : P.cp[0] * (T - P.T_ref);
}
inline R temp_from_h(const Params& P, u8 ph, R h) {
  return ph == V ? P.T_ref + (h - P.h_fg) / P.cp[1] : P.T_ref + h / P.cp[0];
}

// ------------------------------------------------------------------ grid
// spatial_hash.hpp:14-56, spatial_hash.cu:15-46
struct Grid {
  V3 origin{0, 0, 0};
  R cell = 1;
  int nx = 1, ny = 1, nz = 1, num_cells = 1;
  std::vector<int> cs, ce;
  std::vector<u32> keys, perm;
  R inv_cell() const { return R(1) / cell; }
  void configure(V3 lo, V3 hi, R cs_) {
    if (!(cs_ > 0)) throw std::invalid_argument("SpatialHash: cell size must be positive");
    cell = cs_;
    origin = lo - V3{cell, cell, cell};
    const V3 ext = hi - lo;
    nx = std::max(1, static_cast<int>(std::ceil(ext.x / cell)) + 2);
    ny = std::max(1, static_cast<int>(std::ceil(ext.y / cell)) + 2);
    nz = std::max(1, static_cast<int>(std::ceil(ext.z / cell)) + 2);
    const long long cells = static_cast<long long>(nx) * ny * nz;
    if (cells > 400000000LL) throw std::runtime_error("SpatialHash: grid too large; increase the cell size");
    num_cells = static_cast<int>(cells);
    cs.assign(static_cast<std::size_t>(num_cells) + 1, -1);
    ce.assign(static_cast<std::size_t>(num_cells) + 1, -1);
  }
  void coords(V3 p, int& cx, int& cy, int& cz) const {
    const R ic = inv_cell();
    cx = clampv(static_cast<int>(std::floor((p.x - origin.x) * ic)), 0, nx - 1);
    cy = clampv(static_cast<int>(std::floor((p.y - origin.y) * ic)), 0, ny - 1);
    cz = clampv(static_cast<int>(std::floor((p.z - origin.z) * ic)), 0, nz - 1);
  }
  int index(V3 p) const {
    int cx, cy, cz;
    coords(p, cx, cy, cz);
    return (cz * ny + cy) * nx + cx;
  }
  // Enumeration contract: 27 cells z->y->x, index ascending in each cell.
  template <class Fn>
  void each(V3 p, Fn&& f) const {
    int cx, cy, cz;
    coords(p, cx, cy, cz);
    const int z0 = cz > 0 ? cz - 1 : 0, z1 = cz < nz - 1 ? cz + 1 : cz;
    const int y0 = cy > 0 ? cy - 1 : 0, y1 = cy < ny - 1 ? cy + 1 : cy;
    const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx < nx - 1 ? cx + 1 : cx;
    for (int z = z0; z <= z1; ++z)
      for (int y = y0; y <= y1; ++y) {
        const int row = (z * ny + y) * nx;
        for (int x = x0; x <= x1; ++x) {
          const int c = row + x;
          const int s = cs[c];
          if (s < 0) continue;
          const int e = ce[c];
          for (int j = s; j < e; ++j) f(j);
        }
      }
  }
};

// ----------------------------------------------------------- particle state
// particle_system.hpp:37-142, particle_system.cpp:62-87
struct Particles {
  int n = 0, capacity = 0;
  u32 next_id = 0;
  std::vector<V3> x, v, vt, nrm, a, ab, ns, vw, fw, mw;
  std::vector<R> m, m0, rho, T, p, drho, dTdt, sigma, kappa, nu_t, alpha, wall_w, heat_w, out_m, out_h, w_norm, p_wv,
      rmin;
  std::vector<u8> phase, flags;
  std::vector<u32> id;
  template <class Fn>
  void each_array(Fn&& f) {
    f(x); f(v); f(vt); f(nrm); f(a); f(ab); f(ns); f(vw); f(fw); f(mw);
    f(m); f(m0); f(rho); f(T); f(p); f(drho); f(dTdt); f(sigma); f(kappa); f(nu_t); f(alpha); f(wall_w);
    f(heat_w); f(out_m); f(out_h); f(w_norm); f(p_wv); f(rmin);
    f(phase); f(flags); f(id);
  }
  void allocate(std::size_t cap) {
    each_array([cap](auto& arr) { arr.assign(cap, {}); });
    capacity = static_cast<int>(cap);
    n = 0;
  }
  void grow(std::size_t cap) {  // particle_system.cpp:62-67 (contents kept)
    if (static_cast<int>(cap) <= capacity) return;
    each_array([cap](auto& arr) { arr.resize(cap); });
    capacity = static_cast<int>(cap);
  }
};

// Gather new[i] = old[perm[i]] of the 11 persistent arrays (spatial_hash.cu:117-138).
template <class T>
void gather(std::vector<T>& arr, const std::vector<u32>& perm, int n) {
  std::vector<T> out(arr.size());
  for (int i = 0; i < n; ++i) out[i] = arr[perm[i]];
  arr.swap(out);
}

// SpatialHash::build, CPU branch (spatial_hash.cu:48-114): keys, serial stable
// counting sort, cell ranges, gather, dead particles dropped.
int grid_build(Grid& g, Particles& ps) {
  const int n = ps.n;
  if (static_cast<int>(g.keys.size()) < ps.capacity) {
    g.keys.resize(ps.capacity);
    g.perm.resize(ps.capacity);
  }
  const int sentinel = g.num_cells;
  pfor(n, [&](int i) {
    g.keys[i] = ps.phase[i] == D ? static_cast<u32>(sentinel) : static_cast<u32>(g.index(ps.x[i]));
    g.perm[i] = static_cast<u32>(i);
  });
  std::vector<int> counts(static_cast<std::size_t>(g.num_cells) + 2, 0);
  for (int i = 0; i < n; ++i) ++counts[g.keys[i] + 1];
  for (int c = 0; c <= g.num_cells; ++c) counts[c + 1] += counts[c];
  for (int c = 0; c < g.num_cells; ++c) {
    const int s = counts[c], e = counts[c + 1];
    g.cs[c] = e > s ? s : -1;
    g.ce[c] = e > s ? e : -1;
  }
  const int n_alive = counts[g.num_cells];
  std::vector<u32> sorted(static_cast<std::size_t>(n));
  for (int i = 0; i < n; ++i) sorted[static_cast<std::size_t>(counts[g.keys[i]]++)] = static_cast<u32>(i);
  std::copy(sorted.begin(), sorted.end(), g.perm.begin());
  gather(ps.x, g.perm, n_alive);
  gather(ps.v, g.perm, n_alive);
  gather(ps.vt, g.perm, n_alive);
  gather(ps.nrm, g.perm, n_alive);
  gather(ps.m, g.perm, n_alive);
  gather(ps.m0, g.perm, n_alive);
  gather(ps.rho, g.perm, n_alive);
  gather(ps.T, g.perm, n_alive);
  gather(ps.phase, g.perm, n_alive);
  gather(ps.flags, g.perm, n_alive);
  gather(ps.id, g.perm, n_alive);
  ps.n = n_alive;
  return n - n_alive;
}

// ------------------------------------------------------------------ passes
// SPH_kernels.cu:6-34
void density_summation(Particles& ps, const Grid& g, const Params& P) {
  pfor(ps.n, [&](int i) {
    const u8 ph = ps.phase[i];
    if (!is_fluid(ph)) {
      ps.sigma[i] = R(1) / P.vol0;
      ps.rmin[i] = P.kernel.support;
      return;
    }
    const V3 xi = ps.x[i];
    R sum = 0, r2min = P.support2;
    g.each(xi, [&](int j) {
      const R r2 = norm2(xi - ps.x[j]);
      if (r2 >= P.support2) return;
      sum += kW(P.kernel, std::sqrt(r2));
      if (j != i && r2 < r2min) r2min = r2;
    });
    ps.sigma[i] = sum;
    ps.rmin[i] = std::sqrt(r2min);
    if (P.density_mode == 0) ps.rho[i] = (ph == L ? ps.m0[i] : ps.m[i]) * sum;
  });
}

// SPH_kernels.cu:36-54
void normalize_initial_masses(Particles& ps, const Grid& g, const Params& P) {
  pfor(ps.n, [&](int i) {
    if (!is_fluid(ps.phase[i])) return;
    const V3 xi = ps.x[i];
    R sum = 0;
    g.each(xi, [&](int j) {
      const R r2 = norm2(xi - ps.x[j]);
      if (r2 < P.support2) sum += kW(P.kernel, std::sqrt(r2));
    });
    if (sum > R(0)) {
      ps.m[i] = ps.rho[i] / sum;
      ps.m0[i] = ps.m[i];
    }
  });
}

// navier_stokes.cu:5-14
void equation_of_state(Particles& ps, const Params& P) {
  pfor(ps.n, [&](int i) {
    const u8 ph = ps.phase[i];
    if (is_fluid(ph)) ps.p[i] = eos_pressure(P, ph, ps.rho[i], ps.T[i]);
  });
}

// navier_stokes.cu:16-69
void wall_boundary(Particles& ps, const Grid& g, const Params& P) {
  pfor(ps.n, [&](int i) {
    if (ps.phase[i] != W) return;
    const V3 xi = ps.x[i];
    R sw[2] = {0, 0}, sp[2] = {0, 0}, sgr[2] = {0, 0};
    R srho = 0;
    V3 sv{0, 0, 0};
    g.each(xi, [&](int j) {
      const u8 phj = ps.phase[j];
      if (!is_fluid(phj)) return;
      const V3 rwf = xi - ps.x[j];
      const R r2 = norm2(rwf);
      if (r2 >= P.support2) return;
      const R w = kW(P.kernel, std::sqrt(r2));
      const R rhoj = ps.rho[j];
      const int k = pidx(phj);
      sw[k] += w;
      sp[k] += ps.p[j] * w;
      sgr[k] += dot(P.gravity, rwf) * rhoj * w;
      srho += rhoj * w;
      sv += ps.v[j] * w;
    });
    R pw[2];
    for (int k = 0; k < 2; ++k) {
      pw[k] = sw[k] > R(0) ? (sp[k] + sgr[k]) / sw[k] : P.p_bg;
      if (P.clamp_neg && pw[k] < R(0)) pw[k] = 0;
    }
    if (!(sw[0] > R(0)) && sw[1] > R(0)) pw[0] = pw[1];
    if (!(sw[1] > R(0)) && sw[0] > R(0)) pw[1] = pw[0];
    ps.p[i] = pw[0];
    ps.p_wv[i] = pw[1];
    ps.alpha[i] = (sw[0] + sw[1]) > R(0) ? sw[0] / (sw[0] + sw[1]) : R(0);
    const R swt = sw[0] + sw[1];
    if (swt > R(0)) {
      ps.rho[i] = srho / swt;
      ps.vw[i] = sv * (R(-1) / swt);
    } else {
      ps.rho[i] = P.rho0[0];
      ps.vw[i] = V3{0, 0, 0};
    }
    ps.a[i] = V3{0, 0, 0};
  });
}

// surface_tension.cu:5-85
void interface_fields(Particles& ps, const Grid& g, const Params& P) {
  pfor(ps.n, [&](int i) {
    const u8 ph = ps.phase[i];
    if (!is_fluid(ph)) return;
    const V3 xi = ps.x[i];
    const R ci = ph == L ? R(1) : R(0);
    V3 gradc{0, 0, 0}, wall_n{0, 0, 0}, akn{0, 0, 0}, gradphi{0, 0, 0};
    R s_all = 0, s_other = 0, ww = 0, hw = 0;
    g.each(xi, [&](int j) {
      const V3 rij = xi - ps.x[j];
      const R r2 = norm2(rij);
      if (r2 >= P.support2) return;
      const R r = std::sqrt(r2);
      const R w = kW(P.kernel, r);
      const u8 phj = ps.phase[j];
      if (phj == W) {
        const R vw = P.vol0 * w;
        ww += vw;
        if (ps.flags[j] & F_HEATED) hw += vw;
        wall_n += ps.nrm[j] * vw;
        gradphi += kgradW(P.kernel, rij, r) * P.vol0;
        if (P.wall_colour) gradc += kgradW(P.kernel, rij, r) * (P.vol0 * (ps.alpha[j] - ci));
        return;
      }
      const R Vj = ps.m[j] / ps.rho[j];
      s_all += Vj * w;
      if (phj != ph) s_other += Vj * w;
      if (j == i) return;
      const V3 gw = kgradW(P.kernel, rij, r);
      const R cj = phj == L ? R(1) : R(0);
      gradc += gw * (Vj * (cj - ci));
      if (phj == L) akn += gw * Vj;
    });
    ps.alpha[i] = s_all > R(0) ? s_other / s_all : R(0);
    ps.wall_w[i] = ww;
    ps.heat_w[i] = hw;
    if (P.st_model == 2) {
      ps.nrm[i] = ph == L ? akn * (-P.kernel.h) : V3{0, 0, 0};
      return;
    }
    const R gmag = norm(gradc);
    ps.vw[i] = V3{0, 0, 0};
    if (P.wetting != 0 && ww > R(0) && gmag > P.normal_eps && norm2(wall_n) > R(0)) {
      const V3 nw = normalized(wall_n);
      const V3 nh = gradc / gmag;
      const V3 tan_in = nh - nw * dot(nh, nw);
      const R tm = norm(tan_in);
      if (tm > R(1e-6)) {
        const R cos_d = -dot(nh, nw);
        const R delta_w = R(2) * norm(gradphi);
        const R coeff = R(2) * P.sigma / (P.rho0[0] + P.rho0[1]);
        ps.vw[i] = tan_in * (-(P.cos_theta - cos_d) * coeff * gmag * delta_w / tm);
      }
    }
    if (P.wetting != 1 && ww > R(0) && gmag > P.normal_eps) {
      const V3 nw = normalized(wall_n);
      const V3 nh = gradc / gmag;
      V3 t = nh - nw * dot(nh, nw);
      const R tmag = norm(t);
      if (norm2(nw) > R(0) && tmag > R(1e-6)) {
        t = t / tmag;
        const V3 target = nw * (-P.cos_theta) + t * P.sin_theta;
        const R beta = clampv(ww / P.wall_adh_w, R(0), R(1));
        const V3 blended = normalized(nh * (R(1) - beta) + target * beta);
        gradc = blended * gmag;
      }
    }
    ps.nrm[i] = gradc;
  });
}

// SPH_kernels.cu:56-104
void transport_acceleration(Particles& ps, const Grid& g, const Params& P) {
  pfor(ps.n, [&](int i) {
    const u8 ph = ps.phase[i];
    if (!is_fluid(ph) || !P.tvf || ps.sigma[i] < P.tvf_sigma_min ||
        (P.tvf_interface_only && !(ps.alpha[i] > R(0)))) {
      ps.ab[i] = V3{0, 0, 0};
      return;
    }
    const V3 xi = ps.x[i];
    const R rhoi = ps.rho[i];
    const R Vi = R(1) / ps.sigma[i];
    const R Vi2 = Vi * Vi;
    const R mi = rhoi * Vi;
    const R ci = sound_speed(P, ph, ps.T[i]);
    const R pb = P.tvf_factor * P.rho0[pidx(ph)] * ci * ci;
    const R inv_gamma = R(1) / P.tait_gamma;
    V3 sum{0, 0, 0};
    g.each(xi, [&](int j) {
      if (j == i) return;
      const V3 rij = xi - ps.x[j];
      const R r2 = norm2(rij);
      if (r2 >= P.support2) return;
      const u8 phj = ps.phase[j];
      R Vj;
      if (phj == W) {
        R rho_w;
        if (ph == V) {
          rho_w = (ps.p_wv[j] - (P.p_bg - P.vapor_K)) / (P.gas_R * ps.T[i]);
        } else {
          const R ratio = (ps.p[j] - P.p_bg) / P.tait_B + R(1);
          rho_w = P.rho0[0] * std::pow(ratio > R(0.5) ? ratio : R(0.5), inv_gamma);
        }
        Vj = rho_w > R(0) ? Vi * rhoi / rho_w : Vi;
      } else {
        Vj = R(1) / ps.sigma[j];
      }
      sum += kgradW(P.kernel, rij, std::sqrt(r2)) * (Vi2 + Vj * Vj);
    });
    ps.ab[i] = sum * (-pb / mi);
  });
}

// SPH_kernels.cu:106-160
void particle_shifting(Particles& ps, const Grid& g, const Params& P, R mach) {
  const R h = P.kernel.h;
  const R w_dx = kW(P.kernel, P.dx);
  const R coeff = P.shift_strength * mach * R(4) * h * h;
  const R cap = R(0.1) * P.dx;
  pfor(ps.n, [&](int i) {
    ps.ns[i] = V3{0, 0, 0};
    const u8 ph = ps.phase[i];
    if (!is_fluid(ph) || ps.sigma[i] < P.tvf_sigma_min) return;
    const V3 xi = ps.x[i];
    V3 sum{0, 0, 0};
    g.each(xi, [&](int j) {
      if (j == i) return;
      const u8 phj = ps.phase[j];
      if (phj == D) return;
      const V3 rij = xi - ps.x[j];
      const R r2 = norm2(rij);
      if (r2 >= P.support2) return;
      const R r = std::sqrt(r2);
      const R Vj = phj == W ? R(1) / ps.sigma[i] : R(1) / ps.sigma[j];
      const R q = kW(P.kernel, r) / w_dx;
      const R q2 = q * q;
      sum += kgradW(P.kernel, rij, r) * (Vj * (R(1) + R(0.2) * q2 * q2));
    });
    V3 dr = sum * (-coeff);
    if (ps.alpha[i] > R(0)) {
      const R gm = norm(ps.nrm[i]);
      if (gm > P.normal_eps) {
        const V3 nn = ps.nrm[i] / gm;
        dr -= nn * dot(dr, nn);
      } else {
        dr = V3{0, 0, 0};
      }
    }
    const R len = norm(dr);
    if (len > cap) dr *= cap / len;
    ps.ns[i] = dr;
  });
  pfor(ps.n, [&](int i) {
    if (is_fluid(ps.phase[i])) ps.x[i] += ps.ns[i];
  });
}

// SPH_kernels.cu:162-203
void density_rate(Particles& ps, const Grid& g, const Params& P) {
  pfor(ps.n, [&](int i) {
    const u8 ph = ps.phase[i];
    if (!is_fluid(ph)) {
      ps.drho[i] = 0;
      return;
    }
    const V3 xi = ps.x[i];
    const std::vector<V3>& vel = P.tvf ? ps.vt : ps.v;
    const V3 vi = vel[i];
    const R rhoi = ps.rho[i];
    const R h = P.kernel.h;
    const R eta2 = R(0.01) * h * h;
    R div = 0, diff = 0;
    g.each(xi, [&](int j) {
      if (j == i) return;
      const V3 rij = xi - ps.x[j];
      const R r2 = norm2(rij);
      if (r2 >= P.support2) return;
      const R r = std::sqrt(r2);
      const V3 gw = kgradW(P.kernel, rij, r);
      const u8 phj = ps.phase[j];
      if (phj == W) {
        div += P.vol0 * dot(vi, gw);
        return;
      }
      const R Vj = ps.m[j] / ps.rho[j];
      div += Vj * dot(vi - vel[j], gw);
      if (phj == ph) diff += Vj * R(2) * (ps.rho[j] - rhoi) * dot(-rij, gw) / (r2 + eta2);
    });
    ps.drho[i] = rhoi * div + P.delta_sph * h * P.c0 * diff;
  });
}

// navier_stokes.cu:71-119
void sps_viscosity(Particles& ps, const Grid& g, const Params& P) {
  pfor(ps.n, [&](int i) {
    const u8 ph = ps.phase[i];
    if (!is_fluid(ph)) {
      ps.nu_t[i] = 0;
      return;
    }
    const V3 xi = ps.x[i];
    const V3 vi = ps.v[i];
    R Lm[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    g.each(xi, [&](int j) {
      if (j == i) return;
      const V3 rij = xi - ps.x[j];
      const R r2 = norm2(rij);
      if (r2 >= P.support2) return;
      const V3 gw = kgradW(P.kernel, rij, std::sqrt(r2));
      const u8 phj = ps.phase[j];
      V3 vj;
      R Vj;
      if (phj == W) {
        if (!P.no_slip) return;
        vj = ps.vw[j];
        Vj = P.vol0;
      } else {
        vj = ps.v[j];
        Vj = ps.m[j] / ps.rho[j];
      }
      const V3 dv = (vj - vi) * Vj;
      const R dvc[3] = {dv.x, dv.y, dv.z};
      const R gwc[3] = {gw.x, gw.y, gw.z};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Lm[a][b] += dvc[a] * gwc[b];
    });
    R ss = 0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        const R sab = R(0.5) * (Lm[a][b] + Lm[b][a]);
        ss += sab * sab;
      }
    const R smag = std::sqrt(R(2) * ss);
    const R l = P.smag_cs * P.dx;
    ps.nu_t[i] = l * l * smag;
  });
}

// navier_stokes.cu:121-245
void momentum(Particles& ps, const Grid& g, const Params& P) {
  pfor(ps.n, [&](int i) {
    const u8 ph = ps.phase[i];
    if (!is_fluid(ph)) {
      ps.a[i] = V3{0, 0, 0};
      ps.fw[i] = V3{0, 0, 0};
      ps.mw[i] = V3{0, 0, 0};
      return;
    }
    const int pi_idx = pidx(ph);
    const V3 xi = ps.x[i];
    const V3 vi = ps.v[i];
    const R mi = ps.m[i];
    const R rhoi = ps.rho[i];
    const R Vi = mi / rhoi;
    const R Vi2 = Vi * Vi;
    const R pi = ps.p[i];
    const R ci = sound_speed(P, ph, ps.T[i]);
    const R etai = P.mu[pi_idx] + (P.sps ? rhoi * ps.nu_t[i] : R(0));
    const R alpha_i = ph == V ? P.alpha_av_vap : P.alpha_av;
    const R h = P.kernel.h;
    const R eta2 = R(0.01) * h * h;
    V3 force{0, 0, 0}, acc_av{0, 0, 0};
    const R r_guard = R(0.6) * P.dx;
    const bool tvf_stress = P.tvf && P.tvf_stress;
    const V3 dvi = ps.vt[i] - vi;
    V3 fwall{0, 0, 0}, mwall{0, 0, 0};
    g.each(xi, [&](int j) {
      if (j == i) return;
      const V3 rij = xi - ps.x[j];
      const R r2 = norm2(rij);
      if (r2 >= P.support2) return;
      const R r = std::sqrt(r2);
      const R dW = kdW(P.kernel, r);
      const bool separated = r > R(1e-12) * h;
      const V3 gw = separated ? rij * (dW / r) : V3{0, 0, 0};
      const u8 phj = ps.phase[j];
      if (phj == W) {
        const V3 f_before = force, a_before = acc_av;
        const R Vj = P.vol0;
        const R vsum = Vi2 + Vj * Vj;
        const R pw = ph == V ? ps.p_wv[j] : ps.p[j];
        force -= gw * (vsum * R(0.5) * (pi + pw));
        if (tvf_stress) force += vi * (R(0.5) * vsum * rhoi * dot(dvi, gw));
        if (r < r_guard && separated) {
          const R q = R(1) - r / r_guard;
          const V3 nn = rij / r;
          acc_av += nn * (R(4) * ci * ci * q * q / r_guard);
          const R vn = dot(vi, nn);
          if (vn < R(0) && P.dt > R(0)) acc_av -= nn * (q * vn / P.dt);
        }
        if (P.no_slip) {
          const V3 vij = vi - ps.vw[j];
          if (separated) force += vij * (vsum * etai * dW / r);
          const R vr = dot(vij, rij);
          if (alpha_i > R(0) && vr < R(0)) {
            const R mu_ij = h * vr / (r2 + eta2);
            const R Pi = -alpha_i * ci * mu_ij / rhoi;
            acc_av -= gw * (rhoi * P.vol0 * Pi);
          }
        }
        const V3 dF = (force - f_before) + (acc_av - a_before) * mi;
        fwall += dF;
        mwall += cross(ps.x[j] - P.frame_origin, dF);
        return;
      }
      const R mj = ps.m[j];
      const R rhoj = ps.rho[j];
      const R Vj = mj / rhoj;
      const R vsum = Vi2 + Vj * Vj;
      const R pt = (rhoj * pi + rhoi * ps.p[j]) / (rhoi + rhoj);
      force -= gw * (vsum * pt);
      if (tvf_stress) {
        const V3 vj = ps.v[j];
        const V3 Ag = vi * (rhoi * dot(dvi, gw)) + vj * (rhoj * dot(ps.vt[j] - vj, gw));
        force += Ag * (R(0.5) * vsum);
      }
      const V3 vij = vi - ps.v[j];
      const R etaj = P.mu[pidx(phj)] + (P.sps ? rhoj * ps.nu_t[j] : R(0));
      const R eta_sum = etai + etaj;
      if (eta_sum > R(0) && separated) {
        const R eta_ij = R(2) * etai * etaj / eta_sum;
        force += vij * (vsum * eta_ij * dW / r);
      }
      if (alpha_i > R(0) && phj == ph) {
        const R vr = dot(vij, rij);
        if (vr < R(0)) {
          const R cj = sound_speed(P, phj, ps.T[j]);
          const R mu_ij = h * vr / (r2 + eta2);
          const R Pi = -alpha_i * R(0.5) * (ci + cj) * mu_ij / (R(0.5) * (rhoi + rhoj));
          acc_av -= gw * (mj * Pi);
        }
      }
    });
    V3 acc = P.gravity + force / mi + acc_av;
    if (P.frame_rot) {
      const V3 r = xi - P.frame_origin;
      acc -= cross(P.omega_dot, r) + cross(P.omega, cross(P.omega, r)) + cross(P.omega, vi) * R(2);
    }
    ps.a[i] = acc;
    ps.fw[i] = fwall;
    ps.mw[i] = mwall;
  });
}

// surface_tension.cu:87-118
void smooth_normals(Particles& ps, const Grid& g, const Params& P) {
  pfor(ps.n, [&](int i) {
    if (!is_fluid(ps.phase[i])) return;
    const V3 gi = ps.nrm[i];
    const R gmag = norm(gi);
    if (!(gmag > P.normal_eps)) {
      ps.ns[i] = V3{0, 0, 0};
      return;
    }
    if (!P.csf_smooth) {
      ps.ns[i] = gi / gmag;
      return;
    }
    const V3 xi = ps.x[i];
    V3 acc{0, 0, 0};
    g.each(xi, [&](int j) {
      if (!is_fluid(ps.phase[j])) return;
      const V3 gj = ps.nrm[j];
      const R gjm = norm(gj);
      if (!(gjm > P.normal_eps)) return;
      const R r2 = norm2(xi - ps.x[j]);
      if (r2 >= P.support2) return;
      acc += gj * ((ps.m[j] / ps.rho[j]) * kW(P.kernel, std::sqrt(r2)) / gjm);
    });
    ps.ns[i] = normalized(acc);
  });
}

// surface_tension.cu:120-177
void csf(Particles& ps, const Grid& g, const Params& P) {
  const R coeff = R(2) * P.sigma / (P.rho0[0] + P.rho0[1]);
  pfor(ps.n, [&](int i) {
    const u8 ph = ps.phase[i];
    if (!is_fluid(ph)) return;
    const V3 ni = ps.ns[i];
    if (!(norm2(ni) > R(0.25))) {
      ps.kappa[i] = 0;
      return;
    }
    const V3 xi = ps.x[i];
    const R rhoi = ps.rho[i];
    const R mi = ps.m[i];
    const R Vi = mi / rhoi;
    const R ci = ph == L ? R(1) : R(0);
    R num = 0, den = 0;
    V3 G{0, 0, 0};
    g.each(xi, [&](int j) {
      if (j == i) return;
      const u8 phj = ps.phase[j];
      if (!is_fluid(phj)) return;
      const V3 rij = xi - ps.x[j];
      const R r2 = norm2(rij);
      if (r2 >= P.support2) return;
      const R r = std::sqrt(r2);
      const R dW = kdW(P.kernel, r);
      const V3 gw = rij * (dW / r);
      const R rhoj = ps.rho[j];
      const R Vj = ps.m[j] / rhoj;
      const V3 nj = ps.ns[j];
      if (norm2(nj) > R(0.25)) {
        num += Vj * dot(ni - nj, gw);
        den += Vj * r * (-dW);
      }
      if (P.csf_balanced) {
        const R cj = phj == L ? R(1) : R(0);
        G += gw * ((Vi * Vi + Vj * Vj) * (rhoj * ci + rhoi * cj) / (rhoi + rhoj));
      }
    });
    const R kappa = den > R(0) ? R(3) * num / den : R(0);
    ps.kappa[i] = kappa;
    if (P.wetting != 0) ps.a[i] += ps.vw[i];
    if (P.csf_balanced) ps.a[i] += G * (P.sigma * kappa / mi);
    else ps.a[i] += ps.nrm[i] * (coeff * kappa);
  });
}

// surface_tension.cu:179-213
void akinci(Particles& ps, const Grid& g, const Params& P) {
  pfor(ps.n, [&](int i) {
    if (ps.phase[i] != L) return;
    const V3 xi = ps.x[i];
    const V3 ni = ps.nrm[i];
    const R mi = ps.m[i];
    const R rhoi = ps.rho[i];
    const R c = P.kernel.support;
    const R rho0 = P.rho0[0];
    const R psi = rho0 * P.vol0;
    V3 acc{0, 0, 0};
    g.each(xi, [&](int j) {
      if (j == i) return;
      const V3 rij = xi - ps.x[j];
      const R r2 = norm2(rij);
      if (r2 >= c * c || r2 <= R(0)) return;
      const R r = std::sqrt(r2);
      const V3 rhat = rij / r;
      const u8 phj = ps.phase[j];
      if (phj == L) {
        const R K = R(2) * rho0 / (rhoi + ps.rho[j]);
        const V3 f_coh = rhat * (-P.ak_gamma * mi * ps.m[j] * akinci_coh(r, c));
        const V3 f_curv = (ni - ps.nrm[j]) * (-P.ak_gamma * mi);
        acc += (f_coh + f_curv) * (K / mi);
      } else if (phj == W) {
        acc -= rhat * (P.ak_beta * psi * akinci_adh(r, c));
      }
    });
    ps.a[i] += acc;
  });
}

// phase_change.cu:31-75
void thermal_rates(Particles& ps, const Grid& g, const Params& P) {
  pfor(ps.n, [&](int i) {
    const u8 ph = ps.phase[i];
    if (!is_fluid(ph)) {
      ps.dTdt[i] = 0;
      return;
    }
    const int ii = pidx(ph);
    const V3 xi = ps.x[i];
    const R Ti = ps.T[i];
    const R mi = ps.m[i];
    const R Vi = mi / ps.rho[i];
    const R ki = P.k[ii];
    const R eta2 = R(0.01) * P.kernel.h * P.kernel.h;
    const bool iso = P.wall_bc == 1;
    R q = 0;
    g.each(xi, [&](int j) {
      if (j == i) return;
      const u8 phj = ps.phase[j];
      R Vj, kij;
      if (phj == W) {
        if (!iso) return;
        Vj = P.vol0;
        kij = R(2) * ki;
      } else {
        const R kj = P.k[pidx(phj)];
        Vj = ps.m[j] / ps.rho[j];
        kij = R(4) * ki * kj / (ki + kj);
      }
      const V3 rij = xi - ps.x[j];
      const R r2 = norm2(rij);
      if (r2 >= P.support2) return;
      const R r = std::sqrt(r2);
      const R rgw = r * kdW(P.kernel, r);
      q += Vi * Vj * kij * (Ti - ps.T[j]) * rgw / (r2 + eta2);
    });
    q += P.heat_power * ps.heat_w[i] * P.heat_inv_norm;
    ps.dTdt[i] = q / (mi * P.cp[ii]);
  });
}

// phase_change.cu:77-137
void pc_sources(Particles& ps, const Params& P) {
  pfor(ps.n, [&](int i) {
    const u8 ph = ps.phase[i];
    u8 fl = static_cast<u8>(ps.flags[i] & ~F_SCRATCH);
    ps.out_m[i] = 0;
    ps.out_h[i] = 0;
    ps.w_norm[i] = 0;
    if (!is_fluid(ph)) {
      ps.flags[i] = fl;
      return;
    }
    const int ii = pidx(ph);
    const R m = ps.m[i];
    const R T = ps.T[i];
    if (m < P.merge_ratio * P.mref[ii]) {
      ps.out_m[i] = m;
      ps.out_h[i] = enthalpy(P, ph, T);
      fl |= F_SOURCE | F_DYING;
      if (ph == V) fl |= F_TARGET_VAPOR;
      ps.flags[i] = fl;
      return;
    }
    const R alpha = ps.alpha[i];
    if (P.phase_change && alpha > R(0) && P.dt > R(0)) {
      if (ph == L && T > P.T_sat) {
        const R h_out = P.h_fg + P.cp[1] * (P.T_sat - P.T_ref);
        const R latent = h_out - P.cp[0] * (P.T_sat - P.T_ref);
        const R rate = P.lee_evap * m * alpha * (T - P.T_sat) / P.T_sat;
        R dm = rate * P.dt;
        const R dm_energy = m * P.cp[0] * (T - P.T_sat) / latent;
        dm = dm < dm_energy ? dm : dm_energy;
        const R dm_cap = P.max_transfer * m;
        dm = dm < dm_cap ? dm : dm_cap;
        if (dm > R(0)) {
          ps.out_m[i] = dm;
          ps.out_h[i] = h_out;
          fl |= F_SOURCE | F_TARGET_VAPOR;
        }
      } else if (ph == V && T < P.T_sat) {
        const R rate = P.lee_cond * m * alpha * (P.T_sat - T) / P.T_sat;
        R dm = rate * P.dt;
        R dm_cap = P.max_transfer * m;
        if (P.pc_density_limit > R(0) && P.pc_density_limit * m < dm_cap) dm_cap = P.pc_density_limit * m;
        dm = dm < dm_cap ? dm : dm_cap;
        if (dm > R(0)) {
          ps.out_m[i] = dm;
          ps.out_h[i] = enthalpy(P, V, T);
          fl |= F_SOURCE;
        }
      }
    }
    ps.flags[i] = fl;
  });
}

// phase_change.cu:139-172
void pc_normalize(Particles& ps, const Grid& g, const Params& P) {
  pfor(ps.n, [&](int i) {
    const u8 fl = ps.flags[i];
    if (!(fl & F_SOURCE)) return;
    const u8 target = (fl & F_TARGET_VAPOR) ? V : L;
    const V3 xi = ps.x[i];
    R wsum = 0, wr_max = 0;
    g.each(xi, [&](int j) {
      if (j == i || ps.phase[j] != target || (ps.flags[j] & F_DYING)) return;
      const R r2 = norm2(xi - ps.x[j]);
      if (r2 >= P.support2) return;
      const R w = kW(P.kernel, std::sqrt(r2));
      const R rhoj = ps.rho[j];
      wsum += (ps.m[j] / rhoj) * w;
      const R wr = w / rhoj;
      wr_max = wr > wr_max ? wr : wr_max;
    });
    ps.w_norm[i] = wsum;
    if (!(wsum > R(0))) {
      ps.out_m[i] = 0;
      return;
    }
    if (target == V && P.pc_density_limit > R(0) && !(fl & F_DYING) && wr_max > R(0)) {
      const R cap = P.pc_density_limit * wsum / wr_max;
      if (ps.out_m[i] > cap) ps.out_m[i] = cap;
    }
  });
}

// phase_change.cu:174-215
void pc_apply(Particles& ps, const Grid& g, const Params& P) {
  // Gather first (reads out_m/phase of neighbours), then write: the reference
  // body is per-particle race-free because out_m/out_h/w_norm/flags are only
  // read and each particle writes its own m/T/phase; phase[j] of a neighbour
  // is never read here.
  pfor(ps.n, [&](int i) {
    const u8 ph = ps.phase[i];
    if (!is_fluid(ph)) return;
    const u8 fl = ps.flags[i];
    const R mi = ps.m[i];
    const R Vi = mi / ps.rho[i];
    const V3 xi = ps.x[i];
    R in_m = 0, in_H = 0;
    if (!(fl & F_DYING)) {
      g.each(xi, [&](int j) {
        if (j == i) return;
        const R om = ps.out_m[j];
        if (!(om > R(0))) return;
        const u8 target = (ps.flags[j] & F_TARGET_VAPOR) ? V : L;
        if (target != ph) return;
        const R r2 = norm2(xi - ps.x[j]);
        if (r2 >= P.support2) return;
        const R share = om * Vi * kW(P.kernel, std::sqrt(r2)) / ps.w_norm[j];
        in_m += share;
        in_H += share * ps.out_h[j];
      });
    }
    const R om_i = ps.out_m[i];
    if ((fl & F_DYING) && om_i > R(0)) {
      ps.m[i] = 0;
      ps.phase[i] = D;
      return;
    }
    if (!(om_i > R(0)) && !(in_m > R(0))) return;
    const R m_new = mi - om_i + in_m;
    const R H_new = mi * enthalpy(P, ph, ps.T[i]) - om_i * ps.out_h[i] + in_H;
    ps.T[i] = temp_from_h(P, ph, H_new / m_new);
    ps.m[i] = m_new;
  });
}

// phase_change.cu:217-268
void rebalance_vapor(Particles& ps, const Grid& g, const Params& P) {
  const R beta = P.vapor_rebalance;
  // Pass 1 reads m/T/v of neighbours and writes only out_m/out_h/ns.
  pfor(ps.n, [&](int i) {
    ps.out_m[i] = 0;
    ps.out_h[i] = 0;
    ps.ns[i] = V3{0, 0, 0};
    if (ps.phase[i] != V) return;
    const V3 xi = ps.x[i];
    const R mi = ps.m[i];
    const R hi = enthalpy(P, V, ps.T[i]);
    const V3 vi = ps.v[i];
    R dm = 0, dH = 0;
    V3 dP{0, 0, 0};
    g.each(xi, [&](int j) {
      if (j == i || ps.phase[j] != V) return;
      const R r2 = norm2(xi - ps.x[j]);
      if (r2 >= P.support2) return;
      const R flow = beta * P.vol0 * kW(P.kernel, std::sqrt(r2)) * (ps.m[j] - mi);
      const bool from_j = flow > R(0);
      dm += flow;
      dH += flow * (from_j ? enthalpy(P, V, ps.T[j]) : hi);
      dP += (from_j ? ps.v[j] : vi) * flow;
    });
    ps.out_m[i] = dm;
    ps.out_h[i] = dH;
    ps.ns[i] = dP;
  });
  pfor(ps.n, [&](int i) {
    if (ps.phase[i] != V) return;
    const R dm = ps.out_m[i];
    if (dm == R(0)) return;
    const R m = ps.m[i];
    const R m_new = m + dm;
    const R H = m * enthalpy(P, V, ps.T[i]) + ps.out_h[i];
    ps.T[i] = temp_from_h(P, V, H / m_new);
    ps.v[i] = (ps.v[i] * m + ps.ns[i]) / m_new;
    ps.vt[i] = ps.v[i];
    ps.m[i] = m_new;
    ps.out_m[i] = 0;
    ps.out_h[i] = 0;
  });
}

// phase_change.cu:7-27
inline u32 hash_u32(u32 x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}
inline R hash_unit(u32 x) { return static_cast<R>(hash_u32(x) & 0x00FFFFFFu) / static_cast<R>(0x01000000u); }
inline V3 random_direction(u32 id, u32 seed) {
  const R u = R(2) * hash_unit(id * 2654435761u ^ seed) - R(1);
  const R phi = R(2) * kPi * hash_unit(hash_u32(id + 0x9e3779b9u) ^ (seed * 747796405u));
  const R s = std::sqrt(R(1) - u * u > R(0) ? R(1) - u * u : R(0));
  return V3{s * std::cos(phi), s * std::sin(phi), u};
}

// phase_change.cu:270-335 (flags, exclusive-scan slots, split)
void split_heavy(Particles& ps, const Params& P, const std::vector<unsigned>& flag,
                 const std::vector<unsigned>& offset, u32 id_base, u32 seed) {
  const int n = ps.n;
  pfor(n, [&](int i) {
    if (!flag[i]) return;
    const u8 ph = ps.phase[i];
    const R m = ps.m[i];
    const unsigned k = offset[i];
    const int dst = n + static_cast<int>(k);
    if (dst >= ps.capacity) return;
    const R vol = m / ps.rho[i];
    V3 e = random_direction(ps.id[i], seed);
    const V3 xi = ps.x[i];
    const V3 gc = ps.nrm[i];
    const R gm = norm(gc);
    if (gm > P.normal_eps) {
      const V3 nn = gc / gm;
      const V3 t = e - nn * dot(e, nn);
      e = normalized(t - nn * R(0.5));
      ps.x[dst] = xi + e * (R(0.5) * std::cbrt(vol));
    } else {
      const R delta = R(0.25) * std::cbrt(vol);
      ps.x[dst] = xi + e * delta;
      ps.x[i] = xi - e * delta;
    }
    ps.v[dst] = ps.v[i];
    ps.vt[dst] = ps.vt[i];
    ps.ab[dst] = ps.ab[i];
    ps.a[dst] = ps.a[i];
    ps.nrm[dst] = ps.nrm[i];
    ps.m[dst] = R(0.5) * m;
    ps.m0[i] *= R(0.5);
    ps.m0[dst] = ps.m0[i];
    ps.m[i] = R(0.5) * m;
    ps.rho[dst] = ps.rho[i];
    ps.T[dst] = ps.T[i];
    ps.phase[dst] = ph;
    ps.flags[dst] = static_cast<u8>(ps.flags[i] & ~F_SCRATCH);
    ps.id[dst] = id_base + k;
    ps.p[dst] = ps.p[i];
    ps.sigma[i] *= R(2);
    ps.sigma[dst] = ps.sigma[i];
    ps.dTdt[dst] = 0;
    ps.out_m[dst] = 0;
    ps.out_h[dst] = 0;
  });
}

// phase_change.cu:337-358
void venting(Particles& ps, const Params& P, R relax) {
  pfor(ps.n, [&](int i) {
    ps.out_m[i] = 0;
    ps.out_h[i] = 0;
    if (ps.phase[i] != V) return;
    const R m = ps.m[i];
    const R vol = m / ps.rho[i];
    const R target = P.rho0[1] * vol;
    if (!(m > target)) return;
    const R dm = relax * (m - target);
    ps.out_m[i] = dm;
    ps.out_h[i] = enthalpy(P, V, ps.T[i]);
    if (P.density_mode == 1) ps.rho[i] *= (m - dm) / m;
    ps.m[i] = m - dm;
  });
}

// ------------------------------------------------------------ configuration
struct Fluid {
  std::string name = "custom";
  R rho_liquid = 1000, rho_vapor = 1, mu_liquid = R(1e-3), mu_vapor = R(1e-5), k_liquid = R(0.6),
    k_vapor = R(0.025), cp_liquid = 4180, cp_vapor = 2000, sigma = R(0.072), h_fg = R(2.26e6),
    T_sat = R(373.15), p_sat = 101325, R_gas = R(461.5);
  // fluid_properties.cpp:7-14
  R saturation_temperature(R p) const {
    const double inv_T = 1.0 / T_sat - (double(R_gas) / h_fg) * std::log(double(p) / p_sat);
    return static_cast<R>(1.0 / inv_T);
  }
  R saturation_pressure(R T) const {
    return static_cast<R>(p_sat * std::exp(double(h_fg) / R_gas * (1.0 / T_sat - 1.0 / double(T))));
  }
};

struct Maneuver {  // sph_solver.hpp:96-105, sph_solver.cu:234-268
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
    const double dt = time[k + 1] - time[k];
    const R f = static_cast<R>((t - time[k]) / dt);
    a = acc[k] * (R(1) - f) + acc[k + 1] * f;
    if (rot) {
      w = omega[k] * (R(1) - f) + omega[k + 1] * f;
      wd = (omega[k + 1] - omega[k]) / static_cast<R>(dt);
    } else {
      w = V3{0, 0, 0};
    }
  }
};

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

// ------------------------------------------------------------------ solver
// sph_solver.cu:273-748
struct Solver {
  cryo_config cfg;
  Fluid fluid;
  Params P;
  Particles ps;
  Grid grid;
  std::vector<unsigned> split_flag, split_offset;
  bool initialized = false, has_vapor = false, has_maneuver = false;
  double time = 0, last_dt = 0;
  long long step_n = 0;
  R char_length = 0;
  R v_max = 0, a_max = 0, Tv_max = 0, nu_t_max = 0, spacing_ratio = 1;
  double heat_input = 0, evaporated = 0, condensed = 0, vented = 0, vented_h = 0, p_ullage = 0;
  Maneuver maneuver;
  V3 frame_acc{0, 0, 0};
  struct TapeEntry {
    Particles ps;
    std::vector<u32> perm;
    R T_sat = 0;  // the step's saturation temperature (a closed tank's changes every step)
  };
  TapeEntry* record = nullptr;
  double fluid_volume0 = 0;
  int splits = 0, merges = 0;

  V3 cv(cryo_vec3 a) const { return V3{R(a.x), R(a.y), R(a.z)}; }

  Solver(const cryo_config& c, const Fluid& f) : cfg(c), fluid(f) {  // sph_solver.cu:273-279
    if (!(R(cfg.dx) > 0)) throw std::invalid_argument("SolverConfig.dx must be > 0");
    if (R(cfg.contact_angle_deg) < 0 || R(cfg.contact_angle_deg) > 180)
      throw std::invalid_argument("contact angle must be within [0, 180] degrees");
    p_ullage = fluid.p_sat;
  }

  void init_params() {  // sph_solver.cu:292-405
    const cryo_config& c = cfg;
    const R dx = c.dx;
    const R h = R(c.h_factor) * dx;
    P.kernel = Kern::make(c.kernel, h);
    P.dx = dx;
    P.support2 = P.kernel.support * P.kernel.support;
    {
      double sigma0 = 0;
      const int r = static_cast<int>(std::ceil(P.kernel.support / dx));
      for (int k = -r; k <= r; ++k)
        for (int j = -r; j <= r; ++j)
          for (int i = -r; i <= r; ++i) sigma0 += kW(P.kernel, dx * std::sqrt(R(i * i + j * j + k * k)));
      P.vol0 = static_cast<R>(1.0 / sigma0);
    }
    P.rho0[0] = fluid.rho_liquid;
    P.rho0[1] = fluid.rho_vapor;
    P.mu[0] = fluid.mu_liquid;
    P.mu[1] = fluid.mu_vapor;
    P.k[0] = fluid.k_liquid * R(c.conductivity_scale);
    P.k[1] = fluid.k_vapor * R(c.conductivity_scale);
    P.cp[0] = fluid.cp_liquid;
    P.cp[1] = fluid.cp_vapor;
    P.mref[0] = P.rho0[0] * P.vol0;
    P.mref[1] = P.rho0[1] * P.vol0;
    const bool st_on = c.surface_tension != CRYO_ST_NONE;
    const R g = norm(cv(c.gravity));
    R vref = c.velocity_scale;
    if (!(vref > 0)) {
      vref = std::sqrt(g * char_length);
      if (st_on) vref = std::max(vref, std::sqrt(fluid.sigma / (fluid.rho_liquid * h)));
      vref = std::max(vref, R(1e-3));
    }
    P.c0 = R(c.sound_speed) > 0 ? R(c.sound_speed) : R(10) * vref;
    P.tait_gamma = c.tait_gamma;
    P.tait_B = P.rho0[0] * P.c0 * P.c0 / P.tait_gamma;
    P.p_bg = R(c.background_pressure) >= 0 ? R(c.background_pressure)
                                           : (has_vapor ? R(0.1) * P.rho0[1] * P.c0 * P.c0 : R(0));
    P.vapor_K = R(c.vapor_stiffness) > 0 ? R(c.vapor_stiffness) : (has_vapor ? P.rho0[1] * P.c0 * P.c0 : P.p_bg);
    P.gas_R = P.vapor_K > 0 ? P.vapor_K / (P.rho0[1] * fluid.T_sat) : fluid.R_gas;
    P.clamp_neg = c.clamp_negative_pressure ? 1 : 0;
    P.gravity = cv(c.gravity) - frame_acc;
    P.alpha_av = c.artificial_viscosity;
    P.alpha_av_vap = c.artificial_viscosity_vapor;
    P.delta_sph = c.delta_sph;
    P.density_mode = c.density_mode;
    P.no_slip = c.no_slip ? 1 : 0;
    P.sps = c.sps_turbulence ? 1 : 0;
    P.smag_cs = c.smagorinsky_cs;
    P.tvf = c.transport_velocity ? 1 : 0;
    P.tvf_stress = c.transport_stress ? 1 : 0;
    P.tvf_factor = c.transport_background_factor;
    P.tvf_interface_only = c.transport_interface_only ? 1 : 0;
    P.shift_strength = c.particle_shifting ? R(c.shifting_strength) : R(0);
    P.tvf_sigma_min = has_vapor ? R(0) : R(0.95) / P.vol0;
    P.st_model = c.surface_tension;
    P.sigma = st_on ? fluid.sigma : R(0);
    const R theta = R(c.contact_angle_deg) * kPi / R(180);
    P.cos_theta = std::cos(theta);
    P.sin_theta = std::sin(theta);
    P.wall_adh_w = c.contact_angle_wall_weight;
    P.wall_colour = c.wall_colour ? 1 : 0;
    P.wetting = c.wetting_model;
    P.csf_smooth = c.csf_smooth_normals ? 1 : 0;
    P.csf_balanced = c.csf_balanced ? 1 : 0;
    P.normal_eps = R(0.1) / h;
    {
      const double sup = P.kernel.support;
      const double i_coh = radial_moment4([&](double r) { return double(akinci_coh(R(r), R(sup))); }, sup);
      const double i_adh = radial_moment4([&](double r) { return double(akinci_adh(R(r), R(sup))); }, sup);
      const double rho = fluid.rho_liquid;
      const double gamma = R(c.akinci_gamma) >= 0 ? double(R(c.akinci_gamma))
                                                  : (i_coh > 0 ? 8.0 * fluid.sigma / (kPi * rho * rho * i_coh) : 0.0);
      P.ak_gamma = static_cast<R>(gamma);
      P.ak_beta = R(c.akinci_beta) >= 0
                      ? R(c.akinci_beta)
                      : static_cast<R>(i_adh > 0 ? gamma * (i_coh / i_adh) * 0.5 * (1.0 + P.cos_theta) : 0.0);
    }
    P.thermal = c.thermal ? 1 : 0;
    P.phase_change = c.phase_change ? 1 : 0;
    P.wall_bc = c.wall_thermal_bc;
    P.T_wall = fluid.T_sat;
    P.T_ref = fluid.T_sat;
    P.T_sat = fluid.T_sat;
    P.h_fg = fluid.h_fg;
    P.lee_evap = c.lee_evaporation;
    P.lee_cond = c.lee_condensation;
    P.max_transfer = c.max_transfer_fraction;
    P.pc_density_limit = c.phase_change_density_limit;
    P.vapor_rebalance = c.vapor_mass_rebalance;
    P.merge_ratio = c.merge_ratio;
    P.split_ratio = c.split_ratio;
    P.heat_power = c.heater_power;
    P.heat_inv_norm = 0;
    P.dt = 0;
  }

  // sph_solver.cu:407-428 with ParticleSystem::upload (particle_system.cpp:74-87)
  void set_particles(int n, const float* x, const float* v, const float* nrm, const float* m, const float* rho,
                     const float* T, const u8* phase, const u8* flags, const u32* id, const float* lo,
                     const float* hi) {
    if (n <= 0) throw std::invalid_argument("Solver::set_particles: no particles");
    has_vapor = false;
    for (int i = 0; i < n; ++i) has_vapor = has_vapor || phase[i] == V;
    if (has_vapor && cfg.surface_tension == CRYO_ST_AKINCI)
      throw std::invalid_argument(
          "Solver: the Akinci surface-tension model is a free-surface (liquid-only) model; use CSF when vapour "
          "particles are present (FillSpec.include_vapor = false for Akinci runs)");
    const V3 lo3{R(lo[0]), R(lo[1]), R(lo[2])}, hi3{R(hi[0]), R(hi[1]), R(hi[2])};
    const V3 ext = hi3 - lo3;
    char_length = std::max({ext.x, ext.y, ext.z});
    init_params();
    const std::size_t cap =
        std::max<std::size_t>(static_cast<std::size_t>(static_cast<double>(n) * 2.0), static_cast<std::size_t>(n) + 1024);
    if (static_cast<int>(cap) > ps.capacity) ps.allocate(cap);
    else {
      // Scratch fields start from a clean state.
      auto zero = [](auto& a) { std::fill(a.begin(), a.end(), typename std::decay_t<decltype(a)>::value_type{}); };
      zero(ps.a); zero(ps.ab); zero(ps.ns); zero(ps.vw); zero(ps.fw); zero(ps.mw); zero(ps.p); zero(ps.drho);
      zero(ps.dTdt); zero(ps.sigma); zero(ps.kappa); zero(ps.nu_t); zero(ps.alpha); zero(ps.wall_w);
      zero(ps.heat_w); zero(ps.out_m); zero(ps.out_h); zero(ps.w_norm); zero(ps.p_wv);
    }
    ps.next_id = 0;
    for (int i = 0; i < n; ++i) {
      ps.x[i] = V3{R(x[3 * i]), R(x[3 * i + 1]), R(x[3 * i + 2])};
      ps.v[i] = v ? V3{R(v[3 * i]), R(v[3 * i + 1]), R(v[3 * i + 2])} : V3{0, 0, 0};
      ps.vt[i] = ps.v[i];
      ps.nrm[i] = nrm ? V3{R(nrm[3 * i]), R(nrm[3 * i + 1]), R(nrm[3 * i + 2])} : V3{0, 0, 0};
      ps.m[i] = R(m[i]);
      ps.m0[i] = ps.m[i];
      ps.rho[i] = R(rho[i]);
      ps.T[i] = R(T[i]);
      ps.phase[i] = phase[i];
      ps.flags[i] = flags ? flags[i] : 0;
      ps.id[i] = id ? id[i] : static_cast<u32>(i);
      ps.next_id = std::max(ps.next_id, ps.id[i] + 1);
    }
    ps.n = n;
    grid.configure(lo3, hi3, P.kernel.support);
    initialized = false;
    time = 0;
    step_n = 0;
    last_dt = 0;
    heat_input = evaporated = condensed = vented = vented_h = 0;
    splits = merges = 0;
    p_ullage = fluid.p_sat;
    fluid_volume0 = 0;
  }

  void ensure_initialized() {  // sph_solver.cu:439-447
    if (initialized) return;
    if (ps.n == 0) throw std::runtime_error("Solver: no particles loaded");
    grid_build(grid, ps);
    if (cfg.normalize_initial_masses) normalize_initial_masses(ps, grid, P);
    evaluate_rates();
    update_saturation();
    initialized = true;
  }

  void evaluate_rates() {  // sph_solver.cu:449-480
    density_summation(ps, grid, P);
    equation_of_state(ps, P);
    wall_boundary(ps, grid, P);
    interface_fields(ps, grid, P);
    transport_acceleration(ps, grid, P);
    if (P.sps) sps_viscosity(ps, grid, P);
    momentum(ps, grid, P);
    if (cfg.surface_tension == CRYO_ST_CSF && has_vapor) {
      smooth_normals(ps, grid, P);
      csf(ps, grid, P);
    } else if (cfg.surface_tension == CRYO_ST_AKINCI) {
      akinci(ps, grid, P);
    }
    if (cfg.density_mode == CRYO_DENSITY_CONTINUITY) density_rate(ps, grid, P);
    if (P.thermal) {
      if (P.heat_power != R(0)) {
        const double S = reduce_sum<double>(
            ps.n, [&](int i) { return is_fluid(ps.phase[i]) ? double(ps.heat_w[i]) : 0.0; }, 0.0);
        P.heat_inv_norm = S > 0 ? static_cast<R>(1.0 / S) : R(0);
      } else {
        P.heat_inv_norm = 0;
      }
      thermal_rates(ps, grid, P);
    }
    const int n = ps.n;
    // sph_solver.cu:62-99
    v_max = reduce_max<R>(n, [&](int i) { return is_fluid(ps.phase[i]) ? norm(ps.v[i]) : R(0); }, R(0));
    spacing_ratio = reduce_max<R>(
        n,
        [&](int i) {
          if (!is_fluid(ps.phase[i])) return R(0);
          const R r = ps.rmin[i] > R(0.05) * P.dx ? ps.rmin[i] : R(0.05) * P.dx;
          return R(0.5) * P.dx / r;
        },
        R(1));
    a_max = reduce_max<R>(
        n,
        [&](int i) {
          if (!is_fluid(ps.phase[i])) return R(0);
          const R a = norm(ps.a[i]), b = norm(ps.ab[i]);
          return a > b ? a : b;
        },
        R(0));
    Tv_max = has_vapor ? reduce_max<R>(n, [&](int i) { return ps.phase[i] == V ? ps.T[i] : R(0); }, R(0)) : R(0);
    nu_t_max = P.sps ? reduce_max<R>(n, [&](int i) { return is_fluid(ps.phase[i]) ? ps.nu_t[i] : R(0); }, R(0)) : R(0);
  }

  R compute_dt() const {  // sph_solver.cu:482-504
    if (R(cfg.fixed_dt) > 0) return cfg.fixed_dt;
    const R h = P.kernel.h;
    R c_max = P.c0;
    if (has_vapor) c_max = std::max(c_max, std::sqrt(P.gas_R * std::max(Tv_max, P.T_ref)));
    R dt = R(cfg.cfl) * h / ((c_max + v_max) * std::max(R(1), spacing_ratio));
    const R nu = std::max(P.mu[0] / P.rho0[0], P.mu[1] / P.rho0[1]) + nu_t_max;
    if (nu > 0) dt = std::min(dt, R(cfg.viscous_number) * h * h / nu);
    if (a_max > 0) dt = std::min(dt, R(cfg.force_number) * std::sqrt(h / a_max));
    if (P.sigma > 0) {
      const R rho_avg = R(0.5) * (P.rho0[0] + (has_vapor ? P.rho0[1] : P.rho0[0]));
      dt = std::min(dt, R(cfg.capillary_number) * std::sqrt(rho_avg * h * h * h / (R(2) * kPi * P.sigma)));
    }
    if (P.thermal) {
      R alpha = P.k[0] / (P.rho0[0] * P.cp[0]);
      if (has_vapor) alpha = std::max(alpha, P.k[1] / (P.rho0[1] * P.cp[1]));
      if (alpha > 0) dt = std::min(dt, R(cfg.thermal_number) * h * h / alpha);
    }
    return std::clamp(dt, R(cfg.dt_min), R(cfg.dt_max));
  }

  void step() {  // sph_solver.cu:511-533
    ensure_initialized();
    const R dt = compute_dt();
    P.dt = dt;
    // kick_drift, sph_solver.cu:20-34
    {
      const R half = R(0.5) * dt;
      const int cont = cfg.density_mode == CRYO_DENSITY_CONTINUITY;
      const int tvf = P.tvf;
      pfor(ps.n, [&](int i) {
        if (!is_fluid(ps.phase[i])) return;
        const V3 v = ps.v[i] + ps.a[i] * half;
        ps.v[i] = v;
        const V3 vt = tvf ? v + ps.ab[i] * half : v;
        ps.vt[i] = vt;
        ps.x[i] += vt * dt;
        if (g_round_positions) ps.x[i] = V3{R(float(ps.x[i].x)), R(float(ps.x[i].y)), R(float(ps.x[i].z))};
        if (cont) ps.rho[i] += ps.drho[i] * dt;
      });
    }
    grid_build(grid, ps);
    if (record) {  // the adjoint's tape: the sorted, drifted state and the sort permutation
      record->ps = ps;
      record->perm.assign(grid.perm.begin(), grid.perm.begin() + ps.n);
      record->T_sat = P.T_sat;
    }
    if (has_maneuver) {
      V3 a, w, wd;
      maneuver.evaluate(time + dt, a, w, wd);
      apply_frame(a, w, wd, maneuver.origin);
    }
    evaluate_rates();
    // kick, sph_solver.cu:36-46
    {
      const R half = R(0.5) * dt;
      const int thermal = P.thermal;
      pfor(ps.n, [&](int i) {
        if (!is_fluid(ps.phase[i])) return;
        ps.v[i] += ps.a[i] * half;
        if (thermal) ps.T[i] += ps.dTdt[i] * dt;
      });
    }
    if (P.shift_strength > R(0)) particle_shifting(ps, grid, P, v_max / P.c0);
    if (P.thermal && P.heat_inv_norm > 0) heat_input += double(P.heat_power) * dt;
    if (cfg.phase_change) phase_change_step();
    update_saturation();
    time += dt;
    ++step_n;
    last_dt = dt;
  }

  void phase_change_step() {  // sph_solver.cu:545-584
    pc_sources(ps, P);
    pc_normalize(ps, grid, P);
    auto sum_pc = [&](bool evaporation) {  // sph_solver.cu:108-118
      return reduce_sum<double>(
          ps.n,
          [&](int i) {
            const u8 fl = ps.flags[i];
            if (!(fl & F_SOURCE) || (fl & F_DYING)) return 0.0;
            const bool to_vapor = (fl & F_TARGET_VAPOR) != 0;
            return to_vapor == evaporation ? double(ps.out_m[i]) : 0.0;
          },
          0.0);
    };
    evaporated += sum_pc(true);
    condensed += sum_pc(false);
    merges += reduce_sum<int>(
        ps.n, [&](int i) { return ((ps.flags[i] & F_DYING) && ps.out_m[i] > R(0)) ? 1 : 0; }, 0);
    pc_apply(ps, grid, P);
    if (P.vapor_rebalance > R(0) && has_vapor) rebalance_vapor(ps, grid, P);
    if (static_cast<std::size_t>(ps.n) * 4 > static_cast<std::size_t>(ps.capacity) * 3)
      ps.grow(static_cast<std::size_t>(ps.capacity) * 3 / 2 + 1024);
    if (static_cast<int>(split_flag.size()) < ps.capacity) {
      split_flag.resize(ps.capacity);
      split_offset.resize(ps.capacity);
    }
    pfor(ps.n, [&](int i) {
      const u8 ph = ps.phase[i];
      split_flag[i] = (is_fluid(ph) && ps.m[i] > P.split_ratio * P.mref[pidx(ph)]) ? 1u : 0u;
    });
    unsigned requested = 0;
    for (int i = 0; i < ps.n; ++i) {
      split_offset[i] = requested;
      requested += split_flag[i];
    }
    if (requested > 0) {
      split_heavy(ps, P, split_flag, split_offset, ps.next_id, static_cast<u32>(step_n));
      const std::size_t room = static_cast<std::size_t>(ps.capacity - ps.n);
      const std::size_t added = std::min<std::size_t>(requested, room);
      ps.n += static_cast<int>(added);
      ps.next_id += requested;
      splits += static_cast<int>(added);
    }
    if (cfg.thermo_mode == CRYO_VENTED && has_vapor) {
      venting(ps, P, cfg.vent_relaxation);
      vented += reduce_sum<double>(ps.n, [&](int i) { return double(ps.out_m[i]); }, 0.0);
      vented_h += reduce_sum<double>(ps.n, [&](int i) { return double(ps.out_m[i]) * double(ps.out_h[i]); }, 0.0);
    }
  }

  void update_saturation() {  // sph_solver.cu:586-607
    if (cfg.thermo_mode == CRYO_CLOSED && has_vapor) {
      const int n = ps.n;
      const double m_liq = reduce_sum<double>(n, [&](int i) { return ps.phase[i] == L ? double(ps.m[i]) : 0.0; }, 0.0);
      const double m_vap = reduce_sum<double>(n, [&](int i) { return ps.phase[i] == V ? double(ps.m[i]) : 0.0; }, 0.0);
      const double mT_vap =
          reduce_sum<double>(n, [&](int i) { return ps.phase[i] == V ? double(ps.m[i]) * ps.T[i] : 0.0; }, 0.0);
      if (!(fluid_volume0 > 0)) fluid_volume0 = m_liq / fluid.rho_liquid + m_vap / fluid.rho_vapor;
      const double v_ullage = fluid_volume0 - m_liq / fluid.rho_liquid;
      const double ratio =
          v_ullage > 0 && m_vap > 0 ? (mT_vap / v_ullage) / (double(fluid.rho_vapor) * fluid.T_sat) : 0.0;
      if (std::isfinite(ratio) && ratio > 0) {
        p_ullage = fluid.p_sat * ratio;
        P.T_sat = fluid.saturation_temperature(static_cast<R>(p_ullage));
        return;
      }
    }
    p_ullage = fluid.p_sat;
    P.T_sat = fluid.T_sat;
  }

  void apply_frame(V3 a, V3 w, V3 wd, V3 origin) {  // sph_solver.cu:660-667
    frame_acc = a;
    P.gravity = cv(cfg.gravity) - a;
    P.omega = w;
    P.omega_dot = wd;
    P.frame_origin = origin;
    P.frame_rot = (norm2(w) > R(0) || norm2(wd) > R(0)) ? 1 : 0;
  }

  void scale_velocities(R f) {  // sph_solver.cu:48-57
    pfor(ps.n, [&](int i) {
      if (!is_fluid(ps.phase[i])) return;
      ps.v[i] *= f;
      ps.vt[i] *= f;
    });
  }

  void relax(double duration, R damping) {  // sph_solver.cu:707-736
    ensure_initialized();
    const int thermal = P.thermal;
    const int pc = cfg.phase_change;
    P.thermal = 0;
    cfg.phase_change = 0;
    const double t_end = time + duration;
    try {
      while (time < t_end) {
        step();
        scale_velocities(damping);
      }
    } catch (...) {
      P.thermal = thermal;
      cfg.phase_change = pc;
      throw;
    }
    P.thermal = thermal;
    cfg.phase_change = pc;
    scale_velocities(R(0));
    time = 0;
    step_n = 0;
    last_dt = 0;
    heat_input = evaporated = condensed = vented = vented_h = 0;
    splits = merges = 0;
    evaluate_rates();
    update_saturation();
  }

  void diagnostics(cryo_diagnostics* d) {  // sph_solver.cu:609-653
    ensure_initialized();
    const int n = ps.n;
    std::memset(d, 0, sizeof(*d));
    d->time = time;
    d->step = step_n;
    d->dt = last_dt;
    d->n_total = n;
    auto count = [&](u8 ph) { return reduce_sum<int>(n, [&](int i) { return ps.phase[i] == ph ? 1 : 0; }, 0); };
    auto psum = [&](u8 ph, int what) {  // sph_solver.cu:151-174
      return reduce_sum<double>(
          n,
          [&](int i) -> double {
            if (ps.phase[i] != ph) return 0.0;
            const double m = ps.m[i];
            switch (what) {
              case 0: return m;
              case 1: return m * ps.T[i];
              case 2: return 0.5 * m * double(norm2(ps.v[i]));
              case 3: return m * ps.x[i].x;
              case 4: return m * ps.x[i].y;
              case 5: return m * ps.x[i].z;
              case 6: return m * ps.v[i].x;
              case 7: return m * ps.v[i].y;
              case 8: return m * ps.v[i].z;
              case 9: return m * double(enthalpy(P, ph, ps.T[i]));
              default: return 0.0;
            }
          },
          0.0);
    };
    d->n_liquid = count(L);
    d->n_vapor = count(V);
    d->n_wall = count(W);
    d->mass_liquid = psum(L, 0);
    d->mass_vapor = psum(V, 0);
    d->mass_total = d->mass_liquid + d->mass_vapor;
    d->kinetic_energy = psum(L, 2) + psum(V, 2);
    d->enthalpy = psum(L, 9) + psum(V, 9);
    d->heat_input = heat_input;
    d->evaporated_mass = evaporated;
    d->condensed_mass = condensed;
    d->vented_mass = vented;
    d->vented_enthalpy = vented_h;
    d->T_sat = P.T_sat;
    d->p_ullage = p_ullage;
    d->max_speed = reduce_max<R>(n, [&](int i) { return is_fluid(ps.phase[i]) ? norm(ps.v[i]) : R(0); }, R(0));
    auto fv = [](V3 a) { return cryo_vec3{float(a.x), float(a.y), float(a.z)}; };
    if (d->mass_liquid > 0) {
      d->T_liquid_mean = psum(L, 1) / d->mass_liquid;
      d->T_liquid_max = reduce_max<R>(n, [&](int i) { return ps.phase[i] == L ? ps.T[i] : R(0); }, R(0));
      d->liquid_com = fv(V3{static_cast<R>(psum(L, 3) / d->mass_liquid), static_cast<R>(psum(L, 4) / d->mass_liquid),
                            static_cast<R>(psum(L, 5) / d->mass_liquid)});
      d->liquid_momentum =
          fv(V3{static_cast<R>(psum(L, 6)), static_cast<R>(psum(L, 7)), static_cast<R>(psum(L, 8))});
    }
    if (d->mass_vapor > 0) d->T_vapor_mean = psum(V, 1) / d->mass_vapor;
    d->splits = splits;
    d->merges = merges;
    auto fsum = [&](const std::vector<V3>& f) {  // sph_solver.cu:218-232
      double c[3];
      for (int k = 0; k < 3; ++k)
        c[k] = reduce_sum<double>(
            n,
            [&](int i) -> double {
              if (!is_fluid(ps.phase[i])) return 0.0;
              const V3 v = f[i];
              return k == 0 ? double(v.x) : (k == 1 ? double(v.y) : double(v.z));
            },
            0.0);
      return V3{static_cast<R>(c[0]), static_cast<R>(c[1]), static_cast<R>(c[2])};
    };
    d->slosh_force = fv(-fsum(ps.fw));
    d->slosh_moment = fv(-fsum(ps.mw));
    d->frame_acceleration = fv(frame_acc);
    d->frame_angular_velocity = fv(P.omega);
  }

  // sph_solver.cu:197-216
  void probe(int np, const float* pts, R* out) {
    ensure_initialized();
    for (int k = 0; k < np; ++k) {
      const V3 xk{R(pts[3 * k]), R(pts[3 * k + 1]), R(pts[3 * k + 2])};
      R num = 0, den = 0;
      grid.each(xk, [&](int j) {
        if (ps.phase[j] != L) return;
        const R r2 = norm2(xk - ps.x[j]);
        if (r2 >= P.support2) return;
        const R w = (ps.m[j] / ps.rho[j]) * kW(P.kernel, std::sqrt(r2));
        num += ps.p[j] * w;
        den += w;
      });
      out[k] = den > R(0) ? num / den : std::numeric_limits<R>::quiet_NaN();
    }
  }

  // ------------------------------------------------------------ adjoint
  // Reverse-mode adjoint of the hydrodynamic step (no counterpart in the
  // reference): loss = sum_k dt (S_x^2 + S_y^2), S_k = sum of fw over fluid
  // particles after step k's rates, with the frame acceleration of step k
  // acc[3k..3k+2] as the parameters (g_eff = g - acc).  The pair derivatives
  // come from csrc/adjoint_pair.hpp; this restatement drives them in the
  // scatter form (each pair once, both members updated), serially, and is
  // itself checked against central differences of this file's forward step.
  void check_adjoint_scope(bool explicit_rotation = false) const {
    const cryo_config& c = cfg;
    if (!(c.fixed_dt > 0)) throw std::invalid_argument("adjoint: needs fixed_dt > 0");
    if (c.density_mode != CRYO_DENSITY_SUMMATION) throw std::invalid_argument("adjoint: density summation only");
    if (c.surface_tension == CRYO_ST_AKINCI) throw std::invalid_argument("adjoint: Akinci surface tension not differentiated");
    if (c.sps_turbulence || c.particle_shifting)
      throw std::invalid_argument("adjoint: SPS and shifting not differentiated");
    if (has_maneuver) throw std::invalid_argument("adjoint: manoeuvre table not allowed (the frame is per-step input)");
    if (P.frame_rot && !explicit_rotation)
      throw std::invalid_argument("adjoint: a rotating frame is differentiated only with per-step rotation inputs");
  }

  // Surface tension's share of the rates' adjoint: a += nrm coeff kappa
  // (surface_tension.cu:120-177) with nrm the contact-angle-rotated colour
  // gradient (surface_tension.cu:5-85).  At: cotangent of a.  Adds to xr
  // (positions), Vb (V = m / rho) and alphab (wall colour); returns
  // d loss / d contact angle (degrees) of this evaluation.
  double csf_adjoint(const cryo_adj::Ctx<R>& c, const std::vector<cryo_adj::Vec<R>>& At,
                     std::vector<cryo_adj::Vec<R>>& xr, std::vector<R>& Vb, std::vector<R>& alphab,
                     std::vector<R>& rb, std::vector<R>& mr) {
    using namespace cryo_adj;
    using VR = Vec<R>;
    const int n = ps.n;
    auto vr = [](V3 a) { return VR{a.x, a.y, a.z}; };
    double theta_b = 0;
    const R coeff = R(2) * P.sigma / (P.rho0[0] + P.rho0[1]);
    auto live = [&](int i) { return norm2(ps.ns[i]) > R(0.25); };
    auto Vof = [&](int j) { return ps.m[j] / ps.rho[j]; };
    std::vector<R> numb(n, R(0)), denb(n, R(0));
    std::vector<VR> nsb(n, VR{0, 0, 0}), nrmb(n, VR{0, 0, 0});
    // balanced force a += G sigma kappa / m: Gbal is the cotangent of G
    const bool bal = P.csf_balanced != 0;
    std::vector<VR> Gbal(bal ? n : 0, VR{0, 0, 0});
    auto colour = [&](int k) { return ps.phase[k] == L ? R(1) : R(0); };
    for (int i = 0; i < n; ++i) {
      if (!is_fluid(ps.phase[i]) || !live(i)) continue;
      const R kappa = ps.kappa[i];
      R kb;
      if (bal) {
        const R rhoi = ps.rho[i], mi = ps.m[i], Vi = mi / rhoi, ci = colour(i);
        V3 G{0, 0, 0};  // the forward's sum, recomputed
        grid.each(ps.x[i], [&](int j) {
          if (j == i || !is_fluid(ps.phase[j])) return;
          const V3 rij = ps.x[i] - ps.x[j];
          const R r2 = norm2(rij);
          if (r2 >= P.support2) return;
          const R r = std::sqrt(r2);
          const V3 gw = rij * (kdW(P.kernel, r) / r);
          const R rhoj = ps.rho[j], Vj = ps.m[j] / rhoj;
          G += gw * ((Vi * Vi + Vj * Vj) * (rhoj * ci + rhoi * colour(j)) / (rhoi + rhoj));
        });
        const R AG = dot(At[i], vr(G));
        kb = P.sigma / mi * AG;
        Gbal[i] = At[i] * (P.sigma * kappa / mi);
        mr[i] -= P.sigma * kappa * AG / (mi * mi);
      } else {
        kb = coeff * dot(At[i], vr(ps.nrm[i]));
        nrmb[i] += At[i] * (coeff * kappa);
      }
      R den = 0;
      grid.each(ps.x[i], [&](int j) {
        if (j == i || !is_fluid(ps.phase[j]) || !live(j)) return;
        const R r2 = norm2(ps.x[i] - ps.x[j]);
        if (r2 >= P.support2) return;
        const R r = std::sqrt(r2);
        den += Vof(j) * r * (-kdW(P.kernel, r));
      });
      if (den > R(0)) {
        numb[i] = R(3) * kb / den;
        denb[i] = -kappa * kb / den;
      }
    }
    for (int i = 0; i < n; ++i) {
      if (!is_fluid(ps.phase[i]) || !live(i)) continue;
      grid.each(ps.x[i], [&](int j) {
        if (j == i || !is_fluid(ps.phase[j]) || !live(j)) return;
        const CsfGrad<R> o = csf_vjp(c, vr(ps.x[i] - ps.x[j]), vr(ps.ns[i]), vr(ps.ns[j]), Vof(j), numb[i], denb[i]);
        xr[i] += o.r;
        xr[j] -= o.r;
        nsb[i] += o.ns_i;
        nsb[j] += o.ns_j;
        Vb[j] += o.V_j;
      });
    }
    if (bal) {
      for (int i = 0; i < n; ++i) {
        if (!is_fluid(ps.phase[i]) || !live(i)) continue;
        const R Vi = ps.m[i] / ps.rho[i];
        grid.each(ps.x[i], [&](int j) {
          if (j == i || !is_fluid(ps.phase[j])) return;
          const BalGrad<R> o = bal_vjp(c, vr(ps.x[i] - ps.x[j]), Vi, ps.m[j] / ps.rho[j], ps.rho[i], ps.rho[j],
                                       colour(i), colour(j), Gbal[i]);
          xr[i] += o.r;
          xr[j] -= o.r;
          Vb[i] += o.V_o;
          Vb[j] += o.V_n;
          rb[i] += o.rho_o;
          rb[j] += o.rho_n;
        });
      }
    }
    // smoothed normals (smooth_normals above): ns_i = acc_i / |acc_i|,
    // acc_i = sum_j V_j W_ij nrm_j / |nrm_j| over fluid j (i itself included)
    // with a live colour gradient; ghb is the cotangent of nrm_j / |nrm_j|
    std::vector<VR> ghb(P.csf_smooth ? n : 0, VR{0, 0, 0});
    if (P.csf_smooth) {
      for (int i = 0; i < n; ++i) {
        if (!is_fluid(ps.phase[i]) || !(norm(ps.nrm[i]) > P.normal_eps)) continue;
        auto each_live = [&](auto&& f) {
          grid.each(ps.x[i], [&](int j) {
            if (!is_fluid(ps.phase[j])) return;
            const R gjm = norm(ps.nrm[j]);
            if (!(gjm > P.normal_eps)) return;
            const R r2 = norm2(ps.x[i] - ps.x[j]);
            if (r2 >= P.support2) return;
            f(j, gjm, std::sqrt(r2));
          });
        };
        V3 acc{0, 0, 0};  // the forward's sum, recomputed
        each_live([&](int j, R gjm, R r) { acc += ps.nrm[j] * (Vof(j) * kW(P.kernel, r) / gjm); });
        const R A = norm(acc);
        if (!(A > R(1e-20))) continue;
        const VR ns = vr(ps.ns[i]);
        const VR ab = (nsb[i] - ns * dot(ns, nsb[i])) * (R(1) / A);
        each_live([&](int j, R gjm, R r) {
          const VR gh = vr(ps.nrm[j]) * (R(1) / gjm);
          const R Vj = Vof(j), w = kW(P.kernel, r), agh = dot(ab, gh);
          ghb[j] += ab * (Vj * w);
          Vb[j] += w * agh;
          if (r > R(0)) {
            const VR d = vr(ps.x[i] - ps.x[j]) * (kdW(P.kernel, r) / r * Vj * agh);
            xr[i] += d;
            xr[j] -= d;
          }
        });
      }
    }
    std::vector<VR> gcb(n, VR{0, 0, 0}), wnb(n, VR{0, 0, 0}), gpb(n, VR{0, 0, 0});
    std::vector<R> wwb(n, R(0));
    double cos_b = 0, sin_b = 0;
    for (int i = 0; i < n; ++i) {
      const u8 ph = ps.phase[i];
      if (!is_fluid(ph)) continue;
      const V3 g0 = ps.nrm[i];
      const R G = norm(g0);
      if (G > P.normal_eps) {
        if (P.csf_smooth) {  // the smoothed normals' nrm / |nrm|
          const VR gh = vr(g0) * (R(1) / G);
          nrmb[i] += (ghb[i] - gh * dot(gh, ghb[i])) * (R(1) / G);
        } else {  // ns = nrm / |nrm|
          const VR ns = vr(ps.ns[i]);
          nrmb[i] += (nsb[i] - ns * dot(ns, nsb[i])) * (R(1) / G);
        }
      }
      // the colour gradient before the rotation, the wall normal sum and weight
      const R ci = ph == L ? R(1) : R(0);
      V3 gradc{0, 0, 0}, wall_n{0, 0, 0}, gradphi{0, 0, 0};
      R ww = 0;
      grid.each(ps.x[i], [&](int j) {
        const V3 rij = ps.x[i] - ps.x[j];
        const R r2 = norm2(rij);
        if (r2 >= P.support2) return;
        const R r = std::sqrt(r2);
        const R w = kW(P.kernel, r);
        if (ps.phase[j] == W) {
          ww += P.vol0 * w;
          wall_n += ps.nrm[j] * (P.vol0 * w);
          gradphi += kgradW(P.kernel, rij, r) * P.vol0;
          if (P.wall_colour) gradc += kgradW(P.kernel, rij, r) * (P.vol0 * (ps.alpha[j] - ci));
          return;
        }
        if (j == i || !is_fluid(ps.phase[j])) return;
        const R cj = ps.phase[j] == L ? R(1) : R(0);
        gradc += kgradW(P.kernel, rij, r) * (Vof(j) * (cj - ci));
      });
      const RotGrad<R> ro = rot_vjp(vr(gradc), vr(wall_n), ww, P.cos_theta, P.sin_theta, P.wall_adh_w, P.normal_eps,
                                    P.wetting != 1 ? 1 : 0, nrmb[i]);
      gcb[i] = ro.g;
      wnb[i] = ro.wall_n;
      wwb[i] = ro.ww;
      cos_b += double(ro.cos_t);
      sin_b += double(ro.sin_t);
      if (P.wetting != 0 && live(i)) {  // a += vw, the contact-line force (interface pass)
        const ClGrad<R> cl = cl_vjp(vr(gradc), vr(wall_n), vr(gradphi), ww, P.cos_theta, coeff, P.normal_eps, At[i]);
        gcb[i] += cl.g;
        wnb[i] += cl.wall_n;
        gpb[i] = cl.gradphi;
        cos_b += double(cl.cos_t);
      }
    }
    theta_b = (cos_b * -double(P.sin_theta) + sin_b * double(P.cos_theta)) * (3.14159265358979323846 / 180.0);
    for (int i = 0; i < n; ++i) {
      const u8 ph = ps.phase[i];
      if (!is_fluid(ph)) continue;
      const R ci = ph == L ? R(1) : R(0);
      grid.each(ps.x[i], [&](int j) {
        if (ps.phase[j] == W) {
          const GcGrad<R> o = gc_fw_vjp(c, vr(ps.x[i] - ps.x[j]), ps.alpha[j], ci, vr(ps.nrm[j]), P.wall_colour,
                                        gcb[i], wnb[i], wwb[i]);
          xr[i] += o.r;
          if (P.wetting != 0) xr[i] += gphi_fw_vjp(c, vr(ps.x[i] - ps.x[j]), gpb[i]);
          alphab[j] += o.V_j;
          return;
        }
        if (j == i || !is_fluid(ps.phase[j])) return;
        const R cj = ps.phase[j] == L ? R(1) : R(0);
        const GcGrad<R> o = gc_ff_vjp(c, vr(ps.x[i] - ps.x[j]), Vof(j), cj - ci, gcb[i]);
        xr[i] += o.r;
        xr[j] -= o.r;
        Vb[j] += o.V_j;
      });
    }
    return theta_b;
  }

  // Phase change's share of one step's adjoint: kick -> Lee sources -> receiver
  // normalisation -> apply -> vapour rebalance -> venting (phase_change.cu:77-268,
  // 337-358, sph_solver.cu:36-46,545-584), re-run from the rates state ps.  In:
  // the cotangents of the step's end state (vb velocities, mb masses, Tb
  // temperatures).  Out: v1b, T1b (the kicked velocity and temperature), m0b (the
  // mass before phase change) and shares for positions (xr), V = m / rho (Vb),
  // rho (rb) and the interface fraction (fab); lee[0], lee[1] += d loss / d the
  // evaporation and condensation coefficients.
  void pc_adjoint(const cryo_adj::Ctx<R>& c, const std::vector<V3>& vb, const std::vector<R>& mb,
                  const std::vector<R>& Tb, std::vector<cryo_adj::Vec<R>>& v1b, std::vector<R>& T1b,
                  std::vector<R>& m0b, std::vector<cryo_adj::Vec<R>>& xr, std::vector<R>& Vb, std::vector<R>& rb,
                  std::vector<R>& fab, double* lee, double& tsatb) {
    using namespace cryo_adj;
    using VR = Vec<R>;
    const int n = ps.n;
    const R dt = P.dt, half = R(0.5) * dt;
    auto vr = [](V3 a) { return VR{a.x, a.y, a.z}; };
    auto first = [n](const auto& v) { return std::vector<typename std::decay_t<decltype(v)>::value_type>(v.begin(), v.begin() + n); };
    // forward re-run on a copy
    Particles w = ps;
    for (int i = 0; i < n; ++i) {
      if (!is_fluid(w.phase[i])) continue;
      w.v[i] += w.a[i] * half;
      if (P.thermal) w.T[i] += w.dTdt[i] * dt;
    }
    const std::vector<V3> v1 = first(w.v);
    const std::vector<R> T1 = first(w.T);
    const bool pc = cfg.phase_change != 0;
    std::vector<R> om_pre(n, R(0)), om(n, R(0)), oh(n, R(0)), wn(n, R(0));
    std::vector<u8> fl(n, 0);
    if (pc) {
      pc_sources(w, P);
      om_pre = first(w.out_m);
      pc_normalize(w, grid, P);
      om = first(w.out_m);
      oh = first(w.out_h);
      wn = first(w.w_norm);
      fl = first(w.flags);
      for (int i = 0; i < n; ++i)
        if ((fl[i] & F_DYING) && om[i] > R(0)) throw std::invalid_argument("adjoint: a merge inside the window");
      pc_apply(w, grid, P);
    }
    const std::vector<R> m2 = first(w.m), T2 = first(w.T);
    const bool reb = pc && P.vapor_rebalance > R(0) && has_vapor;
    if (reb) rebalance_vapor(w, grid, P);
    const std::vector<R> m3 = first(w.m), T3 = first(w.T);
    const std::vector<V3> v3 = first(w.v);
    const bool vent = pc && cfg.thermo_mode == CRYO_VENTED && has_vapor;
    auto Vof = [&](int j) { return ps.m[j] / ps.rho[j]; };
    // venting: m4 = m3 - relax (m3 - rho0_v m3 / rho)
    std::vector<R> m3b(mb.begin(), mb.begin() + n), T3b(Tb.begin(), Tb.begin() + n);
    std::vector<VR> v3b(n);
    for (int i = 0; i < n; ++i) v3b[i] = vr(vb[i]);
    if (vent) {
      const R relax = R(cfg.vent_relaxation), r0 = P.rho0[1];
      for (int i = 0; i < n; ++i) {
        if (ps.phase[i] != V) continue;
        const R m = m3[i], rho = ps.rho[i];
        if (!(m > r0 * (m / rho))) continue;
        m3b[i] = mb[i] * (R(1) - relax + relax * r0 / rho);
        rb[i] += mb[i] * (-relax * r0 * m / (rho * rho));
      }
    }
    // vapour rebalance: flow_ij = beta vol0 W (m_j - m_i)
    std::vector<R> m2b = m3b, T2b = T3b;
    std::vector<VR> v1b2 = v3b;
    if (reb) {
      const R beta = P.vapor_rebalance, cpv = P.cp[1];
      std::vector<R> dmb(n, R(0)), dHb(n, R(0));
      std::vector<VR> dPb(n, VR{0, 0, 0});
      std::vector<char> upd(n, 0);
      for (int i = 0; i < n; ++i) {
        if (ps.phase[i] != V) continue;
        const V3 xi = ps.x[i];
        const R mi = m2[i], hi = enthalpy(P, V, T2[i]);
        R dm = 0, dH = 0;
        grid.each(xi, [&](int j) {
          if (j == i || ps.phase[j] != V) return;
          const R r2 = norm2(xi - ps.x[j]);
          if (r2 >= P.support2) return;
          const R flow = beta * P.vol0 * kW(P.kernel, std::sqrt(r2)) * (m2[j] - mi);
          dm += flow;
          dH += flow * (flow > R(0) ? enthalpy(P, V, T2[j]) : hi);
        });
        if (dm == R(0)) continue;
        upd[i] = 1;
        const R mn = mi + dm, hh = (mi * hi + dH) / mn;
        const R hhb = T3b[i] / cpv, Hb = hhb / mn;
        R mnb = m3b[i] - hhb * hh / mn;
        const VR Vb3 = v3b[i];
        v1b2[i] = Vb3 * (mi / mn);
        R m2bi = dot(Vb3, vr(v1[i])) / mn;
        dPb[i] = Vb3 * (R(1) / mn);
        mnb -= dot(Vb3, vr(v3[i])) / mn;
        m2bi += mnb + Hb * hi;
        dmb[i] = mnb;
        dHb[i] = Hb;
        T2b[i] = Hb * mi * cpv;
        m2b[i] = m2bi;
      }
      for (int i = 0; i < n; ++i) {
        if (!upd[i]) continue;
        const V3 xi = ps.x[i];
        grid.each(xi, [&](int j) {
          if (j == i || ps.phase[j] != V) return;
          const V3 rij = xi - ps.x[j];
          const R r2 = norm2(rij);
          if (r2 >= P.support2) return;
          const R r = std::sqrt(r2), W = kW(P.kernel, r);
          const R flow = beta * P.vol0 * W * (m2[j] - m2[i]);
          const bool from_j = flow > R(0);
          const int src = from_j ? j : i;
          const R fb = dmb[i] + dHb[i] * enthalpy(P, V, T2[src]) + dot(dPb[i], vr(v1[src]));
          T2b[src] += dHb[i] * flow * cpv;
          v1b2[src] += dPb[i] * flow;
          const R k = fb * beta * P.vol0;
          m2b[j] += k * W;
          m2b[i] -= k * W;
          const VR d = vr(rij) * (k * (m2[j] - m2[i]) * cryo_adj::kern_s(c, r));
          xr[i] += d;
          xr[j] -= d;
        });
      }
    }
    for (int i = 0; i < n; ++i) v1b[i] = v1b2[i];
    // apply: m2 = m - om_i + sum_j om_j V_i W / wn_j, H likewise, T2 = T(H / m2)
    std::vector<R> omb(n, R(0)), ohb(n, R(0)), wnb(n, R(0)), inmb(n, R(0)), inHb(n, R(0));
    for (int i = 0; i < n; ++i) {
      const u8 ph = ps.phase[i];
      if (!is_fluid(ph)) continue;
      m0b[i] = m2b[i];
      T1b[i] = T2b[i];
      if (!pc) continue;
      const V3 xi = ps.x[i];
      const R mi = ps.m[i], Vi = Vof(i);
      R in_m = 0, in_H = 0;
      if (!(fl[i] & F_DYING)) grid.each(xi, [&](int j) {
        if (j == i || !(om[j] > R(0))) return;
        if (((fl[j] & F_TARGET_VAPOR) ? V : L) != ph) return;
        const R r2 = norm2(xi - ps.x[j]);
        if (r2 >= P.support2) return;
        const R share = om[j] * Vi * kW(P.kernel, std::sqrt(r2)) / wn[j];
        in_m += share;
        in_H += share * oh[j];
      });
      if (!(om[i] > R(0)) && !(in_m > R(0))) continue;
      const R mn = mi - om[i] + in_m, hT = enthalpy(P, ph, T1[i]);
      const R hh = (mi * hT - om[i] * oh[i] + in_H) / mn;
      const R cpp = P.cp[pidx(ph)];
      const R hhb = T2b[i] / cpp, Hb = hhb / mn;
      const R mnb = m2b[i] - hhb * hh / mn;
      m0b[i] = mnb + Hb * hT;
      T1b[i] = Hb * mi * cpp;
      omb[i] += -mnb - Hb * oh[i];
      ohb[i] += -Hb * om[i];
      inmb[i] = mnb;
      inHb[i] = Hb;
    }
    if (pc) {
      for (int i = 0; i < n; ++i) {
        if (inmb[i] == R(0) && inHb[i] == R(0)) continue;
        const u8 ph = ps.phase[i];
        const V3 xi = ps.x[i];
        const R Vi = Vof(i);
        grid.each(xi, [&](int j) {
          if (j == i || !(om[j] > R(0))) return;
          if (((fl[j] & F_TARGET_VAPOR) ? V : L) != ph) return;
          const V3 rij = xi - ps.x[j];
          const R r2 = norm2(rij);
          if (r2 >= P.support2) return;
          const R r = std::sqrt(r2), W = kW(P.kernel, r);
          const R share = om[j] * Vi * W / wn[j];
          const R shb = inmb[i] + inHb[i] * oh[j];
          ohb[j] += inHb[i] * share;
          omb[j] += shb * Vi * W / wn[j];
          Vb[i] += shb * om[j] * W / wn[j];
          wnb[j] -= shb * share / wn[j];
          const VR d = vr(rij) * (shb * om[j] * Vi / wn[j] * cryo_adj::kern_s(c, r));
          xr[i] += d;
          xr[j] -= d;
        });
      }
      // normalisation (wn_i = sum V_j W over the target phase, vapour density cap) and the Lee sources
      const R Ts = P.T_sat;
      for (int i = 0; i < n; ++i) {
        const u8 ph = ps.phase[i];
        if (!is_fluid(ph) || !(fl[i] & F_SOURCE)) continue;
        const u8 target = (fl[i] & F_TARGET_VAPOR) ? V : L;
        const V3 xi = ps.x[i];
        R wsum = 0, wr_max = 0;
        int jmax = -1;
        grid.each(xi, [&](int j) {
          if (j == i || ps.phase[j] != target || (fl[j] & F_DYING)) return;
          const R r2 = norm2(xi - ps.x[j]);
          if (r2 >= P.support2) return;
          const R W = kW(P.kernel, std::sqrt(r2));
          wsum += Vof(j) * W;
          const R wr = W / ps.rho[j];
          if (wr > wr_max) {
            wr_max = wr;
            jmax = j;
          }
        });
        if (!(wsum > R(0))) continue;  // om = 0
        R ompb = omb[i];
        if (target == V && P.pc_density_limit > R(0) && wr_max > R(0)) {
          const R cap = P.pc_density_limit * wsum / wr_max;
          if (om_pre[i] > cap) {
            wnb[i] += omb[i] * P.pc_density_limit / wr_max;
            const R wrb = -omb[i] * cap / wr_max;
            ompb = 0;
            const V3 rij = xi - ps.x[jmax];
            const R r = std::sqrt(norm2(rij));
            rb[jmax] -= wrb * wr_max / ps.rho[jmax];
            const VR d = vr(rij) * (wrb / ps.rho[jmax] * cryo_adj::kern_s(c, r));
            xr[i] += d;
            xr[jmax] -= d;
          }
        }
        if (wnb[i] != R(0)) {
          grid.each(xi, [&](int j) {
            if (j == i || ps.phase[j] != target || (fl[j] & F_DYING)) return;
            const V3 rij = xi - ps.x[j];
            const R r2 = norm2(rij);
            if (r2 >= P.support2) return;
            const R r = std::sqrt(r2);
            Vb[j] += wnb[i] * kW(P.kernel, r);
            const VR d = vr(rij) * (wnb[i] * Vof(j) * cryo_adj::kern_s(c, r));
            xr[i] += d;
            xr[j] -= d;
          });
        }
        // Lee source: dm = min(rate dt, [energy cap], cap m), rate = lee m alpha |T - Ts| / Ts
        const R m = ps.m[i], T = T1[i], alpha = ps.alpha[i];
        if (!(alpha > R(0)) || !(dt > R(0))) continue;
        if (ph == L && T > Ts) {
          const R h_out = P.h_fg + P.cp[1] * (Ts - P.T_ref);
          const R latent = h_out - P.cp[0] * (Ts - P.T_ref);
          const R dm_r = P.lee_evap * m * alpha * (T - Ts) / Ts * dt;
          const R dm_e = m * P.cp[0] * (T - Ts) / latent;
          const R dm_c = P.max_transfer * m;
          const R d1 = dm_r < dm_e ? dm_r : dm_e;
          tsatb += double(ohb[i] * P.cp[1]);  // out_h = h_v(T_sat)
          if (!(d1 < dm_c)) {
            m0b[i] += ompb * P.max_transfer;
          } else if (dm_r < dm_e) {
            m0b[i] += ompb * dm_r / m;
            fab[i] += ompb * dm_r / alpha;
            T1b[i] += ompb * P.lee_evap * m * alpha / Ts * dt;
            lee[0] += double(ompb * m * alpha * (T - Ts) / Ts * dt);
            tsatb -= double(ompb * P.lee_evap * m * alpha * T / (Ts * Ts) * dt);
          } else {
            m0b[i] += ompb * P.cp[0] * (T - Ts) / latent;
            T1b[i] += ompb * m * P.cp[0] / latent;
            tsatb -= double(ompb * (m * P.cp[0] + dm_e * (P.cp[1] - P.cp[0])) / latent);
          }
        } else if (ph == V && T < Ts) {
          const R dm_r = P.lee_cond * m * alpha * (Ts - T) / Ts * dt;
          R capf = P.max_transfer;
          if (P.pc_density_limit > R(0) && P.pc_density_limit * m < capf * m) capf = P.pc_density_limit;
          T1b[i] += ohb[i] * P.cp[1];  // out_h = h(T) of the condensing vapour
          if (!(dm_r < capf * m)) {
            m0b[i] += ompb * capf;
          } else {
            m0b[i] += ompb * dm_r / m;
            fab[i] += ompb * dm_r / alpha;
            T1b[i] -= ompb * P.lee_cond * m * alpha / Ts * dt;
            lee[1] += double(ompb * m * alpha * (Ts - T) / Ts * dt);
            tsatb += double(ompb * P.lee_cond * m * alpha * T / (Ts * Ts) * dt);
          }
        }
      }
    }
  }

  V3 adjoint_rates(const std::vector<u32>& perm, V3 Ssum, std::vector<V3>& xb, std::vector<V3>& vb,
                   std::vector<V3>& ab, std::vector<V3>& tb, std::vector<R>& mb, std::vector<R>& Tb, double& theta_b,
                   double* lee, bool sat_prev, double* wb, double* wdb) {
    using namespace cryo_adj;
    using VR = Vec<R>;
    const int n = ps.n;
    const R dt = P.dt, half = R(0.5) * dt;
    Ctx<R> c{P.kernel.norm, P.kernel.inv_h, P.kernel.h, P.support2, R(0.01) * P.kernel.h * P.kernel.h,
             R(0.6) * P.dx, dt, P.vol0, P.no_slip, P.kernel.type, P.tvf && P.tvf_stress ? 1 : 0};
    auto vr = [](V3 a) { return VR{a.x, a.y, a.z}; };
    auto v3 = [](VR a) { return V3{a.x, a.y, a.z}; };
    const VR Fb{R(2) * dt * Ssum.x, R(2) * dt * Ssum.y, R(0)};
    const VR g = vr(P.gravity);
    auto side = [&](int i) {
      const u8 ph = ps.phase[i];
      return Side<R>{vr(ps.x[i]), vr(ps.v[i]), ps.m[i], ps.rho[i], ps.p[i], sound_speed(P, ph, ps.T[i]),
                     P.mu[pidx(ph)], ph == V ? P.alpha_av_vap : P.alpha_av, int(ph), vr(ps.vt[i])};
    };
    std::vector<VR> At(n, VR{0, 0, 0}), xr(n, VR{0, 0, 0}), vrb(n, VR{0, 0, 0}), vwb(n, VR{0, 0, 0});
    std::vector<VR> vtr(n, VR{0, 0, 0});  // the transport stress's cotangent of vt
    std::vector<R> rb(n, R(0)), pb(n, R(0)), pwb0(n, R(0)), pwb1(n, R(0));
    std::vector<R> Vb(n, R(0)), alphab(n, R(0));
    // phase change, then the kick: v1 = v_half + a dt/2, T1 = T + dTdt dt
    std::vector<VR> v1b(n, VR{0, 0, 0});
    std::vector<R> T1b(n, R(0)), mr(n, R(0)), Tr(n, R(0)), fab(n, R(0));
    double tsatb = 0;  // cotangent of this step's saturation temperature
    pc_adjoint(c, vb, mb, Tb, v1b, T1b, mr, xr, Vb, rb, fab, lee, tsatb);
    VR gb{0, 0, 0};
    for (int i = 0; i < n; ++i) {
      if (!is_fluid(ps.phase[i])) continue;
      At[i] = vr(ab[i]) + v1b[i] * half;
      gb += At[i];  // a = g_eff + ...
      Tr[i] += T1b[i];
    }
    // rotating frame (sph_solver.cu:62-99 momentum): a -= wd x r + w x (w x r) + 2 w x v, r = x - origin
    {
      const V3 w = P.omega, wd = P.omega_dot;
      for (int i = 0; i < n; ++i) {
        if (!is_fluid(ps.phase[i])) continue;
        const V3 A = v3(At[i]), r = ps.x[i] - P.frame_origin, v = ps.v[i];
        xr[i] += vr(cross(wd, A) - cross(w, cross(w, A)));
        vrb[i] += vr(cross(w, A) * R(2));
        const V3 dwd = cross(A, r);
        const V3 dw = A * (-dot(w, r)) - r * dot(A, w) + w * (R(2) * dot(A, r)) + cross(A, v) * R(2);
        wdb[0] += double(dwd.x), wdb[1] += double(dwd.y), wdb[2] += double(dwd.z);
        wb[0] += double(dw.x), wb[1] += double(dw.y), wb[2] += double(dw.z);
      }
    }
    auto c_to_T = [&](int i, R cb) {  // c = sqrt(R T) for vapour
      if (ps.phase[i] == V && ps.T[i] > R(0) && cb != R(0)) Tr[i] += cb * P.gas_R / (R(2) * sound_speed(P, V, ps.T[i]));
    };
    // surface tension
    if (cfg.surface_tension == CRYO_ST_CSF && has_vapor) theta_b += csf_adjoint(c, At, xr, Vb, alphab, rb, mr);
    // transport acceleration ab = K sum_J gw (Vt^2 + V_J^2), K = -pb / (rho Vt)
    // (SPH_kernels.cu:56-104), read by the next step's drift: tb is its cotangent
    std::vector<R> tsb(n, R(0));  // sigma cotangent of the transport sums
    if (P.tvf) {
      for (int i = 0; i < n; ++i) {
        const u8 ph = ps.phase[i];
        if (!is_fluid(ph) || ps.sigma[i] < P.tvf_sigma_min || (P.tvf_interface_only && !(ps.alpha[i] > R(0))))
          continue;
        const VR Ab = vr(tb[i]);
        const R Vt = R(1) / ps.sigma[i], Vt2 = Vt * Vt, rhoi = ps.rho[i];
        const R ci = sound_speed(P, ph, ps.T[i]);
        const R K = -(P.tvf_factor * P.rho0[pidx(ph)] * ci * ci) / (rhoi * Vt);
        const VR Sb = Ab * K;
        R vt2b = 0, vtb = 0, rhob = 0;
        VR sum{0, 0, 0};
        grid.each(ps.x[i], [&](int j) {
          if (j == i) return;
          const V3 rij = ps.x[i] - ps.x[j];
          const R r2 = norm2(rij);
          if (r2 >= P.support2) return;
          const bool wall = ps.phase[j] == W;
          R Vj, rho_w = 0;
          if (wall) {
            if (ph == V) {
              rho_w = (ps.p_wv[j] - (P.p_bg - P.vapor_K)) / (P.gas_R * ps.T[i]);
            } else {
              const R ratio = (ps.p[j] - P.p_bg) / P.tait_B + R(1);
              rho_w = P.rho0[0] * std::pow(ratio > R(0.5) ? ratio : R(0.5), R(1) / P.tait_gamma);
            }
            Vj = rho_w > R(0) ? Vt * rhoi / rho_w : Vt;
          } else {
            Vj = R(1) / ps.sigma[j];
          }
          sum += vr(kgradW(P.kernel, rij, std::sqrt(r2))) * (Vt2 + Vj * Vj);
          const TvGrad<R> o = tvf_vjp(c, vr(rij), Vt2, Vj, Sb);
          xr[i] += o.r;
          xr[j] -= o.r;
          vt2b += o.vt2;
          if (wall) {
            const R rwb = tvf_wall_vjp(Vt, rhoi, rho_w, o.vj, vtb, rhob);
            if (ph == V) Tr[i] -= rwb * rho_w / ps.T[i];
            const R dp = tvf_wall_drho_dp(ph == V ? 1 : 0, ph == V ? ps.p_wv[j] : ps.p[j], rho_w, P.p_bg, P.tait_B,
                                          P.tait_gamma, P.gas_R, ps.T[i]);
            (ph == V ? pwb1 : pwb0)[j] += rwb * dp;
          } else {
            tsb[j] -= o.vj * Vj * Vj;  // V_J = 1 / sigma_J
          }
        });
        const R Kb = dot(Ab, sum);
        rhob -= K * Kb / rhoi;
        if (ph == V) Tr[i] += K * Kb / ps.T[i];  // c^2 = R T
        vtb -= K * Kb / Vt;
        vtb += R(2) * Vt * vt2b;
        tsb[i] -= vtb * Vt * Vt;  // Vt = 1 / sigma
        rb[i] += rhob;
      }
    }
    // momentum
    for (int i = 0; i < n; ++i) {
      if (!is_fluid(ps.phase[i])) continue;
      const Side<R> I = side(i);
      const VR cF = At[i] * (R(1) / I.m), cA = At[i];
      const VR cFw = cF + Fb, cAw = cA + Fb * I.m;
      VR F{0, 0, 0};
      grid.each(ps.x[i], [&](int j) {
        if (j == i) return;
        if (ps.phase[j] == W) {
          const FWGrad<R> o = fw_vjp(c, I, vr(ps.x[j]), vr(ps.vw[j]), I.ph == V ? ps.p_wv[j] : ps.p[j], cFw, cAw);
          F += o.f;
          mr[i] += o.m_i + dot(Fb, o.acc);  // the wall load carries m_I acc
          c_to_T(i, o.c_i);
          xr[i] += o.r;
          vrb[i] += o.vi + o.vij;
          vtr[i] += o.vti;
          vwb[j] -= o.vij;
          rb[i] += o.rho_i;
          pb[i] += o.p_i;
          (I.ph == V ? pwb1 : pwb0)[j] += o.p_w;
          return;
        }
        if (!is_fluid(ps.phase[j])) return;
        const FFGrad<R> o = ff_vjp(c, I, side(j), cF, cA);
        F += o.f;
        mr[i] += o.m_i;
        mr[j] += o.m_j;
        c_to_T(i, o.c_i);
        c_to_T(j, o.c_j);
        xr[i] += o.r;
        xr[j] -= o.r;
        vrb[i] += o.vij + o.vi;
        vrb[j] += o.vj - o.vij;
        vtr[i] += o.vti;
        vtr[j] += o.vtj;
        rb[i] += o.rho_i;
        rb[j] += o.rho_j;
        pb[i] += o.p_i;
        pb[j] += o.p_j;
      });
      mr[i] -= dot(At[i], F) / (I.m * I.m);  // a = ... + F / m
    }
    // heat: dTdt = (sum_J q_IJ + P hw_I / sum_K hw_K) / (m cp)
    if (P.thermal) {
      const bool iso = P.wall_bc == 1, heat = P.heat_power != R(0) && P.heat_inv_norm > R(0);
      std::vector<R> hwb(n, R(0));
      R Sb = 0;
      for (int i = 0; i < n; ++i) {
        const u8 ph = ps.phase[i];
        if (!is_fluid(ph)) continue;
        const R D = T1b[i] * dt;
        if (D == R(0)) continue;
        const R mi = ps.m[i], cpi = P.cp[pidx(ph)], qb = D / (mi * cpi);
        mr[i] -= D * ps.dTdt[i] / mi;
        const R Vi = mi / ps.rho[i], ki = P.k[pidx(ph)];
        grid.each(ps.x[i], [&](int j) {
          if (j == i) return;
          const u8 phj = ps.phase[j];
          R Vj, kij;
          if (phj == W) {
            if (!iso) return;
            Vj = P.vol0;
            kij = R(2) * ki;
          } else {
            if (!is_fluid(phj)) return;
            const R kj = P.k[pidx(phj)];
            Vj = ps.m[j] / ps.rho[j];
            kij = R(4) * ki * kj / (ki + kj);
          }
          const ThGrad<R> o = th_vjp(c, vr(ps.x[i] - ps.x[j]), Vi, Vj, kij, ps.T[i], ps.T[j], qb);
          xr[i] += o.r;
          Vb[i] += o.Vi;
          Tr[i] += o.Ti;
          if (phj != W) {
            xr[j] -= o.r;
            Vb[j] += o.Vj;
            Tr[j] += o.Tj;
          }
        });
        if (heat) {
          hwb[i] += qb * P.heat_power * P.heat_inv_norm;
          Sb -= qb * P.heat_power * ps.heat_w[i] * P.heat_inv_norm * P.heat_inv_norm;
        }
      }
      if (heat) {  // hw_I = sum over heated walls of vol0 W
        for (int i = 0; i < n; ++i) {
          if (!is_fluid(ps.phase[i])) continue;
          const R hb = hwb[i] + Sb;
          grid.each(ps.x[i], [&](int j) {
            if (ps.phase[j] != W || !(ps.flags[j] & F_HEATED)) return;
            const V3 rij = ps.x[i] - ps.x[j];
            const R r2 = norm2(rij);
            if (r2 >= P.support2) return;
            xr[i] += vr(rij) * (hb * P.vol0 * kern_s(c, std::sqrt(r2)));
          });
        }
      }
    }
    // interface fraction alpha = s_other / s_all (the Lee rate reads it)
    for (int i = 0; i < n; ++i) {
      const u8 ph = ps.phase[i];
      if (!is_fluid(ph) || fab[i] == R(0)) continue;
      R sa = 0, so = 0;
      grid.each(ps.x[i], [&](int j) {
        if (!is_fluid(ps.phase[j])) return;
        const R r2 = norm2(ps.x[i] - ps.x[j]);
        if (r2 >= P.support2) return;
        const R Wv = ps.m[j] / ps.rho[j] * kW(P.kernel, std::sqrt(r2));
        sa += Wv;
        if (ps.phase[j] != ph) so += Wv;
      });
      if (!(sa > R(0))) continue;
      const R sob = fab[i] / sa, sab = -fab[i] * (so / sa) / sa;
      grid.each(ps.x[i], [&](int j) {
        if (!is_fluid(ps.phase[j])) return;
        const V3 rij = ps.x[i] - ps.x[j];
        const R r2 = norm2(rij);
        if (r2 >= P.support2) return;
        const R r = std::sqrt(r2);
        const R k = sab + (ps.phase[j] != ph ? sob : R(0));
        Vb[j] += k * kW(P.kernel, r);
        if (j == i) return;
        const VR d = vr(rij) * (k * (ps.m[j] / ps.rho[j]) * kern_s(c, r));
        xr[i] += d;
        xr[j] -= d;
      });
    }
    // wall boundary values
    for (int w = 0; w < n; ++w) {
      if (ps.phase[w] != W) continue;
      R sw0 = 0, sw1 = 0;
      VR G0{0, 0, 0}, G1{0, 0, 0};
      grid.each(ps.x[w], [&](int j) {
        if (!is_fluid(ps.phase[j])) return;
        const V3 rwf = ps.x[w] - ps.x[j];
        const R r2 = norm2(rwf);
        if (r2 >= P.support2) return;
        const R wk = kW(P.kernel, std::sqrt(r2));
        if (ps.phase[j] == V) {
          sw1 += wk;
          G1 += vr(rwf) * (ps.rho[j] * wk);
        } else {
          sw0 += wk;
          G0 += vr(rwf) * (ps.rho[j] * wk);
        }
      });
      const bool cl0 = P.clamp_neg && !(ps.p[w] > R(0)), cl1 = P.clamp_neg && !(ps.p_wv[w] > R(0));
      const R cp0 = (sw0 > R(0) && !cl0) ? pwb0[w] / sw0 : R(0);
      const R cp1 = (sw1 > R(0) && !cl1) ? pwb1[w] / sw1 : R(0);
      const R swt = sw0 + sw1;
      const VR cvw = swt > R(0) ? vwb[w] * (R(1) / swt) : VR{0, 0, 0};
      const R calpha = swt > R(0) ? alphab[w] / swt : R(0);  // alpha_w = sw0 / swt
      gb += G0 * cp0 + G1 * cp1;
      grid.each(ps.x[w], [&](int f) {
        if (!is_fluid(ps.phase[f])) return;
        const Side<R> F = side(f);
        const bool vap = F.ph == V;
        const WBGrad<R> o =
            wb_vjp(c, F, vr(ps.x[w]), vap ? ps.p_wv[w] : ps.p[w], vr(ps.vw[w]), vap ? cp1 : cp0, cvw, g);
        xr[f] += o.x;
        vrb[f] += o.v;
        pb[f] += o.p;
        rb[f] += o.rho;
        if (calpha != R(0)) {
          const V3 rwf = ps.x[w] - ps.x[f];
          const R r2 = norm2(rwf);
          if (r2 < P.support2)  // d alpha_w / d x_f = (c_f - alpha_w) / swt * dW/dx_f, dW/dx_f = -s r_wf
            xr[f] += vr(rwf) * (-calpha * ((vap ? R(0) : R(1)) - ps.alpha[w]) * kern_s(c, std::sqrt(r2)));
        }
      });
    }
    // equation of state (and V = m / rho of the surface-tension sums), then density summation
    std::vector<R> sb(n, R(0));
    for (int i = 0; i < n; ++i) {
      const u8 ph = ps.phase[i];
      if (!is_fluid(ph)) continue;
      rb[i] -= Vb[i] * (ps.m[i] / ps.rho[i]) / ps.rho[i];
      mr[i] += Vb[i] / ps.rho[i];
      const R dpdr = eos_dp_drho(ph == V ? 1 : 0, ps.rho[i], ps.T[i], ps.p[i], P.rho0[0], P.tait_B, P.tait_gamma,
                                 P.gas_R, P.clamp_neg);
      Tr[i] += pb[i] * eos_dp_dT(ph == V ? 1 : 0, ps.rho[i], ps.p[i], P.gas_R, P.clamp_neg);
      const R rhob = rb[i] + pb[i] * dpdr;
      if (ph == V) mr[i] += rhob * ps.sigma[i];  // rho = m sigma
      sb[i] = rhob * (ph == L ? ps.m0[i] : ps.m[i]) + tsb[i];
    }
    for (int i = 0; i < n; ++i) {
      if (!is_fluid(ps.phase[i])) continue;
      grid.each(ps.x[i], [&](int j) {
        if (j == i) return;
        const V3 rij = ps.x[i] - ps.x[j];
        const R r2 = norm2(rij);
        if (r2 >= P.support2) return;
        const VR d = vr(rij) * (sb[i] * kern_s(c, std::sqrt(r2)));
        xr[i] += d;
        xr[j] -= d;
      });
    }
    // drift and kick, then back through the cell sort: old[perm[i]] <- new[i]
    std::vector<V3> nx(n, V3{0, 0, 0}), nv(n, V3{0, 0, 0}), na(n, V3{0, 0, 0}), nt(n, V3{0, 0, 0});
    std::vector<R> nm(n, R(0)), nT(n, R(0));
    // Closed tank: this step's T_sat came from the previous step's end state
    // (sph_solver.cu:586-607), p = p_sat (sum_V m T / v_ull) / (rho_v T_sat0),
    // v_ull = V0 - sum_L m / rho_l, T_sat = 1 / (1/T_sat0 - (R/h_fg) ln(p/p_sat)).
    // That state's masses and temperatures are this taped state's.
    double mlb = 0, mTb = 0;
    if (sat_prev && tsatb != 0) {
      double m_liq = 0, m_vap = 0, mT = 0;
      for (int i = 0; i < n; ++i) {
        if (ps.phase[i] == L) m_liq += double(ps.m[i]);
        if (ps.phase[i] == V) m_vap += double(ps.m[i]), mT += double(ps.m[i]) * ps.T[i];
      }
      const double v_u = fluid_volume0 - m_liq / fluid.rho_liquid;
      const double ratio = v_u > 0 && m_vap > 0 ? (mT / v_u) / (double(fluid.rho_vapor) * fluid.T_sat) : 0.0;
      if (std::isfinite(ratio) && ratio > 0) {
        const double pu = fluid.p_sat * ratio, Ts = P.T_sat;
        const double g = tsatb * Ts * Ts * (double(fluid.R_gas) / fluid.h_fg) / pu;  // dT_sat/dp
        mTb = g * pu / mT;
        mlb = g * pu / (v_u * fluid.rho_liquid);
      }
    }
    for (int i = 0; i < n; ++i) {
      if (!is_fluid(ps.phase[i])) continue;
      nm[perm[i]] = mr[i];
      nT[perm[i]] = Tr[i];
      if (ps.phase[i] == L) nm[perm[i]] += R(mlb);
      if (ps.phase[i] == V) nm[perm[i]] += R(mTb * ps.T[i]), nT[perm[i]] += R(mTb * ps.m[i]);
      const VR X = vr(xb[i]) + xr[i];
      const VR Vh = v1b[i] + vrb[i] + X * dt + vtr[i];  // vt = v_half + ab dt/2 (tvf)
      nx[perm[i]] = v3(X);
      nv[perm[i]] = v3(Vh);
      na[perm[i]] = v3(Vh * half);
      if (P.tvf) nt[perm[i]] = v3((X * dt + vtr[i]) * half);  // x' = x + (v_half + ab dt/2) dt
    }
    xb.swap(nx);
    vb.swap(nv);
    ab.swap(na);
    tb.swap(nt);
    mb.swap(nm);
    Tb.swap(nT);
    return v3(gb);
  }

  void adjoint_slosh(long long steps, const float* acc, double* loss, double* grad, double* force, double* grad_theta,
                     double* grad_lee, const float* omega = nullptr, const float* omega_dot = nullptr,
                     double* grad_omega = nullptr, double* grad_omega_dot = nullptr) {
    check_adjoint_scope(omega || omega_dot);
    if (steps < 0) throw std::invalid_argument("adjoint: steps must be >= 0");
    const bool fresh = !initialized;  // the start-up rates (and their a) belong to this run
    ensure_initialized();
    const Particles start_state = ps;
    const V3 start_acc = frame_acc;
    const V3 origin = P.frame_origin;
    const V3 zero{0, 0, 0};
    auto acc_of = [&](long long k) { return V3{R(acc[3 * k]), R(acc[3 * k + 1]), R(acc[3 * k + 2])}; };
    auto w_of = [&](long long k) {
      return omega ? V3{R(omega[3 * k]), R(omega[3 * k + 1]), R(omega[3 * k + 2])} : V3{0, 0, 0};
    };
    auto wd_of = [&](long long k) {
      return omega_dot ? V3{R(omega_dot[3 * k]), R(omega_dot[3 * k + 1]), R(omega_dot[3 * k + 2])} : V3{0, 0, 0};
    };
    std::vector<TapeEntry> tape(static_cast<std::size_t>(steps));
    std::vector<V3> S(static_cast<std::size_t>(steps));
    double L = 0;
    const int events0 = merges + splits;
    for (long long k = 0; k < steps; ++k) {
      apply_frame(acc_of(k), w_of(k), wd_of(k), origin);
      record = grad ? &tape[static_cast<std::size_t>(k)] : nullptr;
      step();
      record = nullptr;
      double sx = 0, sy = 0, sz = 0;
      for (int i = 0; i < ps.n; ++i) {
        if (!is_fluid(ps.phase[i])) continue;
        sx += double(ps.fw[i].x);
        sy += double(ps.fw[i].y);
        sz += double(ps.fw[i].z);
      }
      S[static_cast<std::size_t>(k)] = V3{R(sx), R(sy), R(sz)};
      if (force) {
        force[3 * k] = sx;
        force[3 * k + 1] = sy;
        force[3 * k + 2] = sz;
      }
      L += double(P.dt) * (sx * sx + sy * sy);
    }
    *loss = L;
    if (!grad) return;
    if (merges + splits != events0)
      throw std::invalid_argument("adjoint: a phase-change merge or split inside the window is not differentiable");
    // reverse sweep over the taped states; the solver ends in the forward end state
    Particles end_state = ps;
    const R T_sat_end = P.T_sat;
    std::vector<V3> xb(ps.n, zero), vb(ps.n, zero), ab(ps.n, zero), tb(ps.n, zero);
    std::vector<R> mb(ps.n, R(0)), Tb(ps.n, R(0));
    double theta_b = 0, lee[2] = {0, 0};
    for (long long k = steps - 1; k >= 0; --k) {
      TapeEntry& t = tape[static_cast<std::size_t>(k)];
      ps = t.ps;
      grid_build(grid, ps);  // already sorted: the identity permutation, the same cell table
      apply_frame(acc_of(k), w_of(k), wd_of(k), origin);
      P.dt = R(cfg.fixed_dt);
      P.T_sat = t.T_sat;
      evaluate_rates();
      const bool sat_prev = k > 0 && cfg.thermo_mode == CRYO_CLOSED && has_vapor && cfg.phase_change;
      double wbk[3] = {0, 0, 0}, wdbk[3] = {0, 0, 0};
      const V3 gb =
          adjoint_rates(t.perm, S[static_cast<std::size_t>(k)], xb, vb, ab, tb, mb, Tb, theta_b, lee, sat_prev, wbk, wdbk);
      for (int q = 0; q < 3; ++q) {
        if (grad_omega) grad_omega[3 * k + q] = wbk[q];
        if (grad_omega_dot) grad_omega_dot[3 * k + q] = wdbk[q];
      }
      grad[3 * k] = -double(gb.x);
      grad[3 * k + 1] = -double(gb.y);
      grad[3 * k + 2] = -double(gb.z);
    }
    if (fresh && steps > 0 && cfg.surface_tension == CRYO_ST_CSF && has_vapor) {
      // the start-up a depends on the contact angle too: ab now holds its cotangent
      ps = start_state;
      grid_build(grid, ps);
      apply_frame(start_acc, zero, zero, origin);
      evaluate_rates();
      using VR = cryo_adj::Vec<R>;
      const int n = ps.n;
      cryo_adj::Ctx<R> c{P.kernel.norm, P.kernel.inv_h, P.kernel.h, P.support2, R(0.01) * P.kernel.h * P.kernel.h,
                         R(0.6) * P.dx, P.dt, P.vol0, P.no_slip, P.kernel.type, P.tvf && P.tvf_stress ? 1 : 0};
      std::vector<VR> At(n), xr(n, VR{0, 0, 0});
      for (int i = 0; i < n; ++i) At[i] = is_fluid(ps.phase[i]) ? VR{ab[i].x, ab[i].y, ab[i].z} : VR{0, 0, 0};
      std::vector<R> Vb(n, R(0)), alphab(n, R(0)), rb(n, R(0)), mr(n, R(0));
      theta_b += csf_adjoint(c, At, xr, Vb, alphab, rb, mr);
    }
    if (grad_theta) *grad_theta = theta_b;
    if (grad_lee) {
      grad_lee[0] = lee[0];
      grad_lee[1] = lee[1];
    }
    ps = end_state;        // scratch included: a is the end state's
    P.T_sat = T_sat_end;
    grid_build(grid, ps);  // sorted already: the identity permutation
    if (steps > 0) apply_frame(acc_of(steps - 1), w_of(steps - 1), wd_of(steps - 1), origin);
  }
};

thread_local std::string g_err;

Fluid to_fluid(const cryo_fluid& f) {
  Fluid o;
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

template <class T>
void put(void* dst, const std::vector<T>& src, int n) {
  std::memcpy(dst, src.data(), sizeof(T) * static_cast<std::size_t>(n));
}

}  // namespace orc

using orc::Solver;

#define GUARD(...)                                         \
  try {                                                     \
    __VA_ARGS__;                                            \
    return 0;                                               \
  } catch (const std::invalid_argument& e) {                \
    orc::g_err = e.what();                                  \
    return CRYO_ERR_INVALID_ARGUMENT;                       \
  } catch (const std::out_of_range& e) {                    \
    orc::g_err = e.what();                                  \
    return CRYO_ERR_OUT_OF_RANGE;                           \
  } catch (const std::exception& e) {                       \
    orc::g_err = e.what();                                  \
    return CRYO_ERR_RUNTIME;                                \
  }

extern "C" {

const char* orc_last_error(void) { return orc::g_err.c_str(); }
int orc_real_size(void) { return static_cast<int>(sizeof(orc::R)); }

int orc_create(const cryo_config* cfg, const cryo_fluid* fluid, void** out) {
  GUARD(*out = new Solver(*cfg, orc::to_fluid(*fluid)))
}
void orc_destroy(void* s) { delete static_cast<Solver*>(s); }

int orc_set_particles(void* s, int n, const float* x, const float* v, const float* nrm, const float* m,
                      const float* rho, const float* T, const uint8_t* phase, const uint8_t* flags,
                      const uint32_t* id, const float* lo, const float* hi) {
  GUARD(static_cast<Solver*>(s)->set_particles(n, x, v, nrm, m, rho, T, phase, flags, id, lo, hi))
}
int orc_build_tank(void*, int, const cryo_geo_node*, int, const cryo_mesh*, const cryo_fill*) {
  orc::g_err = "oracle: lattice generation is not restated; pass particles with set_particles";
  return CRYO_ERR_RUNTIME;
}
int orc_build_tank_particles(int, const cryo_geo_node*, int, const cryo_mesh*, const cryo_fluid*, double,
                             const cryo_fill*, int*, void*, void*, void*, void*, void*, void*, uint8_t*, uint8_t*,
                             uint32_t*) {
  orc::g_err = "oracle: lattice generation is not restated";
  return CRYO_ERR_RUNTIME;
}
int orc_step(void* s) { GUARD(static_cast<Solver*>(s)->step()) }
int orc_run_steps(void* s, long long n) {
  GUARD(for (long long k = 0; k < n; ++k) static_cast<Solver*>(s)->step())
}
int orc_adjoint_slosh(void* s, long long steps, const float* acc, int checkpoint_every, double* loss, double* grad,
                      double* force, double* grad_contact_angle, double* grad_lee) {
  (void)checkpoint_every;  // the oracle tapes every step
  GUARD(static_cast<Solver*>(s)->adjoint_slosh(steps, acc, loss, grad, force, grad_contact_angle, grad_lee))
}
int orc_adjoint_slosh_frame(void* s, long long steps, const float* acc, const float* omega, const float* omega_dot,
                            int checkpoint_every, double* loss, double* grad, double* grad_omega,
                            double* grad_omega_dot, double* force, double* grad_contact_angle, double* grad_lee) {
  (void)checkpoint_every;
  GUARD(static_cast<Solver*>(s)->adjoint_slosh(steps, acc, loss, grad, force, grad_contact_angle, grad_lee, omega,
                                               omega_dot, grad_omega, grad_omega_dot))
}
int orc_run_until(void* s, double t) {
  GUARD(Solver& S = *static_cast<Solver*>(s); while (S.time < t) S.step())
}
int orc_relax(void* s, double d, double damping) {
  GUARD(static_cast<Solver*>(s)->relax(d, orc::R(damping)))
}
int orc_next_dt(void* s, double* dt) {
  GUARD(Solver& S = *static_cast<Solver*>(s); S.ensure_initialized(); *dt = S.compute_dt())
}
int orc_damp_velocities(void* s, double f) {
  GUARD(Solver& S = *static_cast<Solver*>(s); const orc::R ff = orc::R(f);
        S.scale_velocities(ff); S.v_max *= ff)
}
double orc_time(void* s) { return static_cast<Solver*>(s)->time; }
long long orc_step_count(void* s) { return static_cast<Solver*>(s)->step_n; }
double orc_last_dt(void* s) { return static_cast<Solver*>(s)->last_dt; }
int orc_num_particles(void* s) { return static_cast<Solver*>(s)->ps.n; }

int orc_set_gravity(void* s, const float* g) {
  GUARD(Solver& S = *static_cast<Solver*>(s); S.cfg.gravity = cryo_vec3{g[0], g[1], g[2]};
        S.P.gravity = S.cv(S.cfg.gravity) - S.frame_acc)
}
int orc_set_frame_motion(void* s, const float* a, const float* w, const float* wd, const float* o) {
  GUARD(Solver& S = *static_cast<Solver*>(s); S.has_maneuver = false;
        S.apply_frame(orc::V3{a[0], a[1], a[2]}, orc::V3{w[0], w[1], w[2]}, orc::V3{wd[0], wd[1], wd[2]},
                      orc::V3{o[0], o[1], o[2]}))
}
int orc_set_maneuver(void* s, int n, const double* t, const float* a, const float* w, const float* o) {
  GUARD({
    Solver& S = *static_cast<Solver*>(s);
    orc::Maneuver m;
    m.time.assign(t, t + n);
    for (int k = 0; k < n; ++k) m.acc.push_back(orc::V3{a[3 * k], a[3 * k + 1], a[3 * k + 2]});
    if (w)
      for (int k = 0; k < n; ++k) m.omega.push_back(orc::V3{w[3 * k], w[3 * k + 1], w[3 * k + 2]});
    m.origin = orc::V3{o[0], o[1], o[2]};
    m.validate();
    S.maneuver = m;
    S.has_maneuver = true;
    orc::V3 aa, ww, wd;
    S.maneuver.evaluate(S.time, aa, ww, wd);
    S.apply_frame(aa, ww, wd, S.maneuver.origin);
  })
}
int orc_clear_frame_motion(void* s) {
  GUARD(Solver& S = *static_cast<Solver*>(s); S.has_maneuver = false;
        S.apply_frame(orc::V3{0, 0, 0}, orc::V3{0, 0, 0}, orc::V3{0, 0, 0}, orc::V3{0, 0, 0}))
}
int orc_set_heater_power(void* s, double w) {
  GUARD(Solver& S = *static_cast<Solver*>(s); S.cfg.heater_power = static_cast<float>(w);
        S.P.heat_power = orc::R(w))
}
int orc_set_contact_angle(void* s, double deg) {
  GUARD({
    Solver& S = *static_cast<Solver*>(s);
    const orc::R d = orc::R(deg);
    if (d < 0 || d > 180) throw std::invalid_argument("contact angle must be within [0, 180] degrees");
    S.cfg.contact_angle_deg = static_cast<float>(deg);
    const orc::R theta = d * orc::kPi / orc::R(180);
    S.P.cos_theta = std::cos(theta);
    S.P.sin_theta = std::sin(theta);
  })
}

int orc_params(void* s, cryo_params* o) {
  GUARD({
    const orc::Params& P = static_cast<Solver*>(s)->P;
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
    auto fv = [](orc::V3 a) { return cryo_vec3{float(a.x), float(a.y), float(a.z)}; };
    o->gravity = fv(P.gravity);
    o->omega = fv(P.omega);
    o->omega_dot = fv(P.omega_dot);
    o->frame_origin = fv(P.frame_origin);
    o->frame_rot = P.frame_rot;
    o->sigma = float(P.sigma);
    o->cos_theta = float(P.cos_theta);
    o->sin_theta = float(P.sin_theta);
    o->normal_eps = float(P.normal_eps);
    o->akinci_gamma = float(P.ak_gamma);
    o->akinci_beta = float(P.ak_beta);
    o->tvf_sigma_min = float(P.tvf_sigma_min);
    o->shift_strength = float(P.shift_strength);
    o->T_ref = float(P.T_ref);
    o->T_sat = float(P.T_sat);
    o->h_fg = float(P.h_fg);
    o->heat_inv_norm = float(P.heat_inv_norm);
    o->dt = float(P.dt);
  })
}

int orc_diagnostics(void* s, cryo_diagnostics* o) { GUARD(static_cast<Solver*>(s)->diagnostics(o)) }

int orc_download(void* sp, int field, void* dst, int count) {
  GUARD({
    orc::Particles& ps = static_cast<Solver*>(sp)->ps;
    if (count < 0 || count > ps.n) throw std::out_of_range("orc_download: count");
    using orc::put;
    switch (field) {
      case CRYO_F_X: put(dst, ps.x, count); break;
      case CRYO_F_V: put(dst, ps.v, count); break;
      case CRYO_F_VT: put(dst, ps.vt, count); break;
      case CRYO_F_NRM: put(dst, ps.nrm, count); break;
      case CRYO_F_A: put(dst, ps.a, count); break;
      case CRYO_F_AB: put(dst, ps.ab, count); break;
      case CRYO_F_NS: put(dst, ps.ns, count); break;
      case CRYO_F_VW: put(dst, ps.vw, count); break;
      case CRYO_F_FW: put(dst, ps.fw, count); break;
      case CRYO_F_MW: put(dst, ps.mw, count); break;
      case CRYO_F_M: put(dst, ps.m, count); break;
      case CRYO_F_M0: put(dst, ps.m0, count); break;
      case CRYO_F_RHO: put(dst, ps.rho, count); break;
      case CRYO_F_T: put(dst, ps.T, count); break;
      case CRYO_F_P: put(dst, ps.p, count); break;
      case CRYO_F_DRHO: put(dst, ps.drho, count); break;
      case CRYO_F_DTDT: put(dst, ps.dTdt, count); break;
      case CRYO_F_SIGMA: put(dst, ps.sigma, count); break;
      case CRYO_F_KAPPA: put(dst, ps.kappa, count); break;
      case CRYO_F_NU_T: put(dst, ps.nu_t, count); break;
      case CRYO_F_ALPHA: put(dst, ps.alpha, count); break;
      case CRYO_F_WALL_W: put(dst, ps.wall_w, count); break;
      case CRYO_F_HEAT_W: put(dst, ps.heat_w, count); break;
      case CRYO_F_OUT_M: put(dst, ps.out_m, count); break;
      case CRYO_F_OUT_H: put(dst, ps.out_h, count); break;
      case CRYO_F_W_NORM: put(dst, ps.w_norm, count); break;
      case CRYO_F_P_WV: put(dst, ps.p_wv, count); break;
      case CRYO_F_RMIN: put(dst, ps.rmin, count); break;
      case CRYO_F_PHASE: put(dst, ps.phase, count); break;
      case CRYO_F_FLAGS: put(dst, ps.flags, count); break;
      case CRYO_F_ID: put(dst, ps.id, count); break;
      default: throw std::invalid_argument("orc_download: unknown field");
    }
  })
}

int orc_grid_info(void* s, int* dims, float* origin, float* cell, int* num_cells) {
  GUARD({
    const orc::Grid& g = static_cast<Solver*>(s)->grid;
    dims[0] = g.nx;
    dims[1] = g.ny;
    dims[2] = g.nz;
    origin[0] = float(g.origin.x);
    origin[1] = float(g.origin.y);
    origin[2] = float(g.origin.z);
    *cell = float(g.cell);
    *num_cells = g.num_cells;
  })
}
int orc_grid_tables(void* s, int* cs, int* ce) {
  GUARD({
    const orc::Grid& g = static_cast<Solver*>(s)->grid;
    std::memcpy(cs, g.cs.data(), sizeof(int) * static_cast<std::size_t>(g.num_cells));
    std::memcpy(ce, g.ce.data(), sizeof(int) * static_cast<std::size_t>(g.num_cells));
  })
}

// spatial_hash.cu:140-171
int orc_neighbor_stats(void* sp, double radius, int* counts, unsigned long long* sums) {
  GUARD({
    Solver& S = *static_cast<Solver*>(sp);
    const orc::R r2max = orc::R(radius) * orc::R(radius);
    orc::pfor(S.ps.n, [&](int i) {
      const orc::V3 xi = S.ps.x[i];
      int c = 0;
      unsigned long long q = 0;
      S.grid.each(xi, [&](int j) {
        if (orc::norm2(xi - S.ps.x[j]) < r2max) {
          ++c;
          q += S.ps.id[j];
        }
      });
      counts[i] = c;
      sums[i] = q;
    });
  })
}
int orc_neighbor_csr(void* sp, double radius, long long* offsets, int* indices) {
  GUARD({
    Solver& S = *static_cast<Solver*>(sp);
    const orc::R r2max = orc::R(radius) * orc::R(radius);
    long long k = 0;
    for (int i = 0; i < S.ps.n; ++i) {
      offsets[i] = k;
      const orc::V3 xi = S.ps.x[i];
      S.grid.each(xi, [&](int j) {
        if (orc::norm2(xi - S.ps.x[j]) < r2max) {
          if (indices) indices[k] = j;
          ++k;
        }
      });
    }
    offsets[S.ps.n] = k;
  })
}
int orc_probe_pressure(void* s, int np, const float* pts, void* out) {
  GUARD(static_cast<Solver*>(s)->probe(np, pts, static_cast<orc::R*>(out)))
}

// Standalone sort (SpatialHash::configure/build over a fresh ParticleSystem).
int orc_hash_sort(int n, const float* x, const uint8_t* phase, const float* lo, const float* hi, double cell,
                  uint32_t* perm, int* n_alive, double radius, int* counts, unsigned long long* sums) {
  GUARD({
    orc::Particles ps;
    ps.allocate(std::max<std::size_t>(static_cast<std::size_t>(n) * 2, static_cast<std::size_t>(n) + 1024));
    for (int i = 0; i < n; ++i) {
      ps.x[i] = orc::V3{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
      ps.phase[i] = phase[i];
      ps.id[i] = static_cast<uint32_t>(i);
    }
    ps.n = n;
    orc::Grid g;
    g.configure(orc::V3{lo[0], lo[1], lo[2]}, orc::V3{hi[0], hi[1], hi[2]}, orc::R(cell));
    orc::grid_build(g, ps);
    *n_alive = ps.n;
    for (int i = 0; i < ps.n; ++i) perm[i] = ps.id[i];
    if (counts) {
      const orc::R r2max = orc::R(radius) * orc::R(radius);
      for (int i = 0; i < ps.n; ++i) {
        int c = 0;
        unsigned long long q = 0;
        g.each(ps.x[i], [&](int j) {
          if (orc::norm2(ps.x[i] - ps.x[j]) < r2max) {
            ++c;
            q += ps.id[j];
          }
        });
        counts[i] = c;
        sums[i] = q;
      }
    }
  })
}

int orc_fluid_preset(const char*, cryo_fluid*) {
  orc::g_err = "oracle: presets come from the reference build or the product";
  return CRYO_ERR_RUNTIME;
}
double orc_saturation_temperature(const cryo_fluid* f, double p) {
  return double(orc::to_fluid(*f).saturation_temperature(orc::R(p)));
}
double orc_saturation_pressure(const cryo_fluid* f, double T) {
  return double(orc::to_fluid(*f).saturation_pressure(orc::R(T)));
}

}  // extern "C"

extern "C" {
// Pins the OpenMP thread count: the FP64 block sums depend on it.
void orc_set_threads(int n) { omp_set_num_threads(n); }
// Diagnostic, not part of the reference algorithm: with on != 0 every drift
// rounds the positions to float afterwards (in the FP64 build, the rest stays
// double).  Measures how much of an FP32 run's error comes from storing
// positions in float alone.
void orc_round_positions(int on) { orc::g_round_positions = on != 0; }
// Kernel value and radial derivative (SPH_kernels.cuh:14-58) for closed-form checks.
void orc_kernel_eval(int type, double h, double r, double* W, double* dWdr) {
  const orc::Kern k = orc::Kern::make(type, orc::R(h));
  *W = double(orc::kW(k, orc::R(r)));
  *dWdr = double(orc::kdW(k, orc::R(r)));
}
// The adjoint's kernel functions (csrc/adjoint_pair.hpp): W, s = W'(r) / r and
// s'(r) / r, for checks against differences of orc_kernel_eval.
void orc_adjoint_kernel_eval(int type, double h, double r, double* W, double* s, double* spr) {
  const orc::Kern k = orc::Kern::make(type, orc::R(h));
  cryo_adj::Ctx<orc::R> c{k.norm, k.inv_h, k.h, k.support * k.support, orc::R(0), orc::R(0), orc::R(0), orc::R(0), 0,
                          type, 0};
  *W = double(cryo_adj::kern_w(c, orc::R(r)));
  *s = double(cryo_adj::kern_s(c, orc::R(r)));
  *spr = double(cryo_adj::kern_spr(c, orc::R(r)));
}
}
