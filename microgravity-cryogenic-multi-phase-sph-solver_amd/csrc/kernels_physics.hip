// Neighbour-loop physics on gfx950 (one thread per target particle).
//
// The reference runs one kernel per operator (evaluate_rates,
// sph_solver.cu:449-480).  Here operators whose inputs are ready at the same
// point share one neighbour sweep:
//   k_density    density summation + EOS          (SPH_kernels.cu:6-34, navier_stokes.cu:5-14)
//   k_wall_bc    Adami wall extrapolation          (navier_stokes.cu:16-69)
//   k_interface  colour gradient, alpha, wall/heat weights, contact angle, SPS,
//                pointwise CSF normals             (surface_tension.cu:5-118, navier_stokes.cu:71-119)
//   k_forces     transport velocity, momentum with wall loads, CSF or Akinci,
//                continuity, conduction, dt maxima (SPH_kernels.cu:56-104,162-203,
//                navier_stokes.cu:121-245, surface_tension.cu:120-213,
//                phase_change.cu:31-75, sph_solver.cu:62-99)
// Every accumulator adds its neighbours in the reference order (ascending
// sorted index), with the reference's per-pair expressions, so each field is
// the same float the reference computes (given equal inputs).
#include "kernels.hpp"

namespace cryo {

namespace {

constexpr int kBlock = 256;
__device__ inline int gtid() { return static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); }
inline int blocks_for(long long n) { return static_cast<int>((n + kBlock - 1) / kBlock); }

// Deterministic FP64 block sum: fixed shuffle tree per wave, waves in order.
__device__ inline void block_sum_store(double v, double* out) {
  __shared__ double wsum[kBlock / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wsum[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int w = 0; w < kBlock / 64; ++w) s += wsum[w];
    *out = s;
  }
}

// Max of non-negative floats (NaN ignored, as `v > acc ? v : acc` does).
__device__ inline float wave_max(float v) {
  for (int off = 32; off > 0; off >>= 1) {
    const float o = __shfl_down(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}
__device__ inline void atomic_max_nonneg(unsigned* addr, float v) {
  if (v > 0.0f) atomicMax(addr, __float_as_uint(v));
}

// pow for positive float arguments: evaluated in double and rounded once.
__device__ inline float powf_cr(float x, float y) { return static_cast<float>(pow(static_cast<double>(x), static_cast<double>(y))); }

// ------------------------------------------------------- density + EOS
// Also builds the step's neighbour lists from the same (cell-culled) sweep.
template <bool KS>
__global__ __launch_bounds__(kBlock) void k_density(PV pv, GridDev g, SimParams P, NbrList L, Count nc) {
  const int n = nc.get();
  const int i = gtid();
  if (i >= n) return;
  const uint8_t ph = pv.s.phase[i];
  const V3 xi = ld3(pv.s.x[i]);
  ListWriter lw(L, i);
  if (!is_fluid(ph)) {
    pv.t.sigma[i] = 1.0f / P.vol0;
    pv.t.rmin[i] = P.kernel.support;
    // wall targets only ever need their fluid neighbours: none without a fluid
    // particle in the 27 cells
    if (ph == kWall && cells_any(g, xi, kCellFluid)) {
      for_each_cand_xph(g, xi, P.kernel.support, pv.s.x, pv.s.phase, [&](int j, float4 xj, uint8_t phj) {
        if (norm2(xi - ld3(xj)) >= P.support2) return;
        if (!is_fluid(phj)) return;
        lw.push(j);
      });
    }
    L.count[i] = lw.finish();
    return;
  }
  Acc<KS> sum;
  float r2min = P.support2;
  const V3 li = pos_lo<KS>(pv.s.xlo, i);
  for_each_cand_x(g, xi, P.kernel.support, pv.s.x, [&](int j, float4 xj) {
    // neighbourhood by the float positions (the lists of every later pass);
    // in the precision form the kernel takes the double-float distance
    const float r2h = norm2(xi - ld3(xj));
    if (r2h >= P.support2) return;
    const float r2 = KS ? norm2(pos_diff<KS>(xi, li, ld3(xj), pv.s.xlo, j)) : r2h;
    sum += kW(P.kernel, sqrtf(r2));
    if (j != i && r2 < r2min) r2min = r2;
    lw.push(j);
  });
  L.count[i] = lw.finish();
  pv.t.sigma[i] = float(sum);
  pv.t.rmin[i] = sqrtf(r2min);
  float rho = pv.s.rho[i];
  const float mi = pv.s.m[i];
  const float Ti = pv.s.T[i];
  float p;
  if constexpr (KS) {
    // precision form: the summation density in double, the EOS on it
    double rho_d = rho;
    if (P.density_mode == 0) {
      rho_d = static_cast<double>(ph == kLiquid ? pv.s.m0[i] : mi) * (static_cast<double>(sum.s) + sum.c);
      rho = static_cast<float>(rho_d);
      pv.s.rho[i] = rho;
    }
    p = eos_pressure_hp(P, ph, rho_d, Ti);
  } else {
    if (P.density_mode == 0) {
      rho = (ph == kLiquid ? pv.s.m0[i] : mi) * float(sum);
      pv.s.rho[i] = rho;
    }
    p = eos_pressure(P, ph, rho, Ti);
  }
  pv.t.p[i] = p;
  pv.t.nrec[i] = make_float4(p, mi, rho, Ti);
  pv.t.vol[i] = mi / rho;
}

// SPH_kernels.cu:36-54 (start-up only).
template <bool KS>
__global__ __launch_bounds__(kBlock) void k_normalize_masses(PV pv, GridDev g, SimParams P, int n) {
  const int i = gtid();
  if (i >= n) return;
  if (!is_fluid(pv.s.phase[i])) return;
  const V3 xi = ld3(pv.s.x[i]);
  const V3 li = pos_lo<KS>(pv.s.xlo, i);
  // the same sum the density pass forms (compensated in the precision form),
  // so that the start state sits at rho0 in either form
  Acc<KS> acc;
  for_each_nbr(g, xi, [&](int j) {
    const V3 xj = ld3(pv.s.x[j]);
    if (norm2(xi - xj) >= P.support2) return;
    acc += kW(P.kernel, sqrtf(KS ? norm2(pos_diff<KS>(xi, li, xj, pv.s.xlo, j)) : norm2(xi - xj)));
  });
  const float sum = float(acc);
  if (sum > 0.0f) {
    const float m = pv.s.rho[i] / sum;
    pv.s.m[i] = m;
    pv.s.m0[i] = m;
  }
}

// --------------------------------------------------------------- wall BC
template <bool KS>
__global__ __launch_bounds__(kBlock) void k_wall_bc(PV pv, GridDev g, SimParams P, NbrList L, const DevState* st,
                                                   Count nc) {
  const int n = nc.get();
  const int i = gtid();
  if (i >= n) return;
  if (pv.s.phase[i] != kWall) return;
  const V3 grav = st->gravity;
  const V3 xi = ld3(pv.s.x[i]);
  const V3 li = pos_lo<KS>(pv.s.xlo, i);
  Acc<KS> asw[2], asp[2], asgr[2], srho;
  AccV3<KS> sv;
  struct Nb {
    float4 x, v;
    float rho, p;
    int ph;
  };
  auto load = [&](int j) { return Nb{pv.s.x[j], pv.s.v[j], pv.s.rho[j], pv.t.p[j], pv.s.phase[j]}; };
  for_each_pair4(L, g, i, xi, load, [&](int j, const Nb d) {
    const int phj = d.ph;
    if (!is_fluid(static_cast<uint8_t>(phj))) return;
    const V3 rwf = pos_diff<KS>(xi, li, ld3(d.x), pv.s.xlo, j);
    const float r2 = norm2(rwf);
    if (r2 >= P.support2) return;
    const float w = kW(P.kernel, sqrtf(r2));
    const float rhoj = d.rho;
    const int k = pidx(static_cast<uint8_t>(phj));
    for (int q = 0; q < 2; ++q) {  // constant indices: the sums stay in registers
      if (q != k) continue;
      asw[q] += w;
      asp[q] += d.p * w;
      asgr[q] += dot(grav, rwf) * rhoj * w;
    }
    srho += rhoj * w;
    sv += ld3(d.v) * w;
  });
  const float sw[2] = {float(asw[0]), float(asw[1])};
  const float sp[2] = {float(asp[0]), float(asp[1])}, sgr[2] = {float(asgr[0]), float(asgr[1])};
  float pw[2];
  for (int k = 0; k < 2; ++k) {
    pw[k] = sw[k] > 0.0f ? (sp[k] + sgr[k]) / sw[k] : P.p_bg;
    if (P.clamp_neg && pw[k] < 0.0f) pw[k] = 0.0f;
  }
  if (!(sw[0] > 0.0f) && sw[1] > 0.0f) pw[0] = pw[1];
  if (!(sw[1] > 0.0f) && sw[0] > 0.0f) pw[1] = pw[0];
  pv.t.p[i] = pw[0];
  pv.t.p_wv[i] = pw[1];
  pv.t.alpha[i] = (sw[0] + sw[1]) > 0.0f ? sw[0] / (sw[0] + sw[1]) : 0.0f;
  const float swt = sw[0] + sw[1];
  float rho_w;
  if (swt > 0.0f) {
    rho_w = float(srho) / swt;
    pv.t.vw[i] = st3(V3(sv) * (-1.0f / swt));
  } else {
    rho_w = P.rho0[0];
    pv.t.vw[i] = st3(V3{0, 0, 0});
  }
  pv.s.rho[i] = rho_w;
  // Liquid targets' transport-velocity wall density (SPH_kernels.cu:56-104)
  // depends on this wall only: evaluated once here, not once per pair.
  float rho_tvf = rho_w;
  if (P.tvf) {
    const float ratio = (pw[0] - P.p_bg) / P.tait_B + 1.0f;
    rho_tvf = P.rho0[0] * powf_cr(ratio > 0.5f ? ratio : 0.5f, 1.0f / P.tait_gamma);
  }
  pv.t.nrec[i] = make_float4(pw[0], pw[1], rho_tvf, pv.s.T[i]);
  pv.t.a[i] = st3(V3{0, 0, 0});
}

// ------------------------------------------------------- interface fields
// What one neighbour contributes to the interface pass.  x.w holds the
// phase in bits 0-7 and the flags in bits 8-15.  Fluid: r.x = m / rho.
// Wall: r = {nrm, alpha} (alpha only when walls carry a colour).  v: the
// fluid velocity or the wall's ghost velocity (SPS only).
struct IfaceNbr {
  float4 x, r, v;
};

template <bool SPS>
__device__ inline IfaceNbr load_iface_nbr(const PV& pv, const SimParams& P, int j) {
  IfaceNbr d;
  const float4 x = pv.s.x[j];
  const int ph = pv.s.phase[j];
  const int fl = pv.s.flags[j];
  d.x = make_float4(x.x, x.y, x.z, __int_as_float(ph | (fl << 8)));
  d.v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (ph == kWall) {
    const float4 nw = pv.s.nrm[j];
    // (fluid alpha is being written by this pass: walls' only)
    d.r = make_float4(nw.x, nw.y, nw.z, P.wall_colour ? pv.t.alpha[j] : 0.0f);
    if (SPS && P.no_slip) d.v = pv.t.vw[j];
  } else {
    const float4 rec = pv.t.nrec[j];
    d.r = make_float4(pv.t.vol[j], 0.0f, 0.0f, 0.0f);  // m / rho (density pass)
    if (SPS) d.v = pv.s.v[j];
  }
  return d;
}

// compute_interface_fields (surface_tension.cu:5-85) + SPS (navier_stokes.cu:
// 71-119) + normals without smoothing (surface_tension.cu:87-118) for particle
// i; returns its heat weight (0 unless fluid).  visit(f): f(j, IfaceNbr) over
// the neighbours in the reference order.
template <bool SPS, bool KS, class Visit>
__device__ inline float interface_particle(const PV& pv, const SimParams& P, int i, int csf_mode, Visit&& visit) {
  const uint8_t ph = pv.s.phase[i];
  if (!is_fluid(ph)) {
    if (SPS) pv.t.nu_t[i] = 0.0f;
    return 0.0f;
  }
  const V3 xi = ld3(pv.s.x[i]);
  const V3 li = pos_lo<KS>(pv.s.xlo, i);
  const float ci = ph == kLiquid ? 1.0f : 0.0f;
  AccV3<KS> agradc, awall_n, aakn, agradphi;
  Acc<KS> as_all, as_other, aww, ahw;
  const V3 vi = ld3(pv.s.v[i]);
  Acc<KS> Lm[3][3];
  visit([&](int j, const IfaceNbr d) {
    const V3 rij = pos_diff<KS>(xi, li, ld3(d.x), pv.s.xlo, j);
    const float r2 = norm2(rij);
    if (r2 >= P.support2) return;
    const float r = sqrtf(r2);
    const float w = kW(P.kernel, r);
    const int bits = __float_as_int(d.x.w);
    const int phj = bits & 0xff;
    const V3 gw = kgradW(P.kernel, rij, r);
    if (SPS && j != i && !(phj == kWall && !P.no_slip)) {
      const V3 vj = ld3(d.v);
      const float Vj = phj == kWall ? P.vol0 : d.r.x;
      const V3 dv = (vj - vi) * Vj;
      const float dvc[3] = {dv.x, dv.y, dv.z};
      const float gwc[3] = {gw.x, gw.y, gw.z};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Lm[a][b] += dvc[a] * gwc[b];
    }
    if (phj == kWall) {
      const float vw = P.vol0 * w;
      aww += vw;
      if ((bits >> 8) & kHeated) ahw += vw;
      awall_n += ld3(d.r) * vw;
      agradphi += gw * P.vol0;
      if (P.wall_colour) agradc += gw * (P.vol0 * (d.r.w - ci));
      return;
    }
    const float Vj = d.r.x;
    as_all += Vj * w;
    if (phj != ph) as_other += Vj * w;
    if (j == i) return;
    const float cj = phj == kLiquid ? 1.0f : 0.0f;
    agradc += gw * (Vj * (cj - ci));
    if (phj == kLiquid) aakn += gw * Vj;
  });
  V3 gradc = agradc;
  const V3 wall_n = awall_n, akn = aakn, gradphi = agradphi;
  const float s_all = as_all, s_other = as_other, ww = aww, hw = ahw;
  if (SPS) {
    float ss = 0.0f;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        const float sab = 0.5f * (float(Lm[a][b]) + float(Lm[b][a]));
        ss += sab * sab;
      }
    const float smag = sqrtf(2.0f * ss);
    const float l = P.smag_cs * P.dx;
    pv.t.nu_t[i] = l * l * smag;
  }
  pv.t.alpha[i] = s_all > 0.0f ? s_other / s_all : 0.0f;
  pv.t.wall_w[i] = ww;
  pv.t.heat_w[i] = hw;
  if (P.st_model == 2) {
    pv.s.nrm[i] = st3(ph == kLiquid ? akn * (-P.kernel.h) : V3{0, 0, 0});
  } else {
    const float gmag = norm(gradc);
    V3 vwi{0, 0, 0};
    if (P.wetting != 0 && ww > 0.0f && gmag > P.normal_eps && norm2(wall_n) > 0.0f) {
      const V3 nw = normalized(wall_n);
      const V3 nh = gradc / gmag;
      const V3 tan_in = nh - nw * dot(nh, nw);
      const float tm = norm(tan_in);
      if (tm > 1e-6f) {
        const float cos_d = -dot(nh, nw);
        const float delta_w = 2.0f * norm(gradphi);
        const float coeff = 2.0f * P.sigma / (P.rho0[0] + P.rho0[1]);
        vwi = tan_in * (-(P.cos_theta - cos_d) * coeff * gmag * delta_w / tm);
      }
    }
    pv.t.vw[i] = st3(vwi);
    if (P.wetting != 1 && ww > 0.0f && gmag > P.normal_eps) {
      const V3 nw = normalized(wall_n);
      const V3 nh = gradc / gmag;
      V3 t = nh - nw * dot(nh, nw);
      const float tmag = norm(t);
      if (norm2(nw) > 0.0f && tmag > 1e-6f) {
        t = t / tmag;
        const V3 target = nw * (-P.cos_theta) + t * P.sin_theta;
        const float beta = clampv(ww / P.wall_adh_w, 0.0f, 1.0f);
        const V3 blended = normalized(nh * (1.0f - beta) + target * beta);
        gradc = blended * gmag;
      }
    }
    pv.s.nrm[i] = st3(gradc);
    if (csf_mode && !P.csf_smooth) {
      const float gm = norm(gradc);
      pv.t.ns[i] = st3(!(gm > P.normal_eps) ? V3{0, 0, 0} : gradc / gm);
    }
  }
  return hw;
}

// The interface pass for runs without SPS or Akinci, skipping what is
// provably zero.  A neighbour of the particle's own phase adds exact zeros
// (c_J - c_I = 0 to the colour gradient, nothing to alpha's numerator), so a
// particle with neither a wall nor the other phase in its 27 cells gets the
// outputs of an empty sum directly.  Without vapour at all, a wall within 2h
// of a liquid particle has colour alpha_w = 1 and so adds zeros too: only wall
// neighbours are summed.  Bitwise the full pass's results (adding a zero to a
// running sum that starts at +0 never changes it).  Resets the dt maxima and
// writes the heat partials as k_interface does.
template <bool KS>
__global__ __launch_bounds__(kBlock) void k_interface_sparse(PV pv, GridDev g, SimParams P, NbrList L, DevState* st,
                                                            Count nc, int csf_mode, double* heat_partials) {
  const int n = nc.get();
  const int i = gtid();
  if (i == 0) {
    for (int k = 0; k < 5; ++k) st->max_bits[k] = 0u;
  }
  double heat_contrib = 0.0;
  const uint8_t ph = i < n ? pv.s.phase[i] : kDead;
  if (is_fluid(ph)) {
    const V3 xi = ld3(pv.s.x[i]);
    if (cells_any(g, xi, kCellWall | (ph == kLiquid ? kCellVapor : kCellLiquid))) {
      if (P.has_vapor) {
        auto all = [&](auto&& f) {
          for_each_pair(L, g, i, xi, [&](int j) { f(j, load_iface_nbr<false>(pv, P, j)); });
        };
        heat_contrib = static_cast<double>(interface_particle<false, KS>(pv, P, i, csf_mode, all));
      } else {
        auto walls = [&](auto&& f) {
          for_each_pair(L, g, i, xi, [&](int j) {
            if (pv.s.phase[j] == kWall) f(j, load_iface_nbr<false>(pv, P, j));
          });
        };
        heat_contrib = static_cast<double>(interface_particle<false, KS>(pv, P, i, csf_mode, walls));
      }
    } else {
      pv.t.alpha[i] = 0.0f;
      pv.t.wall_w[i] = 0.0f;
      pv.t.heat_w[i] = 0.0f;
      pv.t.vw[i] = st3(V3{0, 0, 0});
      pv.s.nrm[i] = st3(V3{0, 0, 0});
      if (csf_mode && !P.csf_smooth) pv.t.ns[i] = st3(V3{0, 0, 0});
    }
  }
  if (heat_partials) block_sum_store(heat_contrib, heat_partials + blockIdx.x);
}

// Global-memory instance.  only_overflow = 0: every particle, resets the dt
// maxima for k_forces and writes FP64 block partials of sum_fluid heat_w
// (sph_solver.cu:101-104) when heat_partials != nullptr.  only_overflow = 1:
// just the particles whose list overflowed (the cell-tile instance does the
// rest).
template <bool SPS, bool KS>
__global__ __launch_bounds__(kBlock) void k_interface(PV pv, GridDev g, SimParams P, NbrList L, DevState* st, Count nc,
                                                     int csf_mode, double* heat_partials, int only_overflow) {
  const int n = nc.get();
  const int i = gtid();
  if (i == 0 && !only_overflow) {
    for (int k = 0; k < 5; ++k) st->max_bits[k] = 0u;
  }
  double heat_contrib = 0.0;
  if (i < n && (!only_overflow || (L.valid && L.count[i] < 0))) {
    const V3 xi = ld3(pv.s.x[i]);
    auto gathered = [&](auto&& f) {
      for_each_pair(L, g, i, xi, [&](int j) { f(j, load_iface_nbr<SPS>(pv, P, j)); });
    };
    heat_contrib = static_cast<double>(interface_particle<SPS, KS>(pv, P, i, csf_mode, gathered));
  }
  if (heat_partials) block_sum_store(heat_contrib, heat_partials + blockIdx.x);
}

// surface_tension.cu:87-118 with neighbour smoothing on.
template <bool KS>
__global__ __launch_bounds__(kBlock) void k_smooth_normals(PV pv, GridDev g, SimParams P, NbrList L, Count nc) {
  const int n = nc.get();
  const int i = gtid();
  if (i >= n) return;
  if (!is_fluid(pv.s.phase[i])) return;
  const V3 gi = ld3(pv.s.nrm[i]);
  const float gmag = norm(gi);
  if (!(gmag > P.normal_eps)) {
    pv.t.ns[i] = st3(V3{0, 0, 0});
    return;
  }
  const V3 xi = ld3(pv.s.x[i]);
  const V3 li = pos_lo<KS>(pv.s.xlo, i);
  AccV3<KS> acc;
  for_each_pair(L, g, i, xi, [&](int j) {
    if (!is_fluid(pv.s.phase[j])) return;
    const V3 gj = ld3(pv.s.nrm[j]);
    const float gjm = norm(gj);
    if (!(gjm > P.normal_eps)) return;
    const float r2 = norm2(pos_diff<KS>(xi, li, ld3(pv.s.x[j]), pv.s.xlo, j));
    if (r2 >= P.support2) return;
    acc += gj * ((pv.s.m[j] / pv.s.rho[j]) * kW(P.kernel, sqrtf(r2)) / gjm);
  });
  pv.t.ns[i] = st3(normalized(V3(acc)));
}

// ------------------------------------------------------------------ forces
// What one neighbour contributes to the forces pass.  For a wall neighbour
// `v` holds its no-slip ghost velocity vw, for a fluid one its velocity; `vt`
// is the fluid transport velocity (read only when the pass needs it).
struct NbrData {
  float4 x;
  int ph;
  float4 rec;  // {p, m, rho, T} fluid; {p, p_wv, rho_tvf, T} wall
  float4 v;
  float4 vt;
  float V;   // m / rho (fluid)
  float Vt;  // 1 / sigma (fluid, transport velocity on)
};

__device__ inline bool forces_need_vt(const SimParams& P, bool continuity) {
  return (P.tvf && P.tvf_stress) || (continuity && P.tvf);
}

// Neighbour data straight from the particle arrays.
__device__ inline NbrData load_nbr(const PV& pv, const SimParams& P, int j, bool need_vt) {
  NbrData d;
  d.x = pv.s.x[j];
  d.ph = pv.s.phase[j];
  d.rec = pv.t.nrec[j];
  d.V = pv.t.vol[j];
  const float4 v = pv.s.v[j];
  float4 vw = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (P.no_slip) vw = pv.t.vw[j];
  d.vt = v;
  if (need_vt) d.vt = pv.s.vt[j];
  float sig = 1.0f;
  if (P.tvf) sig = pv.t.sigma[j];
  // every gather of the pair in flight together, then one wait
  settle(d.x.x, d.x.y, d.x.z, d.ph, d.rec.x, d.rec.y, d.rec.z, d.rec.w, v.x, v.y, v.z, d.V);
  if (P.no_slip) settle(vw.x, vw.y, vw.z);
  if (need_vt) settle(d.vt.x, d.vt.y, d.vt.z);
  if (P.tvf) settle(sig);
  d.v = d.ph == kWall ? vw : v;
  d.Vt = P.tvf && is_fluid(d.ph) ? 1.0f / sig : 0.0f;
  return d;
}

// Momentum, wall loads, transport velocity, CSF / Akinci, continuity and
// conduction rates of particle i, and its dt maxima.  `visit(f)` calls
// f(j, NbrData) for every neighbour in the reference order.
// ST: 0 no tension pass, 1 CSF with vapour present, 2 Akinci.
struct Maxima {
  float v = 0.0f, a = 0.0f, tv = 0.0f, nu = 0.0f, sp = 0.0f;
};
template <int ST, bool CONT, bool THERMAL, bool KS, class Visit>
__device__ inline void forces_particle(const PV& pv, const SimParams& P, DevState* st, int i, Visit&& visit,
                                       Maxima& mx) {
  struct {
    bool csf_mode, akinci, continuity, thermal;
  } constexpr F{ST == 1, ST == 2, CONT, THERMAL};
  {
    const uint8_t ph = pv.s.phase[i];
    if (!is_fluid(ph)) {
      pv.t.ab[i] = st3(V3{0, 0, 0});
      pv.t.a[i] = st3(V3{0, 0, 0});
      pv.t.fw[i] = st3(V3{0, 0, 0});
      pv.t.mw[i] = st3(V3{0, 0, 0});
      if (F.continuity) pv.t.drho[i] = 0.0f;
      if (F.thermal) pv.t.dTdt[i] = 0.0f;
    } else {
      const V3 grav = st->gravity;
      const float dt_now = st->dt;
      const int pi_idx = pidx(ph);
      const V3 xi = ld3(pv.s.x[i]);
      const V3 li = pos_lo<KS>(pv.s.xlo, i);
      const V3 vi = ld3(pv.s.v[i]);
      const float mi = pv.s.m[i];
      const float rhoi = pv.s.rho[i];
      const float Ti = pv.s.T[i];
      const float sigmai = pv.t.sigma[i];
      const float Vi = mi / rhoi;
      const float Vi2 = Vi * Vi;
      const float pi = pv.t.p[i];
      const float ci = sound_speed(P, ph, Ti);
      const float h = P.kernel.h;
      const float eta2 = 0.01f * h * h;
      // transport velocity (SPH_kernels.cu:56-104)
      const bool tvf_on = P.tvf && !(sigmai < P.tvf_sigma_min) &&
                          !(P.tvf_interface_only && !(pv.t.alpha[i] > 0.0f));
      const float Vt = 1.0f / sigmai;
      const float Vt2 = Vt * Vt;
      const float mt = rhoi * Vt;
      const float pb = P.tvf_factor * sel2(P.rho0, pi_idx) * ci * ci;
      const float inv_gamma = 1.0f / P.tait_gamma;
      AccV3<KS> sum_t;
      // momentum (navier_stokes.cu:121-245)
      const float etai = sel2(P.mu, pi_idx) + (P.sps ? rhoi * pv.t.nu_t[i] : 0.0f);
      const float alpha_i = ph == kVapor ? P.alpha_av_vap : P.alpha_av;
      AccV3<KS> force, acc_av;
      const float r_guard = 0.6f * P.dx;
      const bool tvf_stress = P.tvf && P.tvf_stress;
      const V3 dvi = ld3(pv.s.vt[i]) - vi;
      AccV3<KS> fwall, mwall;
      const V3 origin = st->origin;
      // CSF (surface_tension.cu:120-177)
      const V3 ni = F.csf_mode ? ld3(pv.t.ns[i]) : V3{0, 0, 0};
      const bool csf_on = F.csf_mode && norm2(ni) > 0.25f;
      const float cci = ph == kLiquid ? 1.0f : 0.0f;
      Acc<KS> num, den;
      AccV3<KS> G;
      // Akinci (surface_tension.cu:179-213)
      const bool ak_on = F.akinci && ph == kLiquid;
      const V3 nak = ak_on ? ld3(pv.s.nrm[i]) : V3{0, 0, 0};
      const float psi = P.rho0[0] * P.vol0;
      AccV3<KS> acc_ak;
      // continuity (SPH_kernels.cu:162-203)
      const float4* vel = P.tvf ? pv.s.vt : pv.s.v;
      const V3 vci = F.continuity ? ld3(vel[i]) : V3{0, 0, 0};
      Acc<KS> div, diff;
      // conduction (phase_change.cu:31-75)
      const float ki = sel2(P.k, pi_idx);
      const bool iso = P.wall_bc == 1;
      Acc<KS> q;

      visit([&](int j, NbrData d) {
        if (j == i) return;
        const V3 xj = ld3(d.x);
        const int phj = d.ph;
        const float4 rec = d.rec;
        const V3 rij = pos_diff<KS>(xi, li, xj, pv.s.xlo, j);
        const float r2 = norm2(rij);
        if (r2 >= P.support2) return;
        const float r = sqrtf(r2);
        const float dW = kdW(P.kernel, r);
        const bool separated = r > 1e-12f * h;
        const V3 gw = separated ? rij * (dW / r) : V3{0, 0, 0};
        const float Tj = rec.w;
        if (phj == kWall) {
          const float pj = rec.x, pwv = rec.y;
          if (tvf_on) {
            float rho_w;
            if (ph == kVapor) {
              rho_w = (pwv - (P.p_bg - P.vapor_K)) / (P.gas_R * Ti);
            } else {
              rho_w = rec.z;  // the wall pass stored rho0 * max(ratio, 0.5)^(1/gamma) here
            }
            const float Vj = rho_w > 0.0f ? Vt * rhoi / rho_w : Vt;
            sum_t += gw * (Vt2 + Vj * Vj);
          }
          {
            const float Vj = P.vol0;
            const float vsum = Vi2 + Vj * Vj;
            const float pw = ph == kVapor ? pwv : pj;
            auto wall_terms = [&](auto& fo, auto& ac) {
              fo -= gw * (vsum * 0.5f * (pi + pw));
              if (tvf_stress) fo += vi * (0.5f * vsum * rhoi * dot(dvi, gw));
              if (r < r_guard && separated) {
                const float qg = 1.0f - r / r_guard;
                const V3 nn = rij / r;
                ac += nn * (4.0f * ci * ci * qg * qg / r_guard);
                const float vn = dot(vi, nn);
                if (vn < 0.0f && dt_now > 0.0f) ac -= nn * (qg * vn / dt_now);
              }
              if (P.no_slip) {
                const V3 vij = vi - ld3(d.v);
                if (separated) fo += vij * (vsum * etai * dW / r);
                const float vr = dot(vij, rij);
                if (alpha_i > 0.0f && vr < 0.0f) {
                  const float mu_ij = h * vr / (r2 + eta2);
                  const float Pi = -alpha_i * ci * mu_ij / rhoi;
                  ac -= gw * (rhoi * P.vol0 * Pi);
                }
              }
            };
            V3 dF;
            if constexpr (KS) {  // the pair's own terms, added once to each sum
              V3 df{0, 0, 0}, da{0, 0, 0};
              wall_terms(df, da);
              force += df;
              acc_av += da;
              dF = df + da * mi;
            } else {  // the reference's wall load: what the pair changed
              const V3 f_before = force, a_before = acc_av;
              wall_terms(force, acc_av);
              dF = (V3(force) - f_before) + (V3(acc_av) - a_before) * mi;
            }
            fwall += dF;
            mwall += cross(xj - origin, dF);
          }
          if (ak_on && r2 > 0.0f) {
            const float arg = -4.0f * r * r / P.kernel.support + 6.0f * r - 2.0f * P.kernel.support;
            float adh = 0.0f;
            if (!(r >= P.kernel.support || 2.0f * r <= P.kernel.support) && arg > 0.0f)
              adh = P.ak_adh_scale * powf_cr(arg, 0.25f);
            acc_ak -= (rij / r) * (P.ak_beta * psi * adh);
          }
          if (F.continuity) div += P.vol0 * dot(vci, gw);
          if (F.thermal && iso) {
            const float rgw = r * dW;
            q += Vi * P.vol0 * (2.0f * ki) * (Ti - Tj) * rgw / (r2 + eta2);
          }
          return;
        }
        // fluid neighbour
        const float pj = rec.x, mj = rec.y, rhoj = rec.z;
        if (tvf_on) {
          const float Vj = d.Vt;  // 1 / sigma_j
          sum_t += gw * (Vt2 + Vj * Vj);
        }
        const float Vj = d.V;  // mj / rhoj, divided once per particle
        const float vsum = Vi2 + Vj * Vj;
        const float pt = (rhoj * pi + rhoi * pj) / (rhoi + rhoj);
        force -= gw * (vsum * pt);
        const V3 vj = ld3(d.v);
        if (tvf_stress) {
          const V3 Ag = vi * (rhoi * dot(dvi, gw)) + vj * (rhoj * dot(ld3(d.vt) - vj, gw));
          force += Ag * (0.5f * vsum);
        }
        const V3 vij = vi - vj;
        const float etaj = sel2(P.mu, pidx(phj)) + (P.sps ? rhoj * pv.t.nu_t[j] : 0.0f);
        const float eta_sum = etai + etaj;
        if (eta_sum > 0.0f && separated) {
          const float eta_ij = P.sps ? 2.0f * etai * etaj / eta_sum
                                     : (ph == phj ? (ph == kLiquid ? P.eta_ll : P.eta_vv) : P.eta_lv);
          force += vij * (vsum * eta_ij * dW / r);
        }
        if (alpha_i > 0.0f && phj == ph) {
          const float vr = dot(vij, rij);
          if (vr < 0.0f) {
            const float cj = sound_speed(P, phj, Tj);
            const float mu_ij = h * vr / (r2 + eta2);
            const float Pi = -alpha_i * 0.5f * (ci + cj) * mu_ij / (0.5f * (rhoi + rhoj));
            acc_av -= gw * (mj * Pi);
          }
        }
        if (csf_on && is_fluid(phj)) {
          const V3 gwc = rij * (dW / r);
          const V3 nj = ld3(pv.t.ns[j]);
          if (norm2(nj) > 0.25f) {
            num += Vj * dot(ni - nj, gwc);
            den += Vj * r * (-dW);
          }
          if (P.csf_balanced) {
            const float cj = phj == kLiquid ? 1.0f : 0.0f;
            G += gwc * ((Vi * Vi + Vj * Vj) * (rhoj * cci + rhoi * cj) / (rhoi + rhoj));
          }
        }
        if (ak_on && r2 > 0.0f && phj == kLiquid) {
          const V3 rhat = rij / r;
          const float K = 2.0f * P.rho0[0] / (rhoi + rhoj);
          const V3 f_coh = rhat * (-P.ak_gamma * mi * mj * akinci_coh(r, P.kernel.support));
          const V3 f_curv = (nak - ld3(pv.s.nrm[j])) * (-P.ak_gamma * mi);
          acc_ak += (f_coh + f_curv) * (K / mi);
        }
        if (F.continuity) {
          div += Vj * dot(vci - (P.tvf ? ld3(d.vt) : vj), gw);
          if (phj == ph) diff += Vj * 2.0f * (rhoj - rhoi) * dot(-rij, gw) / (r2 + eta2);
        }
        if (F.thermal) {
          const float kj = sel2(P.k, pidx(phj));
          const float kij = 4.0f * ki * kj / (ki + kj);
          const float rgw = r * dW;
          q += Vi * Vj * kij * (Ti - Tj) * rgw / (r2 + eta2);
        }
      });

      const V3 ab = tvf_on ? V3(sum_t) * (-pb / mt) : V3{0, 0, 0};
      pv.t.ab[i] = st3(ab);
      V3 acc = grav + V3(force) / mi + V3(acc_av);
      if (st->frame_rot) {
        const V3 om = st->omega, omd = st->omega_dot;
        const V3 rr = xi - origin;
        acc -= cross(omd, rr) + cross(om, cross(om, rr)) + cross(om, vi) * 2.0f;
      }
      if (F.csf_mode) {
        if (csf_on) {
          const float kappa = float(den) > 0.0f ? 3.0f * float(num) / float(den) : 0.0f;
          pv.t.kappa[i] = kappa;
          if (P.wetting != 0) acc += ld3(pv.t.vw[i]);
          if (P.csf_balanced) {
            acc += V3(G) * (P.sigma * kappa / mi);
          } else {
            const float coeff = 2.0f * P.sigma / (P.rho0[0] + P.rho0[1]);
            acc += ld3(pv.s.nrm[i]) * (coeff * kappa);
          }
        } else {
          pv.t.kappa[i] = 0.0f;
        }
      }
      if (ak_on) acc += V3(acc_ak);
      pv.t.a[i] = st3(acc);
      pv.t.fw[i] = st3(V3(fwall));
      pv.t.mw[i] = st3(V3(mwall));
      if (F.continuity) pv.t.drho[i] = rhoi * float(div) + P.delta_sph * h * P.c0 * float(diff);
      if (F.thermal) {
        const float qt = float(q) + P.heat_power * pv.t.heat_w[i] * st->heat_inv_norm;
        pv.t.dTdt[i] = qt / (mi * sel2(P.cp, pi_idx));
      }
      // evaluate_rates maxima (sph_solver.cu:62-99)
      mx.v = norm(vi);
      const float na = norm(acc), nb = norm(ab);
      mx.a = na > nb ? na : nb;
      if (ph == kVapor) mx.tv = Ti;
      if (P.sps) mx.nu = pv.t.nu_t[i];
      const float rm = pv.t.rmin[i] > 0.05f * P.dx ? pv.t.rmin[i] : 0.05f * P.dx;
      mx.sp = 0.5f * P.dx / rm;
    }
  }
}

__device__ inline void publish_maxima(DevState* st, Maxima mx) {
  mx.v = wave_max(mx.v);
  mx.a = wave_max(mx.a);
  mx.tv = wave_max(mx.tv);
  mx.nu = wave_max(mx.nu);
  mx.sp = wave_max(mx.sp);
  if ((threadIdx.x & 63) == 0) {
    atomic_max_nonneg(&st->max_bits[0], mx.v);
    atomic_max_nonneg(&st->max_bits[1], mx.a);
    atomic_max_nonneg(&st->max_bits[2], mx.tv);
    atomic_max_nonneg(&st->max_bits[3], mx.nu);
    atomic_max_nonneg(&st->max_bits[4], mx.sp);
  }
}

// SWEEP = false: particles whose neighbour list is complete (count >= 0),
// neighbour data gathered from the particle arrays.  SWEEP = true: the
// overflowed ones (count < 0) by the cell sweep, which visits the same
// neighbours in the same order.  Two instances keep the sweep's code and
// registers out of the list loop.
template <int ST, bool CONT, bool THERMAL, bool KS, bool SWEEP>
__global__ __launch_bounds__(kBlock) void k_forces(PV pv, GridDev g, SimParams P, NbrList L, DevState* st, Count nc) {
  const int n = nc.get();
  const int i = gtid();
  Maxima mx;
  const int cnt = i < n ? L.count[i] : 0;
  if (i < n && (cnt < 0) == SWEEP) {
    const bool need_vt = forces_need_vt(P, CONT);
    const V3 xi = ld3(pv.s.x[i]);
    auto gathered = [&](auto&& f) {
      auto one = [&](int j) { f(j, load_nbr(pv, P, j, need_vt)); };
      if (SWEEP) for_each_nbr(g, xi, one);
      else for_each_listed(L, i, cnt, one);
    };
    forces_particle<ST, CONT, THERMAL, KS>(pv, P, st, i, gathered, mx);
  }
  publish_maxima(st, mx);
}

// ------------------------------------------------ forces, LDS-staged tiles
// One workgroup per tile of 3 x 3 x 3 cells.  Every neighbour of the tile's
// particles lies in the 5 x 5 x 5 cells around it: up to 25 (z, y) rows, each
// one contiguous index range of the cell-sorted arrays.  The workgroup copies
// those rows' neighbour data into LDS once (coalesced), then each thread
// walks its particle's neighbour list reading LDS.  List entries ascend, so a
// forward-moving row cursor maps index j to its LDS slot.  A tile whose rows
// hold more than kTileCap particles gathers from global memory instead;
// overflowed lists (count < 0) are left to the sweep instance of k_forces.
// The tile shape (g.tx x g.ty x g.tz cells, 3 x 3 x 3 by default) is a launch
// parameter: any shape with (ty + 2)(tz + 2) <= kTileRows staged rows and
// ty tz <= kTileTargetRows target rows (tile_shape_ok) gives the same results.
constexpr int kTileThreads = 512;
constexpr int kTileCap = 2304;
constexpr int kTileRows = 25;
constexpr int kTileTargetRows = 9;

struct TileRows {
  int start[kTileRows];     // first sorted index of each staged row
  int base[kTileRows + 1];  // its first LDS slot (exclusive prefix of row lengths)
  int rend[kTileRows];      // start + length: first sorted index after the row
  int roff[kTileRows];      // base - start: slot = j + roff for j in the row
  int nrows;
  int tstart[kTileTargetRows];  // the tile's own rows (targets)
  int tbase[kTileTargetRows + 1];
  int ntrows;
};

// Block b of an nblocks grid -> the tile it runs.  The hardware deals
// workgroups round-robin to the 8 XCDs (XCD b % 8); within each group of
// 8 x kTileChunk blocks, XCD x runs the chunk of kTileChunk consecutive tiles
// [x kTileChunk, (x + 1) kTileChunk), so neighbouring tiles, whose staged halos
// overlap, meet in one L2 while the chunks stay spread over the XCDs.  The
// last, partial group keeps the identity.  A bijection of [0, nblocks).
constexpr int kTileChunk = 16;
__device__ inline int xcd_tile(int b, int nblocks) {
  constexpr int span = 8 * kTileChunk;
  if (b >= nblocks / span * span) return b;
  return b / span * span + (b % 8) * kTileChunk + (b / 8) % kTileChunk;
}

// Number of cell tiles (workgroups) of the tiled passes over grid g.
inline long long tile_count(const GridDev& g) {
  if (!tile_shape_ok(g.tx, g.ty, g.tz)) throw std::logic_error("cell-tile shape out of range");
  return static_cast<long long>((g.nx + g.tx - 1) / g.tx) * ((g.ny + g.ty - 1) / g.ty) * ((g.nz + g.tz - 1) / g.tz);
}

// Row tables of tile `b`; called by the whole block, ends with a barrier.
__device__ inline void tile_rows(const GridDev& g, int b, TileRows& T) {
  const int tiles_x = (g.nx + g.tx - 1) / g.tx;
  const int tiles_y = (g.ny + g.ty - 1) / g.ty;
  const int cx0 = (b % tiles_x) * g.tx;
  const int cy0 = ((b / tiles_x) % tiles_y) * g.ty;
  const int cz0 = (b / (tiles_x * tiles_y)) * g.tz;
  const int cx1 = min(cx0 + g.tx, g.nx), cy1 = min(cy0 + g.ty, g.ny), cz1 = min(cz0 + g.tz, g.nz);
  const int sx0 = max(cx0 - 1, 0), sx1 = min(cx1, g.nx - 1);
  const int sy0 = max(cy0 - 1, 0), sy1 = min(cy1, g.ny - 1);
  const int sz0 = max(cz0 - 1, 0), sz1 = min(cz1, g.nz - 1);
  const int nyr = sy1 - sy0 + 1;
  const int nrows = (sz1 - sz0 + 1) * nyr;
  const int tyr = cy1 - cy0;
  const int ntrows = (cz1 - cz0) * tyr;
  const int t = threadIdx.x;
  // row lengths in parallel: threads [0, nrows) staged rows (nrows <= 25),
  // [32, 32 + ntrows) target rows (ntrows <= 9)
  if (t < nrows) {
    const int row = ((sz0 + t / nyr) * g.ny + (sy0 + t % nyr)) * g.nx;
    const int s = g.off[row + sx0];
    T.start[t] = s;
    T.base[t + 1] = g.off[row + sx1 + 1] - s;
  } else if (t >= 32 && t < 32 + ntrows) {
    const int k = t - 32;
    const int row = ((cz0 + k / tyr) * g.ny + (cy0 + k % tyr)) * g.nx;
    const int s = g.off[row + cx0];
    T.tstart[k] = s;
    T.tbase[k + 1] = g.off[row + cx1] - s;
  }
  __syncthreads();
  if (t == 0) {
    T.base[0] = 0;
    for (int r = 0; r < nrows; ++r) {
      const int len = T.base[r + 1];
      T.base[r + 1] += T.base[r];
      T.rend[r] = T.start[r] + len;
      T.roff[r] = T.base[r] - T.start[r];
    }
    T.nrows = nrows;
  } else if (t == 64) {
    T.tbase[0] = 0;
    for (int k = 0; k < ntrows; ++k) T.tbase[k + 1] += T.tbase[k];
    T.ntrows = ntrows;
  }
  __syncthreads();
}

// Maps a particle's ascending list entries to LDS slots: the row of the
// current entry is kept with its end and offset, and only an entry past the
// row's end moves it forward.
struct RowCursor {
  const TileRows& T;
  int r, rend, roff;
  __device__ explicit RowCursor(const TileRows& t) : T(t), r(0), rend(t.rend[0]), roff(t.roff[0]) {}
  __device__ inline int slot(int j) {
    while (j >= rend && r < T.nrows - 1) {
      ++r;
      rend = T.rend[r];
      roff = T.roff[r];
    }
    return j + roff;
  }
};

// Sorted index of LDS slot s of a tile (s < the tile's staged count).
__device__ inline int tile_slot_index(const TileRows& T, int s) {
  int lo = 0, hi = T.nrows - 1;  // last row with base <= s
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (T.base[mid] <= s) lo = mid;
    else hi = mid - 1;
  }
  return T.start[lo] + (s - T.base[lo]);
}
// Slots a thread stages: every load of all of them is issued before the
// first one is waited for (one memory round trip per staging pass).
constexpr int kStageSlots = (kTileCap + kTileThreads - 1) / kTileThreads;

// One round of a tile's targets (round * kTileThreads + threadIdx.x): the
// fluid particles with a complete list go, in target order, to
// work[0, returned count); `other(i)` is called for the rest with a complete
// list (non-fluid: their zero outputs).  Overflowed lists are skipped (their
// own instance handles them).  Block-wide: every thread calls it.
template <class Other>
__device__ inline int tile_round_fluid(const PV& pv, const NbrList& L, const TileRows& T, int ntarget, int round,
                                       int* work, int* wave_tot, Other&& other) {
  const int t = round * kTileThreads + static_cast<int>(threadIdx.x);
  int i = -1;
  bool keep = false;
  if (t < ntarget) {
    int k = 0;
    while (t >= T.tbase[k + 1]) ++k;
    i = T.tstart[k] + (t - T.tbase[k]);
    if (L.count[i] >= 0) {
      if (is_fluid(pv.s.phase[i])) keep = true;
      else other(i);
    }
  }
  const unsigned long long mask = __ballot(keep);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  if (lane == 0) wave_tot[wave] = __popcll(mask);
  __syncthreads();
  int base = 0, total = 0;
  for (int w = 0; w < kTileThreads / 64; ++w) {
    const int c = wave_tot[w];
    base += w < wave ? c : 0;
    total += c;
  }
  if (keep) work[base + __popcll(mask & lt)] = i;
  __syncthreads();
  return total;
}

template <int ST, bool CONT, bool THERMAL, bool KS>
__global__ __launch_bounds__(kTileThreads) void k_forces_tiled(PV pv, GridDev g, SimParams P, NbrList L,
                                                             DevState* st) {
  __shared__ TileRows T;
  __shared__ float4 s_x[kTileCap], s_rec[kTileCap], s_v[kTileCap], s_vt[kTileCap];
  __shared__ float s_V[kTileCap];
  __shared__ int s_work[kTileThreads], s_wave[kTileThreads / 64];
  tile_rows(g, xcd_tile(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x)), T);
  const int ntarget = T.tbase[T.ntrows];
  if (ntarget == 0) return;  // uniform over the block
  const int M = T.base[T.nrows];
  const bool staged = M <= kTileCap;
  const bool need_vt = forces_need_vt(P, CONT);
  Maxima mx;
  // Fluid targets packed onto consecutive threads (walls only write zeros):
  // no lane idles through a pair loop for a wall.  The first round is
  // packed before staging, so a tile of walls alone stages nothing.
  auto none = [](auto&&) {};
  auto other = [&](int i) { forces_particle<ST, CONT, THERMAL, KS>(pv, P, st, i, none, mx); };
  const int nfl0 = tile_round_fluid(pv, L, T, ntarget, 0, s_work, s_wave, other);
  if (nfl0 == 0 && ntarget <= kTileThreads) return;  // uniform over the block
  if (staged) {
    // Named slots, not arrays (arrays of them are not kept in registers).
    struct Slot {
      float4 x, r, v, t;
      int ph;
      float V;
      float sig;
    };
    auto load = [&](int u) {
      Slot q{};
      const int s = threadIdx.x + u * kTileThreads;
      if (s < M) {
        const int j = tile_slot_index(T, s);
        q.x = pv.s.x[j];
        q.ph = pv.s.phase[j];
        q.r = pv.t.nrec[j];
        q.V = pv.t.vol[j];
        const float4 v = pv.s.v[j];
        float4 w = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (P.no_slip) w = pv.t.vw[j];
        float4 t = v;
        if (need_vt) t = pv.s.vt[j];
        q.sig = 1.0f;
        if (P.tvf) q.sig = pv.t.sigma[j];
        q.v = q.ph == kWall ? w : v;
        q.t = t;
      }
      return q;
    };
    auto store = [&](int u, const Slot q) {
      const int s = threadIdx.x + u * kTileThreads;
      if (s < M) {
        s_x[s] = make_float4(q.x.x, q.x.y, q.x.z, __int_as_float(q.ph));
        s_rec[s] = q.r;
        // the velocity's free fourth component carries the neighbour's
        // transport volume 1 / sigma, divided once per staged particle
        s_v[s] = make_float4(q.v.x, q.v.y, q.v.z, P.tvf && is_fluid(q.ph) ? 1.0f / q.sig : 0.0f);
        s_V[s] = q.V;
        if (need_vt) s_vt[s] = q.t;
      }
    };
    static_assert(kStageSlots == 5, "staging is written out for five slots per thread");
    const Slot q0 = load(0), q1 = load(1), q2 = load(2), q3 = load(3), q4 = load(4);
    store(0, q0);
    store(1, q1);
    store(2, q2);
    store(3, q3);
    store(4, q4);
  }
  __syncthreads();
  for (int round = 0; round * kTileThreads < ntarget; ++round) {
    const int nfl = round == 0 ? nfl0 : tile_round_fluid(pv, L, T, ntarget, round, s_work, s_wave, other);
    if (static_cast<int>(threadIdx.x) >= nfl) continue;
    const int i = s_work[threadIdx.x];
    const int cnt = L.count[i];
    if (staged) {
      // Four neighbours at a time: their LDS records are read into registers
      // (occupancy is set by LDS here, so registers are free) before the four
      // pair bodies run, and the next four list entries are in flight.
      auto from_lds = [&](auto&& f) {
        RowCursor rc(T);
        const int4* lp = list_chunks(L, i);
        const int nq = (cnt + 3) >> 2;
        int4 cur = nq > 0 ? lp[0] : make_int4(0, 0, 0, 0);
        for (int q = 0; q < nq; ++q) {
          int4 nxt = cur;
          if (q + 1 < nq) nxt = lp[(q + 1) * kListGroup];
          const int m = cnt - (q << 2);
          const int js[4] = {cur.x, cur.y, cur.z, cur.w};
          // the Akinci bodies hold more per-pair state: two records at a
          // time (one with compensated sums) keep them within the register
          // file (no scratch)
          constexpr int kAhead = ST == 2 ? (KS ? 1 : 2) : 4;
#pragma unroll
          for (int h0 = 0; h0 < 4; h0 += kAhead) {
            NbrData d[kAhead];
#pragma unroll
            for (int u = 0; u < kAhead; ++u) {
              if (h0 + u < m) {
                const int s = rc.slot(js[h0 + u]);
                const float4 xs = s_x[s], rs = s_rec[s], vs = s_v[s];
                float4 vts = vs;
                if (need_vt) vts = s_vt[s];
                d[u] = NbrData{xs, __float_as_int(xs.w), rs, vs, vts, s_V[s], vs.w};
              }
            }
#pragma unroll
            for (int u = 0; u < kAhead; ++u)
              if (h0 + u < m) f(js[h0 + u], d[u]);
          }
          cur = nxt;
        }
      };
      forces_particle<ST, CONT, THERMAL, KS>(pv, P, st, i, from_lds, mx);
    } else {
      auto gathered = [&](auto&& f) {
        for_each_listed(L, i, cnt, [&](int j) { f(j, load_nbr(pv, P, j, need_vt)); });
      };
      forces_particle<ST, CONT, THERMAL, KS>(pv, P, st, i, gathered, mx);
    }
  }
  publish_maxima(st, mx);
}

// ---------------------------------------------- interface, LDS-staged tiles
// Same tiles and row tables as k_forces_tiled.  Staged per particle: x with
// phase and flags bits, and {m/rho} (fluid) or {nrm, alpha} (wall): 32 B, so
// two workgroups fit a CU (48 B and one workgroup with SPS).  Wall normals
// and colours are staged for walls only: the pass writes the fluid ones.
// Writes heat_w, not its sum (k_heat_partials sums it).
template <bool SPS, bool KS>
__global__ __launch_bounds__(kTileThreads) void k_interface_tiled(PV pv, GridDev g, SimParams P, NbrList L,
                                                                DevState* st, int csf_mode) {
  __shared__ TileRows T;
  __shared__ float4 s_x[kTileCap], s_r[kTileCap], s_v[SPS ? kTileCap : 1];
  __shared__ int s_work[kTileThreads], s_wave[kTileThreads / 64];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int k = 0; k < 5; ++k) st->max_bits[k] = 0u;
  }
  tile_rows(g, blockIdx.x, T);
  const int ntarget = T.tbase[T.ntrows];
  if (ntarget == 0) return;  // uniform over the block
  const int M = T.base[T.nrows];
  const bool staged = M <= kTileCap;
  // the first round's fluid targets are packed before staging: a tile of
  // walls alone stages nothing
  auto none = [](auto&&) {};
  auto other = [&](int i) { interface_particle<SPS, KS>(pv, P, i, csf_mode, none); };
  const int nfl0 = tile_round_fluid(pv, L, T, ntarget, 0, s_work, s_wave, other);
  if (nfl0 == 0 && ntarget <= kTileThreads) return;  // uniform over the block
  if (staged) {
    // Named slots, not arrays (arrays of them are not kept in registers).
    // First pass: position, phase, flags; second: the record the phase
    // selects (a wall's normal and colour, a fluid particle's m and rho).
    struct Slot {
      float4 x, r, v;
      int j, bits;
      float a;
    };
    auto load1 = [&](int u) {
      Slot q{};
      const int s = threadIdx.x + u * kTileThreads;
      if (s < M) {
        q.j = tile_slot_index(T, s);
        q.x = pv.s.x[q.j];
        q.bits = static_cast<int>(pv.s.phase[q.j]) | (static_cast<int>(pv.s.flags[q.j]) << 8);
      }
      return q;
    };
    auto load2 = [&](int u, Slot q) {
      const int s = threadIdx.x + u * kTileThreads;
      if (s < M) {
        const bool wall = (q.bits & 0xff) == kWall;
        // (fluid normals and colours are being written by this pass)
        const float4* src = wall ? pv.s.nrm : pv.t.nrec;  // two global arrays: a pointer select is fine
        q.r = src[q.j];
        q.a = (wall && P.wall_colour) ? pv.t.alpha[q.j] : 0.0f;
        q.v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (SPS && (!wall || P.no_slip)) q.v = (wall ? pv.t.vw : pv.s.v)[q.j];
      }
      return q;
    };
    auto store = [&](int u, const Slot q) {
      const int s = threadIdx.x + u * kTileThreads;
      if (s < M) {
        const bool wall = (q.bits & 0xff) == kWall;
        s_x[s] = make_float4(q.x.x, q.x.y, q.x.z, __int_as_float(q.bits));
        s_r[s] = wall ? make_float4(q.r.x, q.r.y, q.r.z, q.a) : make_float4(q.r.y / q.r.z, 0.0f, 0.0f, 0.0f);
        if (SPS) s_v[s] = q.v;
      }
    };
    static_assert(kStageSlots == 5, "staging is written out for five slots per thread");
    Slot q0 = load1(0), q1 = load1(1), q2 = load1(2), q3 = load1(3), q4 = load1(4);
    q0 = load2(0, q0);
    q1 = load2(1, q1);
    q2 = load2(2, q2);
    q3 = load2(3, q3);
    q4 = load2(4, q4);
    store(0, q0);
    store(1, q1);
    store(2, q2);
    store(3, q3);
    store(4, q4);
  }
  __syncthreads();
  for (int round = 0; round * kTileThreads < ntarget; ++round) {
    const int nfl = round == 0 ? nfl0 : tile_round_fluid(pv, L, T, ntarget, round, s_work, s_wave, other);
    if (static_cast<int>(threadIdx.x) >= nfl) continue;
    const int i = s_work[threadIdx.x];
    const int cnt = L.count[i];
    if (staged) {
      auto from_lds = [&](auto&& f) {
        RowCursor rc(T);
        for_each_listed(L, i, cnt, [&](int j) {
          const int s = rc.slot(j);
          const float4 xs = s_x[s], rs = s_r[s];
          float4 vs = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          if (SPS) vs = s_v[s];
          f(j, IfaceNbr{xs, rs, vs});
        });
      };
      interface_particle<SPS, KS>(pv, P, i, csf_mode, from_lds);
    } else {
      auto gathered = [&](auto&& f) {
        for_each_listed(L, i, cnt, [&](int j) { f(j, load_iface_nbr<SPS>(pv, P, j)); });
      };
      interface_particle<SPS, KS>(pv, P, i, csf_mode, gathered);
    }
  }
}

// FP64 block partials of sum_fluid heat_w (sph_solver.cu:101-104).
__global__ __launch_bounds__(kBlock) void k_heat_partials(PV pv, Count nc, double* partials) {
  const int n = nc.get();
  const int i = gtid();
  double v = 0.0;
  if (i < n && is_fluid(pv.s.phase[i])) v = static_cast<double>(pv.t.heat_w[i]);
  block_sum_store(v, partials + blockIdx.x);
}

// Publishes the maxima (spacing ratio starts from 1, sph_solver.cu:62-71).
__global__ void k_publish_maxima(DevState* st, int has_vapor, int sps) {
  st->v_max = __uint_as_float(st->max_bits[0]);
  st->a_max = __uint_as_float(st->max_bits[1]);
  st->Tv_max = has_vapor ? __uint_as_float(st->max_bits[2]) : 0.0f;
  st->nu_t_max = sps ? __uint_as_float(st->max_bits[3]) : 0.0f;
  const float sp = __uint_as_float(st->max_bits[4]);
  st->spacing_ratio = sp > 1.0f ? sp : 1.0f;
}

// sph_solver.cu:466-472: heat_inv_norm from the FP64 partials.
__global__ __launch_bounds__(1024) void k_heat_norm(DevState* st, const double* partials, int nb, int active) {
  __shared__ double s[1024];
  double acc = 0.0;
  for (int k = threadIdx.x; k < nb; k += 1024) acc += partials[k];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (static_cast<int>(threadIdx.x) < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double S = s[0];
    st->heat_inv_norm = active ? (S > 0 ? static_cast<float>(1.0 / S) : 0.0f) : 0.0f;
  }
}

// probe_liquid_pressure (sph_solver.cu:197-216)
__global__ __launch_bounds__(kBlock) void k_probe(PV pv, GridDev g, SimParams P, const float* pts, float* out,
                                                 int np) {
  const int k = gtid();
  if (k >= np) return;
  const V3 x{pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]};
  float num = 0.0f, den = 0.0f;
  for_each_nbr(g, x, [&](int j) {
    if (pv.s.phase[j] != kLiquid) return;
    const float r2 = norm2(x - ld3(pv.s.x[j]));
    if (r2 >= P.support2) return;
    const float w = (pv.s.m[j] / pv.s.rho[j]) * kW(P.kernel, sqrtf(r2));
    num += pv.t.p[j] * w;
    den += w;
  });
  out[k] = den > 0.0f ? num / den : __builtin_nanf("");
}

// SPH_kernels.cu:106-160 (delta+ shifting), pass 1 and pass 2.
template <bool KS>
__global__ __launch_bounds__(kBlock) void k_shift_compute(PV pv, GridDev g, SimParams P, NbrList L,
                                                         const DevState* st, Count nc) {
  const int n = nc.get();
  const int i = gtid();
  if (i >= n) return;
  pv.t.ns[i] = st3(V3{0, 0, 0});
  const uint8_t ph = pv.s.phase[i];
  if (!is_fluid(ph) || pv.t.sigma[i] < P.tvf_sigma_min) return;
  const float h = P.kernel.h;
  const float w_dx = kW(P.kernel, P.dx);
  const float mach = st->v_max / P.c0;
  const float coeff = P.shift_strength * mach * 4.0f * h * h;
  const float cap = 0.1f * P.dx;
  const V3 xi = ld3(pv.s.x[i]);
  const V3 li = pos_lo<KS>(pv.s.xlo, i);
  AccV3<KS> sum;
  for_each_pair(L, g, i, xi, [&](int j) {
    if (j == i) return;
    const uint8_t phj = pv.s.phase[j];
    if (phj == kDead) return;
    const V3 rij = pos_diff<KS>(xi, li, ld3(pv.s.x[j]), pv.s.xlo, j);
    const float r2 = norm2(rij);
    if (r2 >= P.support2) return;
    const float r = sqrtf(r2);
    const float Vj = phj == kWall ? 1.0f / pv.t.sigma[i] : 1.0f / pv.t.sigma[j];
    const float qq = kW(P.kernel, r) / w_dx;
    const float q2 = qq * qq;
    sum += kgradW(P.kernel, rij, r) * (Vj * (1.0f + 0.2f * q2 * q2));
  });
  V3 dr = V3(sum) * (-coeff);
  if (pv.t.alpha[i] > 0.0f) {
    const V3 gn = ld3(pv.s.nrm[i]);
    const float gm = norm(gn);
    if (gm > P.normal_eps) {
      const V3 nn = gn / gm;
      dr -= nn * dot(dr, nn);
    } else {
      dr = V3{0, 0, 0};
    }
  }
  const float len = norm(dr);
  if (len > cap) dr *= cap / len;
  pv.t.ns[i] = st3(dr);
}

__global__ __launch_bounds__(kBlock) void k_shift_apply(PV pv, Count nc) {
  const int n = nc.get();
  const int i = gtid();
  if (i >= n) return;
  if (!is_fluid(pv.s.phase[i])) return;
  if (pv.s.xlo) {  // precision form: the shift into the double-float position
    V3 x = ld3(pv.s.x[i]), lo = ld3(pv.s.xlo[i]);
    const V3 d = ld3(pv.t.ns[i]);
    df_add(x.x, lo.x, d.x);
    df_add(x.y, lo.y, d.y);
    df_add(x.z, lo.z, d.z);
    pv.s.x[i] = st3(x);
    pv.s.xlo[i] = st3(lo);
  } else {
    pv.s.x[i] = st3(ld3(pv.s.x[i]) + ld3(pv.t.ns[i]));
  }
}

}  // namespace

void launch_density(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, Count n, hipStream_t s) {
  if (n.ub <= 0) return;
  if (P.compensated) k_density<true><<<blocks_for(n.ub), kBlock, 0, s>>>(pv, g, P, L, n);
  else k_density<false><<<blocks_for(n.ub), kBlock, 0, s>>>(pv, g, P, L, n);
}
void launch_normalize_masses(const PV& pv, const GridDev& g, const SimParams& P, int n, hipStream_t s) {
  if (n <= 0) return;
  if (P.compensated) k_normalize_masses<true><<<blocks_for(n), kBlock, 0, s>>>(pv, g, P, n);
  else k_normalize_masses<false><<<blocks_for(n), kBlock, 0, s>>>(pv, g, P, n);
}
void launch_wall_bc(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, const DevState* st, Count n,
                    hipStream_t s) {
  if (n.ub <= 0) return;
  if (P.compensated) k_wall_bc<true><<<blocks_for(n.ub), kBlock, 0, s>>>(pv, g, P, L, st, n);
  else k_wall_bc<false><<<blocks_for(n.ub), kBlock, 0, s>>>(pv, g, P, L, st, n);
}
template <bool SPS, bool KS>
static void interface_variant(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, DevState* st,
                              Count n, int cm, double* heat_partials, bool tiled, hipStream_t s) {
  const int nb = blocks_for(n.ub > 0 ? n.ub : 1);
  if (!tiled) {
    k_interface<SPS, KS><<<nb, kBlock, 0, s>>>(pv, g, P, L, st, n, cm, heat_partials, 0);
    return;
  }
  const long long tiles = tile_count(g);
  if (tiles > 0x7fffffffLL) throw std::runtime_error("interface pass: too many cell tiles");
  k_interface_tiled<SPS, KS><<<static_cast<int>(tiles), kTileThreads, 0, s>>>(pv, g, P, L, st, cm);
  k_interface<SPS, KS><<<nb, kBlock, 0, s>>>(pv, g, P, L, st, n, cm, nullptr, 1);
  if (heat_partials) k_heat_partials<<<nb, kBlock, 0, s>>>(pv, n, heat_partials);
}
int launch_interface(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, DevState* st, Count n,
                     bool csf_mode, double* heat_partials, bool tiled, hipStream_t s) {
  const int nb = blocks_for(n.ub > 0 ? n.ub : 1);
  const int cm = csf_mode ? 1 : 0;
  if (P.st_model != 2 && !P.sps && L.valid && g.flags) {
    if (P.compensated) k_interface_sparse<true><<<nb, kBlock, 0, s>>>(pv, g, P, L, st, n, cm, heat_partials);
    else k_interface_sparse<false><<<nb, kBlock, 0, s>>>(pv, g, P, L, st, n, cm, heat_partials);
    return nb;
  }
  tiled = tiled && L.valid;
  if (P.compensated) {
    if (P.sps) interface_variant<true, true>(pv, g, P, L, st, n, cm, heat_partials, tiled, s);
    else interface_variant<false, true>(pv, g, P, L, st, n, cm, heat_partials, tiled, s);
  } else {
    if (P.sps) interface_variant<true, false>(pv, g, P, L, st, n, cm, heat_partials, tiled, s);
    else interface_variant<false, false>(pv, g, P, L, st, n, cm, heat_partials, tiled, s);
  }
  return nb;
}
void launch_smooth_normals(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, Count n,
                           hipStream_t s) {
  if (n.ub <= 0) return;
  if (P.compensated) k_smooth_normals<true><<<blocks_for(n.ub), kBlock, 0, s>>>(pv, g, P, L, n);
  else k_smooth_normals<false><<<blocks_for(n.ub), kBlock, 0, s>>>(pv, g, P, L, n);
}
void launch_heat_norm(DevState* st, const double* partials, int nb, bool active, hipStream_t s) {
  k_heat_norm<<<1, 1024, 0, s>>>(st, partials, nb, active ? 1 : 0);
}

template <int ST, bool CONT, bool THERMAL, bool KS>
static void forces_variant(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, DevState* st, Count n,
                           bool tiled, hipStream_t s) {
  if (!L.valid) throw std::logic_error("forces pass: neighbour lists not built");
  if (tiled) {
    const long long tiles = tile_count(g);
    if (tiles > 0x7fffffffLL) throw std::runtime_error("forces pass: too many cell tiles");
    k_forces_tiled<ST, CONT, THERMAL, KS><<<static_cast<int>(tiles), kTileThreads, 0, s>>>(pv, g, P, L, st);
  } else {
    k_forces<ST, CONT, THERMAL, KS, false><<<blocks_for(n.ub > 0 ? n.ub : 1), kBlock, 0, s>>>(pv, g, P, L, st, n);
  }
  k_forces<ST, CONT, THERMAL, KS, true><<<blocks_for(n.ub > 0 ? n.ub : 1), kBlock, 0, s>>>(pv, g, P, L, st, n);
}
template <int ST, bool KS>
static void forces_st(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, DevState* st, Count n,
                      bool continuity, bool thermal, bool tiled, hipStream_t s) {
  if (continuity) {
    if (thermal) forces_variant<ST, true, true, KS>(pv, g, P, L, st, n, tiled, s);
    else forces_variant<ST, true, false, KS>(pv, g, P, L, st, n, tiled, s);
  } else {
    if (thermal) forces_variant<ST, false, true, KS>(pv, g, P, L, st, n, tiled, s);
    else forces_variant<ST, false, false, KS>(pv, g, P, L, st, n, tiled, s);
  }
}
template <bool KS>
static void forces_ks(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, DevState* st, Count n,
                      bool csf_mode, bool akinci, bool continuity, bool thermal, bool tiled, hipStream_t s) {
  if (akinci) forces_st<2, KS>(pv, g, P, L, st, n, continuity, thermal, tiled, s);
  else if (csf_mode) forces_st<1, KS>(pv, g, P, L, st, n, continuity, thermal, tiled, s);
  else forces_st<0, KS>(pv, g, P, L, st, n, continuity, thermal, tiled, s);
}
void launch_forces(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, DevState* st, Count n,
                   bool csf_mode, bool akinci, bool continuity, bool thermal, bool tiled, hipStream_t s) {
  if (P.compensated) forces_ks<true>(pv, g, P, L, st, n, csf_mode, akinci, continuity, thermal, tiled, s);
  else forces_ks<false>(pv, g, P, L, st, n, csf_mode, akinci, continuity, thermal, tiled, s);
  k_publish_maxima<<<1, 1, 0, s>>>(st, P.has_vapor, P.sps);
}
void launch_probe(const PV& pv, const GridDev& g, const SimParams& P, const float* pts, float* out, int np,
                  hipStream_t s) {
  if (np > 0) k_probe<<<blocks_for(np), kBlock, 0, s>>>(pv, g, P, pts, out, np);
}
void launch_shifting(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, const DevState* st, Count n,
                     hipStream_t s) {
  if (n.ub <= 0) return;
  if (P.compensated) k_shift_compute<true><<<blocks_for(n.ub), kBlock, 0, s>>>(pv, g, P, L, st, n);
  else k_shift_compute<false><<<blocks_for(n.ub), kBlock, 0, s>>>(pv, g, P, L, st, n);
  k_shift_apply<<<blocks_for(n.ub), kBlock, 0, s>>>(pv, n);
}

}  // namespace cryo
