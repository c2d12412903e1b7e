// Step orchestration kernels on gfx950: time step and frame on the device,
// kick, phase change (Lee sources, receiver normalisation, apply, vapour
// rebalance, split, vent), FP64 bookkeeping and the ullage update.
//
// The reference synchronises with the host about 13 times per step
// (sph_solver.cu:449-607); here dt, the frame, the bookkeeping sums and the
// ullage state stay in a device-resident DevState, and every FP64 sum is a
// fixed-order two-level reduction (block partials, then one ordered block),
// so runs are bitwise reproducible.
#include "kernels.hpp"

namespace cryo {

namespace {

constexpr int kBlock = 256;
__device__ inline int gtid() { return static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); }
inline int blocks_for(long long n) { return static_cast<int>((n + kBlock - 1) / kBlock); }

template <int NQ>
__device__ inline void block_sums_store(double (&v)[NQ], double* partials, int stride) {
  __shared__ double wsum[NQ][kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double x = v[q];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) wsum[q][wave] = x;
  }
  __syncthreads();
  if (threadIdx.x < NQ) {
    double s = 0;
    for (int w = 0; w < kBlock / 64; ++w) s += wsum[threadIdx.x][w];
    partials[threadIdx.x * stride + blockIdx.x] = s;
  }
}

// Ordered sum of partials[q*stride + 0..nb) by one block of 1024 threads.
__device__ inline double ordered_sum(const double* p, int nb, double* sh) {
  double acc = 0.0;
  for (int k = threadIdx.x; k < nb; k += 1024) acc += p[k];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (static_cast<int>(threadIdx.x) < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// phase_change.cu:7-27, with cos/sin/cbrt evaluated in double and rounded once.
__device__ inline uint32_t hash_u32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}
__device__ inline float hash_unit(uint32_t x) {
  return static_cast<float>(hash_u32(x) & 0x00FFFFFFu) / static_cast<float>(0x01000000u);
}
__device__ inline V3 random_direction(uint32_t id, uint32_t seed) {
  const float u = 2.0f * hash_unit(id * 2654435761u ^ seed) - 1.0f;
  const float phi = 2.0f * kPi * hash_unit(hash_u32(id + 0x9e3779b9u) ^ (seed * 747796405u));
  const float s = sqrtf(1.0f - u * u > 0.0f ? 1.0f - u * u : 0.0f);
  const float c = static_cast<float>(cos(static_cast<double>(phi)));
  const float sn = static_cast<float>(sin(static_cast<double>(phi)));
  return V3{s * c, s * sn, u};
}
__device__ inline float cbrtf_cr(float x) { return static_cast<float>(cbrt(static_cast<double>(x))); }

// ------------------------------------------------------------- begin step
// dt from the previous maxima (sph_solver.cu:482-504) and the frame at the new
// time level (sph_solver.cu:518-522, ManeuverProfile::evaluate :244-268).
__global__ void k_begin_step(DevState* st, SimParams P, DtLimits L, Maneuver M) {
  const float dt = compute_dt(P, L, st->v_max, st->a_max, st->Tv_max, st->nu_t_max, st->spacing_ratio);
  st->dt = dt;
  st->n_sort = st->n;
  if (M.n > 0) {
    const double t = st->time + static_cast<double>(dt);
    V3 a, w, wd{0, 0, 0};
    const bool rot = M.omega != nullptr;
    if (t <= M.time[0] || M.n == 1) {
      a = M.acc[0];
      w = rot ? M.omega[0] : V3{0, 0, 0};
    } else if (t >= M.time[M.n - 1]) {
      a = M.acc[M.n - 1];
      w = rot ? M.omega[M.n - 1] : V3{0, 0, 0};
    } else {
      int k = 0;  // last index with time[k] <= t (upper_bound - 1)
      int lo = 0, hi = M.n;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (t < M.time[mid]) hi = mid;
        else lo = mid + 1;
      }
      k = lo - 1;
      const double span = M.time[k + 1] - M.time[k];
      const float f = static_cast<float>((t - M.time[k]) / span);
      a = M.acc[k] * (1.0f - f) + M.acc[k + 1] * f;
      if (rot) {
        w = M.omega[k] * (1.0f - f) + M.omega[k + 1] * f;
        wd = (M.omega[k + 1] - M.omega[k]) / static_cast<float>(span);
      } else {
        w = V3{0, 0, 0};
      }
    }
    st->frame_acc = a;
    st->gravity = st->cfg_gravity - a;
    st->omega = w;
    st->omega_dot = wd;
    st->origin = M.origin;
    st->frame_rot = (norm2(w) > 0.0f || norm2(wd) > 0.0f) ? 1 : 0;
  }
}

// ----------------------------------------------- kick (+ Lee sources)
// sph_solver.cu:36-46 and phase_change.cu:77-137 (sources read the kicked T).
__global__ __launch_bounds__(kBlock) void k_kick_sources(PV pv, SimParams P, DevState* st, Count nc, int do_kick,
                                                        int thermal, int sources) {
  const int n = nc.get();
  const int i = gtid();
  if (i == 0 && do_kick && thermal && st->heat_inv_norm > 0.0f)
    st->heat_input += static_cast<double>(P.heat_power) * static_cast<double>(st->dt);
  if (i >= n) return;
  const float dt = st->dt;
  const uint8_t ph = pv.s.phase[i];
  if (do_kick && is_fluid(ph)) {
    const float half = 0.5f * dt;
    pv.s.v[i] = st3(ld3(pv.s.v[i]) + ld3(pv.t.a[i]) * half);
    if (thermal) pv.s.T[i] += pv.t.dTdt[i] * dt;
  }
  if (!sources) return;
  uint8_t fl = static_cast<uint8_t>(pv.s.flags[i] & ~kScratchMask);
  pv.t.out_m[i] = 0.0f;
  pv.t.out_h[i] = 0.0f;
  pv.t.w_norm[i] = 0.0f;
  if (!is_fluid(ph)) {
    pv.s.flags[i] = fl;
    return;
  }
  const int ii = pidx(ph);
  const float m = pv.s.m[i];
  const float T = pv.s.T[i];
  const float T_sat = st->T_sat;
  if (m < P.merge_ratio * sel2(P.mref, ii)) {
    pv.t.out_m[i] = m;
    pv.t.out_h[i] = enthalpy(P, ph, T);
    fl |= kSource | kDying;
    if (ph == kVapor) fl |= kTargetVapor;
    pv.s.flags[i] = fl;
    return;
  }
  const float alpha = pv.t.alpha[i];
  if (P.phase_change && alpha > 0.0f && dt > 0.0f) {
    if (ph == kLiquid && T > T_sat) {
      const float h_out = P.h_fg + P.cp[1] * (T_sat - P.T_ref);
      const float latent = h_out - P.cp[0] * (T_sat - P.T_ref);
      const float rate = P.lee_evap * m * alpha * (T - T_sat) / T_sat;
      float dm = rate * dt;
      const float dm_energy = m * P.cp[0] * (T - T_sat) / latent;
      dm = dm < dm_energy ? dm : dm_energy;
      const float dm_cap = P.max_transfer * m;
      dm = dm < dm_cap ? dm : dm_cap;
      if (dm > 0.0f) {
        pv.t.out_m[i] = dm;
        pv.t.out_h[i] = h_out;
        fl |= kSource | kTargetVapor;
      }
    } else if (ph == kVapor && T < T_sat) {
      const float rate = P.lee_cond * m * alpha * (T_sat - T) / T_sat;
      float dm = rate * dt;
      float dm_cap = P.max_transfer * m;
      if (P.pc_density_limit > 0.0f && P.pc_density_limit * m < dm_cap) dm_cap = P.pc_density_limit * m;
      dm = dm < dm_cap ? dm : dm_cap;
      if (dm > 0.0f) {
        pv.t.out_m[i] = dm;
        pv.t.out_h[i] = enthalpy(P, kVapor, T);
        fl |= kSource;
      }
    }
  }
  pv.s.flags[i] = fl;
}

// phase_change.cu:139-172, plus the partials of sph_solver.cu:108-124
// (evaporated, condensed, merges) over every particle.
template <bool KS>
__global__ __launch_bounds__(kBlock) void k_pc_normalize(PV pv, GridDev g, SimParams P, NbrList L, Count nc,
                                                        double* partials, int stride) {
  const int n = nc.get();
  const int i = gtid();
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < n) {
    const uint8_t fl = pv.s.flags[i];
    if (fl & kSource) {
      const uint8_t target = (fl & kTargetVapor) ? kVapor : kLiquid;
      const V3 xi = ld3(pv.s.x[i]);
      const V3 li = pos_lo<KS>(pv.s.xlo, i);
      Acc<KS> awsum;
      float wr_max = 0.0f;
      struct Nb {
        float4 x;
        float m, rho;
        int ph, fl;
      };
      auto load = [&](int j) {
        return Nb{pv.s.x[j], pv.s.m[j], pv.s.rho[j], pv.s.phase[j], pv.s.flags[j]};
      };
      for_each_pair4(L, g, i, xi, load, [&](int j, const Nb d) {
        if (j == i || d.ph != target || (d.fl & kDying)) return;
        const float r2 = norm2(pos_diff<KS>(xi, li, ld3(d.x), pv.s.xlo, j));
        if (r2 >= P.support2) return;
        const float w = kW(P.kernel, sqrtf(r2));
        const float rhoj = d.rho;
        awsum += (d.m / rhoj) * w;
        const float wr = w / rhoj;
        wr_max = wr > wr_max ? wr : wr_max;
      });
      const float wsum = awsum;
      pv.t.w_norm[i] = wsum;
      float om = pv.t.out_m[i];
      if (!(wsum > 0.0f)) {
        om = 0.0f;
        pv.t.out_m[i] = 0.0f;
      } else if (target == kVapor && P.pc_density_limit > 0.0f && !(fl & kDying) && wr_max > 0.0f) {
        const float cap = P.pc_density_limit * wsum / wr_max;
        if (om > cap) {
          om = cap;
          pv.t.out_m[i] = cap;
        }
      }
      if (!(fl & kDying)) {
        if (fl & kTargetVapor) acc[0] = static_cast<double>(om);
        else acc[1] = static_cast<double>(om);
      }
      if ((fl & kDying) && om > 0.0f) acc[2] = 1.0;
    }
  }
  block_sums_store<3>(acc, partials, stride);
}

// Marks the cells holding a mass source (out_m > 0 after the normalisation)
// by the phase its receivers have; the kind bits stay.
__global__ __launch_bounds__(kBlock) void k_cell_sources(PV pv, GridDev g) {
  const int c = gtid();
  if (c >= g.num_cells) return;
  int f = g.flags[c] & (kCellWall | kCellFluid);
  for (int j = g.off[c], e = g.off[c + 1]; j < e; ++j)
    if (pv.t.out_m[j] > 0.0f) f |= (pv.s.flags[j] & kTargetVapor) ? kCellSrcV : kCellSrcL;
  g.flags[c] = static_cast<uint8_t>(f);
}

// phase_change.cu:174-215
template <bool KS>
__global__ __launch_bounds__(kBlock) void k_pc_apply(PV pv, GridDev g, SimParams P, NbrList L, Count nc) {
  const int n = nc.get();
  const int i = gtid();
  if (i >= n) return;
  const uint8_t ph = pv.s.phase[i];
  if (!is_fluid(ph)) return;
  const uint8_t fl = pv.s.flags[i];
  const float mi = pv.s.m[i];
  const float Vi = mi / pv.s.rho[i];
  const V3 xi = ld3(pv.s.x[i]);
  const V3 li = pos_lo<KS>(pv.s.xlo, i);
  Acc<KS> ain_m, ain_H;
  // receivers only near a source giving mass to this phase
  if (!(fl & kDying) && cells_any(g, xi, ph == kVapor ? kCellSrcV : kCellSrcL)) {
    struct Nb {
      float4 x;
      float om, oh, wn;
      int fl;
    };
    auto load = [&](int j) {
      return Nb{pv.s.x[j], pv.t.out_m[j], pv.t.out_h[j], pv.t.w_norm[j], pv.s.flags[j]};
    };
    for_each_pair4(L, g, i, xi, load, [&](int j, const Nb d) {
      if (j == i) return;
      const float om = d.om;
      if (!(om > 0.0f)) return;
      const uint8_t target = (d.fl & kTargetVapor) ? kVapor : kLiquid;
      if (target != ph) return;
      const float r2 = norm2(pos_diff<KS>(xi, li, ld3(d.x), pv.s.xlo, j));
      if (r2 >= P.support2) return;
      const float share = om * Vi * kW(P.kernel, sqrtf(r2)) / d.wn;
      ain_m += share;
      ain_H += share * d.oh;
    });
  }
  const float in_m = ain_m, in_H = ain_H;
  const float om_i = pv.t.out_m[i];
  if ((fl & kDying) && om_i > 0.0f) {
    pv.s.m[i] = 0.0f;
    pv.s.phase[i] = kDead;
    return;
  }
  if (!(om_i > 0.0f) && !(in_m > 0.0f)) return;
  const float m_new = mi - om_i + in_m;
  const float H_new = mi * enthalpy(P, ph, pv.s.T[i]) - om_i * pv.t.out_h[i] + in_H;
  pv.s.T[i] = temp_from_h(P, ph, H_new / m_new);
  pv.s.m[i] = m_new;
}

// phase_change.cu:217-249 (pass 1)
template <bool KS>
__global__ __launch_bounds__(kBlock) void k_rebalance(PV pv, GridDev g, SimParams P, NbrList L, Count nc) {
  const int n = nc.get();
  const int i = gtid();
  if (i >= n) return;
  pv.t.out_m[i] = 0.0f;
  pv.t.out_h[i] = 0.0f;
  pv.t.ns[i] = st3(V3{0, 0, 0});
  if (pv.s.phase[i] != kVapor) return;
  const float beta = P.vapor_rebalance;
  const V3 xi = ld3(pv.s.x[i]);
  const V3 li = pos_lo<KS>(pv.s.xlo, i);
  const float mi = pv.s.m[i];
  const float hi = enthalpy(P, kVapor, pv.s.T[i]);
  const V3 vi = ld3(pv.s.v[i]);
  Acc<KS> dm, dH;
  AccV3<KS> dP;
  struct Nb {
    float4 x, v;
    float m, T;
    int ph;
  };
  auto load = [&](int j) { return Nb{pv.s.x[j], pv.s.v[j], pv.s.m[j], pv.s.T[j], pv.s.phase[j]}; };
  for_each_pair4(L, g, i, xi, load, [&](int j, const Nb d) {
    if (j == i || d.ph != kVapor) return;
    const float r2 = norm2(pos_diff<KS>(xi, li, ld3(d.x), pv.s.xlo, j));
    if (r2 >= P.support2) return;
    const float flow = beta * P.vol0 * kW(P.kernel, sqrtf(r2)) * (d.m - mi);
    const bool from_j = flow > 0.0f;
    dm += flow;
    dH += flow * (from_j ? enthalpy(P, kVapor, d.T) : hi);
    dP += (from_j ? ld3(d.v) : vi) * flow;
  });
  pv.t.out_m[i] = float(dm);
  pv.t.out_h[i] = float(dH);
  pv.t.ns[i] = st3(V3(dP));
}

// phase_change.cu:250-268 (pass 2) fused with mark_heavy_particles
// (phase_change.cu:270-279) and a per-block flag count for the slot scan.
__global__ __launch_bounds__(kBlock) void k_rebalance_apply_mark(PV pv, SimParams P, Count nc, int rebalance,
                                                                uint32_t* flag) {
  const int n = nc.get();
  const int i = gtid();
  if (i >= n) return;
  const uint8_t ph = pv.s.phase[i];
  float m = pv.s.m[i];
  if (rebalance && ph == kVapor) {
    const float dm = pv.t.out_m[i];
    if (dm != 0.0f) {
      const float m_new = m + dm;
      const float H = m * enthalpy(P, kVapor, pv.s.T[i]) + pv.t.out_h[i];
      pv.s.T[i] = temp_from_h(P, kVapor, H / m_new);
      const V3 v = (ld3(pv.s.v[i]) * m + ld3(pv.t.ns[i])) / m_new;
      pv.s.v[i] = st3(v);
      pv.s.vt[i] = st3(v);
      pv.s.m[i] = m_new;
      pv.t.out_m[i] = 0.0f;
      pv.t.out_h[i] = 0.0f;
      m = m_new;
    }
  }
  flag[i] = (is_fluid(ph) && m > P.split_ratio * sel2(P.mref, pidx(ph))) ? 1u : 0u;
}

// Exclusive scan of the split flags: per-block scan, block totals scanned by
// one block, then offsets added (deterministic slots, never atomics).
__global__ __launch_bounds__(1024) void k_scan_blocks(const uint32_t* in, uint32_t* out, Count nc, uint32_t* sums) {
  const int n = nc.get();
  __shared__ uint32_t s[1024];
  const int i = blockIdx.x * 1024 + threadIdx.x;
  const uint32_t v = i < n ? in[i] : 0u;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t t = threadIdx.x >= static_cast<unsigned>(off) ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = s[threadIdx.x] - v;
  if (threadIdx.x == 1023) sums[blockIdx.x] = s[1023];
}
__global__ __launch_bounds__(1024) void k_scan_add(uint32_t* out, Count nc, const uint32_t* sums) {
  const int n = nc.get();
  const int i = blockIdx.x * 1024 + threadIdx.x;
  if (i < n && blockIdx.x > 0) out[i] += sums[blockIdx.x];
}

// phase_change.cu:281-335
__global__ __launch_bounds__(kBlock) void k_split(PV pv, SimParams P, const DevState* st, Count nc,
                                                 const uint32_t* flag, const uint32_t* offset) {
  const int n = nc.get();
  const int i = gtid();
  if (i >= n) return;
  if (!flag[i]) return;
  const uint32_t id_base = st->next_id;
  const uint32_t seed = static_cast<uint32_t>(st->step);
  const uint8_t ph = pv.s.phase[i];
  const float m = pv.s.m[i];
  const uint32_t k = offset[i];
  const int dst = n + static_cast<int>(k);
  if (dst >= pv.capacity) return;
  const float vol = m / pv.s.rho[i];
  V3 e = random_direction(pv.s.id[i], seed);
  const V3 xi = ld3(pv.s.x[i]);
  const V3 gc = ld3(pv.s.nrm[i]);
  const float gm = norm(gc);
  // the child's (and, split symmetrically, the parent's) offset from the
  // parent position; the precision form adds it to the double-float position
  auto place = [&](int k, V3 off) {
    if (pv.s.xlo) {
      V3 x = xi, lo = ld3(pv.s.xlo[i]);
      df_add(x.x, lo.x, off.x);
      df_add(x.y, lo.y, off.y);
      df_add(x.z, lo.z, off.z);
      pv.s.x[k] = st3(x);
      pv.s.xlo[k] = st3(lo);
    } else {
      pv.s.x[k] = st3(xi + off);
    }
  };
  if (gm > P.normal_eps) {
    const V3 nn = gc / gm;
    const V3 t = e - nn * dot(e, nn);
    e = normalized(t - nn * 0.5f);
    place(dst, e * (0.5f * cbrtf_cr(vol)));
  } else {
    const float delta = 0.25f * cbrtf_cr(vol);
    place(dst, e * delta);
    place(i, -(e * delta));
  }
  pv.s.v[dst] = pv.s.v[i];
  pv.s.vt[dst] = pv.s.vt[i];
  pv.t.ab[dst] = pv.t.ab[i];
  pv.t.a[dst] = pv.t.a[i];
  pv.s.nrm[dst] = pv.s.nrm[i];
  pv.s.m[dst] = 0.5f * m;
  const float m0h = pv.s.m0[i] * 0.5f;
  pv.s.m0[i] = m0h;
  pv.s.m0[dst] = m0h;
  pv.s.m[i] = 0.5f * m;
  pv.s.rho[dst] = pv.s.rho[i];
  pv.s.T[dst] = pv.s.T[i];
  pv.s.phase[dst] = ph;
  pv.s.flags[dst] = static_cast<uint8_t>(pv.s.flags[i] & ~kScratchMask);
  pv.s.id[dst] = id_base + k;
  pv.t.p[dst] = pv.t.p[i];
  const float sg = pv.t.sigma[i] * 2.0f;
  pv.t.sigma[i] = sg;
  pv.t.sigma[dst] = sg;
  pv.t.dTdt[dst] = 0.0f;
  pv.t.out_m[dst] = 0.0f;
  pv.t.out_h[dst] = 0.0f;
}

// Split bookkeeping (sph_solver.cu:567-576): runs on one thread after k_split.
// `capacity` is the arrays' capacity, or with a device count the launch bound
// of the queued steps: splits past it are flagged (the host then re-runs
// those steps one at a time with the count in hand).
__global__ void k_split_commit(DevState* st, Count nc, int capacity, const uint32_t* scan_total) {
  const int n = nc.get();
  const uint32_t requested = *scan_total;
  const long long room = static_cast<long long>(capacity) - n;
  const long long added = requested < room ? static_cast<long long>(requested) : room;
  if (nc.dev && static_cast<long long>(requested) > room) st->count_overflow = 1;
  st->n = n + static_cast<int>(added);
  st->next_id += requested;
  st->splits += static_cast<int>(added);
  st->split_requested = requested;
}

// phase_change.cu:337-358 with the partials of sph_solver.cu:126-133.
__global__ __launch_bounds__(kBlock) void k_venting(PV pv, SimParams P, const DevState* st, float relax,
                                                   double* partials, int stride) {
  const int i = gtid();
  const int n = st->n;
  double acc[2] = {0.0, 0.0};
  if (i < n) {
    pv.t.out_m[i] = 0.0f;
    pv.t.out_h[i] = 0.0f;
    if (pv.s.phase[i] == kVapor) {
      const float m = pv.s.m[i];
      const float vol = m / pv.s.rho[i];
      const float target = P.rho0[1] * vol;
      if (m > target) {
        const float dm = relax * (m - target);
        const float hv = enthalpy(P, kVapor, pv.s.T[i]);
        pv.t.out_m[i] = dm;
        pv.t.out_h[i] = hv;
        if (P.density_mode == 1) pv.s.rho[i] *= (m - dm) / m;
        pv.s.m[i] = m - dm;
        acc[0] = static_cast<double>(dm);
        acc[1] = static_cast<double>(dm) * static_cast<double>(hv);
      }
    }
  }
  block_sums_store<2>(acc, partials, stride);
}

// Phase inventory (sph_solver.cu:137-145): m_liquid, m_vapour, sum m T vapour.
__global__ __launch_bounds__(kBlock) void k_inventory(PV pv, const DevState* st, double* partials, int stride) {
  const int i = gtid();
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < st->n) {
    const uint8_t ph = pv.s.phase[i];
    const double m = pv.s.m[i];
    if (ph == kLiquid) acc[0] = m;
    if (ph == kVapor) {
      acc[1] = m;
      acc[2] = m * static_cast<double>(pv.s.T[i]);
    }
  }
  block_sums_store<3>(acc, partials, stride);
}

// End of step: fold the partial sums into the running totals, the ullage
// update (sph_solver.cu:586-607), then the clock (sph_solver.cu:530-532).
// Partials layout: [0..2] pc (evap, cond, merges), [3..4] vent (m, mh),
// [5..7] inventory.
__global__ __launch_bounds__(1024) void k_end_step(DevState* st, const double* partials, int stride, int nb_pc,
                                                  int nb_vent, int nb_inv, FluidConst fl, int closed,
                                                  int advance_clock) {
  __shared__ double sh[1024];
  if (nb_pc > 0) {
    const double ev = ordered_sum(partials + 0 * stride, nb_pc, sh);
    const double co = ordered_sum(partials + 1 * stride, nb_pc, sh);
    const double mg = ordered_sum(partials + 2 * stride, nb_pc, sh);
    if (threadIdx.x == 0) {
      st->evaporated += ev;
      st->condensed += co;
      st->merges += static_cast<int>(mg);
    }
  }
  if (nb_vent > 0) {
    const double vm = ordered_sum(partials + 3 * stride, nb_vent, sh);
    const double vh = ordered_sum(partials + 4 * stride, nb_vent, sh);
    if (threadIdx.x == 0) {
      st->vented += vm;
      st->vented_h += vh;
    }
  }
  double m_liq = 0, m_vap = 0, mT_vap = 0;
  if (closed && nb_inv > 0) {
    m_liq = ordered_sum(partials + 5 * stride, nb_inv, sh);
    m_vap = ordered_sum(partials + 6 * stride, nb_inv, sh);
    mT_vap = ordered_sum(partials + 7 * stride, nb_inv, sh);
  }
  if (threadIdx.x != 0) return;
  bool set = false;
  if (closed) {
    if (!(st->fluid_volume0 > 0)) st->fluid_volume0 = m_liq / fl.rho_liquid + m_vap / fl.rho_vapor;
    const double v_ullage = st->fluid_volume0 - m_liq / fl.rho_liquid;
    const double ratio =
        v_ullage > 0 && m_vap > 0 ? (mT_vap / v_ullage) / (static_cast<double>(fl.rho_vapor) * fl.T_sat) : 0.0;
    if (isfinite(ratio) && ratio > 0) {
      st->p_ullage = fl.p_sat * ratio;
      // FluidProperties::saturation_temperature (fluid_properties.cpp:7-10)
      const float p = static_cast<float>(st->p_ullage);
      const double inv_T =
          1.0 / fl.T_sat - (static_cast<double>(fl.R_gas) / fl.h_fg) * log(static_cast<double>(p) / fl.p_sat);
      st->T_sat = static_cast<float>(1.0 / inv_T);
      set = true;
    }
  }
  if (!set) {
    st->p_ullage = fl.p_sat;
    st->T_sat = fl.T_sat;
  }
  if (advance_clock) {
    st->time += static_cast<double>(st->dt);
    st->step += 1;
    st->last_dt = static_cast<double>(st->dt);
  }
}

// The phase-change totals of k_end_step (the same ordered sums), folded in
// before the particle arrays grow mid-step: growing re-sizes the partials.
__global__ __launch_bounds__(1024) void k_pc_fold(DevState* st, const double* partials, int stride, int nb_pc) {
  __shared__ double sh[1024];
  const double ev = ordered_sum(partials + 0 * stride, nb_pc, sh);
  const double co = ordered_sum(partials + 1 * stride, nb_pc, sh);
  const double mg = ordered_sum(partials + 2 * stride, nb_pc, sh);
  if (threadIdx.x == 0) {
    st->evaporated += ev;
    st->condensed += co;
    st->merges += static_cast<int>(mg);
  }
}

// sph_solver.cu:48-57
__global__ __launch_bounds__(kBlock) void k_scale_velocities(PV pv, int n, float f) {
  const int i = gtid();
  if (i >= n) return;
  if (!is_fluid(pv.s.phase[i])) return;
  pv.s.v[i] = st3(ld3(pv.s.v[i]) * f);
  pv.s.vt[i] = st3(ld3(pv.s.vt[i]) * f);
}

// ------------------------------------------------------------ diagnostics
// Every sum of Solver::diagnostics (sph_solver.cu:609-653) in one pass.
// Quantities (double): 0-2 counts L,V,W; 3,4 mass L,V; 5,6 kinetic L,V;
// 7,8 enthalpy L,V; 9,10 m T L,V; 11-13 m x L; 14-16 m v L; 17-19 fw fluid;
// 20-22 mw fluid.  Maxima (float bits): speed (fluid), T liquid.
constexpr int kDiagQ = 23;
__global__ __launch_bounds__(kBlock) void k_diag(PV pv, SimParams P, int n, double* partials, int stride,
                                                unsigned* maxbits) {
  const int i = gtid();
  double acc[kDiagQ];
#pragma unroll
  for (int q = 0; q < kDiagQ; ++q) acc[q] = 0.0;
  float vmax = 0.0f, tlmax = 0.0f;
  if (i < n) {
    const uint8_t ph = pv.s.phase[i];
    const double m = pv.s.m[i];
    const V3 v = ld3(pv.s.v[i]);
    if (ph == kLiquid || ph == kVapor) {
      // constant indices only (a runtime index would put acc in scratch)
      const double ke = 0.5 * m * static_cast<double>(norm2(v));
      const double mh = m * static_cast<double>(enthalpy(P, ph, pv.s.T[i]));
      const double mT = m * static_cast<double>(pv.s.T[i]);
      if (ph == kLiquid) {
        acc[0] = 1.0, acc[3] = m, acc[5] = ke, acc[7] = mh, acc[9] = mT;
      } else {
        acc[1] = 1.0, acc[4] = m, acc[6] = ke, acc[8] = mh, acc[10] = mT;
      }
      if (ph == kLiquid) {
        const V3 x = ld3(pv.s.x[i]);
        if (pv.s.xlo) {  // precision form: the double-float position
          const V3 lo = ld3(pv.s.xlo[i]);
          acc[11] = m * (static_cast<double>(x.x) + lo.x);
          acc[12] = m * (static_cast<double>(x.y) + lo.y);
          acc[13] = m * (static_cast<double>(x.z) + lo.z);
        } else {
          acc[11] = m * x.x;
          acc[12] = m * x.y;
          acc[13] = m * x.z;
        }
        acc[14] = m * v.x;
        acc[15] = m * v.y;
        acc[16] = m * v.z;
        tlmax = pv.s.T[i];
      }
      const V3 fw = ld3(pv.t.fw[i]), mw = ld3(pv.t.mw[i]);
      acc[17] = fw.x;
      acc[18] = fw.y;
      acc[19] = fw.z;
      acc[20] = mw.x;
      acc[21] = mw.y;
      acc[22] = mw.z;
      vmax = norm(v);
    } else if (ph == kWall) {
      acc[2] = 1.0;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const float a = __shfl_down(vmax, off, 64), b = __shfl_down(tlmax, off, 64);
    vmax = a > vmax ? a : vmax;
    tlmax = b > tlmax ? b : tlmax;
  }
  if ((threadIdx.x & 63) == 0) {
    if (vmax > 0.0f) atomicMax(&maxbits[0], __float_as_uint(vmax));
    if (tlmax > 0.0f) atomicMax(&maxbits[1], __float_as_uint(tlmax));
  }
  block_sums_store<kDiagQ>(acc, partials, stride);
}
__global__ __launch_bounds__(1024) void k_diag_finalize(const double* partials, int stride, int nb, double* out) {
  __shared__ double sh[1024];
  for (int q = 0; q < kDiagQ; ++q) {
    const double s = ordered_sum(partials + q * stride, nb, sh);
    if (threadIdx.x == 0) out[q] = s;
  }
}

// Telemetry record (telemetry.cu:70-84): x, y, z, T, p, speed, phase.
__global__ __launch_bounds__(kBlock) void k_telemetry_pack(PV pv, int n, float* rec) {
  const int i = gtid();
  if (i >= n) return;
  const V3 x = ld3(pv.s.x[i]);
  float* r = rec + 7 * static_cast<long long>(i);
  r[0] = x.x;
  r[1] = x.y;
  r[2] = x.z;
  r[3] = pv.s.T[i];
  r[4] = pv.t.p[i];
  r[5] = norm(ld3(pv.s.v[i]));
  r[6] = __uint_as_float(static_cast<uint32_t>(pv.s.phase[i]));
}

__global__ void k_zero_u32(unsigned* p, int n) {
  const int i = gtid();
  if (i < n) p[i] = 0u;
}

}  // namespace

int diag_quantities() { return kDiagQ; }

void launch_begin_step(DevState* st, const SimParams& P, const DtLimits& L, const Maneuver& M, hipStream_t s) {
  k_begin_step<<<1, 1, 0, s>>>(st, P, L, M);
}
void launch_kick_sources(const PV& pv, const SimParams& P, DevState* st, Count n, bool do_kick, bool thermal,
                         bool sources, hipStream_t s) {
  k_kick_sources<<<blocks_for(n.ub > 0 ? n.ub : 1), kBlock, 0, s>>>(pv, P, st, n, do_kick ? 1 : 0, thermal ? 1 : 0,
                                                              sources ? 1 : 0);
}
int launch_pc_normalize(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, Count n,
                        double* partials, int stride, hipStream_t s) {
  const int nb = blocks_for(n.ub > 0 ? n.ub : 1);
  if (P.compensated) k_pc_normalize<true><<<nb, kBlock, 0, s>>>(pv, g, P, L, n, partials, stride);
  else k_pc_normalize<false><<<nb, kBlock, 0, s>>>(pv, g, P, L, n, partials, stride);
  return nb;
}
void launch_pc_apply(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, Count n, hipStream_t s) {
  if (n.ub <= 0) return;
  if (g.num_cells > 0) k_cell_sources<<<blocks_for(g.num_cells), kBlock, 0, s>>>(pv, g);
  if (P.compensated) k_pc_apply<true><<<blocks_for(n.ub), kBlock, 0, s>>>(pv, g, P, L, n);
  else k_pc_apply<false><<<blocks_for(n.ub), kBlock, 0, s>>>(pv, g, P, L, n);
}
void launch_rebalance(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, Count n, hipStream_t s) {
  if (n.ub <= 0) return;
  if (P.compensated) k_rebalance<true><<<blocks_for(n.ub), kBlock, 0, s>>>(pv, g, P, L, n);
  else k_rebalance<false><<<blocks_for(n.ub), kBlock, 0, s>>>(pv, g, P, L, n);
}
void launch_rebalance_apply_mark(const PV& pv, const SimParams& P, Count n, bool rebalance, uint32_t* flag,
                                 hipStream_t s) {
  if (n.ub > 0) k_rebalance_apply_mark<<<blocks_for(n.ub), kBlock, 0, s>>>(pv, P, n, rebalance ? 1 : 0, flag);
}
void launch_exclusive_scan(const uint32_t* in, uint32_t* out, Count n, uint32_t* block_sums, uint32_t* total,
                           hipStream_t s) {
  const int nb = (n.ub + 1023) / 1024;
  if (n.ub > 0) {
    k_scan_blocks<<<nb, 1024, 0, s>>>(in, out, n, block_sums);
    launch_scan_one_block(block_sums, nb, total, s);
    k_scan_add<<<nb, 1024, 0, s>>>(out, n, block_sums);
  } else {
    k_zero_u32<<<1, 1, 0, s>>>(total, 1);
  }
}
void launch_split(const PV& pv, const SimParams& P, const DevState* st, Count n, const uint32_t* flag,
                  const uint32_t* offset, hipStream_t s) {
  if (n.ub > 0) k_split<<<blocks_for(n.ub), kBlock, 0, s>>>(pv, P, st, n, flag, offset);
}
void launch_split_commit(DevState* st, Count n, int capacity, const uint32_t* total, hipStream_t s) {
  k_split_commit<<<1, 1, 0, s>>>(st, n, capacity, total);
}
int launch_venting(const PV& pv, const SimParams& P, const DevState* st, int n_ub, float relax, double* partials,
                   int stride, hipStream_t s) {
  const int nb = blocks_for(n_ub > 0 ? n_ub : 1);
  k_venting<<<nb, kBlock, 0, s>>>(pv, P, st, relax, partials, stride);
  return nb;
}
int launch_inventory(const PV& pv, const DevState* st, int n_ub, double* partials, int stride, hipStream_t s) {
  const int nb = blocks_for(n_ub > 0 ? n_ub : 1);
  k_inventory<<<nb, kBlock, 0, s>>>(pv, st, partials, stride);
  return nb;
}
void launch_end_step(DevState* st, const double* partials, int stride, int nb_pc, int nb_vent, int nb_inv,
                     const FluidConst& fl, bool closed, bool advance_clock, hipStream_t s) {
  k_end_step<<<1, 1024, 0, s>>>(st, partials, stride, nb_pc, nb_vent, nb_inv, fl, closed ? 1 : 0,
                                advance_clock ? 1 : 0);
}
void launch_pc_fold(DevState* st, const double* partials, int stride, int nb_pc, hipStream_t s) {
  if (nb_pc > 0) k_pc_fold<<<1, 1024, 0, s>>>(st, partials, stride, nb_pc);
}
void launch_scale_velocities(const PV& pv, int n, float f, hipStream_t s) {
  if (n > 0) k_scale_velocities<<<blocks_for(n), kBlock, 0, s>>>(pv, n, f);
}
void launch_diagnostics(const PV& pv, const SimParams& P, int n, double* partials, int stride, unsigned* maxbits,
                        double* out, hipStream_t s) {
  const int nb = blocks_for(n > 0 ? n : 1);
  k_zero_u32<<<1, 2, 0, s>>>(maxbits, 2);
  k_diag<<<nb, kBlock, 0, s>>>(pv, P, n, partials, stride, maxbits);
  k_diag_finalize<<<1, 1024, 0, s>>>(partials, stride, nb, out);
}
void launch_telemetry_pack(const PV& pv, int n, float* rec, hipStream_t s) {
  if (n > 0) k_telemetry_pack<<<blocks_for(n), kBlock, 0, s>>>(pv, n, rec);
}

}  // namespace cryo
