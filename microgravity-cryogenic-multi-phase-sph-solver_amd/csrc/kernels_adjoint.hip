// Reverse-mode adjoint of the hydrodynamic step on gfx950 (no counterpart in
// the reference; the forward passes it differentiates are wall_boundary,
// momentum and density summation, navier_stokes.cu:16-245, SPH_kernels.cu:6-34,
// with the kick-drift-kick of sph_solver.cu:20-46).
//
// With CSF surface tension on, also the interface and CSF passes
// (surface_tension.cu:5-177: colour gradient with the wall colour and the
// contact-angle rotation, unsmoothed normals, curvature, classic CSF force).
//
// Loss: sum over steps k of dt (S_x^2 + S_y^2), S_k = sum of fw over the fluid
// particles after step k's rates; parameters: the frame acceleration of each
// step (g_eff = g - acc) and the contact angle.  One reverse step runs after
// the rates of step k (all passes but the forces pass, whose outputs the
// reverse step does not read) have been re-evaluated on the taped post-sort
// state, in that step's sorted order:
//   k_adj_prep     a's total cotangent  at = a_bar + (dt/2) v_bar  (+ its g share)
//   k_adj_csf1     (CSF) cotangents of each particle's curvature sums
//   k_adj_csf2     (CSF) both sides of the curvature pairs, the normalisation
//                  and the rotation: colour-gradient, wall-normal and wall-weight
//                  cotangents (+ the contact angle's share)
//   k_adj_wall     per wall: cotangents of its Shepard pressures, ghost
//                  velocity and wall colour, divided by the Shepard weights
//                  (+ its g share)
//   k_adj_fluid    per fluid particle, gathered over its neighbours: both
//                  members' share of every fluid pair, its own share of every
//                  wall pair, the wall values' share, the colour-gradient
//                  pairs (CSF), the EOS and V = m / rho -> sigma_bar
//   k_adj_density  per fluid particle: density summation, drift and kick,
//                  then scattered back through the step's sort permutation
// With conduction and Lee phase change on, masses and temperatures are state
// too.  The step's kick and the phase-change passes up to apply are re-run
// forward on the taped state first (their values are kept), then, ahead of
// k_adj_prep:
//   k_adj_reb1/2   venting and vapour rebalance (phase_change.cu:217-268,337-358)
//   k_adj_app1/2   apply: both sides of each source-receiver pair (:174-215)
//   k_adj_src      receiver normalisation with the vapour density cap and the
//                  Lee sources (:77-172), d loss / d Lee coefficients
//   k_adj_nrm      the normalisation and interface-fraction pairs, both sides
// and k_adj_fluid adds conduction pairs (phase_change.cu:31-75), the heater
// weights and the masses' and temperatures' shares of every rates term.
// Every pair is evaluated by each member from its own side (a gather, no
// atomics), so the result is bitwise reproducible.  The per-pair derivatives
// are in adjoint_pair.hpp.
#include "adjoint_pair.hpp"
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

using AV = cryo_adj::Vec<float>;
__device__ inline AV av(float4 q) { return AV{q.x, q.y, q.z}; }
__device__ inline AV av(V3 q) { return AV{q.x, q.y, q.z}; }
__device__ inline float4 f4(AV a) { return make_float4(a.x, a.y, a.z, 0.0f); }

__device__ inline cryo_adj::Ctx<float> adj_ctx(const SimParams& P, float dt) {
  return cryo_adj::Ctx<float>{P.kernel.norm, P.kernel.inv_h, P.kernel.h, P.support2, 0.01f * P.kernel.h * P.kernel.h,
                              0.6f * P.dx, dt, P.vol0, P.no_slip, P.kernel.type, P.tvf && P.tvf_stress ? 1 : 0};
}
// VT: the transport velocity is read (by the momentum stress term).
template <bool VT = false>
__device__ inline cryo_adj::Side<float> side(const PV& pv, const SimParams& P, int i) {
  const uint8_t ph = pv.s.phase[i];
  return cryo_adj::Side<float>{av(pv.s.x[i]),     av(pv.s.v[i]),
                               pv.s.m[i],         pv.s.rho[i],
                               pv.t.p[i],         sound_speed(P, ph, pv.s.T[i]),
                               sel2(P.mu, pidx(ph)), ph == kVapor ? P.alpha_av_vap : P.alpha_av,
                               static_cast<int>(ph), VT ? av(pv.s.vt[i]) : AV{0.0f, 0.0f, 0.0f}};
}
// Cotangent of the wall loads: d loss / d fw_i = 2 dt (S_x, S_y, 0).
__device__ inline AV load_cot(const double* S, float dt) {
  return AV{2.0f * dt * static_cast<float>(S[0]), 2.0f * dt * static_cast<float>(S[1]), 0.0f};
}

__global__ __launch_bounds__(kBlock) void k_adj_fill(float4* a, float4* b, float4* c, int n) {
  const int i = gtid();
  if (i >= n) return;
  const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  a[i] = z;
  b[i] = z;
  c[i] = z;
}

// The step's frame (apply_frame, sph_solver.cu:660-667): acceleration and,
// when given, angular velocity and its rate.
__global__ __launch_bounds__(kBlock) void k_adj_set_frame(DevState* st, const float* acc, const float* omega,
                                                         const float* omega_dot) {
  const V3 a{acc[0], acc[1], acc[2]};
  st->frame_acc = a;
  st->gravity = st->cfg_gravity - a;
  if (omega || omega_dot) {
    const V3 w = omega ? V3{omega[0], omega[1], omega[2]} : V3{0, 0, 0};
    const V3 wd = omega_dot ? V3{omega_dot[0], omega_dot[1], omega_dot[2]} : V3{0, 0, 0};
    st->omega = w;
    st->omega_dot = wd;
    st->frame_rot = (norm2(w) > 0.0f || norm2(wd) > 0.0f) ? 1 : 0;
  }
}

// Sum of fw over fluid particles: block partials in rows [row, row + 3).
__global__ __launch_bounds__(kBlock) void k_adj_load_sum(PV pv, int n, double* partials, int stride) {
  const int i = gtid();
  double v[3] = {0.0, 0.0, 0.0};
  if (i < n && is_fluid(pv.s.phase[i])) {
    const float4 f = pv.t.fw[i];
    v[0] = f.x;
    v[1] = f.y;
    v[2] = f.z;
  }
  block_sums_store<3>(v, partials, stride);
}

// out[q] = scale * (sum rows r0+q [+ sum rows r1+q]), q = 0..nq-1.
__global__ __launch_bounds__(1024) void k_adj_finish(const double* partials, int stride, int nb, int r0, int r1,
                                                    double scale, double* out, int nq) {
  __shared__ double sh[1024];
  for (int q = 0; q < nq; ++q) {
    double s = ordered_sum(partials + (r0 + q) * stride, nb, sh);
    if (r1 >= 0) s += ordered_sum(partials + (r1 + q) * stride, nb, sh);
    if (threadIdx.x == 0) out[q] = scale * s;
  }
}

__device__ inline float colour(uint8_t ph) { return ph == kLiquid ? 1.0f : 0.0f; }

// Transport acceleration (SPH_kernels.cu:56-104): ab = K sum_J gw (Vt^2 + V_J^2),
// Vt = 1 / sigma, K = -pb / (rho Vt); active where the forward pass computes it.
__device__ inline bool tvf_active(const PV& pv, const SimParams& P, int i) {
  return P.tvf && !(pv.t.sigma[i] < P.tvf_sigma_min) && !(P.tvf_interface_only && !(pv.t.alpha[i] > 0.0f));
}
__device__ inline float tvf_K(const PV& pv, const SimParams& P, int i) {
  const uint8_t ph = pv.s.phase[i];
  const float ci = sound_speed(P, ph, pv.s.T[i]);
  const float Vt = 1.0f / pv.t.sigma[i];
  return -(P.tvf_factor * sel2(P.rho0, pidx(ph)) * ci * ci) / (pv.s.rho[i] * Vt);
}
// The wall density a transport owner of phase ph uses for wall j.
__device__ inline float tvf_rho_w(const PV& pv, const SimParams& P, uint8_t ph, float Ti, int j) {
  if (ph == kVapor) return (pv.t.p_wv[j] - (P.p_bg - P.vapor_K)) / (P.gas_R * Ti);
  return pv.t.nrec[j].z;  // rho0 max(ratio, 0.5)^(1/gamma), evaluated by the wall pass
}
__global__ __launch_bounds__(kBlock) void k_adj_prep(PV pv, SimParams P, AdjArgs A, const DevState* st, int n) {
  const int i = gtid();
  double g[3] = {0.0, 0.0, 0.0};
  double hw = 0.0;
  double wg[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // d loss / d (omega, omega_dot)
  if (i < n) {
    float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (is_fluid(pv.s.phase[i])) {
      const float half = 0.5f * st->dt;
      const float4 a = A.ca[i], v = A.pc ? A.v1b[i] : A.cv[i];
      r = make_float4(a.x + v.x * half, a.y + v.y * half, a.z + v.z * half, 0.0f);
      g[0] = r.x;  // a = g_eff + ...
      g[1] = r.y;
      g[2] = r.z;
      if (A.wout) {  // a -= wd x rr + w x (w x rr) + 2 w x v, rr = x - origin
        const V3 At{r.x, r.y, r.z}, w = st->omega;
        const V3 rr = ld3(pv.s.x[i]) - st->origin, vv = ld3(pv.s.v[i]);
        const V3 dw = At * (-dot(w, rr)) - rr * dot(At, w) + w * (2.0f * dot(At, rr)) + cross(At, vv) * 2.0f;
        const V3 dwd = cross(At, rr);
        wg[0] = dw.x, wg[1] = dw.y, wg[2] = dw.z, wg[3] = dwd.x, wg[4] = dwd.y, wg[5] = dwd.z;
      }
    }
    A.at[i] = r;
    if (A.heat) {  // T1 = T + dTdt dt, dTdt = q / (m cp)
      float q = 0.0f;
      const uint8_t ph = pv.s.phase[i];
      if (is_fluid(ph)) {
        const float T1b = A.pc ? A.pa[i].y : A.cT[i];
        q = T1b * st->dt / (pv.s.m[i] * sel2(P.cp, pidx(ph)));
        hw = static_cast<double>(q) * static_cast<double>(pv.t.heat_w[i]);
      }
      A.qb[i] = q;
    }
    if (A.tvf) {
      float4 t = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (is_fluid(pv.s.phase[i]) && tvf_active(pv, P, i)) {
        const float K = tvf_K(pv, P, i);
        const float4 c = A.ct[i];
        t = make_float4(c.x * K, c.y * K, c.z * K, 1.0f);
      }
      A.tvs[i] = t;
    }
  }
  block_sums_store<3>(g, A.partials + 3 * A.stride, A.stride);
  if (A.wout) block_sums_store<6>(wg, A.partials + 15 * A.stride, A.stride);
  if (A.heat) {
    double h1[1] = {hw};
    block_sums_store<1>(h1, A.partials + 8 * A.stride, A.stride);
  }
}

__device__ inline bool live_normal(const PV& pv, int i) { return norm2(ld3(pv.t.ns[i])) > 0.25f; }

// Curvature sums' cotangents: kappa = 3 num / den; a += nrm coeff kappa
// (classic force) or a += G sigma kappa / m (balanced force, with G's
// cotangent and the 1 / m factor's mass cotangent in A.gbal).
__global__ __launch_bounds__(kBlock) void k_adj_csf1(PV pv, GridDev g, SimParams P, NbrList L, AdjArgs A, int n) {
  const int i = gtid();
  if (i >= n) return;
  float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float4 gbal = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const bool bal = P.csf_balanced != 0;
  if (is_fluid(pv.s.phase[i]) && live_normal(pv, i)) {
    const float coeff = 2.0f * P.sigma / (P.rho0[0] + P.rho0[1]);
    const float4 at = A.at[i], nr = pv.s.nrm[i];
    const V3 xi = ld3(pv.s.x[i]), ni = ld3(pv.t.ns[i]);
    const float rhoi = pv.s.rho[i], mi = pv.s.m[i], Vi = mi / rhoi, ci = colour(pv.s.phase[i]);
    float num = 0.0f, den = 0.0f;  // the curvature sums (the forward pass's, recomputed)
    V3 G{0.0f, 0.0f, 0.0f};        // and the balanced force's sum
    for_each_pair(L, g, i, xi, [&](int j) {
      if (j == i || !is_fluid(pv.s.phase[j])) return;
      const bool livej = live_normal(pv, j);
      if (!livej && !bal) return;
      const V3 rij = xi - ld3(pv.s.x[j]);
      const float r2 = norm2(rij);
      if (!(r2 < P.support2)) return;
      const float r = sqrtf(r2);
      const float dW = kdW(P.kernel, r);
      const float rhoj = pv.s.rho[j], Vj = pv.s.m[j] / rhoj;
      const V3 gw = rij * (dW / r);
      if (livej) {
        num += Vj * dot(ni - ld3(pv.t.ns[j]), gw);
        den += Vj * r * (-dW);
      }
      if (bal) G += gw * ((Vi * Vi + Vj * Vj) * (rhoj * ci + rhoi * colour(pv.s.phase[j])) / (rhoi + rhoj));
    });
    if (den > 0.0f) {
      const float kappa = 3.0f * num / den;
      float kb;
      if (bal) {
        const float AG = at.x * G.x + at.y * G.y + at.z * G.z;
        kb = P.sigma / mi * AG;
        const float f = P.sigma * kappa / mi;
        gbal = make_float4(at.x * f, at.y * f, at.z * f, -P.sigma * kappa * AG / (mi * mi));
      } else {
        kb = coeff * (at.x * nr.x + at.y * nr.y + at.z * nr.z);
      }
      out = make_float4(3.0f * kb / den, -kappa * kb / den, kappa, 0.0f);
    }
  }
  A.cs[i] = out;
  if (bal) A.gbal[i] = gbal;
}

// Per fluid particle: both sides of the curvature pairs, the normalisation,
// then the contact-angle rotation of its colour gradient (recomputed).
// EXT: the smoothed-normal or contact-line terms are compiled in (launched so
// when A.sacc or A.gpb is set).
template <bool EXT>
__global__ __launch_bounds__(kBlock) void k_adj_csf2(PV pv, GridDev g, SimParams P, NbrList L, const DevState* st,
                                                    AdjArgs A, int n, int row) {
  const int k = gtid();
  double th[2] = {0.0, 0.0};
  if (k < n && is_fluid(pv.s.phase[k])) {
    const cryo_adj::Ctx<float> c = adj_ctx(P, st->dt);
    const float coeff = 2.0f * P.sigma / (P.rho0[0] + P.rho0[1]);
    const V3 xk = ld3(pv.s.x[k]);
    const AV nsk = av(pv.t.ns[k]);
    const bool livek = cryo_adj::dot(nsk, nsk) > 0.25f;
    const float4 csk = A.cs[k];
    const float Vk = pv.s.m[k] / pv.s.rho[k];
    const float ck = colour(pv.s.phase[k]);
    AV xr{0.0f, 0.0f, 0.0f}, nsb{0.0f, 0.0f, 0.0f};
    float Vb = 0.0f, ww = 0.0f;
    V3 gradc{0.0f, 0.0f, 0.0f}, wall_n{0.0f, 0.0f, 0.0f};
    const bool smooth = EXT && A.sacc != nullptr;
    V3 sacc{0.0f, 0.0f, 0.0f};  // the smoothed normal's sum (k_smooth_normals), recomputed
    const bool cl = EXT && A.gpb != nullptr && !smooth;  // the contact-line force (k_adj_csf3 with smoothing)
    V3 gradphi{0.0f, 0.0f, 0.0f};
    for_each_pair(L, g, k, xk, [&](int j) {
      const V3 xj = ld3(pv.s.x[j]);
      const V3 rkj = xk - xj;
      const float r2 = norm2(rkj);
      if (!(r2 < P.support2)) return;
      const float r = sqrtf(r2);
      const uint8_t phj = pv.s.phase[j];
      if (phj == kWall) {  // the interface pass's wall sums (surface_tension.cu:5-85)
        const float w = kW(P.kernel, r);
        ww += P.vol0 * w;
        wall_n += ld3(pv.s.nrm[j]) * (P.vol0 * w);
        if (cl) gradphi += kgradW(P.kernel, rkj, r) * P.vol0;
        if (P.wall_colour) gradc += kgradW(P.kernel, rkj, r) * (P.vol0 * (pv.t.alpha[j] - ck));
        return;
      }
      if (!is_fluid(phj)) return;
      const float Vj = pv.s.m[j] / pv.s.rho[j];
      if (smooth) {
        const V3 gj = ld3(pv.s.nrm[j]);
        const float gjm = norm(gj);
        if (gjm > P.normal_eps) sacc += gj * (Vj * kW(P.kernel, r) / gjm);
      }
      if (j == k) return;
      gradc += kgradW(P.kernel, rkj, r) * (Vj * (colour(phj) - ck));
      const AV nsj = av(pv.t.ns[j]);
      if (!livek || !(cryo_adj::dot(nsj, nsj) > 0.25f)) return;
      const AV rv = av(rkj);
      const cryo_adj::CsfGrad<float> o = cryo_adj::csf_vjp(c, rv, nsk, nsj, Vj, csk.x, csk.y);  // k's sums
      xr += o.r;
      nsb += o.ns_i;
      const float4 csj = A.cs[j];
      const cryo_adj::CsfGrad<float> o2 =
          cryo_adj::csf_vjp(c, AV{-rv.x, -rv.y, -rv.z}, nsj, nsk, Vk, csj.x, csj.y);  // j's sums
      xr -= o2.r;
      nsb += o2.ns_j;
      Vb += o2.V_j;
    });
    A.csv[k] = Vb;
    A.xc[k] = f4(xr);
    if (smooth) {  // ns = sacc / |sacc|: k_adj_csf3 goes on from its cotangent
      const float Am = norm(sacc);
      A.sacc[k] = Am > 1e-20f ? f4((nsb - nsk * cryo_adj::dot(nsk, nsb)) * (1.0f / Am))
                              : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else {
      const float4 at = A.at[k];
      AV nrmb = livek && !P.csf_balanced ? AV{at.x, at.y, at.z} * (coeff * csk.z) : AV{0.0f, 0.0f, 0.0f};
      const AV nrk = av(pv.s.nrm[k]);
      const float G = sqrtf(cryo_adj::dot(nrk, nrk));
      if (G > P.normal_eps) nrmb += (nsb - nsk * cryo_adj::dot(nsk, nsb)) * (1.0f / G);  // ns = nrm / |nrm|
      const cryo_adj::RotGrad<float> ro = cryo_adj::rot_vjp(av(gradc), av(wall_n), ww, P.cos_theta, P.sin_theta,
                                                            P.wall_adh_w, P.normal_eps, P.wetting != 1 ? 1 : 0, nrmb);
      AV gcb = ro.g, wnb = ro.wall_n;
      th[0] = ro.cos_t;
      th[1] = ro.sin_t;
      if (cl) {  // a += vw, the contact-line force of the interface pass
        cryo_adj::ClGrad<float> o{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, 0.0f};
        if (livek)
          o = cryo_adj::cl_vjp(av(gradc), av(wall_n), av(gradphi), ww, P.cos_theta, coeff, P.normal_eps,
                               AV{at.x, at.y, at.z});
        gcb += o.g;
        wnb += o.wall_n;
        th[0] += o.cos_t;
        A.gpb[k] = f4(o.gradphi);
      }
      A.gcb[k] = make_float4(gcb.x, gcb.y, gcb.z, ro.ww);
      A.wnb[k] = f4(wnb);
    }
  } else if (k < n) {
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    A.gcb[k] = z;
    A.wnb[k] = z;
    A.csv[k] = 0.0f;
    A.xc[k] = z;
    if (EXT && A.sacc) A.sacc[k] = z;
    if (EXT && A.gpb) A.gpb[k] = z;
  }
  block_sums_store<2>(th, A.partials + row * A.stride, A.stride);
}

// Smoothed normals (k_smooth_normals): ns_i = acc_i / |acc_i| with acc_i =
// sum_j V_j W_ij nrm_j / |nrm_j| over fluid j with a live colour gradient (i
// itself included); A.sacc holds each particle's cotangent of acc_i.  Per
// fluid particle k, over both roles of its pairs: the cotangents of k's unit
// colour gradient, V_k and x_k, then the normal cotangent and the
// contact-angle rotation of k_adj_csf2 (its colour gradient recomputed).
__global__ __launch_bounds__(kBlock) void k_adj_csf3(PV pv, GridDev g, SimParams P, NbrList L, AdjArgs A, int n,
                                                    int row) {
  const int k = gtid();
  double th[2] = {0.0, 0.0};
  if (k < n && is_fluid(pv.s.phase[k])) {
    const float coeff = 2.0f * P.sigma / (P.rho0[0] + P.rho0[1]);
    const V3 xk = ld3(pv.s.x[k]);
    const AV gk = av(pv.s.nrm[k]);
    const float G = sqrtf(cryo_adj::dot(gk, gk));
    const bool live_g = G > P.normal_eps;
    const AV ghk = live_g ? gk * (1.0f / G) : AV{0.0f, 0.0f, 0.0f};
    const float Vk = pv.s.m[k] / pv.s.rho[k];
    const float ck = colour(pv.s.phase[k]);
    const AV sk = av(A.sacc[k]);
    AV ghb{0.0f, 0.0f, 0.0f}, xr{0.0f, 0.0f, 0.0f};
    float Vb = 0.0f, ww = 0.0f;
    V3 gradc{0.0f, 0.0f, 0.0f}, wall_n{0.0f, 0.0f, 0.0f}, gradphi{0.0f, 0.0f, 0.0f};
    const bool cl = A.gpb != nullptr;
    for_each_pair(L, g, k, xk, [&](int j) {
      const V3 rkj = xk - ld3(pv.s.x[j]);
      const float r2 = norm2(rkj);
      if (!(r2 < P.support2)) return;
      const float r = sqrtf(r2);
      const uint8_t phj = pv.s.phase[j];
      if (phj == kWall) {  // as k_adj_csf2
        const float w = kW(P.kernel, r);
        ww += P.vol0 * w;
        wall_n += ld3(pv.s.nrm[j]) * (P.vol0 * w);
        if (cl) gradphi += kgradW(P.kernel, rkj, r) * P.vol0;
        if (P.wall_colour) gradc += kgradW(P.kernel, rkj, r) * (P.vol0 * (pv.t.alpha[j] - ck));
        return;
      }
      if (!is_fluid(phj)) return;
      const float Vj = pv.s.m[j] / pv.s.rho[j];
      if (j != k) gradc += kgradW(P.kernel, rkj, r) * (Vj * (colour(phj) - ck));
      if (!live_g) return;
      const AV gj = av(pv.s.nrm[j]);
      const float gjm = sqrtf(cryo_adj::dot(gj, gj));
      if (!(gjm > P.normal_eps)) return;
      const AV ghj = gj * (1.0f / gjm);
      const AV sj = av(A.sacc[j]);
      const float w = kW(P.kernel, r);
      ghb += sj * (Vk * w);                   // j's sum holds V_k W nrm_k / |nrm_k|
      Vb += w * cryo_adj::dot(sj, ghk);
      if (r > 0.0f) {                         // W_kj's position derivative, both sums
        const float s = kdW(P.kernel, r) / r;
        xr += av(rkj) * (s * (Vj * cryo_adj::dot(sk, ghj) + Vk * cryo_adj::dot(sj, ghk)));
      }
    });
    const float4 at = A.at[k], csk = A.cs[k];
    const bool livek = live_normal(pv, k);
    AV nrmb = livek && !P.csf_balanced ? AV{at.x, at.y, at.z} * (coeff * csk.z) : AV{0.0f, 0.0f, 0.0f};
    if (live_g) nrmb += (ghb - ghk * cryo_adj::dot(ghk, ghb)) * (1.0f / G);
    const cryo_adj::RotGrad<float> ro = cryo_adj::rot_vjp(av(gradc), av(wall_n), ww, P.cos_theta, P.sin_theta,
                                                          P.wall_adh_w, P.normal_eps, P.wetting != 1 ? 1 : 0, nrmb);
    AV gcb = ro.g, wnb = ro.wall_n;
    th[0] = ro.cos_t;
    th[1] = ro.sin_t;
    if (cl) {  // a += vw, the contact-line force of the interface pass
      cryo_adj::ClGrad<float> o{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, 0.0f};
      if (livek)
        o = cryo_adj::cl_vjp(av(gradc), av(wall_n), av(gradphi), ww, P.cos_theta, coeff, P.normal_eps,
                             AV{at.x, at.y, at.z});
      gcb += o.g;
      wnb += o.wall_n;
      th[0] += o.cos_t;
      A.gpb[k] = f4(o.gradphi);
    }
    A.gcb[k] = make_float4(gcb.x, gcb.y, gcb.z, ro.ww);
    A.wnb[k] = f4(wnb);
    A.csv[k] += Vb;
    const float4 x0 = A.xc[k];
    A.xc[k] = make_float4(x0.x + xr.x, x0.y + xr.y, x0.z + xr.z, 0.0f);
  }
  block_sums_store<2>(th, A.partials + row * A.stride, A.stride);
}

template <bool TV>  // the transport-velocity terms compiled in (launched so when A.tvf is set)
__global__ __launch_bounds__(kBlock) void k_adj_wall(PV pv, GridDev g, SimParams P, NbrList L, const DevState* st,
                                                    AdjArgs A, int n) {
  const int w = gtid();
  double gs[3] = {0.0, 0.0, 0.0};
  if (w < n && pv.s.phase[w] == kWall) {
    const float dt = st->dt;
    cryo_adj::Ctx<float> c = adj_ctx(P, dt);
    c.tvf_stress = 0;  // the stress term reaches no wall-side cotangent (p_w, vw)
    const AV Fb = load_cot(A.S, dt);
    const AV xw = av(pv.s.x[w]), vw = av(pv.t.vw[w]);
    const float pw0 = pv.t.p[w], pw1 = pv.t.p_wv[w];
    float sw0 = 0.0f, sw1 = 0.0f, pb0 = 0.0f, pb1 = 0.0f, alb = 0.0f;
    AV G0{0.0f, 0.0f, 0.0f}, G1{0.0f, 0.0f, 0.0f}, vwb{0.0f, 0.0f, 0.0f};
    const float alpha_w = pv.t.alpha[w];
    const AV nrm_w = av(pv.s.nrm[w]);
    for_each_pair(L, g, w, ld3(pv.s.x[w]), [&](int j) {
      const uint8_t ph = pv.s.phase[j];
      if (!is_fluid(ph)) return;  // the lists hold fluid only, the sweep fallback does not
      const AV rwf = xw - av(pv.s.x[j]);
      const float r2 = cryo_adj::dot(rwf, rwf);
      if (!(r2 < P.support2)) return;
      const float wk = kW(P.kernel, sqrtf(r2));
      const float rho = pv.s.rho[j];
      const cryo_adj::Side<float> I = side(pv, P, j);
      const float4 at = A.at[j];
      const AV cA{at.x, at.y, at.z};
      const AV cF = cA * (1.0f / I.m);
      const bool vap = ph == kVapor;
      const cryo_adj::FWGrad<float> o = cryo_adj::fw_vjp(c, I, xw, vw, vap ? pw1 : pw0, cF + Fb, cA + Fb * I.m);
      if (vap) {
        sw1 += wk;
        G1 += rwf * (rho * wk);
        pb1 += o.p_w;
      } else {
        sw0 += wk;
        G0 += rwf * (rho * wk);
        pb0 += o.p_w;
      }
      vwb -= o.vij;
      if (TV) {  // a transport owner's wall volume reads this wall's pressure
        const float4 tv = A.tvs[j];
        if (tv.w > 0.0f) {
          const float Ti = pv.s.T[j];
          const float Vt = 1.0f / pv.t.sigma[j];
          const float rho_w = tvf_rho_w(pv, P, ph, Ti, w);
          const float Vj = rho_w > 0.0f ? Vt * rho / rho_w : Vt;
          const cryo_adj::TvGrad<float> ot =
              cryo_adj::tvf_vjp(c, AV{-rwf.x, -rwf.y, -rwf.z}, Vt * Vt, Vj, AV{tv.x, tv.y, tv.z});
          float dummy_vt = 0.0f, dummy_rho = 0.0f;
          const float rwb = cryo_adj::tvf_wall_vjp(Vt, rho, rho_w, ot.vj, dummy_vt, dummy_rho);
          const float dp = cryo_adj::tvf_wall_drho_dp(vap ? 1 : 0, vap ? pw1 : pw0, rho_w, P.p_bg, P.tait_B,
                                                      P.tait_gamma, P.gas_R, Ti);
          if (vap) pb1 += rwb * dp;
          else pb0 += rwb * dp;
        }
      }
      if (A.csf) {  // the fluid's colour gradient reads the wall colour alpha_w
        const float4 gb = A.gcb[j], wb = A.wnb[j];
        const cryo_adj::GcGrad<float> og = cryo_adj::gc_fw_vjp(c, AV{-rwf.x, -rwf.y, -rwf.z}, alpha_w,
                                                               colour(ph), nrm_w, P.wall_colour,
                                                               AV{gb.x, gb.y, gb.z}, AV{wb.x, wb.y, wb.z}, gb.w);
        alb += og.V_j;
      }
    });
    const bool cl0 = P.clamp_neg && !(pw0 > 0.0f), cl1 = P.clamp_neg && !(pw1 > 0.0f);
    const float cp0 = (sw0 > 0.0f && !cl0) ? pb0 / sw0 : 0.0f;
    const float cp1 = (sw1 > 0.0f && !cl1) ? pb1 / sw1 : 0.0f;
    const float swt = sw0 + sw1;
    const AV cvw = swt > 0.0f ? vwb * (1.0f / swt) : AV{0.0f, 0.0f, 0.0f};
    A.wcp[w] = make_float4(cp0, cp1, 0.0f, 0.0f);
    A.wcv[w] = make_float4(cvw.x, cvw.y, cvw.z, swt > 0.0f ? alb / swt : 0.0f);  // w: alpha_w = sw0 / swt
    const AV gw = G0 * cp0 + G1 * cp1;  // d pw / d g = sum rho r_wf W / sum W
    gs[0] = gw.x;
    gs[1] = gw.y;
    gs[2] = gw.z;
  }
  block_sums_store<3>(gs, A.partials, A.stride);
}

// TV: the transport-velocity terms are compiled in (launched so when A.tvf
// is set); without them the kernel holds fewer registers.
// EX: the transport stress, the balanced CSF force, the contact-line force or
// the cubic-spline kernel is on; the other instances fix Wendland C2 at
// compile time and hold none of those terms' registers.
template <bool TV, bool EX>
__global__ __launch_bounds__(kBlock) void k_adj_fluid(PV pv, GridDev g, SimParams P, NbrList L, const DevState* st,
                                                     AdjArgs A, int n) {
  const int i = gtid();
  if (i >= n) return;
  const uint8_t phi = pv.s.phase[i];
  if (!is_fluid(phi)) return;
  const float dt = st->dt;
  cryo_adj::Ctx<float> c = adj_ctx(P, dt);
  if constexpr (!EX) {  // the common instance: Wendland C2, no stress, no balanced force
    c.tvf_stress = 0;
    c.type = 1;
  }
  const AV Fb = load_cot(A.S, dt);
  const AV grav = av(st->gravity);
  const cryo_adj::Side<float> I = side<EX>(pv, P, i);
  const bool vap = phi == kVapor;
  const float4 ati = A.at[i];
  const AV cA{ati.x, ati.y, ati.z};
  const AV cF = cA * (1.0f / I.m);
  const AV cFw = cF + Fb, cAw = cA + Fb * I.m;
  AV xr{0.0f, 0.0f, 0.0f}, vr{0.0f, 0.0f, 0.0f}, vtr{0.0f, 0.0f, 0.0f};
  float rb = 0.0f, pb = 0.0f;
  const float ci = colour(phi), Vk = I.m / I.rho;
  AV gbk{0.0f, 0.0f, 0.0f}, wbk{0.0f, 0.0f, 0.0f};
  float wwbk = 0.0f, Vbk = 0.0f;
  // transport acceleration: own sums (owner active) and the neighbours' sums that read V_i = 1 / sigma_i
  const float4 tvi = TV ? A.tvs[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const bool tv_own = TV && tvi.w > 0.0f;
  const AV Sbi{tvi.x, tvi.y, tvi.z};
  const float Vti = TV ? 1.0f / pv.t.sigma[i] : 0.0f, Vti2 = Vti * Vti;
  float vt2b = 0.0f, vtb = 0.0f, trhob = 0.0f, tsb = 0.0f, tTb = 0.0f;
  AV tsum{0.0f, 0.0f, 0.0f};
  // masses, temperatures (sound speeds), the force sum for d(F / m) / dm, conduction
  float mb = 0.0f, Tb = 0.0f, cb = 0.0f;
  AV F{0.0f, 0.0f, 0.0f};
  const float Ti = pv.s.T[i];
  const float qi = A.heat ? A.qb[i] : 0.0f;
  const float ki = sel2(P.k, pidx(phi));
  const bool iso = P.wall_bc == 1;
  float hwb = 0.0f;  // heater: hw_i's cotangent, q_i P hinv + (sum's share)
  if (A.heat && P.heat_power != 0.0f && st->heat_inv_norm > 0.0f) {
    const float hinv = st->heat_inv_norm;
    hwb = qi * P.heat_power * hinv - static_cast<float>(A.hsum[0]) * P.heat_power * hinv * hinv;
  }
  const bool bal = EX && A.csf && P.csf_balanced;
  const float4 gbi4 = bal ? A.gbal[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const AV Gbi{gbi4.x, gbi4.y, gbi4.z};
  AV gpbk{0.0f, 0.0f, 0.0f};
  if (EX && A.gpb) gpbk = av(A.gpb[i]);
  if (A.csf) {
    const float4 gb = A.gcb[i], wb = A.wnb[i];
    gbk = AV{gb.x, gb.y, gb.z};
    wwbk = gb.w;
    wbk = AV{wb.x, wb.y, wb.z};
    Vbk = A.csv[i];
    xr = av(A.xc[i]);
  }
  for_each_pair(L, g, i, ld3(pv.s.x[i]), [&](int j) {
    if (j == i) return;
    const uint8_t phj = pv.s.phase[j];
    if (phj == kWall) {
      const AV xw = av(pv.s.x[j]), vw = av(pv.t.vw[j]);
      const float pw = vap ? pv.t.p_wv[j] : pv.t.p[j];
      const cryo_adj::FWGrad<float> o = cryo_adj::fw_vjp(c, I, xw, vw, pw, cFw, cAw);
      xr += o.r;
      vr += o.vi + o.vij;
      if (EX) vtr += o.vti;
      rb += o.rho_i;
      pb += o.p_i;
      if (A.mT) {
        F += o.f;
        mb += o.m_i + cryo_adj::dot(Fb, o.acc);  // the wall load carries m_i acc
        cb += o.c_i;
        if (A.heat) {
          const AV rkw = I.x - xw;
          if (iso) {
            const cryo_adj::ThGrad<float> oh = cryo_adj::th_vjp(c, rkw, Vk, P.vol0, 2.0f * ki, Ti, pv.s.T[j], qi);
            xr += oh.r;
            Vbk += oh.Vi;
            Tb += oh.Ti;
          }
          const float r2 = cryo_adj::dot(rkw, rkw);
          if (hwb != 0.0f && (pv.s.flags[j] & kHeated) && r2 < P.support2)
            xr += rkw * (hwb * P.vol0 * cryo_adj::kern_s(c, sqrtf(r2)));
        }
      }
      const float4 cp = A.wcp[j], cv = A.wcv[j];
      const cryo_adj::WBGrad<float> b =
          cryo_adj::wb_vjp(c, I, xw, pw, vw, vap ? cp.y : cp.x, AV{cv.x, cv.y, cv.z}, grav);
      xr += b.x;
      vr += b.v;
      pb += b.p;
      rb += b.rho;
      if (tv_own) {
        const AV rkw = I.x - xw;
        const float r2 = cryo_adj::dot(rkw, rkw);
        if (r2 < P.support2) {
          const float rho_w = tvf_rho_w(pv, P, phi, pv.s.T[i], j);
          const float Vj = rho_w > 0.0f ? Vti * I.rho / rho_w : Vti;
          tsum += av(kgradW(P.kernel, V3{rkw.x, rkw.y, rkw.z}, sqrtf(r2))) * (Vti2 + Vj * Vj);
          const cryo_adj::TvGrad<float> ot = cryo_adj::tvf_vjp(c, rkw, Vti2, Vj, Sbi);
          xr += ot.r;
          vt2b += ot.vt2;
          const float rwb = cryo_adj::tvf_wall_vjp(Vti, I.rho, rho_w, ot.vj, vtb, trhob);  // wall: k_adj_wall
          if (vap) tTb -= rwb * rho_w / pv.s.T[i];  // vapour wall density (p_wv - p0) / (R T_i)
        }
      }
      if (A.csf) {
        const float alpha_w = pv.t.alpha[j];
        const cryo_adj::GcGrad<float> og = cryo_adj::gc_fw_vjp(c, I.x - xw, alpha_w, ci, av(pv.s.nrm[j]),
                                                               P.wall_colour, gbk, wbk, wwbk);
        xr += og.r;
        if (EX && A.gpb) xr += cryo_adj::gphi_fw_vjp(c, I.x - xw, gpbk);  // the contact-line force's gradphi
        const AV rwk = xw - I.x;
        const float r2 = cryo_adj::dot(rwk, rwk);
        if (cv.w != 0.0f && r2 < P.support2)  // alpha_w = sw0 / swt: d/dx_k = (c_k - alpha_w) / swt (-s r_wk)
          xr += rwk * (-cv.w * (ci - alpha_w) * cryo_adj::kern_s(c, sqrtf(r2)));
      }
      return;
    }
    if (!is_fluid(phj)) return;
    if (A.csf) {  // colour-gradient pairs, both sides
      const float cj = colour(phj);
      const float4 gbj4 = A.gcb[j];
      const AV rkj = I.x - av(pv.s.x[j]);
      const cryo_adj::GcGrad<float> og = cryo_adj::gc_ff_vjp(c, rkj, pv.s.m[j] / pv.s.rho[j], cj - ci, gbk);
      const cryo_adj::GcGrad<float> og2 =
          cryo_adj::gc_ff_vjp(c, AV{-rkj.x, -rkj.y, -rkj.z}, Vk, ci - cj, AV{gbj4.x, gbj4.y, gbj4.z});
      xr += og.r;
      xr -= og2.r;
      Vbk += og2.V_j;
    }
    if (TV) {
      const AV rkj = I.x - av(pv.s.x[j]);
      const float r2 = cryo_adj::dot(rkj, rkj);
      if (r2 < P.support2) {
        const float Vj = 1.0f / pv.t.sigma[j];
        if (tv_own) {
          tsum += av(kgradW(P.kernel, V3{rkj.x, rkj.y, rkj.z}, sqrtf(r2))) * (Vti2 + Vj * Vj);
          const cryo_adj::TvGrad<float> ot = cryo_adj::tvf_vjp(c, rkj, Vti2, Vj, Sbi);
          xr += ot.r;
          vt2b += ot.vt2;
        }
        const float4 tvj = A.tvs[j];
        if (tvj.w > 0.0f) {  // j's sum reads V_i
          const cryo_adj::TvGrad<float> o2 =
              cryo_adj::tvf_vjp(c, AV{-rkj.x, -rkj.y, -rkj.z}, Vj * Vj, Vti, AV{tvj.x, tvj.y, tvj.z});
          xr -= o2.r;
          tsb -= o2.vj * Vti * Vti;
        }
      }
    }
    const cryo_adj::Side<float> J = side<EX>(pv, P, j);
    if (bal) {  // balanced CSF: i's G sum and j's G sum, which reads i
      const AV rkj = I.x - J.x;
      const float Vj = J.m / J.rho;
      const float4 gj = A.gbal[j];
      const cryo_adj::BalGrad<float> b1 = cryo_adj::bal_vjp(c, rkj, Vk, Vj, I.rho, J.rho, ci, colour(phj), Gbi);
      const cryo_adj::BalGrad<float> b2 = cryo_adj::bal_vjp(c, AV{-rkj.x, -rkj.y, -rkj.z}, Vj, Vk, J.rho, I.rho,
                                                            colour(phj), ci, AV{gj.x, gj.y, gj.z});
      xr += b1.r;
      xr -= b2.r;
      Vbk += b1.V_o + b2.V_n;
      rb += b1.rho_o + b2.rho_n;
    }
    const float4 atj = A.at[j];
    const AV cAj{atj.x, atj.y, atj.z};
    const cryo_adj::FFGrad<float> o1 = cryo_adj::ff_vjp(c, I, J, cF, cA);         // i's pair term
    const cryo_adj::FFGrad<float> o2 = cryo_adj::ff_vjp(c, J, I, cAj * (1.0f / J.m), cAj);  // j's
    xr += o1.r;
    xr -= o2.r;
    vr += o1.vij;
    vr -= o2.vij;
    if (EX) {  // the stress term's own shares
      vr += o1.vi + o2.vj;
      vtr += o1.vti + o2.vtj;
    }
    rb += o1.rho_i + o2.rho_j;
    pb += o1.p_i + o2.p_j;
    if (A.mT) {
      F += o1.f;
      mb += o1.m_i + o2.m_j;
      cb += o1.c_i + o2.c_j;
      if (A.heat) {  // conduction pairs, both sides
        const AV rkj = I.x - J.x;
        const float kj = sel2(P.k, pidx(phj));
        const float kij = 4.0f * ki * kj / (ki + kj);
        const float Vj = J.m / J.rho;
        const cryo_adj::ThGrad<float> h1 = cryo_adj::th_vjp(c, rkj, Vk, Vj, kij, Ti, pv.s.T[j], qi);
        const cryo_adj::ThGrad<float> h2 =
            cryo_adj::th_vjp(c, AV{-rkj.x, -rkj.y, -rkj.z}, Vj, Vk, kij, pv.s.T[j], Ti, A.qb[j]);
        xr += h1.r;
        xr -= h2.r;
        Vbk += h1.Vi + h2.Vj;
        Tb += h1.Ti + h2.Tj;
      }
    }
  });
  const float dpdr = cryo_adj::eos_dp_drho(vap ? 1 : 0, I.rho, pv.s.T[i], I.p, P.rho0[0], P.tait_B, P.tait_gamma,
                                           P.gas_R, P.clamp_neg);
  if (A.pc) {  // phase change's shares of V, rho and x
    Vbk += A.pV[i];
    rb += A.pr[i];
    xr += av(A.xp[i]);
  }
  rb -= Vbk * Vk / I.rho;  // V = m / rho in the curvature, colour, conduction and phase-change sums
  if (tv_own) {  // K = -pb / (rho Vt), Vt = 1 / sigma
    const float K = tvf_K(pv, P, i);
    const float4 ct = A.ct[i];
    const float Kb = ct.x * tsum.x + ct.y * tsum.y + ct.z * tsum.z;
    trhob -= K * Kb / I.rho;
    vtb -= K * Kb / Vti;
    vtb += 2.0f * Vti * vt2b;
    tsb -= vtb * Vti * Vti;
    rb += trhob;
    if (vap) tTb += K * Kb / pv.s.T[i];  // c^2 = R T
  }
  const float rhob = rb + pb * dpdr;
  A.sb[i] = rhob * (vap ? I.m : pv.s.m0[i]) + tsb;  // rho = m0 sigma (liquid), m sigma (vapour)
  if (st->frame_rot) {  // a -= wd x r + w x (w x r) + 2 w x v
    const V3 At{cA.x, cA.y, cA.z}, w = st->omega, wd = st->omega_dot;
    xr += av(cross(wd, At) - cross(w, cross(w, At)));
    vr += av(cross(w, At) * 2.0f);
  }
  if (A.mT) {
    const float4 pa = A.pc ? A.pa[i] : make_float4(A.cm[i], A.cT[i], 0.0f, 0.0f);  // (own shares)
    mb += pa.x + Vbk / I.rho;
    if (bal) mb += gbi4.w;  // the balanced force's 1 / m
    mb -= (cA.x * F.x + cA.y * F.y + cA.z * F.z) / (I.m * I.m);  // a = g + F / m + ...
    if (A.heat) mb -= pa.y * dt * pv.t.dTdt[i] / I.m;      // dTdt = q / (m cp)
    if (vap) mb += rhob * pv.t.sigma[i];
    Tb += pa.y + tTb + pb * cryo_adj::eos_dp_dT(vap ? 1 : 0, I.rho, I.p, P.gas_R, P.clamp_neg);
    if (vap && Ti > 0.0f) Tb += cb * P.gas_R / (2.0f * I.c);  // c = sqrt(R T)
    A.mr[i] = mb;
    A.Tr[i] = Tb;
  }
  A.xr[i] = f4(xr);
  A.vr[i] = f4(vr);
  if (EX && A.vtr) A.vtr[i] = f4(vtr);
}

__global__ __launch_bounds__(kBlock) void k_adj_density(PV pv, GridDev g, SimParams P, NbrList L, const DevState* st,
                                                       AdjArgs A, int n) {
  const int i = gtid();
  if (i >= n) return;
  const uint32_t dst = A.perm[i];
  if (!is_fluid(pv.s.phase[i])) {
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    A.nx[dst] = z;
    A.nv[dst] = z;
    A.na[dst] = z;
    if (A.tvf) A.nt[dst] = z;
    if (A.mT) A.nm[dst] = A.nT[dst] = 0.0f;
    return;
  }
  const float dt = st->dt, half = 0.5f * dt;
  const cryo_adj::Ctx<float> c = adj_ctx(P, dt);
  const float sbi = A.sb[i];
  const AV xi = av(pv.s.x[i]);
  AV xr = av(A.xr[i]);
  for_each_pair(L, g, i, ld3(pv.s.x[i]), [&](int j) {
    if (j == i) return;
    const AV rij = xi - av(pv.s.x[j]);
    const float r2 = cryo_adj::dot(rij, rij);
    if (!(r2 < P.support2)) return;
    const float sj = is_fluid(pv.s.phase[j]) ? A.sb[j] : 0.0f;  // wall sigma is a constant
    xr += rij * ((sbi + sj) * cryo_adj::kern_s(c, sqrtf(r2)));
  });
  // x' = x + vt dt, vt = v_half (+ ab dt/2 with the transport velocity),
  // v' = v_half + a' dt/2, v_half = v + a dt/2; vtb: the rates' share of vt's
  // cotangent (transport stress)
  const AV X = av(A.cx[i]) + xr;
  const AV vtb = A.vtr ? av(A.vtr[i]) : AV{0.0f, 0.0f, 0.0f};
  const AV Vh = av(A.pc ? A.v1b[i] : A.cv[i]) + av(A.vr[i]) + X * dt + vtb;
  if (A.mT) {
    float nm = A.mr[i], nT = A.Tr[i];
    if (A.sat) {  // the closed tank's T_sat feedback (k_adj_sat)
      const uint8_t ph = pv.s.phase[i];
      if (ph == kLiquid) nm += static_cast<float>(A.satv[0]);
      if (ph == kVapor) {
        nm += static_cast<float>(A.satv[1] * pv.s.T[i]);
        nT += static_cast<float>(A.satv[1] * pv.s.m[i]);
      }
    }
    A.nm[dst] = nm;
    A.nT[dst] = nT;
  }
  A.nx[dst] = f4(X);
  A.nv[dst] = f4(Vh);
  A.na[dst] = f4(Vh * half);
  if (A.tvf) A.nt[dst] = f4(A.vtr ? (X * dt + vtb) * half : X * (half * dt));  // vt = v_half + ab dt/2
}

// ------------------------------------------------------ phase change
__device__ inline uint8_t pc_target(uint8_t fl) { return (fl & kTargetVapor) ? kVapor : kLiquid; }

// Venting (m4 = m3 - relax (m3 - rho0_v m3 / rho)) and the owner's part of the
// vapour rebalance (m3 = m2 + dm, T3 = T(H / m3), v3 = (v1 m2 + dP) / m3, with
// dm, dH, dP recomputed from the applied masses, temperatures and kicked velocities).
__global__ __launch_bounds__(kBlock) void k_adj_reb1(PV pv, GridDev g, SimParams P, NbrList L, AdjArgs A, int n) {
  const int i = gtid();
  if (i >= n) return;
  const uint8_t ph = pv.s.phase[i];
  float m2b = A.cm[i], T2b = A.cT[i], dmb = 0.0f, dHb = 0.0f, rhob = 0.0f;
  AV v1b = av(A.cv[i]), dPb{0.0f, 0.0f, 0.0f};
  if (ph == kVapor) {
    const float beta = P.vapor_rebalance, cpv = P.cp[1];
    const V3 xi = ld3(pv.s.x[i]);
    const float mi = A.m2[i], hi = enthalpy(P, kVapor, A.T2[i]);
    const AV vi = av(A.v1[i]);
    float dm = 0.0f, dH = 0.0f;
    AV dP{0.0f, 0.0f, 0.0f};
    if (A.reb) {
      for_each_pair(L, g, i, xi, [&](int j) {
        if (j == i || pv.s.phase[j] != kVapor) return;
        const float r2 = norm2(xi - ld3(pv.s.x[j]));
        if (r2 >= P.support2) return;
        const float flow = beta * P.vol0 * kW(P.kernel, sqrtf(r2)) * (A.m2[j] - mi);
        const bool from_j = flow > 0.0f;
        dm += flow;
        dH += flow * (from_j ? enthalpy(P, kVapor, A.T2[j]) : hi);
        dP += (from_j ? av(A.v1[j]) : vi) * flow;
      });
    }
    const bool upd = A.reb && dm != 0.0f;
    const float m3 = upd ? mi + dm : mi;
    float m3b = A.cm[i];
    if (A.vent) {
      const float rho = pv.s.rho[i], r0 = P.rho0[1];
      if (m3 > r0 * (m3 / rho)) {
        m3b = A.cm[i] * (1.0f - A.relax + A.relax * r0 / rho);
        rhob = A.cm[i] * (-A.relax * r0 * m3 / (rho * rho));
      }
    }
    m2b = m3b;
    if (upd) {
      const float mn = m3, hh = (mi * hi + dH) / mn;
      const AV v3 = (vi * mi + dP) * (1.0f / mn);
      const float hhb = A.cT[i] / cpv, Hb = hhb / mn;
      float mnb = m3b - hhb * hh / mn;
      const AV Vb3 = av(A.cv[i]);
      v1b = Vb3 * (mi / mn);
      float m2bi = cryo_adj::dot(Vb3, vi) / mn;
      dPb = Vb3 * (1.0f / mn);
      mnb -= cryo_adj::dot(Vb3, v3) / mn;
      m2b = m2bi + mnb + Hb * hi;
      dmb = mnb;
      dHb = Hb;
      T2b = Hb * mi * cpv;
    }
  }
  A.pb4[i] = make_float4(dmb, dHb, m2b, T2b);
  A.pdp[i] = make_float4(dPb.x, dPb.y, dPb.z, rhob);
  A.v1b[i] = f4(v1b);
}

// Both sides of each rebalance pair: flow_ij = beta vol0 W (m2_j - m2_i).
__global__ __launch_bounds__(kBlock) void k_adj_reb2(PV pv, GridDev g, SimParams P, NbrList L, AdjArgs A, int n) {
  const int i = gtid();
  if (i >= n) return;
  const float4 own = A.pb4[i];
  float m2b = own.z, T2b = own.w;
  AV xr{0.0f, 0.0f, 0.0f};
  if (A.reb && pv.s.phase[i] == kVapor) {
    const float beta = P.vapor_rebalance, cpv = P.cp[1];
    const cryo_adj::Ctx<float> c = adj_ctx(P, 0.0f);
    const V3 xi = ld3(pv.s.x[i]);
    const float mi = A.m2[i], hi = enthalpy(P, kVapor, A.T2[i]);
    const AV vi = av(A.v1[i]);
    const float4 dpi = A.pdp[i];
    const AV dPi{dpi.x, dpi.y, dpi.z};
    AV v1b = av(A.v1b[i]);
    for_each_pair(L, g, i, xi, [&](int j) {
      if (j == i || pv.s.phase[j] != kVapor) return;
      const V3 rv = xi - ld3(pv.s.x[j]);
      const float r2 = norm2(rv);
      if (r2 >= P.support2) return;
      const float r = sqrtf(r2), W = kW(P.kernel, r), sW = cryo_adj::kern_s(c, r);
      const AV rij = av(rv);
      const float mj = A.m2[j], hj = enthalpy(P, kVapor, A.T2[j]);
      const AV vj = av(A.v1[j]);
      // i's own flow from j
      const float flow = beta * P.vol0 * W * (mj - mi);
      const bool from_j = flow > 0.0f;
      const float fb = own.x + own.y * (from_j ? hj : hi) + cryo_adj::dot(dPi, from_j ? vj : vi);
      if (!from_j) {
        T2b += own.y * flow * cpv;
        v1b += dPi * flow;
      }
      const float k = fb * beta * P.vol0;
      m2b -= k * W;
      xr += rij * (k * (mj - mi) * sW);
      // j's flow from i
      const float4 oj = A.pb4[j], dpj = A.pdp[j];
      const AV dPj{dpj.x, dpj.y, dpj.z};
      const float flow2 = beta * P.vol0 * W * (mi - mj);
      const bool from_i = flow2 > 0.0f;
      const float fb2 = oj.x + oj.y * (from_i ? hi : hj) + cryo_adj::dot(dPj, from_i ? vi : vj);
      if (from_i) {
        T2b += oj.y * flow2 * cpv;
        v1b += dPj * flow2;
      }
      const float k2 = fb2 * beta * P.vol0;
      m2b += k2 * W;
      xr += rij * (k2 * (mi - mj) * sW);
    });
    A.v1b[i] = f4(v1b);
  }
  A.pm2[i] = make_float2(m2b, T2b);
  A.xp[i] = f4(xr);
}

// Apply, the receiver's own part: m2 = m - om + in_m, H2 = m h(T1) - om oh + in_H,
// T2 = T(H2 / m2) (in_m, in_H recomputed).
__global__ __launch_bounds__(kBlock) void k_adj_app1(PV pv, GridDev g, SimParams P, NbrList L, AdjArgs A, int n) {
  const int i = gtid();
  if (i >= n) return;
  const uint8_t ph = pv.s.phase[i];
  const float2 b2 = A.pm2[i];
  float m0b = b2.x, T1b = b2.y, inmb = 0.0f, inHb = 0.0f, omb = 0.0f, ohb = 0.0f;
  if (is_fluid(ph)) {
    const V3 xi = ld3(pv.s.x[i]);
    const float mi = pv.s.m[i], Vi = mi / pv.s.rho[i];
    float in_m = 0.0f, in_H = 0.0f;
    // receivers only near a source giving mass to this phase (the cells' source bits of the re-run apply)
    if (!(A.fl[i] & kDying) && cells_any(g, xi, ph == kVapor ? kCellSrcV : kCellSrcL)) {
      for_each_pair(L, g, i, xi, [&](int j) {
        if (j == i || !(A.om[j] > 0.0f) || pc_target(A.fl[j]) != ph) return;
        const float r2 = norm2(xi - ld3(pv.s.x[j]));
        if (r2 >= P.support2) return;
        const float share = A.om[j] * Vi * kW(P.kernel, sqrtf(r2)) / pv.t.w_norm[j];
        in_m += share;
        in_H += share * A.oh[j];
      });
    }
    const float om = A.om[i];
    if (om > 0.0f || in_m > 0.0f) {
      const float mn = mi - om + in_m, hT = enthalpy(P, ph, A.T1[i]);
      const float hh = (mi * hT - om * A.oh[i] + in_H) / mn;
      const float cpp = sel2(P.cp, pidx(ph));
      const float hhb = b2.y / cpp, Hb = hhb / mn;
      const float mnb = b2.x - hhb * hh / mn;
      m0b = mnb + Hb * hT;
      T1b = Hb * mi * cpp;
      omb = -mnb - Hb * A.oh[i];
      ohb = -Hb * om;
      inmb = mnb;
      inHb = Hb;
    }
  }
  A.pa[i] = make_float4(m0b, T1b, inmb, inHb);
  A.po[i] = make_float4(omb, ohb, 0.0f, 0.0f);
}

// Apply pairs, both sides: share = om_j V_i W / wn_j.
__global__ __launch_bounds__(kBlock) void k_adj_app2(PV pv, GridDev g, SimParams P, NbrList L, AdjArgs A, int n) {
  const int i = gtid();
  if (i >= n) return;
  const uint8_t ph = pv.s.phase[i];
  if (!is_fluid(ph)) {
    A.pV[i] = 0.0f;
    return;
  }
  const cryo_adj::Ctx<float> c = adj_ctx(P, 0.0f);
  const float4 pai = A.pa[i];
  float4 po = A.po[i];
  const bool recv = pai.z != 0.0f || pai.w != 0.0f;
  const float om = A.om[i], oh = A.oh[i], wn = pv.t.w_norm[i];
  if (!recv && !(om > 0.0f)) {  // neither side of any apply pair
    A.pV[i] = 0.0f;
    return;
  }
  const uint8_t tgt = pc_target(A.fl[i]);
  const V3 xi = ld3(pv.s.x[i]);
  const float Vi = pv.s.m[i] / pv.s.rho[i];
  float Vb = 0.0f;
  AV xr{0.0f, 0.0f, 0.0f};
  for_each_pair(L, g, i, xi, [&](int j) {
    if (j == i) return;
    const uint8_t phj = pv.s.phase[j];
    if (!is_fluid(phj)) return;
    const V3 rv = xi - ld3(pv.s.x[j]);
    const float r2 = norm2(rv);
    if (r2 >= P.support2) return;
    const float r = sqrtf(r2), W = kW(P.kernel, r), sW = cryo_adj::kern_s(c, r);
    const float omj = A.om[j];
    if (recv && omj > 0.0f && pc_target(A.fl[j]) == ph) {  // i receives from j
      const float wnj = pv.t.w_norm[j];
      const float shb = pai.z + pai.w * A.oh[j];
      Vb += shb * omj * W / wnj;
      xr += av(rv) * (shb * omj * Vi / wnj * sW);
    }
    if (om > 0.0f && phj == tgt && !(A.fl[j] & kDying)) {  // j receives from i
      const float4 paj = A.pa[j];
      if (paj.z != 0.0f || paj.w != 0.0f) {
        const float Vj = pv.s.m[j] / pv.s.rho[j];
        const float share = om * Vj * W / wn;
        const float shb = paj.z + paj.w * oh;
        po.y += paj.w * share;
        po.x += shb * Vj * W / wn;
        po.z -= shb * share / wn;
        xr += av(rv) * (shb * om * Vj / wn * sW);
      }
    }
  });
  A.po[i] = po;
  A.pV[i] = Vb;
  A.xp[i] = f4(av(A.xp[i]) + xr);
}

// Receiver normalisation (with the vapour density cap) and the Lee sources of
// each source; the interface-fraction cotangent's sums for the next kernel.
__global__ __launch_bounds__(kBlock) void k_adj_src(PV pv, GridDev g, SimParams P, NbrList L, const DevState* st,
                                                   AdjArgs A, int n) {
  const int i = gtid();
  double lee[3] = {0.0, 0.0, 0.0};  // d loss / d lee_evap, lee_cond, T_sat
  if (i < n) {
    const uint8_t ph = pv.s.phase[i];
    float4 pa = A.pa[i];
    const float4 po = A.po[i];
    float wnb = po.z, wrb = 0.0f, fab = 0.0f, sab = 0.0f, sob = 0.0f;
    int amax = -1;
    const uint8_t fl = A.fl[i];
    const V3 xi = ld3(pv.s.x[i]);
    if (is_fluid(ph) && (fl & kSource) && pv.t.w_norm[i] > 0.0f) {
      const float dt = st->dt, Ts = st->T_sat;
      const float m = pv.s.m[i], T = A.T1[i], alpha = pv.t.alpha[i];
      // the source's out_m before the cap (k_kick_sources' arithmetic)
      float om_pre = 0.0f, dm_r = 0.0f, dm_e = 0.0f;
      int br = 0;  // 1 rate, 2 energy cap, 3 mass cap
      if (ph == kLiquid) {
        const float h_out = P.h_fg + P.cp[1] * (Ts - P.T_ref);
        const float latent = h_out - P.cp[0] * (Ts - P.T_ref);
        const float rate = P.lee_evap * m * alpha * (T - Ts) / Ts;
        dm_r = rate * dt;
        dm_e = m * P.cp[0] * (T - Ts) / latent;
        float dm = dm_r < dm_e ? dm_r : dm_e;
        br = dm_r < dm_e ? 1 : 2;
        const float cap = P.max_transfer * m;
        if (!(dm < cap)) br = 3, dm = cap;
        om_pre = dm;
      } else {
        const float rate = P.lee_cond * m * alpha * (Ts - T) / Ts;
        dm_r = rate * dt;
        float cap = P.max_transfer * m;
        if (P.pc_density_limit > 0.0f && P.pc_density_limit * m < cap) cap = P.pc_density_limit * m;
        br = dm_r < cap ? 1 : 3;
        om_pre = dm_r < cap ? dm_r : cap;
      }
      float ompb = po.x;
      if (A.om[i] != om_pre) {  // the vapour density cap: om = limit wn / wr_max
        const uint8_t tgt = pc_target(fl);
        float wr_max = 0.0f;
        for_each_pair(L, g, i, xi, [&](int j) {
          if (j == i || pv.s.phase[j] != tgt || (A.fl[j] & kDying)) return;
          const float r2 = norm2(xi - ld3(pv.s.x[j]));
          if (r2 >= P.support2) return;
          const float wr = kW(P.kernel, sqrtf(r2)) / pv.s.rho[j];
          if (wr > wr_max) {
            wr_max = wr;
            amax = j;
          }
        });
        if (wr_max > 0.0f) {
          wnb += ompb * P.pc_density_limit / wr_max;
          wrb = -ompb * A.om[i] / wr_max;
        }
        ompb = 0.0f;
      }
      const float capf = om_pre / m;
      if (ph == kLiquid) {
        const float h_out = P.h_fg + P.cp[1] * (Ts - P.T_ref);
        const float latent = h_out - P.cp[0] * (Ts - P.T_ref);
        lee[2] = static_cast<double>(po.y * P.cp[1]);  // out_h = h_v(T_sat)
        if (br == 3) {
          pa.x += ompb * capf;
        } else if (br == 1) {
          pa.x += ompb * dm_r / m;
          fab = ompb * dm_r / alpha;
          pa.y += ompb * P.lee_evap * m * alpha / Ts * dt;
          lee[0] = static_cast<double>(ompb * m * alpha * (T - Ts) / Ts * dt);
          lee[2] -= static_cast<double>(ompb * P.lee_evap * m * alpha * T / (Ts * Ts) * dt);
        } else {
          pa.x += ompb * P.cp[0] * (T - Ts) / latent;
          pa.y += ompb * m * P.cp[0] / latent;
          lee[2] -= static_cast<double>(ompb * (m * P.cp[0] + dm_e * (P.cp[1] - P.cp[0])) / latent);
        }
      } else {
        pa.y += po.y * P.cp[1];  // out_h = h(T) of the condensing vapour
        if (br == 3) {
          pa.x += ompb * capf;
        } else {
          pa.x += ompb * dm_r / m;
          fab = ompb * dm_r / alpha;
          pa.y -= ompb * P.lee_cond * m * alpha / Ts * dt;
          lee[1] = static_cast<double>(ompb * m * alpha * (Ts - T) / Ts * dt);
          lee[2] += static_cast<double>(ompb * P.lee_cond * m * alpha * T / (Ts * Ts) * dt);
        }
      }
      A.pa[i] = pa;
    }
    if (fab != 0.0f) {  // alpha = s_other / s_all over fluid neighbours (self included in s_all)
      float sa = 0.0f, so = 0.0f;
      for_each_pair(L, g, i, xi, [&](int j) {
        if (!is_fluid(pv.s.phase[j])) return;
        const float r2 = norm2(xi - ld3(pv.s.x[j]));
        if (r2 >= P.support2) return;
        const float Wv = pv.s.m[j] / pv.s.rho[j] * kW(P.kernel, sqrtf(r2));
        sa += Wv;
        if (pv.s.phase[j] != ph) so += Wv;
      });
      if (sa > 0.0f) {
        sob = fab / sa;
        sab = -fab * (so / sa) / sa;
      }
    }
    A.pn[i] = make_float4(wnb, wrb, sab, sob);
    A.amax[i] = amax;
  }
  block_sums_store<3>(lee, A.partials + 9 * A.stride, A.stride);
}

// Closed tank (sph_solver.cu:586-607): this step's T_sat came from the previous
// step's end state, whose masses and temperatures are this taped state's:
// p = p_sat (S_mT / v_ull) / (rho_v T_sat0), v_ull = V0 - S_l / rho_l,
// T_sat = 1 / (1/T_sat0 - (R/h_fg) ln(p/p_sat)).  Rows 12..14: the inventory
// partials (S_l, S_v, S_mT); the T_sat cotangent is row 11.  Writes the
// cotangents of S_l and S_mT.
__global__ __launch_bounds__(1024) void k_adj_sat(const DevState* st, AdjArgs A, int nb, int nb_inv) {
  __shared__ double sh[1024];
  const double tsb = ordered_sum(A.partials + 11 * A.stride, nb, sh);
  const double m_liq = ordered_sum(A.partials + 12 * A.stride, nb_inv, sh);
  const double m_vap = ordered_sum(A.partials + 13 * A.stride, nb_inv, sh);
  const double mT = ordered_sum(A.partials + 14 * A.stride, nb_inv, sh);
  if (threadIdx.x != 0) return;
  double mlb = 0.0, mTb = 0.0;
  const FluidConst fl = A.fluid;
  const double v_u = st->fluid_volume0 - m_liq / fl.rho_liquid;
  const double ratio = v_u > 0 && m_vap > 0 ? (mT / v_u) / (static_cast<double>(fl.rho_vapor) * fl.T_sat) : 0.0;
  if (tsb != 0.0 && isfinite(ratio) && ratio > 0) {
    const double pu = fl.p_sat * ratio, Ts = st->T_sat;
    const double g = tsb * Ts * Ts * (static_cast<double>(fl.R_gas) / fl.h_fg) / pu;  // dT_sat / dp
    mTb = g * pu / mT;
    mlb = g * pu / (v_u * fl.rho_liquid);
  }
  A.satv[0] = mlb;
  A.satv[1] = mTb;
}

// Normalisation pairs (wn_i = sum V_j W over the target phase; the cap's wr_max
// pair) and interface-fraction pairs, both sides.
__global__ __launch_bounds__(kBlock) void k_adj_nrm(PV pv, GridDev g, SimParams P, NbrList L, AdjArgs A, int n) {
  const int i = gtid();
  if (i >= n) return;
  const uint8_t ph = pv.s.phase[i];
  if (!is_fluid(ph)) {
    A.pr[i] = 0.0f;
    return;
  }
  const cryo_adj::Ctx<float> c = adj_ctx(P, 0.0f);
  const float4 pni = A.pn[i];
  const uint8_t fli = A.fl[i];
  const bool src = (fli & kSource) && pni.x != 0.0f;
  const uint8_t tgt = pc_target(fli);
  const int ami = A.amax[i];
  const V3 xi = ld3(pv.s.x[i]);
  // every pair here has a source with out_m > 0 on one side (or is i's own)
  if (!src && ami < 0 && pni.z == 0.0f && pni.w == 0.0f && !cells_any(g, xi, kCellSrcV | kCellSrcL)) {
    A.pr[i] = A.pdp[i].w;
    return;
  }
  const float Vi = pv.s.m[i] / pv.s.rho[i], rhoi = pv.s.rho[i];
  float Vb = 0.0f, rhob = 0.0f;
  AV xr{0.0f, 0.0f, 0.0f};
  for_each_pair(L, g, i, xi, [&](int j) {
    const uint8_t phj = pv.s.phase[j];
    if (!is_fluid(phj)) return;
    const V3 rv = xi - ld3(pv.s.x[j]);
    const float r2 = norm2(rv);
    if (r2 >= P.support2) return;
    const float r = sqrtf(r2), W = kW(P.kernel, r);
    if (j == i) {  // s_all holds the owner itself
      Vb += pni.z * W;
      return;
    }
    const float sW = cryo_adj::kern_s(c, r);
    const AV rij = av(rv);
    const float Vj = pv.s.m[j] / pv.s.rho[j];
    if (src && phj == tgt && !(A.fl[j] & kDying)) xr += rij * (pni.x * Vj * sW);
    if (ami == j) xr += rij * (pni.y / pv.s.rho[j] * sW);
    const float4 pnj = A.pn[j];
    const uint8_t flj = A.fl[j];
    if ((flj & kSource) && pnj.x != 0.0f && ph == pc_target(flj) && !(fli & kDying)) {
      Vb += pnj.x * W;
      xr += rij * (pnj.x * Vi * sW);
    }
    if (A.amax[j] == i) {
      rhob -= pnj.y * (W / rhoi) / rhoi;
      xr += rij * (pnj.y / rhoi * sW);
    }
    const float k1 = pni.z + (phj != ph ? pni.w : 0.0f);  // i's interface fraction
    if (k1 != 0.0f) xr += rij * (k1 * Vj * sW);
    const float k2 = pnj.z + (phj != ph ? pnj.w : 0.0f);  // j's
    if (k2 != 0.0f) {
      Vb += k2 * W;
      xr += rij * (k2 * Vi * sW);
    }
  });
  A.pV[i] += Vb;
  A.pr[i] = rhob + A.pdp[i].w;
  A.xp[i] = f4(av(A.xp[i]) + xr);
}

constexpr int kCopyTable = 32;
struct CopyTable {
  CopyItem it[kCopyTable];
};
// blockIdx.y: the copy; 16-byte words over the x blocks, the byte tail by block 0.
__global__ __launch_bounds__(kBlock) void k_copy_batch(CopyTable t) {
  const CopyItem c = t.it[blockIdx.y];
  const unsigned long long nv = c.bytes / 16;
  const uint4* src = static_cast<const uint4*>(c.src);
  uint4* dst = static_cast<uint4*>(c.dst);
  for (unsigned long long k = blockIdx.x * static_cast<unsigned long long>(blockDim.x) + threadIdx.x; k < nv;
       k += static_cast<unsigned long long>(gridDim.x) * blockDim.x)
    dst[k] = src[k];
  const unsigned long long done = nv * 16;
  if (blockIdx.x == 0 && threadIdx.x < c.bytes - done)
    static_cast<uint8_t*>(c.dst)[done + threadIdx.x] = static_cast<const uint8_t*>(c.src)[done + threadIdx.x];
}

}  // namespace

void launch_copy_batch(const CopyItem* items, int count, hipStream_t s) {
  for (int b = 0; b < count; b += kCopyTable) {
    const int m = count - b < kCopyTable ? count - b : kCopyTable;
    CopyTable t{};
    unsigned long long most = 0;
    for (int k = 0; k < m; ++k) {
      t.it[k] = items[b + k];
      if (t.it[k].bytes > most) most = t.it[k].bytes;
    }
    const unsigned long long words = most / 16 + 1;
    const unsigned long long bx = (words + kBlock - 1) / kBlock;
    const unsigned gx = static_cast<unsigned>(bx < 512 ? bx : 512);
    k_copy_batch<<<dim3(gx, static_cast<unsigned>(m)), kBlock, 0, s>>>(t);
  }
}

void launch_adj_fill(float4* a, float4* b, float4* c, int n, hipStream_t s) {
  if (n <= 0) return;
  k_adj_fill<<<blocks_for(n), kBlock, 0, s>>>(a, b, c, n);
}
void launch_adj_set_frame(DevState* st, const float* acc, const float* omega, const float* omega_dot, hipStream_t s) {
  k_adj_set_frame<<<1, 1, 0, s>>>(st, acc, omega, omega_dot);
}
void launch_adj_load_sum(const PV& pv, int n, double* partials, int stride, double* out, hipStream_t s) {
  const int nb = blocks_for(n > 0 ? n : 1);
  k_adj_load_sum<<<nb, kBlock, 0, s>>>(pv, n, partials, stride);
  k_adj_finish<<<1, 1024, 0, s>>>(partials, stride, nb, 0, -1, 1.0, out, 3);
}
void launch_adj_csf_only(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, const DevState* st,
                         const AdjArgs& A, int n, double* theta_out, hipStream_t s) {
  const int nb = blocks_for(n > 0 ? n : 1);
  k_adj_csf1<<<nb, kBlock, 0, s>>>(pv, g, P, L, A, n);
  if (A.sacc || A.gpb) k_adj_csf2<true><<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n, 6);
  else k_adj_csf2<false><<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n, 6);
  if (A.sacc) k_adj_csf3<<<nb, kBlock, 0, s>>>(pv, g, P, L, A, n, 6);
  k_adj_finish<<<1, 1024, 0, s>>>(A.partials, A.stride, nb, 6, -1, 1.0, theta_out, 2);
}
void launch_adj_step(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, const DevState* st,
                     const AdjArgs& A, int n, double* grad_out, double* theta_out, hipStream_t s) {
  const int nb = blocks_for(n > 0 ? n : 1);
  if (A.pc) {
    k_adj_reb1<<<nb, kBlock, 0, s>>>(pv, g, P, L, A, n);
    k_adj_reb2<<<nb, kBlock, 0, s>>>(pv, g, P, L, A, n);
    k_adj_app1<<<nb, kBlock, 0, s>>>(pv, g, P, L, A, n);
    k_adj_app2<<<nb, kBlock, 0, s>>>(pv, g, P, L, A, n);
    k_adj_src<<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n);
    k_adj_finish<<<1, 1024, 0, s>>>(A.partials, A.stride, nb, 9, -1, 1.0, A.lee_out, 2);
    k_adj_nrm<<<nb, kBlock, 0, s>>>(pv, g, P, L, A, n);
  }
  k_adj_prep<<<nb, kBlock, 0, s>>>(pv, P, A, st, n);
  if (A.heat) k_adj_finish<<<1, 1024, 0, s>>>(A.partials, A.stride, nb, 8, -1, 1.0, A.hsum, 1);
  if (A.csf) {
    k_adj_csf1<<<nb, kBlock, 0, s>>>(pv, g, P, L, A, n);
    if (A.sacc || A.gpb) k_adj_csf2<true><<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n, 6);
    else k_adj_csf2<false><<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n, 6);
    if (A.sacc) k_adj_csf3<<<nb, kBlock, 0, s>>>(pv, g, P, L, A, n, 6);
    k_adj_finish<<<1, 1024, 0, s>>>(A.partials, A.stride, nb, 6, -1, 1.0, theta_out, 2);
  }
  if (A.tvf) k_adj_wall<true><<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n);
  else k_adj_wall<false><<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n);
  const bool ex = A.vtr != nullptr || A.gbal != nullptr || A.gpb != nullptr || P.kernel.type != 1;
  if (A.tvf) {
    if (ex) k_adj_fluid<true, true><<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n);
    else k_adj_fluid<true, false><<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n);
  } else {
    if (ex) k_adj_fluid<false, true><<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n);
    else k_adj_fluid<false, false><<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n);
  }
  if (A.sat) {
    const int nb_inv = launch_inventory(pv, st, n, A.partials + 12 * A.stride, A.stride, s);
    k_adj_sat<<<1, 1024, 0, s>>>(st, A, nb, nb_inv);
  }
  k_adj_density<<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n);
  // d loss / d acc = -(d loss / d g_eff): wall rows 0..2 + fluid rows 3..5
  k_adj_finish<<<1, 1024, 0, s>>>(A.partials, A.stride, nb, 0, 3, -1.0, grad_out, 3);
  if (A.wout) k_adj_finish<<<1, 1024, 0, s>>>(A.partials, A.stride, nb, 15, -1, 1.0, A.wout, 6);
}

}  // namespace cryo
