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
                              0.6f * P.dx, dt, P.vol0, P.no_slip};
}
__device__ inline cryo_adj::Side<float> side(const PV& pv, const SimParams& P, int i) {
  const uint8_t ph = pv.s.phase[i];
  return cryo_adj::Side<float>{av(pv.s.x[i]),     av(pv.s.v[i]),
                               pv.s.m[i],         pv.s.rho[i],
                               pv.t.p[i],         sound_speed(P, ph, pv.s.T[i]),
                               sel2(P.mu, pidx(ph)), ph == kVapor ? P.alpha_av_vap : P.alpha_av,
                               static_cast<int>(ph)};
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

__global__ __launch_bounds__(kBlock) void k_adj_set_frame(DevState* st, const float* acc) {
  const V3 a{acc[0], acc[1], acc[2]};
  st->frame_acc = a;
  st->gravity = st->cfg_gravity - a;
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
  if (i < n) {
    float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (is_fluid(pv.s.phase[i])) {
      const float half = 0.5f * st->dt;
      const float4 a = A.ca[i], v = A.cv[i];
      r = make_float4(a.x + v.x * half, a.y + v.y * half, a.z + v.z * half, 0.0f);
      g[0] = r.x;  // a = g_eff + ...
      g[1] = r.y;
      g[2] = r.z;
    }
    A.at[i] = r;
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
}

__device__ inline bool live_normal(const PV& pv, int i) { return norm2(ld3(pv.t.ns[i])) > 0.25f; }

// Curvature sums' cotangents: kappa = 3 num / den, a += nrm coeff kappa.
__global__ __launch_bounds__(kBlock) void k_adj_csf1(PV pv, GridDev g, SimParams P, NbrList L, AdjArgs A, int n) {
  const int i = gtid();
  if (i >= n) return;
  float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (is_fluid(pv.s.phase[i]) && live_normal(pv, i)) {
    const float coeff = 2.0f * P.sigma / (P.rho0[0] + P.rho0[1]);
    const float4 at = A.at[i], nr = pv.s.nrm[i];
    const V3 xi = ld3(pv.s.x[i]), ni = ld3(pv.t.ns[i]);
    float num = 0.0f, den = 0.0f;  // the curvature sums (the forward pass's, recomputed)
    for_each_pair(L, g, i, xi, [&](int j) {
      if (j == i || !is_fluid(pv.s.phase[j]) || !live_normal(pv, j)) return;
      const V3 rij = xi - ld3(pv.s.x[j]);
      const float r2 = norm2(rij);
      if (!(r2 < P.support2)) return;
      const float r = sqrtf(r2);
      const float dW = kdW(P.kernel, r);
      const float Vj = pv.s.m[j] / pv.s.rho[j];
      num += Vj * dot(ni - ld3(pv.t.ns[j]), rij * (dW / r));
      den += Vj * r * (-dW);
    });
    if (den > 0.0f) {
      const float kappa = 3.0f * num / den;
      const float kb = coeff * (at.x * nr.x + at.y * nr.y + at.z * nr.z);
      out = make_float4(3.0f * kb / den, -kappa * kb / den, kappa, 0.0f);
    }
  }
  A.cs[i] = out;
}

// Per fluid particle: both sides of the curvature pairs, the normalisation,
// then the contact-angle rotation of its colour gradient (recomputed).
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
        if (P.wall_colour) gradc += kgradW(P.kernel, rkj, r) * (P.vol0 * (pv.t.alpha[j] - ck));
        return;
      }
      if (j == k || !is_fluid(phj)) return;
      const float Vj = pv.s.m[j] / pv.s.rho[j];
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
    const float4 at = A.at[k];
    AV nrmb = livek ? AV{at.x, at.y, at.z} * (coeff * csk.z) : AV{0.0f, 0.0f, 0.0f};
    const AV nrk = av(pv.s.nrm[k]);
    const float G = sqrtf(cryo_adj::dot(nrk, nrk));
    if (G > P.normal_eps) nrmb += (nsb - nsk * cryo_adj::dot(nsk, nsb)) * (1.0f / G);  // ns = nrm / |nrm|
    const cryo_adj::RotGrad<float> ro = cryo_adj::rot_vjp(av(gradc), av(wall_n), ww, P.cos_theta, P.sin_theta,
                                                          P.wall_adh_w, P.normal_eps, P.wetting != 1 ? 1 : 0, nrmb);
    A.gcb[k] = make_float4(ro.g.x, ro.g.y, ro.g.z, ro.ww);
    A.wnb[k] = f4(ro.wall_n);
    A.csv[k] = Vb;
    A.xc[k] = f4(xr);
    th[0] = ro.cos_t;
    th[1] = ro.sin_t;
  } else if (k < n) {
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    A.gcb[k] = z;
    A.wnb[k] = z;
    A.csv[k] = 0.0f;
    A.xc[k] = z;
  }
  block_sums_store<2>(th, A.partials + row * A.stride, A.stride);
}

__global__ __launch_bounds__(kBlock) void k_adj_wall(PV pv, GridDev g, SimParams P, NbrList L, const DevState* st,
                                                    AdjArgs A, int n) {
  const int w = gtid();
  double gs[3] = {0.0, 0.0, 0.0};
  if (w < n && pv.s.phase[w] == kWall) {
    const float dt = st->dt;
    const cryo_adj::Ctx<float> c = adj_ctx(P, dt);
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
      if (A.tvf) {  // a transport owner's wall volume reads this wall's pressure
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

__global__ __launch_bounds__(kBlock) void k_adj_fluid(PV pv, GridDev g, SimParams P, NbrList L, const DevState* st,
                                                     AdjArgs A, int n) {
  const int i = gtid();
  if (i >= n) return;
  const uint8_t phi = pv.s.phase[i];
  if (!is_fluid(phi)) return;
  const float dt = st->dt;
  const cryo_adj::Ctx<float> c = adj_ctx(P, dt);
  const AV Fb = load_cot(A.S, dt);
  const AV grav = av(st->gravity);
  const cryo_adj::Side<float> I = side(pv, P, i);
  const bool vap = phi == kVapor;
  const float4 ati = A.at[i];
  const AV cA{ati.x, ati.y, ati.z};
  const AV cF = cA * (1.0f / I.m);
  const AV cFw = cF + Fb, cAw = cA + Fb * I.m;
  AV xr{0.0f, 0.0f, 0.0f}, vr{0.0f, 0.0f, 0.0f};
  float rb = 0.0f, pb = 0.0f;
  const float ci = colour(phi), Vk = I.m / I.rho;
  AV gbk{0.0f, 0.0f, 0.0f}, wbk{0.0f, 0.0f, 0.0f};
  float wwbk = 0.0f, Vbk = 0.0f;
  // transport acceleration: own sums (owner active) and the neighbours' sums that read V_i = 1 / sigma_i
  const float4 tvi = A.tvf ? A.tvs[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const bool tv_own = tvi.w > 0.0f;
  const AV Sbi{tvi.x, tvi.y, tvi.z};
  const float Vti = A.tvf ? 1.0f / pv.t.sigma[i] : 0.0f, Vti2 = Vti * Vti;
  float vt2b = 0.0f, vtb = 0.0f, trhob = 0.0f, tsb = 0.0f;
  AV tsum{0.0f, 0.0f, 0.0f};
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
      rb += o.rho_i;
      pb += o.p_i;
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
          (void)cryo_adj::tvf_wall_vjp(Vti, I.rho, rho_w, ot.vj, vtb, trhob);  // the wall's share: k_adj_wall
        }
      }
      if (A.csf) {
        const float alpha_w = pv.t.alpha[j];
        const cryo_adj::GcGrad<float> og = cryo_adj::gc_fw_vjp(c, I.x - xw, alpha_w, ci, av(pv.s.nrm[j]),
                                                               P.wall_colour, gbk, wbk, wwbk);
        xr += og.r;
        const AV rwk = xw - I.x;
        const float r2 = cryo_adj::dot(rwk, rwk);
        if (cv.w != 0.0f && r2 < P.support2)  // alpha_w = sw0 / swt: d/dx_k = (c_k - alpha_w) / swt (-s r_wk)
          xr += rwk * (-cv.w * (ci - alpha_w) * cryo_adj::wendland_s(c, sqrtf(r2)));
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
    if (A.tvf) {
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
    const cryo_adj::Side<float> J = side(pv, P, j);
    const float4 atj = A.at[j];
    const AV cAj{atj.x, atj.y, atj.z};
    const cryo_adj::FFGrad<float> o1 = cryo_adj::ff_vjp(c, I, J, cF, cA);         // i's pair term
    const cryo_adj::FFGrad<float> o2 = cryo_adj::ff_vjp(c, J, I, cAj * (1.0f / J.m), cAj);  // j's
    xr += o1.r;
    xr -= o2.r;
    vr += o1.vij;
    vr -= o2.vij;
    rb += o1.rho_i + o2.rho_j;
    pb += o1.p_i + o2.p_j;
  });
  const float dpdr = cryo_adj::eos_dp_drho(vap ? 1 : 0, I.rho, pv.s.T[i], I.p, P.rho0[0], P.tait_B, P.tait_gamma,
                                           P.gas_R, P.clamp_neg);
  rb -= Vbk * Vk / I.rho;  // V = m / rho in the curvature and colour sums
  if (tv_own) {  // K = -pb / (rho Vt), Vt = 1 / sigma
    const float K = tvf_K(pv, P, i);
    const float4 ct = A.ct[i];
    const float Kb = ct.x * tsum.x + ct.y * tsum.y + ct.z * tsum.z;
    trhob -= K * Kb / I.rho;
    vtb -= K * Kb / Vti;
    vtb += 2.0f * Vti * vt2b;
    tsb -= vtb * Vti * Vti;
    rb += trhob;
  }
  A.sb[i] = (rb + pb * dpdr) * (vap ? I.m : pv.s.m0[i]) + tsb;  // rho = m0 sigma (liquid), m sigma (vapour)
  A.xr[i] = f4(xr);
  A.vr[i] = f4(vr);
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
    xr += rij * ((sbi + sj) * cryo_adj::wendland_s(c, sqrtf(r2)));
  });
  // x' = x + v_half dt, v' = v_half + a' dt/2, v_half = v + a dt/2
  const AV X = av(A.cx[i]) + xr;
  const AV Vh = av(A.cv[i]) + av(A.vr[i]) + X * dt;
  A.nx[dst] = f4(X);
  A.nv[dst] = f4(Vh);
  A.na[dst] = f4(Vh * half);
  if (A.tvf) A.nt[dst] = f4(X * (half * dt));  // x' = x + (v_half + ab dt/2) dt
}

}  // namespace

void launch_adj_fill(float4* a, float4* b, float4* c, int n, hipStream_t s) {
  if (n <= 0) return;
  k_adj_fill<<<blocks_for(n), kBlock, 0, s>>>(a, b, c, n);
}
void launch_adj_set_frame(DevState* st, const float* acc, hipStream_t s) { k_adj_set_frame<<<1, 1, 0, s>>>(st, acc); }
void launch_adj_load_sum(const PV& pv, int n, double* partials, int stride, double* out, hipStream_t s) {
  const int nb = blocks_for(n > 0 ? n : 1);
  k_adj_load_sum<<<nb, kBlock, 0, s>>>(pv, n, partials, stride);
  k_adj_finish<<<1, 1024, 0, s>>>(partials, stride, nb, 0, -1, 1.0, out, 3);
}
void launch_adj_csf_only(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, const DevState* st,
                         const AdjArgs& A, int n, double* theta_out, hipStream_t s) {
  const int nb = blocks_for(n > 0 ? n : 1);
  k_adj_csf1<<<nb, kBlock, 0, s>>>(pv, g, P, L, A, n);
  k_adj_csf2<<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n, 6);
  k_adj_finish<<<1, 1024, 0, s>>>(A.partials, A.stride, nb, 6, -1, 1.0, theta_out, 2);
}
void launch_adj_step(const PV& pv, const GridDev& g, const SimParams& P, const NbrList& L, const DevState* st,
                     const AdjArgs& A, int n, double* grad_out, double* theta_out, hipStream_t s) {
  const int nb = blocks_for(n > 0 ? n : 1);
  k_adj_prep<<<nb, kBlock, 0, s>>>(pv, P, A, st, n);
  if (A.csf) {
    k_adj_csf1<<<nb, kBlock, 0, s>>>(pv, g, P, L, A, n);
    k_adj_csf2<<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n, 6);
    k_adj_finish<<<1, 1024, 0, s>>>(A.partials, A.stride, nb, 6, -1, 1.0, theta_out, 2);
  }
  k_adj_wall<<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n);
  k_adj_fluid<<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n);
  k_adj_density<<<nb, kBlock, 0, s>>>(pv, g, P, L, st, A, n);
  // d loss / d acc = -(d loss / d g_eff): wall rows 0..2 + fluid rows 3..5
  k_adj_finish<<<1, 1024, 0, s>>>(A.partials, A.stride, nb, 0, 3, -1.0, grad_out, 3);
}

}  // namespace cryo
