// Per-pair vector-Jacobian products of the hydrodynamic step: the pieces of the
// reverse-mode adjoint of one step, written out by hand.
//
// The forward pair terms they differentiate are those of
//   wall_boundary  (navier_stokes.cu:16-69)   Shepard wall pressure + ghost velocity
//   momentum       (navier_stokes.cu:121-245) pressure, Morris viscosity, Monaghan
//                                             artificial viscosity, wall pressure,
//                                             near-wall guard, no-slip shear, wall AV
//   density        (SPH_kernels.cu:6-34)      sigma = sum_j W(r_ij)
//   CSF            (surface_tension.cu:5-177) colour gradient, contact-angle
//                                             rotation, curvature, classic or
//                                             balanced force
// with the Wendland C2 or the cubic-spline kernel (SPH_kernels.cuh:14-64).  Each function returns
// d(c . term)/d(inputs) for the cotangent c of the quantity the term adds to.
//
// Templated on the scalar type: float on the GPU, double in the FP64 CPU
// adjoint that checks it.  Plain expressions only, so the same source builds
// with hipcc (device and host) and with g++.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define CRYO_ADJ_HD __host__ __device__
#else
#define CRYO_ADJ_HD
#endif

namespace cryo_adj {

template <class T>
struct Vec {
  T x, y, z;
};
template <class T>
CRYO_ADJ_HD inline Vec<T> operator+(Vec<T> a, Vec<T> b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T>
CRYO_ADJ_HD inline Vec<T> operator-(Vec<T> a, Vec<T> b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T>
CRYO_ADJ_HD inline Vec<T> operator*(Vec<T> a, T s) {
  return {a.x * s, a.y * s, a.z * s};
}
template <class T>
CRYO_ADJ_HD inline Vec<T>& operator+=(Vec<T>& a, Vec<T> b) {
  a.x += b.x;
  a.y += b.y;
  a.z += b.z;
  return a;
}
template <class T>
CRYO_ADJ_HD inline Vec<T>& operator-=(Vec<T>& a, Vec<T> b) {
  a.x -= b.x;
  a.y -= b.y;
  a.z -= b.z;
  return a;
}
template <class T>
CRYO_ADJ_HD inline T dot(Vec<T> a, Vec<T> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}

// Constants of one rates evaluation.
template <class T>
struct Ctx {
  T norm, inv_h, h;  // kernel normalisation, 1 / h, h
  T support2;        // (2h)^2
  T eta2;            // 0.01 h^2 (the AV / viscosity regulariser)
  T r_guard;         // 0.6 dx
  T dt;              // the step's dt (the guard's damping term)
  T vol0;            // wall particle volume dx^3
  int no_slip;
  int type;          // 1 Wendland C2: W = norm (1 - q/2)^4 (2q + 1);
                     // 0 cubic spline: W = norm (1 - 3q^2/2 + 3q^3/4) (q < 1), norm (2 - q)^3 / 4
  int tvf_stress;    // the transport-velocity stress term of momentum is on
};

// W(r), s = W'(r) / r and s' / r = (ds/dr) / r, q = r / h < 2.
template <class T>
CRYO_ADJ_HD inline T kern_w(const Ctx<T>& c, T r) {
  const T q = r * c.inv_h;
  if (c.type == 0) {
    if (q < T(1)) return c.norm * (T(1) - T(1.5) * q * q + T(0.75) * q * q * q);
    const T u = T(2) - q;
    return c.norm * T(0.25) * u * u * u;
  }
  const T t = T(1) - T(0.5) * q;
  return c.norm * t * t * t * t * (T(2) * q + T(1));
}
template <class T>
CRYO_ADJ_HD inline T kern_s(const Ctx<T>& c, T r) {
  if (c.type == 0) {
    const T q = r * c.inv_h;
    if (q < T(1)) return c.norm * c.inv_h * c.inv_h * (T(-3) + T(2.25) * q);
    const T u = T(2) - q;
    return T(-0.75) * c.norm * c.inv_h * u * u / r;
  }
  const T t = T(1) - T(0.5) * r * c.inv_h;
  return T(-5) * c.norm * c.inv_h * c.inv_h * t * t * t;
}
template <class T>
CRYO_ADJ_HD inline T kern_spr(const Ctx<T>& c, T r) {
  if (c.type == 0) {
    const T q = r * c.inv_h;
    if (q < T(1)) return T(2.25) * c.norm * c.inv_h * c.inv_h * c.inv_h / r;
    const T u = T(2) - q;
    return T(0.75) * c.norm * c.inv_h * u * (T(2) * c.inv_h / r + u / (r * r)) / r;
  }
  const T t = T(1) - T(0.5) * r * c.inv_h;
  return T(7.5) * c.norm * c.inv_h * c.inv_h * c.inv_h * t * t / r;
}

// One fluid particle as a member of a pair.
template <class T>
struct Side {
  Vec<T> x, v;
  T m, rho, p;
  T c;      // sound speed
  T eta;    // dynamic viscosity mu[phase]
  T alpha;  // AV coefficient of its phase
  int ph;
  Vec<T> vt;  // transport velocity (read by the stress term only)
};

// Fluid-fluid term of momentum with owner I and neighbour J (j != i):
//   a_I += (T1 + T2) / m_I + T3,
//   T1 = -gw vsum pt,  T2 = v_IJ vsum eta_IJ s,  T3 = -gw m_J Pi (same phase, v_IJ.r_IJ < 0)
// with gw = r_IJ s, vsum = V_I^2 + V_J^2, pt = (rho_J p_I + rho_I p_J) / (rho_I + rho_J).
// cF is the cotangent of the force sum (a_I's cotangent / m_I), cA that of the
// artificial-viscosity sum.  r is the cotangent of r_IJ = x_I - x_J, vij that
// of v_IJ = v_I - v_J; m_* and c_* those of the masses (through V = m / rho
// and m_J in T3) and sound speeds.  f is the value T1 + T2 of the pair (the
// owner needs the force sum for d(F / m_I) / d m_I).
template <class T>
struct FFGrad {
  Vec<T> r, vij;
  T rho_i, rho_j, p_i, p_j, m_i, m_j, c_i, c_j;
  Vec<T> f;
  Vec<T> vi, vj, vti, vtj;  // the stress term's own cotangents of v_I, v_J, vt_I, vt_J
};
template <class T>
CRYO_ADJ_HD inline FFGrad<T> ff_vjp(const Ctx<T>& c, const Side<T>& I, const Side<T>& J, Vec<T> cF, Vec<T> cA) {
  FFGrad<T> o{{T(0), T(0), T(0)}, {T(0), T(0), T(0)}, T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0),
              {T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}};
  const Vec<T> rij = I.x - J.x;
  const T r2 = dot(rij, rij);
  if (!(r2 < c.support2)) return o;
  const T r = std::sqrt(r2);
  if (!(r > T(1e-12) * c.h)) return o;  // gw = 0 and the viscous term is skipped
  const T s = kern_s(c, r), spr = kern_spr(c, r);
  const T Vi = I.m / I.rho, Vj = J.m / J.rho;
  const T vsum = Vi * Vi + Vj * Vj;
  const T rs = I.rho + J.rho;
  const T pt = (J.rho * I.p + I.rho * J.p) / rs;
  const T cFr = dot(cF, rij);
  const T G = s * cFr;  // cF . gw
  // T1
  o.r = (cF * s + rij * (cFr * spr)) * (-vsum * pt);
  o.f = rij * (-s * vsum * pt);
  T vsum_b = -pt * G;
  const T pt_b = -vsum * G;
  T rhoi_s = T(0), rhoj_s = T(0);
  // T4 (transport stress, SPH_kernels.cu:162-203 with tvf_stress):
  //   force += A vsum / 2,  A = v_I rho_I (dv_I . gw) + v_J rho_J (dv_J . gw),  dv = vt - v
  if (c.tvf_stress) {
    const Vec<T> dvi = I.vt - I.v, dvj = J.vt - J.v;
    const Vec<T> gw = rij * s;
    const T di = dot(dvi, gw), dj = dot(dvj, gw);
    const T a = dot(cF, I.v), b = dot(cF, J.v);
    const T h = T(0.5) * vsum;
    o.vi = (cF * di - gw * a) * (h * I.rho);
    o.vj = (cF * dj - gw * b) * (h * J.rho);
    o.vti = gw * (h * a * I.rho);
    o.vtj = gw * (h * b * J.rho);
    rhoi_s = h * a * di;
    rhoj_s = h * b * dj;
    vsum_b += T(0.5) * (a * I.rho * di + b * J.rho * dj);
    const Vec<T> Dg = (dvi * (a * I.rho) + dvj * (b * J.rho)) * h;
    o.r += Dg * s + rij * (dot(Dg, rij) * spr);
    o.f += (I.v * (I.rho * di) + J.v * (J.rho * dj)) * h;
  }
  // T2
  const Vec<T> vij = I.v - J.v;
  const T es = I.eta + J.eta;
  if (es > T(0)) {
    const T eta = T(2) * I.eta * J.eta / es;
    const T cFv = dot(cF, vij);
    o.r += rij * (vsum * eta * cFv * spr);
    o.vij += cF * (vsum * eta * s);
    o.f += vij * (vsum * eta * s);
    vsum_b += eta * s * cFv;
  }
  o.m_i = vsum_b * T(2) * Vi / I.rho;
  o.m_j = vsum_b * T(2) * Vj / J.rho;
  T rhoi_b = vsum_b * (T(-2) * Vi * Vi / I.rho) + pt_b * (J.p - pt) / rs;
  T rhoj_b = vsum_b * (T(-2) * Vj * Vj / J.rho) + pt_b * (I.p - pt) / rs;
  o.p_i = pt_b * J.rho / rs;
  o.p_j = pt_b * I.rho / rs;
  // T3: Pi = -alpha (c_I + c_J)/2 mu / ((rho_I + rho_J)/2), mu = h vr / (r^2 + eta2)
  if (I.alpha > T(0) && I.ph == J.ph) {
    const T vr = dot(vij, rij);
    if (vr < T(0)) {
      const T den = r2 + c.eta2;
      const T mu = c.h * vr / den;
      const T cbar = T(0.5) * (I.c + J.c), rbar = T(0.5) * rs;
      const T Pi = -I.alpha * cbar * mu / rbar;
      const T cAr = dot(cA, rij);
      const T H = s * cAr;  // cA . gw
      o.r += (cA * s + rij * (cAr * spr)) * (-J.m * Pi);
      const T Pi_b = -J.m * H;
      o.m_j -= Pi * H;
      o.c_i = o.c_j = T(0.5) * Pi_b * (-I.alpha * mu / rbar);
      const T mu_b = Pi_b * (-I.alpha * cbar / rbar);
      const T rbar_b = Pi_b * (-Pi / rbar);
      rhoi_b += T(0.5) * rbar_b;
      rhoj_b += T(0.5) * rbar_b;
      const T vr_b = mu_b * c.h / den;
      const T r2_b = -mu_b * mu / den;
      o.vij += rij * vr_b;
      o.r += vij * vr_b + rij * (T(2) * r2_b);
    }
  }
  o.rho_i = c.tvf_stress ? rhoi_b + rhoi_s : rhoi_b;
  o.rho_j = c.tvf_stress ? rhoj_b + rhoj_s : rhoj_b;
  return o;
}

// Fluid-wall term of momentum with fluid owner I and a wall particle at xw
// with ghost velocity vw and wall pressure pw (of I's phase):
//   force += -gw vsum (p_I + pw)/2                       (W1)
//   acc   += n 4 c_I^2 q^2 / r_g, q = 1 - r/r_g          (W2, r < r_g)
//   acc   -= n q (v_I . n) / dt                          (W3, r < r_g, v_I . n < 0)
//   force += v_Iw vsum eta_I s                           (W4, no-slip)
//   acc   += gw vol0 alpha c_I mu                        (W5, no-slip, v_Iw . r < 0)
// with vsum = V_I^2 + vol0^2 and n = r_Iw / r.  cF / cA: cotangents of the
// force and acceleration sums (both carry a_I's and the wall-load cotangent).
// r: cotangent of r_Iw = x_I - xw; vi: of v_I alone (W3); vij: of
// v_Iw = v_I - vw; p_w: of pw; m_i, c_i: of m_I (through V_I) and c_I.
// f and acc are the pair's values W1 + W4 and W2 + W3 + W5.
template <class T>
struct FWGrad {
  Vec<T> r, vi, vij;
  T rho_i, p_i, p_w, m_i, c_i;
  Vec<T> f, acc;
  Vec<T> vti;  // the stress term's cotangent of vt_I
};
template <class T>
CRYO_ADJ_HD inline FWGrad<T> fw_vjp(const Ctx<T>& c, const Side<T>& I, Vec<T> xw, Vec<T> vw, T pw, Vec<T> cF,
                                    Vec<T> cA) {
  FWGrad<T> o{{T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}, T(0), T(0), T(0), T(0), T(0),
              {T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}};
  const Vec<T> rij = I.x - xw;
  const T r2 = dot(rij, rij);
  if (!(r2 < c.support2)) return o;
  const T r = std::sqrt(r2);
  if (!(r > T(1e-12) * c.h)) return o;  // every term vanishes or is skipped
  const T s = kern_s(c, r), spr = kern_spr(c, r);
  const T Vi = I.m / I.rho;
  const T vsum = Vi * Vi + c.vol0 * c.vol0;
  const T cFr = dot(cF, rij);
  const T G = s * cFr;
  // W1
  const T pm = T(0.5) * (I.p + pw);
  o.r = (cF * s + rij * (cFr * spr)) * (-vsum * pm);
  o.f = rij * (-s * vsum * pm);
  o.p_i = T(-0.5) * vsum * G;
  o.p_w = o.p_i;
  T vsum_b = -pm * G;
  T rho_s = T(0);
  // W6 (transport stress): force += v_I (vsum / 2) rho_I (dv_I . gw), dv_I = vt_I - v_I
  if (c.tvf_stress) {
    const Vec<T> dvi = I.vt - I.v;
    const Vec<T> gw = rij * s;
    const T di = dot(dvi, gw), a = dot(cF, I.v), h = T(0.5) * vsum;
    o.vi += (cF * di - gw * a) * (h * I.rho);
    o.vti = gw * (h * a * I.rho);
    rho_s = h * a * di;
    vsum_b += T(0.5) * a * I.rho * di;
    const Vec<T> Dg = dvi * (h * a * I.rho);
    o.r += Dg * s + rij * (dot(Dg, rij) * spr);
    o.f += I.v * (h * I.rho * di);
  }
  // W2, W3
  if (r < c.r_guard) {
    const T q = T(1) - r / c.r_guard;
    const Vec<T> n = rij * (T(1) / r);
    const T a = dot(cA, n);
    const T kap = T(4) * I.c * I.c / c.r_guard;
    o.r += (cA - n * a) * (kap * q * q / r);
    o.r -= n * (kap * T(2) * q * a / c.r_guard);
    o.c_i = a * T(8) * I.c * q * q / c.r_guard;
    o.acc = n * (kap * q * q);
    const T b = dot(I.v, n);
    if (b < T(0) && c.dt > T(0)) {
      const T k = T(-1) / c.dt;
      o.acc += n * (k * q * b);
      o.vi += n * (k * q * a);
      o.r += (n * (-(a * b) / c.r_guard) + (cA - n * a) * (q * b / r) + (I.v - n * b) * (q * a / r)) * k;
    }
  }
  if (c.no_slip) {
    const Vec<T> vij = I.v - vw;
    const T cFv = dot(cF, vij);
    // W4
    o.r += rij * (vsum * I.eta * cFv * spr);
    o.vij += cF * (vsum * I.eta * s);
    o.f += vij * (vsum * I.eta * s);
    vsum_b += I.eta * s * cFv;
    // W5
    const T vr = dot(vij, rij);
    if (I.alpha > T(0) && vr < T(0)) {
      const T den = r2 + c.eta2;
      const T mu = c.h * vr / den;
      const T cAr = dot(cA, rij);
      const T H = s * cAr;
      const T k5 = c.vol0 * I.alpha * I.c;
      o.r += (cA * s + rij * (cAr * spr)) * (k5 * mu);
      o.acc += rij * (s * k5 * mu);
      o.c_i += H * c.vol0 * I.alpha * mu;
      const T mu_b = k5 * H;
      const T vr_b = mu_b * c.h / den;
      const T r2_b = -mu_b * mu / den;
      o.vij += rij * vr_b;
      o.r += vij * vr_b + rij * (T(2) * r2_b);
    }
  }
  o.rho_i = vsum_b * (T(-2) * Vi * Vi / I.rho);
  if (c.tvf_stress) o.rho_i += rho_s;
  o.m_i = vsum_b * T(2) * Vi / I.rho;
  return o;
}

// Shepard wall values seen from one fluid neighbour F of a wall particle at xw
// (navier_stokes.cu:16-69), phase k = phase of F:
//   pw_k = sum_F (p_F + rho_F g . r_wF) W / sum_F W,   vw = -sum_F v_F W / sum_F W
// cp = (cotangent of pw_k) / sum_{F in k} W and cvw = (cotangent of vw) / sum_F W
// are the wall's; pw = pw_k, vw its values.  Returns F's share: cotangents of
// x_F, v_F, p_F, rho_F.  The cotangent of g is accumulated by the wall itself
// (cp * sum rho_F r_wF W).
template <class T>
struct WBGrad {
  Vec<T> x, v;
  T p, rho;
};
template <class T>
CRYO_ADJ_HD inline WBGrad<T> wb_vjp(const Ctx<T>& c, const Side<T>& F, Vec<T> xw, T pw, Vec<T> vw, T cp,
                                    Vec<T> cvw, Vec<T> g) {
  WBGrad<T> o{{T(0), T(0), T(0)}, {T(0), T(0), T(0)}, T(0), T(0)};
  const Vec<T> rwf = xw - F.x;
  const T r2 = dot(rwf, rwf);
  if (!(r2 < c.support2)) return o;
  const T r = std::sqrt(r2);
  const T W = kern_w(c, r), s = kern_s(c, r);
  const T gr = dot(g, rwf);
  o.p = cp * W;
  o.rho = cp * gr * W;
  // dW/dx_F = -s r_wF
  o.x = g * (-cp * F.rho * W) + rwf * (-cp * (F.p + F.rho * gr - pw) * s);
  o.x += rwf * (dot(cvw, F.v + vw) * s);
  o.v = cvw * (-W);
  return o;
}

// dp/drho of the equation of state (sim_params.cuh:100-124); zero where the
// pressure was clamped at 0.
template <class T>
CRYO_ADJ_HD inline T eos_dp_drho(int vapor, T rho, T temp, T p, T rho0, T tait_B, T tait_gamma, T gas_R,
                                 int clamp_neg) {
  if (clamp_neg && !(p > T(0))) return T(0);
  if (vapor) return gas_R * temp;
  const T ratio = rho / rho0;
  if (tait_gamma == T(7)) {
    const T r2 = ratio * ratio;
    return T(7) * tait_B * r2 * r2 * r2 / rho0;
  }
  return tait_gamma * tait_B * std::pow(ratio, tait_gamma - T(1)) / rho0;
}

// dp/dT of the equation of state: rho R for vapour, 0 for liquid (and where
// the pressure was clamped at 0).
template <class T>
CRYO_ADJ_HD inline T eos_dp_dT(int vapor, T rho, T p, T gas_R, int clamp_neg) {
  if (!vapor || (clamp_neg && !(p > T(0)))) return T(0);
  return rho * gas_R;
}

// ------------------------------------------------------------ heat
// Conduction pair of the thermal rates (phase_change.cu:31-75), owner I:
//   q_I += V_I V_J k_IJ (T_I - T_J) f(r),  f = r^2 s / (r^2 + eta2)   (r W'(r) = r^2 s)
// qb: cotangent of q_I.  Returns the cotangents of r_IJ, V_I, V_J, T_I, T_J.
template <class T>
struct ThGrad {
  Vec<T> r;
  T Vi, Vj, Ti, Tj;
};
template <class T>
CRYO_ADJ_HD inline ThGrad<T> th_vjp(const Ctx<T>& c, Vec<T> rij, T Vi, T Vj, T kij, T Ti, T Tj, T qb) {
  ThGrad<T> o{{T(0), T(0), T(0)}, T(0), T(0), T(0), T(0)};
  const T r2 = dot(rij, rij);
  if (!(r2 < c.support2)) return o;
  const T r = std::sqrt(r2);
  const T s = kern_s(c, r);
  const T den = r2 + c.eta2;
  const T f = r2 * s / den;
  const T dT = Ti - Tj;
  const T k = kij * qb;
  o.Vi = k * Vj * dT * f;
  o.Vj = k * Vi * dT * f;
  o.Ti = k * Vi * Vj * f;
  o.Tj = -o.Ti;
  if (r > T(1e-12) * c.h) {
    const T spr = kern_spr(c, r);
    // (1/r) df/dr = ((2 s + r^2 spr) den - 2 r^2 s) / den^2
    const T fr = ((T(2) * s + r2 * spr) * den - T(2) * r2 * s) / (den * den);
    o.r = rij * (k * Vi * Vj * dT * fr);
  }
  return o;
}

// ---------------------------------------------------------------- CSF
// Curvature pair of the CSF pass (surface_tension.cu:120-177), owner I and
// fluid J (j != i), both with |ns|^2 > 0.25:
//   num_I += V_J (ns_I - ns_J) . gw,   den_I += V_J r (-W'(r)) = -V_J r^2 s
// num_b / den_b: cotangents of num_I and den_I.  Returns the cotangents of
// r_IJ, ns_I, ns_J and V_J.
template <class T>
struct CsfGrad {
  Vec<T> r, ns_i, ns_j;
  T V_j;
};
template <class T>
CRYO_ADJ_HD inline CsfGrad<T> csf_vjp(const Ctx<T>& c, Vec<T> rij, Vec<T> ns_i, Vec<T> ns_j, T Vj, T num_b,
                                      T den_b) {
  CsfGrad<T> o{{T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}, T(0)};
  const T r2 = dot(rij, rij);
  if (!(r2 < c.support2)) return o;
  const T r = std::sqrt(r2);
  if (!(r > T(1e-12) * c.h)) return o;
  const T s = kern_s(c, r), spr = kern_spr(c, r);
  const Vec<T> u = ns_i - ns_j;
  const Vec<T> gw = rij * s;
  o.ns_i = gw * (num_b * Vj);
  o.ns_j = gw * (-num_b * Vj);
  o.V_j = num_b * dot(u, gw) - den_b * r2 * s;
  o.r = (u * s + rij * (dot(u, rij) * spr)) * (num_b * Vj) - rij * (den_b * Vj * (T(2) * s + r2 * spr));
  return o;
}

// Colour gradient of the interface pass (surface_tension.cu:5-85) for fluid I:
//   gradc_I += gw V_J (c_J - c_I)                 fluid J != I
//   gradc_I += gw vol0 (alpha_w - c_I)            wall w (wall colour on)
//   wall_n_I += nrm_w vol0 W,  ww_I += vol0 W     wall w
// gb, wnb, wwb: cotangents of gradc_I, wall_n_I, ww_I.
template <class T>
struct GcGrad {
  Vec<T> r;
  T V_j;  // fluid pair: cotangent of V_J; wall pair: of alpha_w
};
template <class T>
CRYO_ADJ_HD inline GcGrad<T> gc_ff_vjp(const Ctx<T>& c, Vec<T> rij, T Vj, T dcol, Vec<T> gb) {
  GcGrad<T> o{{T(0), T(0), T(0)}, T(0)};
  if (dcol == T(0)) return o;
  const T r2 = dot(rij, rij);
  if (!(r2 < c.support2)) return o;
  const T r = std::sqrt(r2);
  if (!(r > T(1e-12) * c.h)) return o;
  const T s = kern_s(c, r), spr = kern_spr(c, r);
  const T gr = dot(gb, rij);
  o.r = (gb * s + rij * (gr * spr)) * (Vj * dcol);
  o.V_j = dcol * s * gr;
  return o;
}
template <class T>
CRYO_ADJ_HD inline GcGrad<T> gc_fw_vjp(const Ctx<T>& c, Vec<T> rij, T alpha_w, T c_i, Vec<T> nrm_w, int wall_colour,
                                       Vec<T> gb, Vec<T> wnb, T wwb) {
  GcGrad<T> o{{T(0), T(0), T(0)}, T(0)};
  const T r2 = dot(rij, rij);
  if (!(r2 < c.support2)) return o;
  const T r = std::sqrt(r2);
  const T s = kern_s(c, r);
  o.r = rij * (c.vol0 * (dot(wnb, nrm_w) + wwb) * s);  // dW/dr_IJ = s r_IJ
  if (wall_colour && r > T(1e-12) * c.h) {
    const T spr = kern_spr(c, r);
    const T gr = dot(gb, rij);
    o.r += (gb * s + rij * (gr * spr)) * (c.vol0 * (alpha_w - c_i));
    o.V_j = c.vol0 * s * gr;
  }
  return o;
}

// The wall sum gradphi_I += gw vol0 (wall w) of the contact-line force:
// gpb its cotangent, the share of r_IJ (walls do not move).
template <class T>
CRYO_ADJ_HD inline Vec<T> gphi_fw_vjp(const Ctx<T>& c, Vec<T> rij, Vec<T> gpb) {
  const T r2 = dot(rij, rij);
  if (!(r2 < c.support2)) return Vec<T>{T(0), T(0), T(0)};
  const T r = std::sqrt(r2);
  if (!(r > T(1e-12) * c.h)) return Vec<T>{T(0), T(0), T(0)};
  const T s = kern_s(c, r), spr = kern_spr(c, r);
  return (gpb * s + rij * (dot(gpb, rij) * spr)) * c.vol0;
}

// Contact-line force of a fluid particle near a wall (surface_tension.cu:5-85):
//   G = |g|, nh = g/G, nw = wall_n/|wall_n|, d = nh.nw, t = nh - nw d,
//   delta = 2 |gradphi|, vw = t f,  f = -(cos(theta) + d) coeff G delta / |t|
// with g the colour gradient before the contact-angle rotation; applied when
// ww > 0 && G > normal_eps && |wall_n| > 0 && |t| > 1e-6, else vw = 0.
// vb: cotangent of vw.
template <class T>
struct ClGrad {
  Vec<T> g, wall_n, gradphi;
  T cos_t;
};
template <class T>
CRYO_ADJ_HD inline ClGrad<T> cl_vjp(Vec<T> g, Vec<T> wall_n, Vec<T> gradphi, T ww, T cos_t, T coeff, T normal_eps,
                                    Vec<T> vb) {
  ClGrad<T> o{{T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}, T(0)};
  const T G = std::sqrt(dot(g, g));
  const T Wn = std::sqrt(dot(wall_n, wall_n));
  if (!(ww > T(0) && G > normal_eps && Wn > T(1e-20))) return o;
  const Vec<T> nw = wall_n * (T(1) / Wn);
  const Vec<T> nh = g * (T(1) / G);
  const T d = dot(nh, nw);
  const Vec<T> t = nh - nw * d;
  const T tm = std::sqrt(dot(t, t));
  if (!(tm > T(1e-6))) return o;
  const T P = std::sqrt(dot(gradphi, gradphi));
  const T delta = T(2) * P;
  const T base = -coeff / tm;                  // f = base (cos_t + d) G delta
  const T f = base * (cos_t + d) * G * delta;
  const T q = dot(vb, t);                      // cotangent of f
  o.cos_t = q * base * G * delta;
  const T db = q * base * G * delta;           // d f / d d, then the t path below
  const T Gb = q * base * (cos_t + d) * delta;
  const T deltab = q * base * (cos_t + d) * G;
  if (P > T(0)) o.gradphi = gradphi * (T(2) * deltab / P);
  // t = nh - nw d with |t| in f: tb = f vb - (f / tm) t (t / tm)
  const Vec<T> tb = vb * f - t * (f * q / (tm * tm));
  const T dbt = db - dot(tb, nw);              // d's cotangent, both paths
  const Vec<T> nhb = tb + nw * dbt;
  const Vec<T> nwb = nh * dbt - tb * d;
  o.g = (nhb - nh * dot(nh, nhb)) * (T(1) / G) + nh * Gb;
  o.wall_n = (nwb - nw * dot(nw, nwb)) * (T(1) / Wn);
  return o;
}

// Contact-angle rotation of the colour gradient (surface_tension.cu:60-80):
//   G = |g|, nh = g/G, nw = wall_n/|wall_n|, t = nh - nw (nh.nw),
//   target = -cos(theta) nw + sin(theta) t/|t|, beta = clamp(ww / adh_w, 0, 1),
//   nrm = G normalize(nh (1 - beta) + target beta)
// applied when rotate && ww > 0 && G > normal_eps && |wall_n| > 1e-20 && |t| > 1e-6,
// else nrm = g.  nb: cotangent of nrm.
template <class T>
struct RotGrad {
  Vec<T> g, wall_n;
  T ww, cos_t, sin_t;
};
template <class T>
CRYO_ADJ_HD inline RotGrad<T> rot_vjp(Vec<T> g, Vec<T> wall_n, T ww, T cos_t, T sin_t, T adh_w, T normal_eps,
                                      int rotate, Vec<T> nb) {
  RotGrad<T> o{nb, {T(0), T(0), T(0)}, T(0), T(0), T(0)};
  const T G = std::sqrt(dot(g, g));
  if (!(rotate && ww > T(0) && G > normal_eps)) return o;
  const T Wn = std::sqrt(dot(wall_n, wall_n));
  if (!(Wn > T(1e-20))) return o;
  const Vec<T> nw = wall_n * (T(1) / Wn);
  const Vec<T> nh = g * (T(1) / G);
  const T d = dot(nh, nw);
  const Vec<T> t = nh - nw * d;
  const T tm = std::sqrt(dot(t, t));
  if (!(tm > T(1e-6))) return o;
  const Vec<T> th = t * (T(1) / tm);
  const Vec<T> target = nw * (-cos_t) + th * sin_t;
  const T q = ww / adh_w;
  const T beta = q < T(0) ? T(0) : (q > T(1) ? T(1) : q);
  const Vec<T> b = nh * (T(1) - beta) + target * beta;
  const T B = std::sqrt(dot(b, b));
  if (!(B > T(1e-20))) {  // blended = 0: nrm = 0 whatever g is
    o.g = Vec<T>{T(0), T(0), T(0)};
    return o;
  }
  const Vec<T> bl = b * (T(1) / B);
  const T Gb = dot(nb, bl);
  const Vec<T> blb = nb * G;
  const Vec<T> bb = (blb - bl * dot(bl, blb)) * (T(1) / B);
  Vec<T> nhb = bb * (T(1) - beta);
  const Vec<T> tgb = bb * beta;
  const T betab = dot(bb, target - nh);
  if (q > T(0) && q < T(1)) o.ww = betab / adh_w;
  Vec<T> nwb = tgb * (-cos_t);
  const Vec<T> thb = tgb * sin_t;
  o.cos_t = -dot(tgb, nw);
  o.sin_t = dot(tgb, th);
  const Vec<T> tb = (thb - th * dot(th, thb)) * (T(1) / tm);
  nhb += tb - nw * dot(nw, tb);
  nwb += tb * (-d) - nh * dot(nw, tb);
  o.wall_n = (nwb - nw * dot(nw, nwb)) * (T(1) / Wn);
  o.g = (nhb - nh * dot(nh, nhb)) * (T(1) / G) + nh * Gb;
  return o;
}

// ---------------------------------------------------- transport velocity
// Balanced CSF sum of an owner O (surface_tension.cu:120-177, csf_balanced),
// over its fluid neighbours N:
//   G_O += gw (V_O^2 + V_N^2) (rho_N c_O + rho_O c_N) / (rho_O + rho_N),  gw = r_ON s
// (c: colour, 1 liquid and 0 vapour), a_O += G_O sigma kappa_O / m_O.  Gb: the
// cotangent of G_O.  Returns the cotangents of r_ON, V_O, V_N, rho_O and rho_N.
template <class T>
struct BalGrad {
  Vec<T> r;
  T V_o, V_n, rho_o, rho_n;
};
template <class T>
CRYO_ADJ_HD inline BalGrad<T> bal_vjp(const Ctx<T>& c, Vec<T> r_on, T Vo, T Vn, T rho_o, T rho_n, T co, T cn,
                                      Vec<T> Gb) {
  BalGrad<T> o{{T(0), T(0), T(0)}, T(0), T(0), T(0), T(0)};
  const T r2 = dot(r_on, r_on);
  if (!(r2 < c.support2)) return o;
  const T r = std::sqrt(r2);
  if (!(r > T(0))) return o;
  const T s = kern_s(c, r), spr = kern_spr(c, r);
  const T rs = rho_o + rho_n;
  const T frac = (rho_n * co + rho_o * cn) / rs;
  const T vs = Vo * Vo + Vn * Vn;
  const T Gr = dot(Gb, r_on);
  o.r = (Gb * s + r_on * (Gr * spr)) * (vs * frac);
  const T Q = s * Gr;  // Gb . gw
  o.V_o = Q * T(2) * Vo * frac;
  o.V_n = Q * T(2) * Vn * frac;
  o.rho_o = Q * vs * (cn - frac) / rs;
  o.rho_n = Q * vs * (co - frac) / rs;
  return o;
}

// Transport acceleration of an active owner I (SPH_kernels.cu:56-104):
//   ab_I = K_I sum_J gw (Vt_I^2 + V_J^2),  K_I = -pb_I / (rho_I Vt_I),  Vt_I = 1 / sigma_I,
//   V_J = 1 / sigma_J (fluid) or the wall volume below.
// Sb = K_I * (cotangent of ab_I) is the cotangent of the sum.  Returns the
// cotangents of r_IJ, Vt_I^2 and V_J.
template <class T>
struct TvGrad {
  Vec<T> r;
  T vt2, vj;
};
template <class T>
CRYO_ADJ_HD inline TvGrad<T> tvf_vjp(const Ctx<T>& c, Vec<T> rij, T Vt2, T Vj, Vec<T> Sb) {
  TvGrad<T> o{{T(0), T(0), T(0)}, T(0), T(0)};
  const T r2 = dot(rij, rij);
  if (!(r2 < c.support2)) return o;
  const T r = std::sqrt(r2);
  if (!(r > T(1e-12) * c.h)) return o;
  const T s = kern_s(c, r), spr = kern_spr(c, r);
  const T sr = dot(Sb, rij);
  const T G = s * sr;  // Sb . gw
  o.r = (Sb * s + rij * (sr * spr)) * (Vt2 + Vj * Vj);
  o.vt2 = G;
  o.vj = T(2) * Vj * G;
  return o;
}
// A wall neighbour's volume for the owner: V_J = rho_w > 0 ? Vt_I rho_I / rho_w : Vt_I.
// Adds the owner's shares to vtb and rhob; returns the cotangent of rho_w.
template <class T>
CRYO_ADJ_HD inline T tvf_wall_vjp(T Vt, T rho_i, T rho_w, T vjb, T& vtb, T& rhob) {
  if (!(rho_w > T(0))) {
    vtb += vjb;
    return T(0);
  }
  vtb += vjb * rho_i / rho_w;
  rhob += vjb * Vt / rho_w;
  return -vjb * (Vt * rho_i / rho_w) / rho_w;
}
// d rho_w / d p of the wall density a TVF owner uses: liquid owners
// rho0 max(ratio, 0.5)^(1/gamma) with ratio = (p_w - p_bg) / B + 1 (the value
// rho_w is passed in); vapour owners (p_wv - (p_bg - K)) / (R T_I).
template <class T>
CRYO_ADJ_HD inline T tvf_wall_drho_dp(int vapor, T p_w, T rho_w, T p_bg, T tait_B, T tait_gamma, T gas_R, T T_i) {
  if (vapor) return T(1) / (gas_R * T_i);
  const T ratio = (p_w - p_bg) / tait_B + T(1);
  if (!(ratio > T(0.5))) return T(0);
  return rho_w / (tait_gamma * ratio * tait_B);
}

}  // namespace cryo_adj
