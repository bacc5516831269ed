// Device physics for the Newton hot path, FP64, evaluated per device per lane.
//
// Each model returns its local contribution vector  out[i] = cf*I_i + cq*Q_i
// over the device's m = T+K local unknowns (terminals, then internals) and,
// when asked, its local Jacobian block  J[i][j] = d out[i] / d u[j].  The
// physics follows the reference element models:
//   Resistor / Capacitor        voltax/elements/passive.py:13-59
//   VoltageSource / CurrentSource voltax/elements/sources.py:58-112
//   EKVMOSFET (EKV 2.6 style)   voltax/elements/mosfet.py:144-205
//   waveforms                   voltax/signals.py:56-189
// Jacobians of the EKV model are forward-mode dual numbers over the four
// terminal voltages, chained by hand (equal to what jax.jvp computes in
// voltax/sparse.py:240-244); the linear elements have closed-form blocks.
#pragma once
#include <hip/hip_runtime.h>
#include "vxm_types.h"

namespace vxm {

// ------------------------------------------------------------ fast math
// 1/x by the hardware reciprocal estimate refined with Newton-Raphson steps
// r <- r + r (1 - x r) (two for FP64, one for FP32): correct to about one
// ulp at a third of the IEEE division's instruction count.
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ float recip(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  return fmaf(r, fmaf(-x, r, 1.0f), r);
}

// log(1 + e) for 0 <= e <= 1 (the softplus of the EKV model only needs this
// range: e = exp(-|a|)), from the series log(1 + e) = 2 atanh(s) =
// 2 (s + s^3/3 + s^5/5 + ...), s = e / (2 + e) <= 1/3: seventeen odd terms
// leave a truncation error below 2^-55 relative, and the terms are positive,
// so the sum is accurate to a few ulp, tiny e included (no cancellation).
// About a fifth of the instructions of the general-range library log1p.
__device__ __forceinline__ double log1p_unit(double e) {
  const double s = e * recip(2.0 + e);
  const double s2 = s * s;
  double p = 1.0 / 33.0;
#pragma unroll
  for (int k = 15; k >= 0; --k) p = fma(p, s2, 1.0 / (2 * k + 1));
  return 2.0 * s * p;
}

// ------------------------------------------------------------------ waveforms
// Wave parameter layouts (per device, inside the per-instance parameter row):
//   CONST  : value
//   STEP   : v0 v1 delay rise
//   PULSE  : v1 v2 delay rise fall width period
//   SINE   : offset amplitude freq delay damping phase_deg
//   PWL    : npts delay repeat t[0..npts) v[0..npts)
__device__ __forceinline__ double floor_mod(double x, double y) {
  // x - y*floor(x/y) computed as fmod + sign fix (the float `%` of numpy/jnp)
  double r = fmod(x, y);
  if (r != 0.0 && ((r < 0.0) != (y < 0.0))) r += y;
  return r;
}

__device__ __forceinline__ double clip01(double x) { return fmin(fmax(x, 0.0), 1.0); }

__device__ inline double wave_value(int kind, const double* w, double t) {
  switch (kind) {
    case WAVE_CONST:
      return w[0];
    case WAVE_STEP: {
      double frac = clip01((t - w[2]) / w[3]);
      return w[0] + (w[1] - w[0]) * frac;
    }
    case WAVE_PULSE: {
      double v1 = w[0], v2 = w[1], delay = w[2], rise = w[3], fall = w[4];
      double width = w[5], period = w[6];
      double tc = floor_mod(t - delay, period);
      double up = clip01(tc / (rise + 1e-30));
      double tfe = rise + width + fall;
      double down = clip01((tfe - tc) / (fall + 1e-30));
      double value = v1 + (v2 - v1) * fmin(up, down);
      return (t < delay) ? v1 : value;
    }
    case WAVE_SINE: {
      double tau = t - w[3];
      const double kPi = 3.14159265358979323846;
      double wave = sin(2.0 * kPi * w[2] * tau + w[5] * (kPi / 180.0));
      double active = (tau >= 0.0) ? exp(-w[4] * tau) * wave : 0.0;
      return w[0] + w[1] * active;
    }
    case WAVE_PWL: {
      int n = (int)w[0];
      double delay = w[1], repeat = w[2];
      const double* ts = w + 3;
      const double* vs = w + 3 + n;
      double tau = t - delay;
      double t_last = ts[n - 1];
      double period = t_last - repeat;
      double safe = period > 0.0 ? period : 1.0;
      double k = fmax(ceil((tau - t_last) / safe - 1e-9), 0.0);
      if (repeat >= 0.0 && period > 0.0) tau = tau - k * safe;
      if (tau <= ts[0]) return vs[0];
      if (tau >= t_last) return vs[n - 1];
      int i = 0;
      while (i + 1 < n && ts[i + 1] < tau) ++i;
      // numpy.interp: right-continuous search (ts[i] <= tau < ts[i+1])
      while (i + 1 < n && ts[i + 1] <= tau) ++i;
      double dt = ts[i + 1] - ts[i];
      double s = (tau - ts[i]) / dt;
      return vs[i] + s * (vs[i + 1] - vs[i]);
    }
  }
  return 0.0;
}

// --------------------------------------------------------------- EKV MOSFET
// Parameter row: log_w log_l dvth sign kp vth n lam vt cox cov.
//
// Evaluated in the mirrored n-frame v' = sign*v (PMOS: sign = -1), outputs
// mapped back by sign.  Because d(sign*f(sign*v))/dv = f'(v'), the Jacobian
// block needs no sign.  Derivatives are chained by hand through the scalar
// intermediates (inversion levels i_f, i_r; x_f, x_r; |v_ds|), which keeps
// the live register set small; they equal the forward-mode tangents of the
// same expressions.
//   res[i] = cf*I_i + cq*Q_i,   jac[4*i+j] = d res[i] / d v_j  (d,g,s,b)
//
// The quantities that depend on the parameters alone (w, l and the
// reciprocals of n and vt enter through them) form the derived row below,
// in the slots of the parameter row; ekv_eval computes it and goes on with
// ekv_core, and the fused Newton kernel, whose parameters are fixed for a
// run, computes it once per run and calls ekv_core directly.  Both do the
// same operations on the same values, so they agree bit for bit.
enum EkvDerived {
  EKD_SIGN = 0,
  EKD_INV_N = 1,   // 1/n
  EKD_VTH = 2,
  EKD_DVTH = 3,
  EKD_H = 4,       // 1/(2 vt)
  EKD_ISPEC = 5,   // 2 n kp (w/l) vt^2
  EKD_LAM = 6,
  EKD_COX = 7,     // cox w l
  EKD_QSCALE = 8,  // cox w l n vt
  EKD_NB = 9,      // -(n-1)/n
  EKD_COV = 10,    // cov w
};
static_assert(EKD_COV < EKV_NPARAM, "the derived EKV row fits in the parameter row");

__device__ __forceinline__ void ekv_derive(const double* p, double* d) {
  const double n = p[EKV_N], vt = p[EKV_VT], kp = p[EKV_KP];
  const double w = exp(p[EKV_LOGW]), l = exp(p[EKV_LOGL]);
  const double inv_n = recip(n);
  const double c_ox = p[EKV_COX] * w * l;
  d[EKD_SIGN] = p[EKV_SIGN];
  d[EKD_INV_N] = inv_n;
  d[EKD_VTH] = p[EKV_VTH];
  d[EKD_DVTH] = p[EKV_DVTH];
  d[EKD_H] = 0.5 * recip(vt);
  d[EKD_ISPEC] = 2.0 * n * kp * (w * recip(l)) * vt * vt;
  d[EKD_LAM] = p[EKV_LAM];
  d[EKD_COX] = c_ox;
  d[EKD_QSCALE] = c_ox * n * vt;
  d[EKD_NB] = -(n - 1.0) * inv_n;
  d[EKD_COV] = p[EKV_COV] * w;
}

__device__ __forceinline__ void ekv_core(const double* dp, const double (&v)[4], double cf,
                                         double cq, double (&res)[4], double* jac) {
  const double s = dp[EKD_SIGN];
  const double lam = dp[EKD_LAM];
  const double vd = s * v[0], vg = s * v[1], vs = s * v[2], vb = s * v[3];
  const double inv_n = dp[EKD_INV_N];
  const double vp = (vg - vb - dp[EKD_VTH] - dp[EKD_DVTH]) * inv_n;
  const double h = dp[EKD_H];
  const double af = (vp - (vs - vb)) * h, ar = (vp - (vd - vb)) * h;
  // softplus and its derivative (sigmoid)
  const double ef = exp(-fabs(af)), er = exp(-fabs(ar));
  const double sf = fmax(af, 0.0) + log1p_unit(ef), sr = fmax(ar, 0.0) + log1p_unit(er);
  const double gf = (af >= 0.0 ? 1.0 : ef) * recip(1.0 + ef);
  const double gr = (ar >= 0.0 ? 1.0 : er) * recip(1.0 + er);
  const double i_f = sf * sf, i_r = sr * sr;
  const double dif = 2.0 * sf * gf * h, dir = 2.0 * sr * gr * h;  // d i / d(2vt*a)
  // d a_f / d v' = (0, 1/n, -1, 1 - 1/n) * h ; d a_r / d v' = (-1, 1/n, 0, 1 - 1/n) * h
  const double Df[4] = {0.0, dif * inv_n, -dif, dif * (1.0 - inv_n)};
  const double Dr[4] = {-dir, dir * inv_n, 0.0, dir * (1.0 - inv_n)};
  const double i_spec = dp[EKD_ISPEC];
  const double vds = vd - vs;
  const double eps = 1e-3;
  const double root = sqrt(vds * vds + eps * eps);
  const double sabs = root - eps;
  const double clm = 1.0 + lam * sabs;
  const double ids = i_spec * (i_f - i_r) * clm;
  const double g_i = i_spec * clm;                           // d ids / d i_f (= -d/d i_r)
  const double g_v = i_spec * (i_f - i_r) * lam * vds * recip(root);  // d ids / d vds
  double dids[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) dids[j] = g_i * (Df[j] - Dr[j]);
  dids[0] += g_v;
  dids[2] -= g_v;
  // currents into (d, g, s, b) = sign * (ids, 0, -ids, 0)
  res[0] = cf * s * ids;
  res[1] = 0.0;
  res[2] = -cf * s * ids;
  res[3] = 0.0;
  if (jac) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      jac[j] = cf * dids[j];
      jac[4 + j] = 0.0;
      jac[8 + j] = -cf * dids[j];
      jac[12 + j] = 0.0;
    }
  }
  if (cq == 0.0) return;
  const double xf = sqrt(0.25 + i_f), xr = sqrt(0.25 + i_r);
  const double c_ox = dp[EKD_COX];
  const double scale = dp[EKD_QSCALE];
  const double sm = xf + xr;
  const double inv_sm2 = recip(sm * sm);
  const double k = 4.0 / 15.0;
  const double xf2 = xf * xf, xr2 = xr * xr;
  const double Nd = 3.0 * xr2 * xr + 6.0 * xr2 * xf + 4.0 * xr * xf2 + 2.0 * xf2 * xf;
  const double Ns = 3.0 * xf2 * xf + 6.0 * xf2 * xr + 4.0 * xf * xr2 + 2.0 * xr2 * xr;
  const double q_d = -scale * (k * (Nd * inv_sm2) - 0.5);
  const double q_s = -scale * (k * (Ns * inv_sm2) - 0.5);
  const double q_i = q_d + q_s;
  const double nb = dp[EKD_NB];
  const double q_b = nb * (c_ox * (vg - vb) + q_i);
  const double q_g = -(q_i + q_b);
  const double cov = dp[EKD_COV];
  const double q_gs = cov * (vg - vs), q_gd = cov * (vg - vd);
  res[0] += cq * s * (q_d - q_gd);
  res[1] += cq * s * (q_g + q_gs + q_gd);
  res[2] += cq * s * (q_s - q_gs);
  res[3] += cq * s * q_b;
  if (!jac) return;
  // partials of q_d, q_s with respect to x_f, x_r
  const double two_sm = 2.0 * recip(sm);
  const double dNd_f = 6.0 * xr2 + 8.0 * xr * xf + 6.0 * xf2;
  const double dNd_r = 9.0 * xr2 + 12.0 * xr * xf + 4.0 * xf2;
  const double dNs_f = 9.0 * xf2 + 12.0 * xf * xr + 4.0 * xr2;
  const double dNs_r = 6.0 * xf2 + 8.0 * xf * xr + 6.0 * xr2;
  const double c = -scale * k * inv_sm2;
  const double qd_f = c * (dNd_f - Nd * two_sm), qd_r = c * (dNd_r - Nd * two_sm);
  const double qs_f = c * (dNs_f - Ns * two_sm), qs_r = c * (dNs_r - Ns * two_sm);
  const double hf = 0.5 * recip(xf), hr = 0.5 * recip(xr);  // d x / d i
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const double dxf = hf * Df[j], dxr = hr * Dr[j];
    const double dqd = qd_f * dxf + qd_r * dxr;
    const double dqs = qs_f * dxf + qs_r * dxr;
    const double dqi = dqd + dqs;
    const double dvgb = (j == 1 ? 1.0 : 0.0) - (j == 3 ? 1.0 : 0.0);
    const double dqb = nb * (c_ox * dvgb + dqi);
    const double dqg = -(dqi + dqb);
    const double dgs = cov * ((j == 1 ? 1.0 : 0.0) - (j == 2 ? 1.0 : 0.0));
    const double dgd = cov * ((j == 1 ? 1.0 : 0.0) - (j == 0 ? 1.0 : 0.0));
    jac[j] += cq * (dqd - dgd);
    jac[4 + j] += cq * (dqg + dgs + dgd);
    jac[8 + j] += cq * (dqs - dgs);
    jac[12 + j] += cq * dqb;
  }
}

__device__ __forceinline__ void ekv_eval(const double* p, const double (&v)[4], double cf,
                                         double cq, double (&res)[4], double* jac) {
  double d[EKV_NPARAM];
  ekv_derive(p, d);
  ekv_core(d, v, cf, cq, res, jac);
}

// ------------------------------------------------------------ generic device
// Evaluates device `kind` with unknown values u[0..m) (ground already 0):
// res[i] = cf*I_i + cq*Q_i and, when JAC, jac[i*m+j] = d res[i] / d u[j].
// `ss` scales every independent source (source stepping homotopy; 1 = off).
// DERIVED: p is the derived row of derive_params (EKV: ekv_derive; R: the
// conductance; C: the capacitance; sources: their waveform row).
template <bool JAC, bool DERIVED = false>
__device__ __forceinline__ void device_eval(int kind, int wave, const double* p,
                                            const double (&u)[4], double t, double cf,
                                            double cq, double (&res)[4], double (&jac)[16],
                                            double ss = 1.0) {
  if (kind == KIND_EKV) {
    if (DERIVED) ekv_core(p, u, cf, cq, res, JAC ? jac : nullptr);
    else ekv_eval(p, u, cf, cq, res, JAC ? jac : nullptr);
  } else if (kind == KIND_R) {
    const double g = DERIVED ? p[0] : exp(-p[0]);
    const double i = (u[0] - u[1]) * g;
    res[0] = cf * i;
    res[1] = -cf * i;
    if (JAC) {
      const double a = cf * g;
      jac[0] = a; jac[1] = -a; jac[2] = -a; jac[3] = a;
    }
  } else if (kind == KIND_C) {
    const double c = DERIVED ? p[0] : exp(p[0]);
    const double q = c * (u[0] - u[1]);
    res[0] = cq * q;
    res[1] = -cq * q;
    if (JAC) {
      const double a = cq * c;
      jac[0] = a; jac[1] = -a; jac[2] = -a; jac[3] = a;
    }
  } else if (kind == KIND_V) {
    // unknowns (p, n, i): I = (i, -i), F = (v_p - v_n) - value(t)
    const double val = ss * wave_value(wave, p, t);
    res[0] = cf * u[2];
    res[1] = -cf * u[2];
    res[2] = cf * ((u[0] - u[1]) - val);
    if (JAC) {
      jac[0] = 0.0; jac[1] = 0.0; jac[2] = cf;
      jac[3] = 0.0; jac[4] = 0.0; jac[5] = -cf;
      jac[6] = cf;  jac[7] = -cf; jac[8] = 0.0;
    }
  } else {  // KIND_I
    const double val = ss * wave_value(wave, p, t);
    res[0] = cf * val;
    res[1] = -cf * val;
    if (JAC) {
      jac[0] = 0.0; jac[1] = 0.0; jac[2] = 0.0; jac[3] = 0.0;
    }
  }
}

// Derived row of one device (see device_eval) into d; rows of other kinds
// are copied by the caller beforehand.
__device__ __forceinline__ void derive_device(int kind, const double* p, double* d) {
  if (kind == KIND_EKV) ekv_derive(p, d);
  else if (kind == KIND_R) d[0] = exp(-p[0]);
  else if (kind == KIND_C) d[0] = exp(p[0]);
}

}  // namespace vxm
