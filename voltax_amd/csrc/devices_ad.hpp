// Parameter tangents of the device residuals, for the adjoint's parameter
// VJP: lam^T d(cf*I + cq*Q)/d p_k for every entry k of a device's parameter
// row (what jax.vjp of the residual with respect to the circuit pytree gives
// in voltax/analysis.py:282-288, including the MOSProcess leaves of
// voltax/elements/mosfet.py:26-87 and the waveform leaves of
// voltax/signals.py:56-189).
//
// The residual equations are those of devices.hpp, restated over a one-tangent
// forward-mode number (value, d/dp_k); the unknowns enter as constants.  The
// kernel seeds one row entry at a time.  Branches on values (softplus, clips,
// the pulse's min, the PWL segment) take the derivative of the branch taken,
// as JAX's jvp of the same expressions does away from the break points.
#pragma once
#include <hip/hip_runtime.h>

#include "devices.hpp"

namespace vxm {

struct Ad {
  double v, d;
};
__device__ __forceinline__ Ad ad(double v) { return {v, 0.0}; }
__device__ __forceinline__ Ad operator+(Ad a, Ad b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Ad operator-(Ad a, Ad b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Ad operator-(Ad a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Ad operator*(Ad a, Ad b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Ad operator/(Ad a, Ad b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Ad operator+(Ad a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ Ad operator+(double a, Ad b) { return {a + b.v, b.d}; }
__device__ __forceinline__ Ad operator-(Ad a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Ad operator-(double a, Ad b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ Ad operator*(Ad a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Ad operator*(double a, Ad b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Ad operator/(Ad a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ Ad operator/(double a, Ad b) {
  const double q = a / b.v;
  return {q, -q * b.d / b.v};
}
__device__ __forceinline__ Ad ad_exp(Ad a) {
  const double e = exp(a.v);
  return {e, e * a.d};
}
__device__ __forceinline__ Ad ad_sqrt(Ad a) {
  const double r = sqrt(a.v);
  return {r, a.d * 0.5 / r};
}
__device__ __forceinline__ Ad ad_sin(Ad a) { return {sin(a.v), cos(a.v) * a.d}; }
__device__ __forceinline__ Ad ad_abs(Ad a) { return a.v < 0.0 ? -a : a; }
__device__ __forceinline__ Ad ad_relu(Ad a) { return a.v > 0.0 ? a : ad(0.0); }
__device__ __forceinline__ Ad ad_log1p_unit(Ad e) {  // e in [0, 1]
  return {log1p_unit(e.v), e.d / (1.0 + e.v)};
}
__device__ __forceinline__ Ad ad_clip01(Ad x) {
  return x.v < 0.0 ? ad(0.0) : (x.v > 1.0 ? ad(1.0) : x);
}
__device__ __forceinline__ Ad ad_min(Ad a, Ad b) { return b.v < a.v ? b : a; }
// x - y floor(x / y), the floor-mod of devices.hpp's floor_mod
__device__ __forceinline__ Ad ad_floor_mod(Ad x, Ad y) {
  const double r = floor_mod(x.v, y.v);
  const double k = (x.v - r) / y.v;  // the integer floor(x / y) the value used
  return {r, x.d - rint(k) * y.d};
}

// EKV residual (ekv_derive + ekv_core of devices.hpp) over a parameter row p.
__device__ inline void ekv_res_ad(const Ad* p, const double (&v)[4], double cf, double cq,
                                  Ad (&res)[4]) {
  const Ad n = p[EKV_N], vt = p[EKV_VT], kp = p[EKV_KP];
  const Ad w = ad_exp(p[EKV_LOGW]), l = ad_exp(p[EKV_LOGL]);
  const Ad inv_n = 1.0 / n;
  const Ad h = 0.5 / vt;
  const Ad i_spec = 2.0 * n * kp * (w / l) * vt * vt;
  const double s = p[EKV_SIGN].v;
  const double vd = s * v[0], vg = s * v[1], vs = s * v[2], vb = s * v[3];
  const Ad vp = (vg - vb - p[EKV_VTH] - p[EKV_DVTH]) * inv_n;
  const Ad af = (vp - (vs - vb)) * h, ar = (vp - (vd - vb)) * h;
  const Ad ef = ad_exp(-ad_abs(af)), er = ad_exp(-ad_abs(ar));
  const Ad sf = ad_relu(af) + ad_log1p_unit(ef), sr = ad_relu(ar) + ad_log1p_unit(er);
  const Ad i_f = sf * sf, i_r = sr * sr;
  const double vds = vd - vs, eps = 1e-3;
  const double sabs = sqrt(vds * vds + eps * eps) - eps;
  const Ad clm = 1.0 + p[EKV_LAM] * sabs;
  const Ad ids = i_spec * (i_f - i_r) * clm;
  res[0] = cf * s * ids;
  res[1] = ad(0.0);
  res[2] = -cf * s * ids;
  res[3] = ad(0.0);
  if (cq == 0.0) return;
  const Ad xf = ad_sqrt(0.25 + i_f), xr = ad_sqrt(0.25 + i_r);
  const Ad c_ox = p[EKV_COX] * w * l;
  const Ad scale = c_ox * n * vt;
  const Ad sm = xf + xr;
  const Ad inv_sm2 = 1.0 / (sm * sm);
  const double k = 4.0 / 15.0;
  const Ad xf2 = xf * xf, xr2 = xr * xr;
  const Ad Nd = 3.0 * xr2 * xr + 6.0 * xr2 * xf + 4.0 * xr * xf2 + 2.0 * xf2 * xf;
  const Ad Ns = 3.0 * xf2 * xf + 6.0 * xf2 * xr + 4.0 * xf * xr2 + 2.0 * xr2 * xr;
  const Ad q_d = -scale * (k * (Nd * inv_sm2) - 0.5);
  const Ad q_s = -scale * (k * (Ns * inv_sm2) - 0.5);
  const Ad q_i = q_d + q_s;
  const Ad nb = -(n - 1.0) * inv_n;
  const Ad q_b = nb * (c_ox * (vg - vb) + q_i);
  const Ad q_g = -(q_i + q_b);
  const Ad cov = p[EKV_COV] * w;
  const Ad q_gs = cov * (vg - vs), q_gd = cov * (vg - vd);
  res[0] = res[0] + cq * s * (q_d - q_gd);
  res[1] = res[1] + cq * s * (q_g + q_gs + q_gd);
  res[2] = res[2] + cq * s * (q_s - q_gs);
  res[3] = res[3] + cq * s * q_b;
}

// Waveform value (wave_value of devices.hpp) over a waveform row w and time t
// (either may carry the tangent).
__device__ inline Ad wave_value_ad(int kind, const Ad* w, Ad t) {
  switch (kind) {
    case WAVE_CONST:
      return w[0];
    case WAVE_STEP: {
      const Ad frac = ad_clip01((t - w[2]) / w[3]);
      return w[0] + (w[1] - w[0]) * frac;
    }
    case WAVE_PULSE: {
      const Ad v1 = w[0], v2 = w[1], delay = w[2], rise = w[3], fall = w[4];
      const Ad width = w[5], period = w[6];
      const Ad tc = ad_floor_mod(t - delay, period);
      const Ad up = ad_clip01(tc / (rise + 1e-30));
      const Ad tfe = rise + width + fall;
      const Ad down = ad_clip01((tfe - tc) / (fall + 1e-30));
      const Ad value = v1 + (v2 - v1) * ad_min(up, down);
      return (t.v < delay.v) ? v1 : value;
    }
    case WAVE_SINE: {
      const Ad tau = t - w[3];
      const double kPi = 3.14159265358979323846;
      const Ad wave = ad_sin(2.0 * kPi * w[2] * tau + w[5] * (kPi / 180.0));
      const Ad active = (tau.v >= 0.0) ? ad_exp(-w[4] * tau) * wave : ad(0.0);
      return w[0] + w[1] * active;
    }
    case WAVE_PWL: {
      const int n = (int)w[0].v;
      const Ad delay = w[1], repeat = w[2];
      const Ad* ts = w + 3;
      const Ad* vs = w + 3 + n;
      Ad tau = t - delay;
      const Ad t_last = ts[n - 1];
      const Ad period = t_last - repeat;
      const Ad safe = period.v > 0.0 ? period : ad(1.0);
      const double k = fmax(ceil((tau.v - t_last.v) / safe.v - 1e-9), 0.0);  // piecewise constant
      if (repeat.v >= 0.0 && period.v > 0.0) tau = tau - k * safe;
      if (tau.v <= ts[0].v) return vs[0];
      if (tau.v >= t_last.v) return vs[n - 1];
      int i = 0;
      while (i + 1 < n && ts[i + 1].v < tau.v) ++i;
      while (i + 1 < n && ts[i + 1].v <= tau.v) ++i;
      const Ad sfrac = (tau - ts[i]) / (ts[i + 1] - ts[i]);
      return vs[i] + sfrac * (vs[i + 1] - vs[i]);
    }
  }
  return ad(0.0);
}

// Entries of a device's parameter row (the row of Circuit.device_table).
__device__ inline int param_row_len(int kind, int wave, const double* p) {
  if (kind == KIND_EKV) return EKV_NPARAM;
  if (kind == KIND_R || kind == KIND_C) return 1;
  switch (wave) {
    case WAVE_CONST: return 1;
    case WAVE_STEP: return 4;
    case WAVE_PULSE: return 7;
    case WAVE_SINE: return 6;
    case WAVE_PWL: return 3 + 2 * (int)p[0];
  }
  return 0;
}

// Largest parameter row a built-in device may have here: EKV_NPARAM, or a
// PWL of (kMaxRowAd - 3) / 2 points (longer rows get no gradient).
constexpr int kMaxRowAd = 67;

// g_k = sum_i lam[i] d res[i] / d p_k of the device's residual cf*I + cq*Q at
// unknowns u and time t, for k = 0 .. len-1 (row entries that are not
// parameters -- the EKV polarity, a PWL's point count -- get 0).  Returns len.
__device__ inline int device_row_vjp(int kind, int wave, const double* p, const double (&u)[4],
                                     double t, double cf, double cq, const double (&lam)[4],
                                     double* g, double ss = 1.0) {
  const int len = param_row_len(kind, wave, p);
  if (len > kMaxRowAd) return 0;
  Ad row[kMaxRowAd];
  for (int e = 0; e < len; ++e) row[e] = ad(p[e]);
  for (int k = 0; k < len; ++k) {
    const bool fixed = (kind == KIND_EKV && k == EKV_SIGN) ||
                       (kind != KIND_EKV && wave == WAVE_PWL && k == 0);
    if (fixed) {
      g[k] = 0.0;
      continue;
    }
    row[k].d = 1.0;
    Ad res[4] = {ad(0.0), ad(0.0), ad(0.0), ad(0.0)};
    if (kind == KIND_EKV) {
      ekv_res_ad(row, u, cf, cq, res);
    } else if (kind == KIND_R) {
      const Ad i = (u[0] - u[1]) * ad_exp(-row[0]);
      res[0] = cf * i;
      res[1] = -cf * i;
    } else if (kind == KIND_C) {
      const Ad q = ad_exp(row[0]) * (u[0] - u[1]);
      res[0] = cq * q;
      res[1] = -cq * q;
    } else if (kind == KIND_V) {
      const Ad val = ss * wave_value_ad(wave, row, ad(t));
      res[2] = cf * ((u[0] - u[1]) - val);
    } else {  // KIND_I
      const Ad val = ss * wave_value_ad(wave, row, ad(t));
      res[0] = cf * val;
      res[1] = -cf * val;
    }
    row[k].d = 0.0;
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) acc += lam[i] * res[i].d;
    g[k] = acc;
  }
  return len;
}

// lam^T d(cf*f)/dt of a device at unknowns u: only independent sources depend
// on time (their waveforms); charges do not.
__device__ inline double device_time_vjp(int kind, int wave, const double* p, double t, double cf,
                                         const double (&lam)[4]) {
  if (kind != KIND_V && kind != KIND_I) return 0.0;
  const int len = param_row_len(kind, wave, p);
  if (len > kMaxRowAd) return 0.0;
  Ad row[kMaxRowAd];
  for (int e = 0; e < len; ++e) row[e] = ad(p[e]);
  const Ad val = wave_value_ad(wave, row, Ad{t, 1.0});
  // V: res[2] = cf*((u0 - u1) - val); I: res = cf*val*(1, -1)
  return kind == KIND_V ? -cf * lam[2] * val.d : cf * (lam[0] - lam[1]) * val.d;
}

}  // namespace vxm
