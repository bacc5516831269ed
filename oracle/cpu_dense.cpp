// Compiled single-thread CPU restatement of the reference's dense path, used
// as the timed CPU baseline of the batched ring workload (test
// infrastructure: only bench.py's cpu_baseline leg and the tests load it).
//
// Per backward-Euler time step (voltax/analysis.py:434-480, residual and
// Jacobian :338-342), damped Newton (:208-233): stamp the residual
//   r = (q(z) - q(z_prev)) / h + f(z, t) + gmin v
// and the Jacobian df/dz + dq/dz / h + gmin on node diagonals with
// forward-mode derivatives of every device (dual numbers, as the reference's
// JVPs, voltax/sparse.py:200-249), solve J dz = -r densely by LU with partial
// pivoting (the reference's `jnp.linalg.solve` below the sparse threshold,
// :201-205, 385-402), scale the update by min(1, max_dv / max|dz[:nv]|) and
// stop when every |dz| <= atol + rtol |z_new| with no damping.
//
// Device equations (the reference's element files):
//   Resistor   voltax/elements/passive.py:13-38   i = (vp - vn) exp(-log_r)
//   Capacitor  voltax/elements/passive.py:41-59   q = exp(log_c) (vp - vn)
//   V source   voltax/elements/sources.py:58-90   rows (i, -i, vp - vn - V(t))
//   I source   voltax/elements/sources.py:93-112  (I(t), -I(t))
//   EKVMOSFET  voltax/elements/mosfet.py:127-205  currents and charges below
//   Pulse      voltax/signals.py:108-114          (t - delay) mod period, clipped ramps
// Inputs use the device-table layout of voltax_amd (kind, wave, four unknowns
// per device, parameter row offsets; EKV rows [log_w log_l dvth sign kp vth
// n lam vt cox cov]).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

enum { R = 0, C = 1, V = 2, I = 3, EKV = 4 };
enum { WCONST = 1, WPULSE = 3 };

struct Dual {  // value and derivatives with respect to a device's 4 unknowns
  double v, d[4];
};
inline Dual cst(double x) { return {x, {0, 0, 0, 0}}; }
inline Dual seed(double x, int k) {
  Dual r = cst(x);
  r.d[k] = 1.0;
  return r;
}
inline Dual operator+(Dual a, Dual b) {
  for (int k = 0; k < 4; ++k) a.d[k] += b.d[k];
  a.v += b.v;
  return a;
}
inline Dual operator-(Dual a, Dual b) {
  for (int k = 0; k < 4; ++k) a.d[k] -= b.d[k];
  a.v -= b.v;
  return a;
}
inline Dual operator-(Dual a) {
  for (int k = 0; k < 4; ++k) a.d[k] = -a.d[k];
  a.v = -a.v;
  return a;
}
inline Dual operator*(Dual a, Dual b) {
  Dual r;
  r.v = a.v * b.v;
  for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
inline Dual operator*(double s, Dual a) {
  for (int k = 0; k < 4; ++k) a.d[k] *= s;
  a.v *= s;
  return a;
}
inline Dual operator+(Dual a, double s) {
  a.v += s;
  return a;
}
inline Dual operator/(Dual a, Dual b) {
  Dual r;
  r.v = a.v / b.v;
  for (int k = 0; k < 4; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
inline Dual dsqrt(Dual a) {
  Dual r;
  r.v = std::sqrt(a.v);
  for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] / (2.0 * r.v);
  return r;
}
// log(1 + exp(x)) as max(x, 0) + log1p(exp(-|x|)); derivative sigmoid(x)
inline Dual softplus(Dual a) {
  const double e = std::exp(-std::fabs(a.v));
  Dual r;
  r.v = std::fmax(a.v, 0.0) + std::log1p(e);
  const double sg = a.v >= 0 ? 1.0 / (1.0 + e) : e / (1.0 + e);
  for (int k = 0; k < 4; ++k) r.d[k] = sg * a.d[k];
  return r;
}

// jnp.remainder: the sign of the divisor
double floor_mod(double x, double p) {
  double r = std::fmod(x, p);
  if (r != 0.0 && ((r < 0.0) != (p < 0.0))) r += p;
  return r;
}

double wave_value(int wave, const double* w, double t) {
  if (wave == WCONST) return w[0];
  // PULSE [v1 v2 delay rise fall width period]
  const double v1 = w[0], v2 = w[1], delay = w[2], rise = w[3], fall = w[4], width = w[5],
               period = w[6];
  const double tc = floor_mod(t - delay, period);
  const double up = std::fmin(std::fmax(tc / (rise + 1e-30), 0.0), 1.0);
  const double t_end = rise + width + fall;
  const double down = std::fmin(std::fmax((t_end - tc) / (fall + 1e-30), 0.0), 1.0);
  const double value = v1 + (v2 - v1) * std::fmin(up, down);
  return t < delay ? v1 : value;
}

// EKV currents (n-frame, mirrored by sign) and charges of one device.
void ekv(const double* p, const double* u, Dual* I, Dual* Q, bool want_q) {
  const double s = p[3], kp = p[4], vth = p[5], n = p[6], lam = p[7], vt = p[8], cox = p[9],
               cov = p[10];
  const double w = std::exp(p[0]), l = std::exp(p[1]), dvth = p[2];
  Dual vd = s * seed(u[0], 0), vg = s * seed(u[1], 1), vs = s * seed(u[2], 2),
       vb = s * seed(u[3], 3);
  const Dual vp = (1.0 / n) * (vg - vb + (-vth - dvth));
  const Dual af = softplus((1.0 / (2 * vt)) * (vp - (vs - vb)));
  const Dual ar = softplus((1.0 / (2 * vt)) * (vp - (vd - vb)));
  const Dual i_f = af * af, i_r = ar * ar;
  const double i_spec = 2 * n * kp * (w / l) * vt * vt;
  const Dual x = vd - vs;
  const Dual sabs = dsqrt(x * x + 1e-6) + (-1e-3);
  const Dual ids = i_spec * ((i_f - i_r) * (cst(1.0) + lam * sabs));
  I[0] = s * ids;
  I[1] = cst(0.0);
  I[2] = -(s * ids);
  I[3] = cst(0.0);
  if (!want_q) return;
  const Dual xf = dsqrt(i_f + 0.25), xr = dsqrt(i_r + 0.25);
  const double c_ox = cox * w * l, scale = c_ox * n * vt;
  const Dual sm = xf + xr, sm2 = sm * sm;
  const Dual xf2 = xf * xf, xr2 = xr * xr;
  const Dual q_d = -scale * ((4.0 / 15.0) * ((3.0 * (xr2 * xr) + 6.0 * (xr2 * xf) +
                                              4.0 * (xr * xf2) + 2.0 * (xf2 * xf)) / sm2) +
                             (-0.5));
  const Dual q_s = -scale * ((4.0 / 15.0) * ((3.0 * (xf2 * xf) + 6.0 * (xf2 * xr) +
                                              4.0 * (xf * xr2) + 2.0 * (xr2 * xr)) / sm2) +
                             (-0.5));
  const Dual q_i = q_d + q_s;
  const Dual q_b = -((n - 1) / n) * (c_ox * (vg - vb) + q_i);
  const Dual q_g = -(q_i + q_b);
  const double cw = cov * w;
  const Dual q_gs = cw * (vg - vs), q_gd = cw * (vg - vd);
  Q[0] = s * (q_d - q_gd);
  Q[1] = s * (q_g + q_gs + q_gd);
  Q[2] = s * (q_s - q_gs);
  Q[3] = s * q_b;
}

struct Circuit {
  int S, nv, D;
  const int *kind, *wave, *unk, *poff;
  const double* params;
};

inline double zval(const double* z, int i) { return i >= 0 ? z[i] : 0.0; }

// f(z, t) (cf = 1) plus dq/h part: res += cf f + cq q; J += cf df/dz + cq dq/dz.
// want_j = false skips the Jacobian (history term).
void stamp(const Circuit& c, const double* z, double t, double cf, double cq, double* res,
           double* J, bool want_j) {
  const int S = c.S;
  for (int d = 0; d < c.D; ++d) {
    const int* u = c.unk + 4 * d;
    const double* p = c.params + c.poff[d];
    const int k = c.kind[d];
    if (k == R || k == C) {
      const double g = k == R ? cf * std::exp(-p[0]) : cq * std::exp(p[0]);
      const double i = g * (zval(z, u[0]) - zval(z, u[1]));
      const int a = u[0], b = u[1];
      if (a >= 0) res[a] += i;
      if (b >= 0) res[b] -= i;
      if (want_j) {
        if (a >= 0) J[a * S + a] += g;
        if (b >= 0) J[b * S + b] += g;
        if (a >= 0 && b >= 0) {
          J[a * S + b] -= g;
          J[b * S + a] -= g;
        }
      }
    } else if (k == V) {
      const int a = u[0], b = u[1], x = u[2];
      const double i = z[x];
      if (a >= 0) res[a] += cf * i;
      if (b >= 0) res[b] -= cf * i;
      res[x] += cf * (zval(z, a) - zval(z, b) - wave_value(c.wave[d], p, t));
      if (want_j) {
        if (a >= 0) {
          J[a * S + x] += cf;
          J[x * S + a] += cf;
        }
        if (b >= 0) {
          J[b * S + x] -= cf;
          J[x * S + b] -= cf;
        }
      }
    } else if (k == I) {
      const double i = cf * wave_value(c.wave[d], p, t);
      if (u[0] >= 0) res[u[0]] += i;
      if (u[1] >= 0) res[u[1]] -= i;
    } else if (k == EKV) {
      double uv[4];
      for (int j = 0; j < 4; ++j) uv[j] = zval(z, u[j]);
      Dual Id[4], Qd[4];
      ekv(p, uv, Id, Qd, cq != 0.0);
      for (int r = 0; r < 4; ++r) {
        if (u[r] < 0) continue;
        Dual tot = cf * Id[r];
        if (cq != 0.0) tot = tot + cq * Qd[r];
        res[u[r]] += tot.v;
        if (want_j)
          for (int j = 0; j < 4; ++j)
            if (u[j] >= 0) J[u[r] * S + u[j]] += tot.d[j];
      }
    }
  }
}

// LU with partial pivoting (LAPACK getrf order: the first largest |a| of
// the column), then the two substitutions; b is overwritten by x.
void lu_solve(int n, double* A, double* b, std::vector<int>& piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[i * n + k]) > best) {
        best = std::fabs(A[i * n + k]);
        p = i;
      }
    piv[k] = p;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
      std::swap(b[k], b[p]);
    }
    const double inv = 1.0 / A[k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const double l = A[i * n + k] * inv;
      A[i * n + k] = l;
      if (l == 0.0) continue;
      for (int j = k + 1; j < n; ++j) A[i * n + j] -= l * A[k * n + j];
      b[i] -= l * b[k];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    double acc = b[k];
    for (int j = k + 1; j < n; ++j) acc -= A[k * n + j] * b[j];
    b[k] = acc / A[k * n + k];
  }
}

}  // namespace

extern "C" {

// Backward-Euler transient of one circuit from the state z0 at ts[0] over
// ts[0..T): out_z [T][S] (row 0 = z0), out_iters [T], out_conv [T].
// Returns 0, or -4 for a device kind / waveform this restatement does not hold.
int vxm_cpu_transient_be(int S, int nv, int D, const int* kind, const int* wave, const int* unk,
                         const int* poff, const double* params, const double* z0,
                         const double* ts, int T, int max_steps, double rtol, double atol,
                         double max_dv, double gmin, double* out_z, int* out_iters,
                         signed char* out_conv) {
  for (int d = 0; d < D; ++d) {
    if (kind[d] < R || kind[d] > EKV) return -4;
    if ((kind[d] == V || kind[d] == I) && wave[d] != WCONST && wave[d] != WPULSE) return -4;
  }
  Circuit c{S, nv, D, kind, wave, unk, poff, params};
  std::vector<double> z(z0, z0 + S), hist(S), r(S), J((size_t)S * S), dz(S), zn(S);
  std::vector<int> piv(S);
  std::memcpy(out_z, z0, sizeof(double) * S);
  out_iters[0] = 0;
  out_conv[0] = 1;
  for (int n = 1; n < T; ++n) {
    const double h = ts[n] - ts[n - 1], cq = 1.0 / h;
    // history: -q(z_prev) / h
    std::fill(hist.begin(), hist.end(), 0.0);
    stamp(c, z.data(), ts[n - 1], 0.0, -cq, hist.data(), nullptr, false);
    bool done = false;
    int it = 0;
    while (it < max_steps && !done) {
      for (int i = 0; i < S; ++i) r[i] = hist[i] + (i < nv ? gmin * z[i] : 0.0);
      std::fill(J.begin(), J.end(), 0.0);
      for (int i = 0; i < nv; ++i) J[(size_t)i * S + i] += gmin;
      stamp(c, z.data(), ts[n], 1.0, cq, r.data(), J.data(), true);
      for (int i = 0; i < S; ++i) dz[i] = -r[i];
      lu_solve(S, J.data(), dz.data(), piv);
      double dv = 0.0;
      for (int i = 0; i < nv; ++i) dv = std::fmax(dv, std::fabs(dz[i]));
      const double scale = std::fmin(1.0, max_dv / (dv + 1e-300));
      bool small = true, finite = true;
      for (int i = 0; i < S; ++i) {
        zn[i] = z[i] + scale * dz[i];
        small = small && std::fabs(dz[i]) <= atol + rtol * std::fabs(zn[i]);
        finite = finite && std::isfinite(zn[i]);
      }
      done = small && scale == 1.0 && finite;
      z.swap(zn);
      ++it;
    }
    std::memcpy(out_z + (size_t)n * S, z.data(), sizeof(double) * S);
    out_iters[n] = it;
    out_conv[n] = done ? 1 : 0;
  }
  return 0;
}

}  // extern "C"
