// Fused batched Newton for small circuits (S <= 64): DC operating point
// with gmin stepping and a fixed-grid BE / trapezoidal transient, entirely
// inside one kernel launch.
//
// Reference semantics: voltax/analysis.py:208-233 (damped Newton and its
// convergence test), :236-266 (gmin stepping), :325-382 (DC / BE / trap
// residuals and stamped Jacobians), :434-480 (transient scan), with the
// reference's dense path (`jnp.linalg.solve`, LU with partial pivoting) for
// systems below the sparse threshold (voltax/analysis.py:201-205, 385-402).
//
// Mapping to CDNA4: one workgroup = one 64-lane wavefront holding 64/L
// instances of one pattern; lane r of a segment owns row r of the Newton
// system and keeps that row in registers.  Per iteration:
//   1. every lane evaluates devices r, r+L, ... (FP64 physics, devices.hpp)
//      and adds their residual / Jacobian entries into the instance's
//      residual vector and column-major matrix in LDS (FP64 LDS atomics);
//   2. lane r reads row r of that matrix into registers (independent loads,
//      conflict-free across lanes);
//   3. right-looking LU with partial pivoting across the lanes of the
//      segment (argmax by butterfly shuffles, pivot row broadcast by lane
//      shuffles), then forward / back substitution;
//   4. damping, convergence test and update, reduced over the segment.
// Instances converge independently; a segment that has converged idles
// until every segment of the wave is done.
//
// precision 1 factors in FP32 and recovers FP64 accuracy by iterative
// refinement against the FP64 row held alongside (residual b - A x in FP64).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "devices.hpp"
#include "kernels.hpp"

namespace vxm {

// ------------------------------------------------------ segment reductions
// All-reductions over the L lanes of a segment without going through LDS:
// DPP butterflies inside rows of 16 lanes (xor 1, xor 2, half mirror,
// mirror), a swizzle across the two rows of 32 lanes, and readlane across
// the two halves of the wave.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  int x = __float_as_int(v);
  x = __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false);
  return __int_as_float(x);
}
__device__ __forceinline__ double swz16_d(double v) {  // lane ^ 16 within 32
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_ds_swizzle(lo, 0x401F);
  hi = __builtin_amdgcn_ds_swizzle(hi, 0x401F);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float swz16_f(float v) {
  return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));
}
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;

template <int L>
__device__ __forceinline__ double seg_max(double v) {
  v = fmax(v, dpp_d<kDppXor1>(v));
  v = fmax(v, dpp_d<kDppXor2>(v));
  if (L >= 8) v = fmax(v, dpp_d<kDppHalfMirror>(v));
  if (L >= 16) v = fmax(v, dpp_d<kDppMirror>(v));
  if (L >= 32) v = fmax(v, swz16_d(v));
  if (L >= 64) {
    int lo0 = __builtin_amdgcn_readlane(__double2loint(v), 0);
    int hi0 = __builtin_amdgcn_readlane(__double2hiint(v), 0);
    int lo1 = __builtin_amdgcn_readlane(__double2loint(v), 32);
    int hi1 = __builtin_amdgcn_readlane(__double2hiint(v), 32);
    v = fmax(__hiloint2double(hi0, lo0), __hiloint2double(hi1, lo1));
  }
  return v;
}
__device__ __forceinline__ float seg_max_f_impl(float v) { return v; }
template <int L>
__device__ __forceinline__ float seg_max(float v) {
  v = fmaxf(v, dpp_f<kDppXor1>(v));
  v = fmaxf(v, dpp_f<kDppXor2>(v));
  if (L >= 8) v = fmaxf(v, dpp_f<kDppHalfMirror>(v));
  if (L >= 16) v = fmaxf(v, dpp_f<kDppMirror>(v));
  if (L >= 32) v = fmaxf(v, swz16_f(v));
  if (L >= 64)
    v = fmaxf(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0)),
              __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32)));
  return v;
}

// bits of a wave-wide ballot that belong to this lane's segment
template <int L>
__device__ __forceinline__ unsigned long long seg_bits(unsigned long long m) {
  if (L == 64) return m;
  const int seg = (threadIdx.x & 63) / L;
  return (m >> (seg * L)) & ((1ull << L) - 1);
}

template <int L>
__device__ __forceinline__ bool seg_all(bool b) {
  return seg_bits<L>(__ballot(!b)) == 0;
}

// Pivot-search key: the high word of |x| (sign cleared) plus one, as an
// unsigned integer; it orders finite magnitudes (to 20 mantissa bits: a
// pivot within a relative 2^-20 of the largest candidate may be chosen,
// which partial pivoting's stability does not distinguish), puts NaN above
// infinity, and leaves 0 for rows that are no longer candidates.
__device__ __forceinline__ unsigned pivot_key(double x) {
  return ((unsigned)__double2hiint(x) & 0x7FFFFFFFu) + 1u;
}
__device__ __forceinline__ unsigned pivot_key(float x) {
  return (((unsigned)__float_as_int(x) & 0x7FFFFFFFu) >> 8) + 1u;
}

// Maximum of an unsigned key over the segment.  Up to 16 lanes: DPP
// butterflies.  32 and 64 lanes: the butterflies give every lane its 16-lane
// row maximum and readlane gathers the four rows into scalar registers, so
// the reduction issues nothing to the LDS queue (where the broadcasts of the
// previous column may still wait); for L = 64 the result is wave-uniform.
template <int L>
__device__ __forceinline__ unsigned seg_max_u(unsigned v) {
  int x = (int)v;
  x = max((unsigned)x, (unsigned)__builtin_amdgcn_update_dpp(x, x, kDppXor1, 0xF, 0xF, false));
  x = max((unsigned)x, (unsigned)__builtin_amdgcn_update_dpp(x, x, kDppXor2, 0xF, 0xF, false));
  if (L >= 8)
    x = max((unsigned)x,
            (unsigned)__builtin_amdgcn_update_dpp(x, x, kDppHalfMirror, 0xF, 0xF, false));
  if (L >= 16)
    x = max((unsigned)x, (unsigned)__builtin_amdgcn_update_dpp(x, x, kDppMirror, 0xF, 0xF, false));
  if (L >= 32) {
    const unsigned r0 = __builtin_amdgcn_readlane(x, 0), r1 = __builtin_amdgcn_readlane(x, 16);
    const unsigned r2 = __builtin_amdgcn_readlane(x, 32), r3 = __builtin_amdgcn_readlane(x, 48);
    const unsigned lo = max(r0, r1), hi = max(r2, r3);
    if (L == 64) return max(lo, hi);
    return (threadIdx.x & 32) ? hi : lo;
  }
  return (unsigned)x;
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}
__device__ __forceinline__ float readlane_f(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double readlane_t(double v, int lane) { return readlane_d(v, lane); }
__device__ __forceinline__ float readlane_t(float v, int lane) { return readlane_f(v, lane); }

// The pivot lane of one elimination step and broadcasts from it.  `lane` is
// segment-relative.  L = 64: it is wave-uniform (a scalar) and values come
// through scalar registers (v_readlane), off the LDS path; narrower
// segments each have their own pivot and use ds_bpermute.
template <int L>
struct Pivot {
  int lane;
  template <typename T>
  __device__ __forceinline__ T get(T v) const {
    if constexpr (L == 64) return readlane_t(v, lane);
    else return __shfl(v, lane, L);
  }
};

// Lowest lane of the segment holding the largest key.
template <int L>
__device__ __forceinline__ Pivot<L> seg_pivot(unsigned key) {
  const unsigned m = seg_max_u<L>(key);
  const unsigned long long hit = __ballot(key == m);
  Pivot<L> p;
  if constexpr (L == 64) {
    p.lane = __ffsll(hit) - 1;
  } else if constexpr (L == 32) {
    const int p0 = __ffs((unsigned)hit) - 1, p1 = __ffs((unsigned)(hit >> 32)) - 1;
    p.lane = (threadIdx.x & 32) ? p1 : p0;
  } else {
    p.lane = __ffsll(seg_bits<L>(hit)) - 1;
  }
  return p;
}

// Pivot of step k recorded by lu_factor (pivl, LDS, one table per segment).
template <int L>
__device__ __forceinline__ Pivot<L> load_pivot(const int* pivl, int k) {
  Pivot<L> p;
  if constexpr (L == 64) p.lane = __builtin_amdgcn_readfirstlane(pivl[k]);
  else p.lane = pivl[k];
  return p;
}

// ------------------------------------------------------------------ LU
// In-place row-per-lane LU with partial pivoting over the lanes of a
// segment.  After the call lane r holds, for columns k < myk, the
// multipliers of the steps that eliminated it, and for columns >= myk its
// row of U; pivl[k] is the lane chosen at step k.
template <int SMAX, int L, typename T, bool RHS>
__device__ __forceinline__ void lu_factor(T (&a)[SMAX], int r, int* pivl, int& myk, T& b,
                                          T& dinv) {
  bool assigned = (r >= SMAX);
  dinv = T(1);
  myk = assigned ? -1 : SMAX;
#pragma unroll
  for (int k = 0; k < SMAX; ++k) {
    const T ak = a[k];
    // pivot lane: the lowest lane holding the largest key of the segment
    const unsigned key = assigned ? 0u : pivot_key(ak);
    const Pivot<L> p = seg_pivot<L>(key);
    const T apk = p.get(ak);
    const T bp = RHS ? p.get(b) : T(0);
    const T rinv = recip(apk);  // reciprocal pivot, as LAPACK's getf2 scales
    if (r == p.lane) {
      assigned = true;
      myk = k;
      dinv = rinv;
    }
    const T l = assigned ? T(0) : ak * rinv;
    if (!assigned) a[k] = l;
#pragma unroll
    for (int j = k + 1; j < SMAX; ++j) {
      const T pj = p.get(a[j]);
      a[j] -= l * pj;
    }
    if (RHS) b -= l * bp;
    if (r == 0) pivl[k] = p.lane;
  }
  __syncthreads();
}

// Back substitution U x = y, y eliminated in place by lu_factor<RHS=true>;
// dinv is 1/U[myk][myk] on each pivot lane.
template <int SMAX, int L, typename T>
__device__ __forceinline__ T lu_backsolve(const T (&a)[SMAX], T b, int r, const int* pivl,
                                          int myk, T dinv) {
  T x = T(0);
#pragma unroll
  for (int k = SMAX - 1; k >= 0; --k) {
    const Pivot<L> p = load_pivot<L>(pivl, k);
    const T mine = b * dinv;
    const T xk = p.get(mine);
    if (myk < k) b -= a[k] * xk;
    if (r == k) x = xk;
  }
  return x;
}

template <int SMAX, int L, typename T>
__device__ __forceinline__ T lu_solve(const T (&a)[SMAX], T b, int r, const int* pivl,
                                      int myk, T dinv) {
#pragma unroll
  for (int k = 0; k < SMAX; ++k) {
    const Pivot<L> p = load_pivot<L>(pivl, k);
    const T bp = p.get(b);
    if (myk > k) b -= a[k] * bp;
  }
  T x = T(0);
#pragma unroll
  for (int k = SMAX - 1; k >= 0; --k) {
    const Pivot<L> p = load_pivot<L>(pivl, k);
    const T mine = b * dinv;  // meaningful on the pivot lane of step k only
    const T xk = p.get(mine);
    if (myk < k) b -= a[k] * xk;
    if (r == k) x = xk;
  }
  return x;
}

// ------------------------------------------------------ static-pivot band LU
// f(integral_constant<int, i>) for i = B .. E-1, unrolled at compile time (the
// broadcasts below need the source lane as an immediate).
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Value held by lane K of this lane's segment, K known at compile time:
// 64-lane segments read it into a scalar register (v_readlane); narrower ones
// use a ds_swizzle broadcast inside each group of 32 lanes (the and-mask
// keeps the segment bits of the lane id, the or-mask selects lane K), which
// needs no address register.
template <int L, int K>
__device__ __forceinline__ double bcast_k(double v) {
  if constexpr (L == 64) {
    return readlane_d(v, K);
  } else {
    constexpr int pat = (0x1F & ~(L - 1)) | (K << 5);
    const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), pat);
    const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), pat);
    return __hiloint2double(hi, lo);
  }
}

// LU of a system numbered for the band LU (band_numbering on the host, the
// numbering in kernels.hpp): step k < KB = SMAX - kBandBorder pivots on lane
// k, whose row has its nonzeros right of the diagonal in columns k+1 ..
// k+kBandWidth and the border columns KB.. only -- fill stays inside that
// profile -- so only those columns are updated; the border block is then
// factored with partial pivoting among the rows not yet used.  Same outputs
// as lu_factor (pivots of the border steps in pivl); lmax: the largest
// |multiplier| this lane produced in the band steps, bad: one was NaN.
template <int SMAX, int L>
__device__ __forceinline__ void lu_factor_band(double (&a)[SMAX], int r, int* pivl, int& myk,
                                               double& b, double& dinv, double& lmax,
                                               bool& bad) {
  constexpr int KB = SMAX - kBandBorder;
  bool assigned = (r >= SMAX);
  dinv = 1.0;
  myk = assigned ? -1 : SMAX;
  lmax = 0.0;
  bad = false;
  static_for<0, KB>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    constexpr int JE = (k + 1 + kBandWidth < KB) ? k + 1 + kBandWidth : KB;
    const double apk = bcast_k<L, k>(a[k]);
    const double bp = bcast_k<L, k>(b);
    const double rinv = recip(apk);
    if (r == k) {
      assigned = true;
      myk = k;
      dinv = rinv;
    }
    const double l = assigned ? 0.0 : a[k] * rinv;
    if (!assigned) a[k] = l;
    lmax = fmax(lmax, fabs(l));
    bad |= (l != l);
    static_for<k + 1, JE>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      a[j] -= l * bcast_k<L, k>(a[j]);
    });
    static_for<KB, SMAX>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      a[j] -= l * bcast_k<L, k>(a[j]);
    });
    b -= l * bp;
  });
#pragma unroll
  for (int k = KB; k < SMAX; ++k) {
    const double ak = a[k];
    const unsigned key = assigned ? 0u : pivot_key(ak);
    const Pivot<L> p = seg_pivot<L>(key);
    const double apk = p.get(ak);
    const double bp = p.get(b);
    const double rinv = recip(apk);
    if (r == p.lane) {
      assigned = true;
      myk = k;
      dinv = rinv;
    }
    const double l = assigned ? 0.0 : ak * rinv;
    if (!assigned) a[k] = l;
#pragma unroll
    for (int j = k + 1; j < SMAX; ++j) a[j] -= l * p.get(a[j]);
    b -= l * bp;
    if (r == 0) pivl[k] = p.lane;
  }
  __syncthreads();
}

// Back substitution after lu_factor_band (pivots of the band steps are the
// lanes k themselves).
template <int SMAX, int L>
__device__ __forceinline__ double lu_backsolve_band(const double (&a)[SMAX], double b, int r,
                                                    const int* pivl, int myk, double dinv) {
  constexpr int KB = SMAX - kBandBorder;
  double x = 0.0;
#pragma unroll
  for (int k = SMAX - 1; k >= KB; --k) {
    const Pivot<L> p = load_pivot<L>(pivl, k);
    const double xk = p.get(b * dinv);
    if (myk < k) b -= a[k] * xk;
    if (r == k) x = xk;
  }
  static_for<0, KB>([&](auto kc) {
    constexpr int k = KB - 1 - decltype(kc)::value;
    const double xk = bcast_k<L, k>(b * dinv);
    if (myk < k) b -= a[k] * xk;
    if (r == k) x = xk;
  });
  return x;
}

// Forward and back substitution with factors of lu_factor_band.
template <int SMAX, int L>
__device__ __forceinline__ double lu_solve_band(const double (&a)[SMAX], double b, int r,
                                                const int* pivl, int myk, double dinv) {
  constexpr int KB = SMAX - kBandBorder;
  static_for<0, KB>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    const double bk = bcast_k<L, k>(b);
    if (myk > k) b -= a[k] * bk;
  });
#pragma unroll
  for (int k = KB; k < SMAX; ++k) {
    const Pivot<L> p = load_pivot<L>(pivl, k);
    const double bp = p.get(b);
    if (myk > k) b -= a[k] * bp;
  }
  return lu_backsolve_band<SMAX, L>(a, b, r, pivl, myk, dinv);
}

// ------------------------------------------------------ per-instance pieces
struct Seg;
template <int SMAX, int L>
__device__ void lu_factor_lds(double (&a)[SMAX], const Seg& s, int r, double gmin, int& myk,
                              double& b, double& dinv);
struct Seg {
  double* zl;       // [SMAX] state values visible to device lanes (by position)
  double* res;      // [sk] residual
  double* mat;      // [SMAX][sk] Jacobian, column-major; columns sk.. are zero
  int* pivl;        // [SMAX]
  const short* dev;     // LDS tables of the pattern
  const short* jpos;
  const short* rpos;
  int sk;           // positions of the system (PatternDev::sk)
  bool real;        // this lane's position holds an unknown (else: padding)
  bool node;        // ... a node voltage (gmin, damping)
};

// v added to an FP64 value in LDS (ds_add_f64).
__device__ __forceinline__ void lds_add(double* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <bool JAC>
__device__ __forceinline__ void eval_one(const short* dv, const double* prow, const Seg& s,
                                         double tt, double cf, double cq, bool store) {
  const int kind = dv[0];
  const int m = kind_m(kind);
  double u[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int idx = dv[2 + i];
    u[i] = idx >= 0 ? s.zl[idx] : 0.0;
  }
  double res[4], jac[16];
  device_eval<JAC, true>(kind, dv[1], prow + dv[6], u, tt, cf, cq, res, jac);
  if (!store) return;
  const short* rp = s.rpos + dv[7];
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (i < m) {
      const int q = rp[i];
      if (q >= 0) lds_add(s.res + q, res[i]);
    }
  if (JAC) {
    const short* jp = s.jpos + dv[8];
#pragma unroll
    for (int i = 0; i < 16; ++i)
      if (i < m * m) {
        const int q = jp[i];
        if (q >= 0) lds_add(s.mat + q, jac[i]);
      }
  }
}

// Devices of the segment, one per lane per pass, after lane r has cleared
// residual entry r and matrix row r.  The device table lists every device
// kind in its own run of passes (padded with kind -1), so the lanes of a
// pass never diverge between the EKV model and the linear elements.  The
// segment's LDS is touched by its own wave only, so the additions of one
// evaluation land in a fixed order (pass by pass, lanes in the LDS unit's
// order within an instruction): repeated runs give identical sums.
// PAIR: two devices per step, so that their dependent chains (exp, log1p,
// sqrt, divisions) interleave; it costs registers and pays where measured
// (the reused-factor kernels of tiles 28 and 32, two waves per SIMD: 1.05x
// on the ring workload; elsewhere, tile 40 included, it was slower).
template <int L, bool JAC, bool PAIR = false>
__device__ __forceinline__ void eval_devices(const TopoDev& t, const double* prow,
                                             const Seg& s, double tt, double cf, double cq,
                                             int r, int n_slots) {
  if (s.real) {
    s.res[r] = 0.0;
    if (JAC)
      for (int j = 0; j < s.sk; ++j) s.mat[j * s.sk + r] = 0.0;
  }
  __syncthreads();
  int d = r;
  if (PAIR) {
    for (; d + L < n_slots; d += 2 * L) {
      const short* dv0 = s.dev + 10 * d;
      const short* dv1 = s.dev + 10 * (d + L);
      if (dv0[0] >= 0) eval_one<JAC>(dv0, prow, s, tt, cf, cq, true);
      if (dv1[0] >= 0) eval_one<JAC>(dv1, prow, s, tt, cf, cq, true);
    }
  }
  for (; d < n_slots; d += L) {
    const short* dv = s.dev + 10 * d;
    if (dv[0] >= 0) eval_one<JAC>(dv, prow, s, tt, cf, cq, true);
  }
}

__device__ __forceinline__ double row_residual(const TopoDev& t, const Seg& s, int r) {
  return s.real ? s.res[r] : 0.0;
}

// Row r of the Jacobian (plus gmin on node diagonals) into registers.
template <int SMAX>
__device__ __forceinline__ void assemble_row(const TopoDev& t, const Seg& s, int r,
                                             double gmin, double (&a)[SMAX]) {
  if (s.real) {
    const double* row = s.mat + r;
    const int sk = s.sk;
#pragma unroll
    for (int j = 0; j < SMAX; ++j) {
      double v = row[j * sk];
      if (j == r && s.node) v += gmin;
      a[j] = v;
    }
  } else {
#pragma unroll
    for (int j = 0; j < SMAX; ++j) a[j] = (j == r) ? 1.0 : 0.0;  // padding rows
  }
}

// dz = A^-1 b for the segment's system held row-per-lane.
template <int SMAX, int L, int PREC>
__device__ __forceinline__ double linear_solve(double (&a)[SMAX], double b, int r,
                                               const Seg& s) {
  int myk;
  if (PREC == 0) {
    double y = b, dinv;
    lu_factor<SMAX, L, double, true>(a, r, s.pivl, myk, y, dinv);
    return lu_backsolve<SMAX, L, double>(a, y, r, s.pivl, myk, dinv);
  }
  // FP32 factorization + FP64 iterative refinement
  float f[SMAX];
#pragma unroll
  for (int j = 0; j < SMAX; ++j) f[j] = (float)a[j];
  float y = (float)b, dinv;
  lu_factor<SMAX, L, float, true>(f, r, s.pivl, myk, y, dinv);
  double x = (double)lu_backsolve<SMAX, L, float>(f, y, r, s.pivl, myk, dinv);
#pragma unroll 1
  for (int it = 0; it < 2; ++it) {
    __syncthreads();
    if (r < SMAX) s.zl[r] = x;
    __syncthreads();
    double ax = 0.0;
#pragma unroll
    for (int j = 0; j < SMAX; ++j) ax += a[j] * s.zl[j];
    double res = (r < SMAX) ? b - ax : 0.0;
    x += (double)lu_solve<SMAX, L, float>(f, (float)res, r, s.pivl, myk, dinv);
  }
  __syncthreads();
  return x;
}

// Partial-pivoting LU of the segment's Jacobian where it lies in LDS
// (column-major, plus gmin on the node diagonals), the fallback of the band
// LU when a static pivot fails its test: loops with run-time bounds keep it
// out of the registers of the unrolled code.  Leaves a, pivl, myk, dinv and
// the eliminated right-hand side b as lu_factor<RHS=true> does; positions
// sk.. are identity rows pivoted on themselves.
template <int SMAX, int L>
__device__ void lu_factor_lds(double (&a)[SMAX], const Seg& s, int r, double gmin, int& myk,
                              double& b, double& dinv) {
  const int sk = s.sk;
  double* m = s.mat;
  __syncthreads();
  if (r < sk) {
    if (s.node) m[r * sk + r] += gmin;
    if (!s.real) m[r * sk + r] = 1.0;  // padding position: identity row
  }
  bool assigned = r >= sk;
  myk = assigned ? r : SMAX;
  dinv = 1.0;
  __syncthreads();
  for (int k = 0; k < sk; ++k) {
    const double ak = assigned ? 0.0 : m[k * sk + r];
    const Pivot<L> p = seg_pivot<L>(assigned ? 0u : pivot_key(ak));
    const double apk = m[k * sk + p.lane];
    const double bp = p.get(b);
    const double rinv = recip(apk);
    if (r == p.lane) {
      assigned = true;
      myk = k;
      dinv = rinv;
    }
    const double l = assigned ? 0.0 : ak * rinv;
    if (!assigned) m[k * sk + r] = l;
    for (int j = k + 1; j < sk; ++j) {
      const double pj = m[j * sk + p.lane];
      if (!assigned) m[j * sk + r] -= l * pj;
    }
    b -= l * bp;
    if (r == 0) s.pivl[k] = p.lane;
    __syncthreads();
  }
  if (r >= sk && r < SMAX) s.pivl[r] = r;
#pragma unroll
  for (int j = 0; j < SMAX; ++j) a[j] = (r < sk && j < sk) ? m[j * sk + r] : (j == r ? 1.0 : 0.0);
  __syncthreads();
}

// Damped Newton on one segment's system (voltax/analysis.py:221-230).
// PREC 0: FP64 LU every iteration; 1: FP32 LU + FP64 refinement; 2: FP64 LU
// whose factors later iterations of the same solve reuse when `reuse` is set
// (simplified Newton: residual-only device passes and a forward/backward
// substitution).  Such an update converges linearly, with rate theta = its
// tolerance-weighted size over the previous one's, so the error it leaves is
// about theta/(1-theta) times its size, where a fresh Newton update leaves
// about the square of its size.  A reused-factor update therefore ends the
// iteration only when it passes the plain convergence test and that error
// estimate is below kReuseTol of the tolerance; one with theta > 1/2 makes
// the next iteration factor afresh (and is dropped if theta >= 1).
template <int SMAX, int L, int PREC, bool BAND>
__device__ __forceinline__ void newton(const TopoDev& t, const double* prow, const Seg& s,
                                       int r, int n_slots, double& z,
                       double hist, double tt, double cf, double cq, double gmin,
                       int max_steps, const NewtonOpts& o, bool active, bool& conv,
                       int& iters, int& facts, bool reuse) {
  constexpr bool KEEP = PREC == 2;
  constexpr bool PAIR = KEEP && !BAND && SMAX >= 28 && SMAX <= 32;
  constexpr double kReuseTol = 1e-4;
  bool done = false;
  int k = 0, nf = 0;
  double a[SMAX];
  int myk = 0;
  double dinv = 1.0;
  bool factored = false;  // wave-uniform: a, myk, dinv and pivl hold this solve's factors
  bool fband = false;     // wave-uniform: ... from lu_factor_band
  double wprev = 0.0;     // weighted size of the last applied update (per segment)
  for (int it = 0; it < max_steps; ++it) {
    if (__all(done || !active)) break;
    __syncthreads();
    if (s.real) s.zl[r] = z;
    __syncthreads();
    const bool chord = KEEP && reuse && factored;
    double dz;
    if (chord) {
      eval_devices<L, false, PAIR>(t, prow, s, tt, cf, cq, r, n_slots);
      __syncthreads();
      const double rv = row_residual(t, s, r) + hist + (s.node ? gmin * z : 0.0);
      if (BAND && fband)
        dz = lu_solve_band<SMAX, L>(a, -rv, r, s.pivl, myk, dinv);
      else
        dz = lu_solve<SMAX, L, double>(a, -rv, r, s.pivl, myk, dinv);
    } else {
      eval_devices<L, true, PAIR>(t, prow, s, tt, cf, cq, r, n_slots);
      __syncthreads();
      const double rv = row_residual(t, s, r) + hist + (s.node ? gmin * z : 0.0);
      assemble_row<SMAX>(t, s, r, gmin, a);
      if (PREC == 1) {
        dz = linear_solve<SMAX, L, PREC>(a, -rv, r, s);
      } else if (BAND) {
        // static pivots; a multiplier above o.band_max_mult anywhere in the
        // wave (a pivot small against its column) sends the factorization
        // back to partial pivoting
        double y = -rv, lmax;
        bool bad;
        lu_factor_band<SMAX, L>(a, r, s.pivl, myk, y, dinv, lmax, bad);
        const bool bf = !__any(bad || !(lmax <= o.band_max_mult));
        if (!bf) {
          y = -rv;
          lu_factor_lds<SMAX, L>(a, s, r, gmin, myk, y, dinv);
        }
        dz = bf ? lu_backsolve_band<SMAX, L>(a, y, r, s.pivl, myk, dinv)
                : lu_backsolve<SMAX, L, double>(a, y, r, s.pivl, myk, dinv);
        factored = true;
        fband = bf;
      } else {
        double y = -rv;
        lu_factor<SMAX, L, double, true>(a, r, s.pivl, myk, y, dinv);
        dz = lu_backsolve<SMAX, L, double>(a, y, r, s.pivl, myk, dinv);
        factored = true;
      }
    }
    if (!s.real) dz = 0.0;
    const double vmax = seg_max<L>(s.node ? fabs(dz) : 0.0);
    const double scale = fmin(1.0, o.max_dv / (vmax + 1e-300));
    const double zn = z + scale * dz;
    const bool small = !s.real || (fabs(dz) <= o.atol + o.rtol * fabs(zn));
    const bool fin = !s.real || isfinite(zn);
    bool ok = seg_all<L>(small && fin) && scale == 1.0;
    bool apply = true;
    if (KEEP) {
      const double w = seg_max<L>(s.real ? fabs(dz) / (o.atol + o.rtol * fabs(z)) : 0.0);
      if (chord) {
        const double theta = w / wprev;
        const bool slow = !(theta <= 0.5);  // also NaN
        ok = ok && !slow && theta * w <= kReuseTol * (1.0 - theta);
        apply = theta < 1.0;                // keep a reduced step, drop a growing one
        factored = !__any(slow && active && !done);
      }
      if (apply) wprev = w;
    }
    if (active && !done) {
      if (apply) z = zn;
      ++k;
      nf += chord ? 0 : 1;
      done = ok;
    }
  }
  conv = done;
  iters = k;
  facts = nf;
}

template <int SMAX, int L, int PREC, bool BAND>
__device__ void run_segment(const PatternDev& pd, const WaveItem& item, const int* item_inst,
                            const int64_t* ipoff,
                            const int64_t* izoff, const double* params, const double* z0,
                            const double* ts, int T, int method, int do_dc, double t_dc,
                            const NewtonOpts& o, double* out_z, int8_t* out_conv,
                            int32_t* out_iters) {
  extern __shared__ double lds[];
  const TopoDev& t = pd.t;
  const int lane = threadIdx.x;
  const int seg = lane / L, r = lane % L;
  const bool live = seg < item.ninst;  // segment holds a real instance
  const int inst = item_inst[live ? seg : 0];
  // pattern tables -> LDS (once per launch), then per-instance buffers
  int* tw = reinterpret_cast<int*>(lds);
  const int* gw = reinterpret_cast<const int*>(pd.blob);
  for (int i = lane; i < pd.blob_words; i += 64) tw[i] = gw[i];
  const short* tb = reinterpret_cast<const short*>(lds);
  const int blob_doubles = (pd.blob_words + 1) / 2;
  double* base = lds + blob_doubles + (long)seg * pd.seg_doubles;
  Seg s;
  s.dev = tb;
  s.jpos = tb + pd.o_jpos;
  s.rpos = tb + pd.o_rpos;
  const int sk = pd.sk;
  s.sk = sk;
  s.real = (pd.realmask >> r) & 1;
  s.node = (pd.nodemask >> r) & 1;
  s.zl = base;
  s.res = base + SMAX;
  s.mat = s.res + sk;
  s.pivl = reinterpret_cast<int*>(s.mat + SMAX * sk);
  // residual and matrix start at zero: padding rows and columns stay so
  for (int q = r; q < sk + SMAX * sk; q += L) s.res[q] = 0.0;
  __syncthreads();
  const double* prow = params + ipoff[inst];
  const int64_t zo = izoff[inst];
  const int S = t.S;
  const int qr = s.real ? tb[pd.o_q + r] : 0;  // the unknown at this lane's position
  double z = s.real ? z0[zo + qr] : 0.0;
  double* oz = out_z + zo * (int64_t)T;
  int8_t* oc = out_conv + (int64_t)inst * T;
  int32_t* oi = out_iters + (int64_t)inst * T;

  // One call site of the Newton loop (code size: the loop body must stay
  // resident in the instruction cache).  Phases: DC direct attempt, gmin
  // stages (only for segments whose direct attempt failed), time steps.
  const int first = o.gmin_steps == 0 ? o.max_steps : o.direct_steps;
  const double guess = z;
  method &= 3;
  // extrapolated starts for backward Euler only: the trapezoidal rule's
  // step-to-step ringing makes extrapolated guesses worse than the last point
  const bool predict = method == 2;
  method &= 1;
  // the last accepted points, newest first (zh[0] is the reference's guess)
  constexpr int kHist = 4;  // cubic: measured best (quadratic 2.34, cubic 2.06, quartic 2.06 iterations)
  constexpr int kPredictSteps = 8;  // iteration budget of a predicted start
  double zh[kHist];
#pragma unroll
  for (int i = 0; i < kHist; ++i) zh[i] = z;
  bool retry = false;     // redo the current step from zh[0]
  int it_pred = 0;        // iterations spent on a failed predicted start
  int nf_pred = 0;        // and their LU factorizations
  // iteration counts as reported: with reused factors (PREC 2) the number of
  // LU factorizations goes into bits 16..30
  auto count = [](int it, int nf) { return PREC == 2 ? (it | (nf << 16)) : it; };
  const double l0 = log10(o.gmin_start), l1 = log10(o.gmin);
  bool conv0 = true, need = false;
  int it0 = 0, nf0 = 0;
  int phase = do_dc ? 0 : o.gmin_steps + 1;  // 0: direct, 1..G: gmin stage, G+1..: steps
  const int G = o.gmin_steps;
  if (!do_dc) {  // the given initial state is time point 0
    if (live) {
      if (s.real) oz[qr] = z;
      if (r == 0) {
        oc[0] = 1;
        oi[0] = 0;
      }
    }
    if (T == 1) return;
  }
  while (true) {
    double hist = 0.0, tt = t_dc, cf = 1.0, cq = 0.0, gmin = o.gmin;
    int max_steps = o.max_steps;
    bool active = true;
    int n = 0;
    if (phase == 0) {
      max_steps = first;
    } else if (phase <= G) {
      const double e = (G == 1) ? l0 : (phase == G ? l1 : l0 + (phase - 1) * ((l1 - l0) / (G - 1)));
      gmin = pow(10.0, e);
      active = need;
      if (phase == 1 && need) z = guess;
    } else {
      n = phase - G;
      if (n >= T) break;
      const double tp = ts[n - 1], tn = ts[n];
      cq = 1.0 / (tn - tp);
      cf = method == 0 ? 1.0 : 0.5;
      tt = tn;
      if (!retry) {
#pragma unroll
        for (int i = kHist - 1; i > 0; --i) zh[i] = zh[i - 1];
        zh[0] = z;
      }
      // history: -q(z_prev)/h (+ f(z_prev, t_prev)/2 for trap)
      __syncthreads();
      if (s.real) s.zl[r] = z;
      __syncthreads();
      eval_devices<L, false>(t, prow, s, tp, method == 0 ? 0.0 : 0.5, -cq, r, pd.dev_slots);
      __syncthreads();
      hist = row_residual(t, s, r);
      // initial guess: the last point, or the value at t_n of the cubic
      // through the last four points (fewer on the first steps; Lagrange
      // form).  A predicted start gets a few iterations; if it does not
      // converge the step is redone from the last point with the full budget.
      if (predict && n >= 2 && !retry) {
        const int np = n < kHist ? n : kHist;
        max_steps = o.max_steps < kPredictSteps ? o.max_steps : kPredictSteps;
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < kHist; ++i) {
          if (i >= np) break;
          const double ti = ts[n - 1 - i];
          double w = 1.0;
#pragma unroll
          for (int j = 0; j < kHist; ++j) {
            if (j >= np || j == i) continue;
            const double tj = ts[n - 1 - j];
            w *= (tn - tj) / (ti - tj);
          }
          acc += w * zh[i];
        }
        z = acc;
      }
    }
    bool ok;
    int it, nf;
    newton<SMAX, L, PREC, BAND>(t, prow, s, r, pd.dev_slots, z, hist, tt, cf, cq, gmin,
                                max_steps, o, active, ok, it, nf, phase > G);
    bool dc_done = false;
    if (phase == 0) {
      conv0 = ok;
      it0 = it;
      nf0 = nf;
      need = !ok && G > 0;
      phase = __any(need) ? 1 : G + 1;
      dc_done = phase == G + 1;
    } else if (phase <= G) {
      if (need) {
        conv0 = ok;
        it0 += it;
        nf0 += nf;
      }
      ++phase;
      dc_done = phase == G + 1;
    } else {
      // a predicted start that did not converge: the same step again from
      // the last point (every segment of the wave takes the same path)
      const bool again = predict && !retry && n >= 2 && __any(!ok && active && live);
      if (again) {
        retry = true;
        z = zh[0];
        it_pred = it;
        nf_pred = nf;
        continue;
      }
      if (live) {
        if (s.real) oz[(int64_t)n * S + qr] = z;
        if (r == 0) {
          oc[n] = ok ? 1 : 0;
          oi[n] = count(it + (retry ? it_pred : 0), nf + (retry ? nf_pred : 0));
        }
      }
      retry = false;
      ++phase;
    }
    if (dc_done) {
      // the DC operating point is time point 0
      if (live) {
        if (s.real) oz[qr] = z;
        if (r == 0) {
          oc[0] = conv0 ? 1 : 0;
          oi[0] = count(it0, nf0);
        }
      }
      if (T == 1) break;
    }
  }
}

// Waves per SIMD the register budget is set for: two for tiles up to 32
// (the compiler then keeps the kernel within 256 registers, spilling some
// values to scratch, and the second wave hides the latency of the
// pivot-search and row-broadcast chains), one above (there the spills of a
// 256-register budget cost more than the second wave gains).  Measured on
// the 4096-ring workload: 1.31x over one wave everywhere, 1.13x over two.
constexpr int dense_waves_per_simd(int smax) { return smax <= 32 ? 2 : 1; }

template <int SMAX, int L, int PREC, bool BAND>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(dense_waves_per_simd(SMAX), dense_waves_per_simd(SMAX))))
k_dense_run(const PatternDev* __restrict__ pats, const WaveItem* __restrict__ items,
            const int64_t* __restrict__ ipoff, const int64_t* __restrict__ izoff,
            const double* __restrict__ params, const double* __restrict__ z0,
            const double* __restrict__ ts, int T, int method, int do_dc, double t_dc,
            NewtonOpts o, double* __restrict__ out_z, int8_t* __restrict__ out_conv,
            int32_t* __restrict__ out_iters) {
  const WaveItem item = items[blockIdx.x];
  const PatternDev pd = pats[item.pattern];
  run_segment<SMAX, L, PREC, BAND>(pd, item, items[blockIdx.x].inst, ipoff, izoff, params, z0,
                                   ts, T, method, do_dc, t_dc, o, out_z, out_conv, out_iters);
}

// Launches the items of one tile class (all with the same smax): with
// VXM_BAND_LU in `method` the first n_band items (patterns with a band
// numbering) run the band-LU kernel, the others the partial-pivoting one,
// one launch after the other on `stream`.  Runs that reuse factors
// (precision 2) take the band LU in 64-lane classes only: in narrower
// segments two instances share a wave, the static-pivot broadcasts go
// through the LDS unit, and with the factors held across iterations the
// kernel measured no faster than partial pivoting.
template <int SMAX, int L>
int launch_dense_class(const PatternDev* pats, const WaveItem* items, int n_items, int n_band,
                       int lds_bytes, const int64_t* ipoff, const int64_t* izoff,
                       const double* params, const double* z0, const double* ts, int T,
                       int method, int do_dc, double t_dc, const NewtonOpts& o, int precision,
                       double* out_z, int8_t* out_conv, int32_t* out_iters,
                       hipStream_t stream) {
  if (n_items <= 0) return 0;
  const bool band = (method & 8) && n_band > 0 &&
                    (precision == 0 || (precision == 2 && L == 64));
  const int nb = band ? n_band : 0;
  if (nb > 0) {
    if constexpr (L == 64) {
      if (precision == 2)
        hipLaunchKernelGGL((k_dense_run<SMAX, L, 2, true>), dim3(nb), dim3(64), lds_bytes,
                           stream, pats, items, ipoff, izoff, params, z0, ts, T, method, do_dc,
                           t_dc, o, out_z, out_conv, out_iters);
    }
    if (precision == 0)
      hipLaunchKernelGGL((k_dense_run<SMAX, L, 0, true>), dim3(nb), dim3(64), lds_bytes, stream,
                         pats, items, ipoff, izoff, params, z0, ts, T, method, do_dc, t_dc, o,
                         out_z, out_conv, out_iters);
  }
  const int nr = n_items - nb;
  const WaveItem* rest = items + nb;
  if (nr > 0) {
    if (precision == 0)
      hipLaunchKernelGGL((k_dense_run<SMAX, L, 0, false>), dim3(nr), dim3(64), lds_bytes, stream,
                         pats, rest, ipoff, izoff, params, z0, ts, T, method, do_dc, t_dc, o,
                         out_z, out_conv, out_iters);
    else if (precision == 2)
      hipLaunchKernelGGL((k_dense_run<SMAX, L, 2, false>), dim3(nr), dim3(64), lds_bytes, stream,
                         pats, rest, ipoff, izoff, params, z0, ts, T, method, do_dc, t_dc, o,
                         out_z, out_conv, out_iters);
    else
      hipLaunchKernelGGL((k_dense_run<SMAX, L, 1, false>), dim3(nr), dim3(64), lds_bytes, stream,
                         pats, rest, ipoff, izoff, params, z0, ts, T, method, do_dc, t_dc, o,
                         out_z, out_conv, out_iters);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

#define VXM_DENSE_CLASS_DECL(SM, LN)                                                        \
  template int launch_dense_class<SM, LN>(                                                 \
      const PatternDev*, const WaveItem*, int, int, int, const int64_t*, const int64_t*,   \
      const double*, const double*, const double*, int, int, int, double, const NewtonOpts&, \
      int, double*, int8_t*, int32_t*, hipStream_t);

}  // namespace vxm
