// Numeric phase of the supernodal multifrontal LU and its triangular solves.
//
// Reference: voltax/sparse.py:549-572 (factor: level by level, pivot blocks
// with partial pivoting inside the block, Schur complements into ancestors)
// and :575-595 (solve, also transposed for the adjoint).
//
// One workgroup per (supernode, instance) of a level; a level is one
// launch.  Supernode s assembles its dense front F (k+r)^2 from the CSR
// values it owns and, child by child in a fixed order, the children's Schur
// complements (extend-add through index maps: no atomics, reproducible
// sums), then eliminates its k pivot columns with partial pivoting among its
// k pivot rows.  After that F holds L11\U11, U12, L21 and, in the trailing
// r x r block, the Schur complement its parent will pull.  Fronts small
// enough live in LDS during the elimination; larger ones are eliminated in
// place in HBM (L2-resident).
//
// Factors may be FP32 (mixed precision: FP32 LU + FP64 refinement) or FP64;
// solves always accumulate in FP64.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "numeric.hpp"
#include "splu.hpp"

namespace vxm {

// Maximum of a pivot magnitude over the 64 lanes of a wave: DPP butterflies
// inside rows of 16 lanes, then the four row maxima through scalar
// registers (no LDS traffic on the panel's critical path).
template <int CTRL>
__device__ __forceinline__ double front_dpp(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float front_dpp(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ double front_readlane(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
__device__ __forceinline__ float front_readlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ cplx front_readlane(cplx v, int l) {
  return {front_readlane(v.re, l), front_readlane(v.im, l)};
}
template <typename R>
__device__ __forceinline__ R front_wave_max(R v) {
  v = fmax(v, front_dpp<0xB1>(v));   // lane ^ 1
  v = fmax(v, front_dpp<0x4E>(v));   // lane ^ 2
  v = fmax(v, front_dpp<0x141>(v));  // half-row mirror
  v = fmax(v, front_dpp<0x140>(v));  // row mirror
  return fmax(fmax(front_readlane(v, 0), front_readlane(v, 16)),
              fmax(front_readlane(v, 32), front_readlane(v, 48)));
}

// Trailing update A22 -= L21 U12 of one 16-column panel (rows and columns
// je..nf-1 of the row-major front F) on the matrix cores: 16 x 16 output
// tiles, one per wavefront in turn, each the sum of four
// v_mfma_{f64,f32}_16x16x4 products over the panel's columns (zero padded
// past je).  Operand lanes: A[i][q] at lane i + 16 q, B[q][c] at lane
// c + 16 q; result lane l, register v: row l/16 + 4 v (FP64) or
// 4 (l/16) + v (FP32), column l % 16.  Waves wave, wave + nwaves, ... of the
// workgroup take the tiles of column blocks tj0 .. tj1-1 (block 0: columns
// je..je+15, the next panel's).
template <typename T>
__device__ __forceinline__ void front_update_mfma(T* F, int nf, int jb, int je, int wave,
                                                  int nwaves, int tj0, int tj1) {
  using V4 = T __attribute__((ext_vector_type(4)));
  const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
  const int m = nf - je, mt = (m + 15) / 16;
  const int nc = min(tj1, mt) - tj0;
  if (nc <= 0) return;
  for (int tile = wave; tile < mt * nc; tile += nwaves) {
    const int ti = tile / nc, tj = tj0 + (tile - ti * nc);
    const int ia = je + 16 * ti + lr, cb = je + 16 * tj + lr;
    V4 acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int q = jb + 4 * kk + lk;
      const T av = (q < je && ia < nf) ? F[ia * nf + q] : T(0);
      const T bv = (q < je && cb < nf) ? F[q * nf + cb] : T(0);
      if constexpr (std::is_same<T, double>::value)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
      else
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
    }
    if (cb < nf) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = je + 16 * ti + (std::is_same<T, double>::value ? lk + 4 * v : 4 * lk + v);
        if (row < nf) F[row * nf + cb] -= acc[v];
      }
    }
  }
}

// Smallest of an int over the 64 lanes of a wave, as a wave-uniform (scalar)
// value: DPP inside rows of 16 lanes, then the four row minima.
template <int CTRL>
__device__ __forceinline__ int front_dpp_int(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ int front_wave_min(int v) {
  v = min(v, front_dpp_int<0xB1>(v));   // lane ^ 1
  v = min(v, front_dpp_int<0x4E>(v));   // lane ^ 2
  v = min(v, front_dpp_int<0x141>(v));  // half-row mirror
  v = min(v, front_dpp_int<0x140>(v));  // row mirror
  return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

template <int B, int E, typename Fn>
__device__ __forceinline__ void panel_static_for(Fn&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    panel_static_for<B + 1, E>(f);
  }
}

// One wavefront factors a panel -- rows jb..nf-1, columns jb..je-1 of the
// row-major front F (LDS or HBM) -- in registers: lane l holds rows
// jb + l + 64 r (r < RPL, nf - jb <= 64 RPL), 16 columns each, and the row
// position each of them has reached (row swaps exchange positions, not
// data).  Step j: the first largest magnitude among positions j..k-1 of
// column j (the lowest position on ties, as LAPACK's getf2), the pivot row's
// entries read from its lane, the positions of rows j and pivot exchanged,
// the multipliers by the reciprocal pivot and the rank-1 update of the
// panel's later columns; each row is stored at its final position.  The
// pivots go to pv (the factor's record) and prow (for the swaps of the other
// columns).
template <typename T, int RPL>
__device__ __forceinline__ void panel_regs(T* F, int nf, int k, int jb, int je, int* pv,
                                           int* prow) {
  using R = typename Num<T>::R;
  constexpr int kNone = 0x7fffffff;
  const int lane = threadIdx.x & 63, w = je - jb;
  T P[RPL][16];
  int pos[RPL];
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int i = jb + lane + 64 * r;
    pos[r] = i < nf ? i : kNone;
#pragma unroll
    for (int q = 0; q < 16; ++q) P[r][q] = (i < nf && q < w) ? F[i * nf + jb + q] : T(0);
  }
  // every step unrolled (compile-time jj), so P stays in registers
  panel_static_for<0, 16>([&](auto jc) {
    constexpr int jj = decltype(jc)::value;
    if (jj >= w) return;
    const int j = jb + jj;
    R best = R(-1);
    int bp = kNone;
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      const R v = pmag(P[r][jj]);
      if (pos[r] >= j && pos[r] < k && (v > best || (v == best && pos[r] < bp))) {
        best = v;
        bp = pos[r];
      }
    }
    const R mx = front_wave_max(best);
    int bi = front_wave_min(best == mx ? bp : kNone);  // wave-uniform
    if (bi >= k) bi = j;  // no comparable magnitude (NaN column): keep the diagonal
    if (lane == 0) {
      pv[j] = bi;
      prow[jj] = bi;
    }
    // the lane and slot holding position bi (one ballot per slot, scalar)
    int lb = 0, rb = 0;
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      const unsigned long long m = __ballot(pos[r] == bi);
      if (m) {
        lb = __ffsll(m) - 1;
        rb = r;
      }
    }
    T rowb[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      T v = P[0][q];
#pragma unroll
      for (int r = 1; r < RPL; ++r) v = rb == r ? P[r][q] : v;
      rowb[q] = front_readlane(v, lb);
    }
#pragma unroll
    for (int r = 0; r < RPL; ++r) pos[r] = pos[r] == j ? bi : (pos[r] == bi ? j : pos[r]);
    const T dr = T(1) / rowb[jj];  // reciprocal pivot (LAPACK's getf2 scaling)
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      if (pos[r] > j && pos[r] < nf) {
        const T l = P[r][jj] * dr;
        P[r][jj] = l;
#pragma unroll
        for (int q = jj + 1; q < 16; ++q)
          if (q < w) P[r][q] -= l * rowb[q];
      }
    }
  });
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (pos[r] < nf && q < w) F[pos[r] * nf + jb + q] = P[r][q];
  }
}

// Fronts larger than 64 unknowns: one 256-thread workgroup per front, blocked
// right-looking LU in panels of 16 pivot columns.  Within a panel: pivot
// search among the k pivot rows, full-row swap, column scaling by the
// reciprocal pivot and rank-1 updates restricted to the panel's columns; then
// the panel's block row U12 = L11^-1 A12 (one thread per column) and one
// rank-16 update of the trailing block (FP64 / FP32 fronts on the matrix
// cores, complex ones in 4 x 4 register tiles per thread), so the trailing
// matrix is read and written once per panel instead of once per pivot.
//
// A front factored in place in HBM (LDS = false) stages each panel in LDS
// when `panel_lds` (its rows x 17 elements fit): the panel's pivot steps then
// read and write LDS only, and its row swaps reach the columns outside the
// panel afterwards, all columns in parallel -- the same operations on the same
// values as swapping whole rows step by step, so the same bits.
template <typename T, bool LDS>
__global__ void __launch_bounds__(256)
k_front_factor(SPlanDev p, int lbeg, const typename Num<T>::D* __restrict__ data,
               T* __restrict__ front, int* __restrict__ piv, int panel_lds) {
  using R = typename Num<T>::R;
  constexpr int NB = 16, PP = NB + 1;  // panel width, LDS pitch of a staged panel
  __shared__ int prow[NB];
  const int s = p.level_sn[lbeg + blockIdx.x];
  const int b = blockIdx.y;
  const SNode sn = p.sn[s];
  const int k = sn.k, nf = sn.k + sn.r;
  T* Fg = front + (int64_t)b * p.front_total + sn.front_off;
  extern __shared__ double smem_d[];
  T* F = LDS ? reinterpret_cast<T*>(smem_d) : Fg;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nf2 = nf * nf;
  for (int e = tid; e < nf2; e += nt) F[e] = T(0);
  __syncthreads();
  const typename Num<T>::D* db = data + (int64_t)b * p.nnz;
  for (int e = tid; e < sn.n_aent; e += nt)
    F[p.apos[sn.aent_off + e]] += cvt<T>(db[p.aslot[sn.aent_off + e]]);
  __syncthreads();
  for (int ci = 0; ci < sn.n_child; ++ci) {
    const SNode cs = p.sn[p.children[sn.child_off + ci]];
    const int kc = cs.k, rc = cs.r, nfc = kc + rc, n2 = rc * rc;
    const T* Fc = front + (int64_t)b * p.front_total + cs.front_off + kc * nfc + kc;
    const int* cm = p.cmap + cs.cmap_off;
    // the child's Schur block, eight entries per thread and round: all eight
    // loads in flight before the adds (each parent entry gets one term per child)
    for (int e0 = tid; e0 < n2; e0 += 8 * nt) {
      T v[8];
      int dst[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int e = e0 + q * nt;
        const int i = e < n2 ? e / rc : 0, j = e < n2 ? e - i * rc : 0;
        v[q] = e < n2 ? Fc[i * nfc + j] : T(0);
        dst[q] = e < n2 ? cm[i] * nf + cm[j] : -1;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (dst[q] >= 0) F[dst[q]] += v[q];
    }
    __syncthreads();
  }
  int* pv = piv + (int64_t)b * p.piv_total + sn.piv_off;
  // FP64 / FP32 fronts: panels and block rows in registers (complex ones in LDS or HBM)
  constexpr bool kRegPanel = std::is_same<T, double>::value || std::is_same<T, float>::value;
  // the last panel's trailing update of the columns beyond the next panel,
  // when it is still owed (ub < ue)
  int ub = 0, ue = 0;
  for (int jb = 0; jb < k; jb += NB) {
    const int je = min(jb + NB, k);
    if (kRegPanel && nf - jb <= 128) {
      // wave 0 factors the panel in registers; the other waves meanwhile
      // finish the last panel's trailing update (columns beyond this panel)
      if (tid < 64) {
        if (nf - jb <= 64)
          panel_regs<T, 1>(F, nf, k, jb, je, pv, prow);
        else
          panel_regs<T, 2>(F, nf, k, jb, je, pv, prow);
      } else if (ub < ue) {
        if constexpr (kRegPanel)
          front_update_mfma<T>(F, nf, ub, ue, (tid >> 6) - 1, (nt >> 6) - 1, 1, nf);
      }
      ub = ue = 0;
      __syncthreads();
      // the panel's row swaps in the columns outside it
      for (int c = tid; c < nf; c += nt) {
        if (c >= jb && c < je) continue;
        for (int j = jb; j < je; ++j) {
          const int q = prow[j - jb];
          if (q != j) {
            const T a = F[j * nf + c];
            F[j * nf + c] = F[q * nf + c];
            F[q * nf + c] = a;
          }
        }
      }
      __syncthreads();
    } else if (!LDS && panel_lds) {
      T* P = reinterpret_cast<T*>(smem_d);  // rows jb..nf-1, columns jb..je-1
      const int np = nf - jb, w = je - jb;
      for (int e = tid; e < np * NB; e += nt) {
        const int i = e / NB, q = e - i * NB;
        if (q < w) P[i * PP + q] = F[(jb + i) * nf + jb + q];
      }
      __syncthreads();
      if (tid < 64) {
        for (int j = jb; j < je; ++j) {
          const int jj = j - jb;
          R best = R(-1);
          int bi = j;
          for (int i = j + tid; i < k; i += 64) {
            const R v = pmag(P[(i - jb) * PP + jj]);
            if (v > best) {
              best = v;
              bi = i;
            }
          }
          const R mx = front_wave_max(best);
          const int wl = __ffsll(__ballot(best == mx)) - 1;
          bi = __builtin_amdgcn_readlane(bi, wl);
          if (tid == 0) {
            pv[j] = bi;
            prow[jj] = bi;
          }
          if (bi != j && tid < w) {
            const T a = P[jj * PP + tid];
            P[jj * PP + tid] = P[(bi - jb) * PP + tid];
            P[(bi - jb) * PP + tid] = a;
          }
          __threadfence_block();
          const T dr = T(1) / P[jj * PP + jj];
          for (int i = j + 1 + tid; i < nf; i += 64) {
            T* Pi = P + (i - jb) * PP;
            const T l = Pi[jj] * dr;
            Pi[jj] = l;
            for (int c = jj + 1; c < w; ++c) Pi[c] -= l * P[jj * PP + c];
          }
          __threadfence_block();
        }
      }
      __syncthreads();
      for (int e = tid; e < np * NB; e += nt) {
        const int i = e / NB, q = e - i * NB;
        if (q < w) F[(jb + i) * nf + jb + q] = P[i * PP + q];
      }
      for (int c = tid; c < nf; c += nt) {
        if (c >= jb && c < je) continue;
        for (int j = jb; j < je; ++j) {
          const int q = prow[j - jb];
          if (q != j) {
            const T a = F[j * nf + c];
            F[j * nf + c] = F[q * nf + c];
            F[q * nf + c] = a;
          }
        }
      }
      __syncthreads();
    } else {
    // the panel (its pivot search, full-row swaps, scaling and rank-1 updates
    // of the panel columns) is sequential in j: one wavefront does it without
    // workgroup barriers, the other waves wait once
    if (tid < 64) {
      for (int j = jb; j < je; ++j) {
        R best = R(-1);
        int bi = j;
        for (int i = j + tid; i < k; i += 64) {
          const R v = pmag(F[i * nf + j]);
          if (v > best) {
            best = v;
            bi = i;
          }
        }
        // the lowest lane holding the largest magnitude supplies the pivot row
        const R mx = front_wave_max(best);
        const int wl = __ffsll(__ballot(best == mx)) - 1;
        bi = __builtin_amdgcn_readlane(bi, wl);
        if (tid == 0) pv[j] = bi;
        if (bi != j)
          for (int c = tid; c < nf; c += 64) {
            const T a = F[j * nf + c];
            F[j * nf + c] = F[bi * nf + c];
            F[bi * nf + c] = a;
          }
        __threadfence_block();  // swapped rows visible to every lane of the wave
        const T dr = T(1) / F[j * nf + j];  // reciprocal pivot (LAPACK's getf2 scaling)
        for (int i = j + 1 + tid; i < nf; i += 64) {
          const T l = F[i * nf + j] * dr;
          F[i * nf + j] = l;
          for (int c = j + 1; c < je; ++c) F[i * nf + c] -= l * F[j * nf + c];
        }
        __threadfence_block();
      }
    }
    __syncthreads();
    }
    // U12 = L11^-1 A12 (L11 unit lower triangular), one column per thread,
    // the column's 16 entries in registers during the substitution
    if constexpr (kRegPanel) {
      const int w = je - jb;
      for (int c = je + tid; c < nf; c += nt) {
        T u[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) u[q] = q < w ? F[(jb + q) * nf + c] : T(0);
#pragma unroll
        for (int r = 1; r < NB; ++r) {
          if (r >= w) continue;
          T acc = u[r];
#pragma unroll
          for (int q = 0; q < NB; ++q)
            if (q < r) acc -= F[(jb + r) * nf + jb + q] * u[q];
          u[r] = acc;
        }
#pragma unroll
        for (int q = 1; q < NB; ++q)
          if (q < w) F[(jb + q) * nf + c] = u[q];
      }
    } else {
      for (int c = je + tid; c < nf; c += nt)
        for (int r = jb + 1; r < je; ++r) {
          T acc = F[r * nf + c];
          for (int q = jb; q < r; ++q) acc -= F[r * nf + q] * F[q * nf + c];
          F[r * nf + c] = acc;
        }
    }
    __syncthreads();
    // A22 -= L21 U12 over rows and columns je..nf-1.  Look-ahead when the
    // next panel is a register panel: only the next panel's columns now; the
    // other columns while wave 0 factors that panel (disjoint columns; the
    // new panel's row swaps reach them after both, as after a full update),
    // so the pivot chain and the matrix-core update overlap.  Every entry
    // gets the same operations on the same values either way.
    if constexpr (kRegPanel) {
      const int nw = nt >> 6;
      const bool ahead = je < k && nf - je <= 128 && nw > 1;
      front_update_mfma<T>(F, nf, jb, je, tid >> 6, nw, 0, ahead ? 1 : nf);
      if (ahead) {
        ub = jb;
        ue = je;
      }
      __syncthreads();
      continue;
    }
    const int m = nf - je, mt = (m + 3) / 4;  // complex: 4 x 4 tiles per thread
    for (int tile = tid; tile < mt * mt; tile += nt) {
      const int ti = tile / mt, tj = tile - ti * mt;
      const int i0 = je + 4 * ti, c0 = je + 4 * tj;
      T acc[4][4];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[a][e] = T(0);
      for (int q = jb; q < je; ++q) {
        T lv[4], uv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) lv[a] = i0 + a < nf ? F[(i0 + a) * nf + q] : T(0);
#pragma unroll
        for (int e = 0; e < 4; ++e) uv[e] = c0 + e < nf ? F[q * nf + c0 + e] : T(0);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int e = 0; e < 4; ++e) acc[a][e] += lv[a] * uv[e];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (i0 + a < nf && c0 + e < nf) F[(i0 + a) * nf + c0 + e] -= acc[a][e];
    }
    __syncthreads();
  }
  if (LDS)
    for (int e = tid; e < nf2; e += nt) Fg[e] = F[e];
}

// The children of the front of an L-lane segment: lane c of the segment
// loads child c's record (one round of independent loads for all children),
// and the loop over children reads child ci's fields from lane ci -- instead
// of two dependent global loads (list entry, then record) per child.  Child
// ci >= L is read directly.
struct ChildRec {
  int k = 0, r = 0, cmap_off = 0, uvec_off = 0;
  int64_t front_off = 0;
};

__device__ __forceinline__ ChildRec child_rec(const SPlanDev& p, const SNode& sn, int n_child,
                                              int c) {
  ChildRec m;
  if (c < n_child) {
    const SNode& s = p.sn[p.children[sn.child_off + c]];
    m.k = s.k;
    m.r = s.r;
    m.cmap_off = s.cmap_off;
    m.uvec_off = s.uvec_off;
    m.front_off = s.front_off;
  }
  return m;
}

template <int L>
__device__ __forceinline__ ChildRec child_at(const SPlanDev& p, const SNode& sn,
                                             const ChildRec& mine, int ci) {
  if (ci >= L) return child_rec(p, sn, ci + 1, ci);
  ChildRec m;
  m.k = __shfl(mine.k, ci, L);
  m.r = __shfl(mine.r, ci, L);
  m.cmap_off = __shfl(mine.cmap_off, ci, L);
  m.uvec_off = __shfl(mine.uvec_off, ci, L);
  const int lo = __shfl((int)(uint32_t)mine.front_off, ci, L);
  const int hi = __shfl((int)(mine.front_off >> 32), ci, L);
  m.front_off = (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
  return m;
}

// Small fronts (k + r <= L, L = 16, 32 or 64): one wavefront factors 64/L
// fronts, one per L-lane segment, and lane c of a segment holds column c of
// its front in registers.  The pivot search of step j is then local to lane j
// (its own column), the row swap is a register swap in every lane, and the
// rank-1 update needs only the multipliers of column j, broadcast from lane j
// inside the segment: no barriers, no LDS traffic during the elimination.
// LDS serves only to assemble the fronts (A entries, then the children's
// Schur complements in child order) before each lane loads its column.  The
// result has exactly the layout of k_front_factor: LAPACK-style swaps in piv,
// L and U in the pivot columns and rows, the Schur complement in the trailing
// block.
//
// FWD (FP64 fronts of fewer than L unknowns): the forward sweep of the solve
// rides along as one more column.  Lane nf holds the front's right-hand side
// -- x at its pivots, then the children's border vectors added child by child,
// as k_front_fwd_w gathers it -- and takes every row swap and every update of
// the elimination like the other columns: that is the same sequence of
// operations on the same values as P b followed by L^-1, so the y (to x) and
// border vector (to uvec) it writes are those of k_front_fwd_w, bit for bit.
// Ordering of a segment's LDS work: a workgroup barrier where the workgroup is
// one wavefront, else only a compiler barrier -- a segment's LDS belongs to its
// own wavefront, whose LDS operations complete in program order.
template <bool WAVE>
__device__ __forceinline__ void seg_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  } else {
    __syncthreads();
  }
}

// The factorization of supernode s (live: the segment holds a real front) by
// the L lanes of one segment (lane c), in the segment's LDS F (pitch L).
template <typename T, int L, bool FWD, bool WAVE>
__device__ __forceinline__ void factor_front_w(const SPlanDev& p, int s, bool live, int b,
                                               const typename Num<T>::D* __restrict__ data,
                                               T* __restrict__ front, int* __restrict__ piv,
                                               typename Num<T>::A* __restrict__ x,
                                               typename Num<T>::A* __restrict__ uvec, T* F,
                                               int c) {
  using R = typename Num<T>::R;
  static_assert(!FWD || std::is_same<T, typename Num<T>::A>::value,
                "the forward sweep rides along only where factors and solves share a type");
  const SNode sn = p.sn[s];
  const int k = live ? sn.k : 0, nf = live ? sn.k + sn.r : 0;
  for (int e = c; e < L * L; e += L) F[e] = T(0);
  seg_sync<WAVE>();
  const typename Num<T>::D* db = data + (int64_t)b * p.nnz;
  for (int e = c; e < (live ? sn.n_aent : 0); e += L) {
    const int pos = p.apos[sn.aent_off + e];
    const int i = pos / nf, j = pos - i * nf;
    F[i * L + j] += cvt<T>(db[p.aslot[sn.aent_off + e]]);
  }
  if constexpr (FWD) {  // right-hand side column: x at the pivots
    if (live && c < k) F[c * L + nf] = x[(int64_t)b * p.S + p.idx[sn.idx_off + c]];
  }
  seg_sync<WAVE>();
  const int nch = live ? sn.n_child : 0;
  const ChildRec mine = child_rec(p, sn, nch, c);
  for (int ci = 0; ci < nch; ++ci) {
    const ChildRec cs = child_at<L>(p, sn, mine, ci);
    const int kc = cs.k, rc = cs.r, nfc = kc + rc;
    const T* Fc = front + (int64_t)b * p.front_total + cs.front_off;
    const int* cm = p.cmap + cs.cmap_off;
    if constexpr (FWD) {  // the child's border vector into the right-hand side
      if (c < rc) F[cm[c] * L + nf] += uvec[(int64_t)b * p.r_total + cs.uvec_off + c];
    }
    // lane c adds column c of the child's Schur block (rc <= nf <= L), eight
    // rows per round so that eight global loads are in flight per lane
    const bool col = c < rc;
    const int dst = col ? cm[c] : 0;
    const T* src = Fc + kc * nfc + kc + (col ? c : 0);
    for (int i0 = 0; i0 < rc; i0 += 8) {
      T v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = (col && i0 + q < rc) ? src[(i0 + q) * nfc] : T(0);
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (col && i0 + q < rc) F[cm[i0 + q] * L + dst] += v[q];
    }
    seg_sync<WAVE>();  // orders the LDS sums child by child
  }
  seg_sync<WAVE>();
  T a[L];
#pragma unroll
  for (int i = 0; i < L; ++i) a[i] = F[i * L + c];
  int* pv = piv + (int64_t)b * p.piv_total + sn.piv_off;
#pragma unroll
  for (int j = 0; j < L; ++j) {
    if (!__any(j < k)) break;
    const bool act = j < k;
    // pivot search in column j (lane j's registers) among rows j..k-1
    R best = R(-1);
    int bi = j;
#pragma unroll
    for (int i = j; i < L; ++i) {
      const R v = pmag(a[i]);
      if (i < k && v > best) {
        best = v;
        bi = i;
      }
    }
    const int pr = __shfl(bi, j, L);
    // swap rows j and pr in every column of the segment
#pragma unroll
    for (int i = j + 1; i < L; ++i) {
      const bool sw = act && i == pr;
      const T t = a[i];
      a[i] = sw ? a[j] : a[i];
      a[j] = sw ? t : a[j];
    }
    if (act && c == 0) pv[j] = pr;
    const T rinv = T(1) / shfl(a[j], j, L);
    const T uj = a[j];
#pragma unroll
    for (int i = j + 1; i < L; ++i) {
      const T l = shfl(a[i], j, L) * rinv;
      if (act && c == j) a[i] = l;
      if (act && c > j) a[i] -= l * uj;
    }
  }
  if constexpr (FWD) {
    if (live && c == nf) {
      const int* id = p.idx + sn.idx_off;
#pragma unroll
      for (int i = 0; i < L; ++i) {
        if (i < k)
          x[(int64_t)b * p.S + id[i]] = a[i];
        else if (i < nf)
          uvec[(int64_t)b * p.r_total + sn.uvec_off + i - k] = a[i];
      }
    }
  }
  if (!live || c >= nf) return;
  T* Fg = front + (int64_t)b * p.front_total + sn.front_off;
#pragma unroll
  for (int i = 0; i < L; ++i)
    if (i < nf) Fg[i * nf + c] = a[i];
}

template <typename T, int L, bool FWD>
__global__ void __launch_bounds__(64)
k_front_factor_w(SPlanDev p, const int* __restrict__ list, int count,
                 const typename Num<T>::D* __restrict__ data, T* __restrict__ front,
                 int* __restrict__ piv, typename Num<T>::A* __restrict__ x,
                 typename Num<T>::A* __restrict__ uvec) {
  extern __shared__ double smem_d[];
  const int lane = threadIdx.x, seg = lane / L, c = lane % L;
  const int fi = blockIdx.x * (64 / L) + seg;
  const bool live = fi < count;
  factor_front_w<T, L, FWD, false>(p, list[live ? fi : count - 1], live, blockIdx.y, data, front,
                                   piv, x, uvec, reinterpret_cast<T*>(smem_d) + seg * L * L, c);
}

// ---------------------------------------------------------------- solves
// Forward sweep of A x = b: y = L^-1 P b, border contributions to parents.
template <typename T>
__global__ void __launch_bounds__(128)
k_front_fwd(SPlanDev p, int lbeg, const T* __restrict__ front, const int* __restrict__ piv,
            typename Num<T>::A* __restrict__ x, typename Num<T>::A* __restrict__ uvec) {
  using A = typename Num<T>::A;
  const int s = p.level_sn[lbeg + blockIdx.x];
  const int b = blockIdx.y;
  const SNode sn = p.sn[s];
  const int k = sn.k, nf = sn.k + sn.r;
  const T* F = front + (int64_t)b * p.front_total + sn.front_off;
  extern __shared__ double smem_v[];
  A* v = reinterpret_cast<A*>(smem_v);
  const int* id = p.idx + sn.idx_off;
  A* xb = x + (int64_t)b * p.S;
  A* ub = uvec + (int64_t)b * p.r_total;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int j = tid; j < nf; j += nt) v[j] = j < k ? xb[id[j]] : A(0);
  __syncthreads();
  for (int ci = 0; ci < sn.n_child; ++ci) {
    const SNode cs = p.sn[p.children[sn.child_off + ci]];
    const int* cm = p.cmap + cs.cmap_off;
    for (int i = tid; i < cs.r; i += nt) v[cm[i]] += ub[cs.uvec_off + i];
    __syncthreads();
  }
  if (tid == 0) {
    const int* pv = piv + (int64_t)b * p.piv_total + sn.piv_off;
    for (int j = 0; j < k; ++j) {
      const int q = pv[j];
      const A a = v[j];
      v[j] = v[q];
      v[q] = a;
    }
  }
  __syncthreads();
  // blocks of 64 pivots: the block's unit lower triangle by wavefront 0
  // (entry t in lane t, each step's value read from its lane), then the rows
  // below by all threads; per entry the same subtractions in the same order
  for (int j0 = 0; j0 < k; j0 += 64) {
    const int jw = min(64, k - j0);
    if (tid < 64) {
      A vt = tid < jw ? v[j0 + tid] : A(0);
      const T* Fr = F + (int64_t)(j0 + min(tid, jw - 1)) * nf + j0;
      for (int j = 0; j < jw; ++j) {
        const A vj = front_readlane(vt, j);
        if (tid > j && tid < jw) vt -= cvt<A>(Fr[j]) * vj;
      }
      if (tid < jw) v[j0 + tid] = vt;
    }
    __syncthreads();
    for (int i = j0 + jw + tid; i < nf; i += nt) {
      A acc = v[i];
      for (int j = j0; j < j0 + jw; ++j) acc -= cvt<A>(F[i * nf + j]) * v[j];
      v[i] = acc;
    }
    __syncthreads();
  }
  for (int j = tid; j < k; j += nt) xb[id[j]] = v[j];
  for (int i = tid; i < sn.r; i += nt) ub[sn.uvec_off + i] = v[k + i];
}

// Backward sweep of A x = b: x_K = U11^-1 (y_K - U12 x_R).
template <typename T>
__global__ void __launch_bounds__(128)
k_front_bwd(SPlanDev p, int lbeg, const T* __restrict__ front, typename Num<T>::A* __restrict__ x) {
  using A = typename Num<T>::A;
  const int s = p.level_sn[lbeg + blockIdx.x];
  const int b = blockIdx.y;
  const SNode sn = p.sn[s];
  const int k = sn.k, nf = sn.k + sn.r;
  const T* F = front + (int64_t)b * p.front_total + sn.front_off;
  extern __shared__ double smem_v[];
  A* v = reinterpret_cast<A*>(smem_v);
  const int* id = p.idx + sn.idx_off;
  A* xb = x + (int64_t)b * p.S;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int j = tid; j < nf; j += nt) v[j] = xb[id[j]];
  __syncthreads();
  for (int j = tid; j < k; j += nt) {
    A acc = A(0);
    for (int i = k; i < nf; ++i) acc += cvt<A>(F[j * nf + i]) * v[i];
    v[j] -= acc;
  }
  __syncthreads();
  // blocks of 64 pivots from the last: the block's upper triangle by
  // wavefront 0, then the rows above by all threads (same order per entry)
  for (int jhi = k; jhi > 0; jhi -= 64) {
    const int jlo = max(0, jhi - 64), jw = jhi - jlo;
    if (tid < 64) {
      A vt = tid < jw ? v[jlo + tid] : A(0);
      const T* Fr = F + (int64_t)(jlo + min(tid, jw - 1)) * nf + jlo;
      for (int jj = jw - 1; jj >= 0; --jj) {
        if (tid == jj) vt /= cvt<A>(Fr[jj]);
        const A vj = front_readlane(vt, jj);
        if (tid < jj) vt -= cvt<A>(Fr[jj]) * vj;
      }
      if (tid < jw) v[jlo + tid] = vt;
    }
    __syncthreads();
    for (int i = tid; i < jlo; i += nt) {
      A acc = v[i];
      for (int j = jhi - 1; j >= jlo; --j) acc -= cvt<A>(F[i * nf + j]) * v[j];
      v[i] = acc;
    }
    __syncthreads();
  }
  for (int j = tid; j < k; j += nt) xb[id[j]] = v[j];
}

// Forward sweep of A^T x = b: U^T w = b (U11^T lower), border via U12^T.
template <typename T>
__global__ void __launch_bounds__(128)
k_front_fwd_t(SPlanDev p, int lbeg, const T* __restrict__ front, typename Num<T>::A* __restrict__ x,
              typename Num<T>::A* __restrict__ uvec) {
  using A = typename Num<T>::A;
  const int s = p.level_sn[lbeg + blockIdx.x];
  const int b = blockIdx.y;
  const SNode sn = p.sn[s];
  const int k = sn.k, nf = sn.k + sn.r;
  const T* F = front + (int64_t)b * p.front_total + sn.front_off;
  extern __shared__ double smem_v[];
  A* v = reinterpret_cast<A*>(smem_v);
  const int* id = p.idx + sn.idx_off;
  A* xb = x + (int64_t)b * p.S;
  A* ub = uvec + (int64_t)b * p.r_total;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int j = tid; j < nf; j += nt) v[j] = j < k ? xb[id[j]] : A(0);
  __syncthreads();
  for (int ci = 0; ci < sn.n_child; ++ci) {
    const SNode cs = p.sn[p.children[sn.child_off + ci]];
    const int* cm = p.cmap + cs.cmap_off;
    for (int i = tid; i < cs.r; i += nt) v[cm[i]] += ub[cs.uvec_off + i];
    __syncthreads();
  }
  // blocks of 64 pivots: U11^T's block by wavefront 0, then the entries
  // below by all threads (same order per entry)
  for (int j0 = 0; j0 < k; j0 += 64) {
    const int jw = min(64, k - j0);
    if (tid < 64) {
      A vt = tid < jw ? v[j0 + tid] : A(0);
      const int c = j0 + min(tid, jw - 1);
      for (int jj = 0; jj < jw; ++jj) {
        const int j = j0 + jj;
        if (tid == jj) vt /= cvt<A>(F[j * nf + j]);
        const A vj = front_readlane(vt, jj);
        if (tid > jj && tid < jw) vt -= cvt<A>(F[j * nf + c]) * vj;
      }
      if (tid < jw) v[j0 + tid] = vt;
    }
    __syncthreads();
    for (int i = j0 + jw + tid; i < nf; i += nt) {
      A acc = v[i];
      for (int j = j0; j < j0 + jw; ++j) acc -= cvt<A>(F[j * nf + i]) * v[j];
      v[i] = acc;
    }
    __syncthreads();
  }
  for (int j = tid; j < k; j += nt) xb[id[j]] = v[j];
  for (int i = tid; i < sn.r; i += nt) ub[sn.uvec_off + i] = v[k + i];
}

// Backward sweep of A^T x = b: x_K = P^T L11^-T (w_K - L21^T x_R), with x_R
// the final values of the ancestors' unknowns.
template <typename T>
__global__ void __launch_bounds__(128)
k_front_bwd_t(SPlanDev p, int lbeg, const T* __restrict__ front, const int* __restrict__ piv,
              typename Num<T>::A* __restrict__ x) {
  using A = typename Num<T>::A;
  const int s = p.level_sn[lbeg + blockIdx.x];
  const int b = blockIdx.y;
  const SNode sn = p.sn[s];
  const int k = sn.k, nf = sn.k + sn.r;
  const T* F = front + (int64_t)b * p.front_total + sn.front_off;
  extern __shared__ double smem_v[];
  A* v = reinterpret_cast<A*>(smem_v);
  const int* id = p.idx + sn.idx_off;
  A* xb = x + (int64_t)b * p.S;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int j = tid; j < nf; j += nt) v[j] = xb[id[j]];
  __syncthreads();
  for (int j = tid; j < k; j += nt) {
    A acc = A(0);
    for (int i = k; i < nf; ++i) acc += cvt<A>(F[i * nf + j]) * v[i];
    v[j] -= acc;
  }
  __syncthreads();
  // blocks of 64 pivots from the last: L11^T's block by wavefront 0, then
  // the entries above by all threads (same order per entry)
  for (int jhi = k; jhi > 0; jhi -= 64) {
    const int jlo = max(0, jhi - 64), jw = jhi - jlo;
    if (tid < 64) {
      A vt = tid < jw ? v[jlo + tid] : A(0);
      const int c = jlo + min(tid, jw - 1);
      for (int jj = jw - 1; jj >= 0; --jj) {
        const A vj = front_readlane(vt, jj);
        if (tid < jj) vt -= cvt<A>(F[(jlo + jj) * nf + c]) * vj;
      }
      if (tid < jw) v[jlo + tid] = vt;
    }
    __syncthreads();
    for (int i = tid; i < jlo; i += nt) {
      A acc = v[i];
      for (int j = jhi - 1; j >= jlo; --j) acc -= cvt<A>(F[j * nf + i]) * v[j];
      v[i] = acc;
    }
    __syncthreads();
  }
  if (tid == 0) {
    const int* pv = piv + (int64_t)b * p.piv_total + sn.piv_off;
    for (int j = k - 1; j >= 0; --j) {
      const int q = pv[j];
      const A a = v[j];
      v[j] = v[q];
      v[q] = a;
    }
  }
  __syncthreads();
  for (int j = tid; j < k; j += nt) xb[id[j]] = v[j];
}

// ------------------------------------------------- register-front solves
// Fronts of at most L unknowns: one L-lane segment per front, 64/L fronts per
// wavefront, lane i holding entry i of the front's vector in a register.
// Each elimination step broadcasts one entry inside the segment; the
// arithmetic and its order are those of the block kernels above, so both
// give the same bits.  LDS (L doubles per front) only gathers the children's
// border contributions and applies the row swaps.
template <typename A>
struct WaveFront {
  SNode sn;
  int k, nf, i;
  bool live;
  A* V;  // LDS vector of the front
};

template <int L, typename A>
__device__ __forceinline__ WaveFront<A> wave_front(const SPlanDev& p, const int* list,
                                                   int count) {
  extern __shared__ double smem_d[];
  WaveFront<A> w;
  const int lane = threadIdx.x, seg = lane / L;
  const int fi = blockIdx.x * (64 / L) + seg;
  w.live = fi < count;
  w.sn = p.sn[list[w.live ? fi : count - 1]];
  w.k = w.live ? w.sn.k : 0;
  w.nf = w.live ? w.sn.k + w.sn.r : 0;
  w.i = lane % L;
  w.V = reinterpret_cast<A*>(smem_d) + seg * L;
  return w;
}

// V[cm[q]] += border vector of each child (child order, as k_front_fwd).
template <int L, typename A>
__device__ __forceinline__ void wave_gather_children(const SPlanDev& p, const WaveFront<A>& w,
                                                     const A* ub) {
  const int nch = w.live ? w.sn.n_child : 0;
  const ChildRec mine = child_rec(p, w.sn, nch, w.i);
  for (int ci = 0; ci < nch; ++ci) {
    const ChildRec cs = child_at<L>(p, w.sn, mine, ci);
    if (w.i < cs.r) w.V[p.cmap[cs.cmap_off + w.i]] += ub[cs.uvec_off + w.i];
    __syncthreads();
  }
}

// The forward sweep of supernode sn by the L lanes of one segment (lane i),
// V: the segment's LDS vector of L entries.
template <typename T, int L, bool WAVE>
__device__ __forceinline__ void fwd_front_w(const SPlanDev& p, int s, bool live, int b,
                                            const T* __restrict__ front,
                                            const int* __restrict__ piv,
                                            typename Num<T>::A* __restrict__ x,
                                            typename Num<T>::A* __restrict__ uvec,
                                            typename Num<T>::A* V, int i) {
  using A = typename Num<T>::A;
  const SNode sn = p.sn[s];
  const int k = live ? sn.k : 0, nf = live ? sn.k + sn.r : 0;
  const T* F = front + (int64_t)b * p.front_total + sn.front_off;
  const int* id = p.idx + sn.idx_off;
  A* xb = x + (int64_t)b * p.S;
  A* ub = uvec + (int64_t)b * p.r_total;
  V[i] = i < k ? xb[id[i]] : A(0);
  seg_sync<WAVE>();
  // V[cm[q]] += border vector of each child (child order, as k_front_fwd)
  const int nch = live ? sn.n_child : 0;
  const ChildRec mine = child_rec(p, sn, nch, i);
  for (int ci = 0; ci < nch; ++ci) {
    const ChildRec cs = child_at<L>(p, sn, mine, ci);
    if (i < cs.r) V[p.cmap[cs.cmap_off + i]] += ub[cs.uvec_off + i];
    seg_sync<WAVE>();
  }
  if (i == 0 && live) {
    const int* pv = piv + (int64_t)b * p.piv_total + sn.piv_off;
    for (int j = 0; j < k; ++j) {
      const int q = pv[j];
      const A a = V[j];
      V[j] = V[q];
      V[q] = a;
    }
  }
  seg_sync<WAVE>();
  A v = V[i];
#pragma unroll
  for (int j = 0; j < L; ++j) {
    if (!__any(j < k)) break;
    const A vj = shfl(v, j, L);
    if (j < k && j < i && i < nf) v -= cvt<A>(F[i * nf + j]) * vj;
  }
  if (i < k)
    xb[id[i]] = v;
  else if (i < nf)
    ub[sn.uvec_off + i - k] = v;
}

template <typename T, int L>
__global__ void __launch_bounds__(64)
k_front_fwd_w(SPlanDev p, const int* __restrict__ list, int count, const T* __restrict__ front,
              const int* __restrict__ piv, typename Num<T>::A* __restrict__ x,
              typename Num<T>::A* __restrict__ uvec) {
  extern __shared__ double smem_d[];
  const int lane = threadIdx.x, seg = lane / L;
  const int fi = blockIdx.x * (64 / L) + seg;
  const bool live = fi < count;
  fwd_front_w<T, L, false>(p, list[live ? fi : count - 1], live, blockIdx.y, front, piv, x,
                           uvec, reinterpret_cast<typename Num<T>::A*>(smem_d) + seg * L,
                           lane % L);
}

// The backward sweep of supernode s by the L lanes of one segment (lane i).
template <typename T, int L>
__device__ __forceinline__ void bwd_front_w(const SPlanDev& p, int s, bool live, int b,
                                            const T* __restrict__ front,
                                            typename Num<T>::A* __restrict__ x, int i) {
  using A = typename Num<T>::A;
  const SNode sn = p.sn[s];
  const int k = live ? sn.k : 0, nf = live ? sn.k + sn.r : 0;
  const T* F = front + (int64_t)b * p.front_total + sn.front_off;
  const int* id = p.idx + sn.idx_off;
  A* xb = x + (int64_t)b * p.S;
  A v = i < nf ? xb[id[i]] : A(0);
  A acc = A(0);
#pragma unroll
  for (int c = 0; c < L; ++c) {
    const A xc = shfl(v, c, L);
    if (c >= k && c < nf && i < k) acc += cvt<A>(F[i * nf + c]) * xc;
  }
  v -= acc;
#pragma unroll
  for (int j = L - 1; j >= 0; --j) {
    if (j < k && i == j) v /= cvt<A>(F[j * nf + j]);
    const A vj = shfl(v, j, L);
    if (j < k && i < j) v -= cvt<A>(F[i * nf + j]) * vj;
  }
  if (i < k) xb[id[i]] = v;
}

template <typename T, int L>
__global__ void __launch_bounds__(64)
k_front_bwd_w(SPlanDev p, const int* __restrict__ list, int count, const T* __restrict__ front,
              typename Num<T>::A* __restrict__ x) {
  const int lane = threadIdx.x;
  const int fi = blockIdx.x * (64 / L) + lane / L;
  const bool live = fi < count;
  bwd_front_w<T, L>(p, list[live ? fi : count - 1], live, blockIdx.y, front, x, lane % L);
}

template <typename T, int L>
__global__ void __launch_bounds__(64)
k_front_fwd_t_w(SPlanDev p, const int* __restrict__ list, int count,
                const T* __restrict__ front, typename Num<T>::A* __restrict__ x, typename Num<T>::A* __restrict__ uvec) {
  using A = typename Num<T>::A;
  const WaveFront<A> w = wave_front<L, A>(p, list, count);
  const int b = blockIdx.y, i = w.i, k = w.k, nf = w.nf;
  const T* F = front + (int64_t)b * p.front_total + w.sn.front_off;
  const int* id = p.idx + w.sn.idx_off;
  A* xb = x + (int64_t)b * p.S;
  A* ub = uvec + (int64_t)b * p.r_total;
  w.V[i] = i < k ? xb[id[i]] : A(0);
  __syncthreads();
  wave_gather_children<L>(p, w, ub);
  A v = w.V[i];
#pragma unroll
  for (int j = 0; j < L; ++j) {
    if (!__any(j < k)) break;
    if (j < k && i == j) v /= cvt<A>(F[j * nf + j]);
    const A vj = shfl(v, j, L);
    if (j < k && j < i && i < nf) v -= cvt<A>(F[j * nf + i]) * vj;
  }
  if (i < k)
    xb[id[i]] = v;
  else if (i < nf)
    ub[w.sn.uvec_off + i - k] = v;
}

template <typename T, int L>
__global__ void __launch_bounds__(64)
k_front_bwd_t_w(SPlanDev p, const int* __restrict__ list, int count,
                const T* __restrict__ front, const int* __restrict__ piv,
                typename Num<T>::A* __restrict__ x) {
  using A = typename Num<T>::A;
  const WaveFront<A> w = wave_front<L, A>(p, list, count);
  const int b = blockIdx.y, i = w.i, k = w.k, nf = w.nf;
  const T* F = front + (int64_t)b * p.front_total + w.sn.front_off;
  const int* id = p.idx + w.sn.idx_off;
  A* xb = x + (int64_t)b * p.S;
  A v = i < nf ? xb[id[i]] : A(0);
  A acc = A(0);
#pragma unroll
  for (int c = 0; c < L; ++c) {
    const A xc = shfl(v, c, L);
    if (c >= k && c < nf && i < k) acc += cvt<A>(F[c * nf + i]) * xc;
  }
  v -= acc;
#pragma unroll
  for (int j = L - 1; j >= 0; --j) {
    const A vj = shfl(v, j, L);
    if (j < k && i < j) v -= cvt<A>(F[j * nf + i]) * vj;
  }
  w.V[i] = v;
  __syncthreads();
  if (i == 0 && w.live) {
    const int* pv = piv + (int64_t)b * p.piv_total + w.sn.piv_off;
    for (int j = k - 1; j >= 0; --j) {
      const int q = pv[j];
      const A a = w.V[j];
      w.V[j] = w.V[q];
      w.V[q] = a;
    }
  }
  __syncthreads();
  if (i < k) xb[id[i]] = w.V[i];
}

// ------------------------------------------------------------ launchers
template <typename T, int L>
void launch_factor_w(const SPlanDev& d, int beg, int cnt, int B, const typename Num<T>::D* data,
                     T* front, int* piv, hipStream_t st) {
  if (cnt <= 0 || B <= 0) return;
  constexpr int FPW = 64 / L;
  hipLaunchKernelGGL((k_front_factor_w<T, L, false>), dim3((cnt + FPW - 1) / FPW, B), dim3(64),
                     (size_t)64 * L * sizeof(T), st, d, d.level_sn + beg, cnt, data, front, piv,
                     nullptr, nullptr);
}

// The top levels of the elimination tree in one launch.  Above the wide
// bottom levels, the levels of a batch hold a few small fronts per instance
// (the BSIM-sized ring: 22, 13, 8, 3, 2, 1, 1, 1 fronts of at most 9
// unknowns), and a launch per level costs more than its work.  Here one
// workgroup per instance (blockIdx.y) walks those levels -- all of one
// register class L -- with 256 / L segments, level after level, a workgroup
// barrier in between (a level's fronts, border vectors and x entries are
// written to global memory by some of its wavefronts and read by others; all
// of them on one compute unit).  Each front gets the per-level kernels' code,
// so the same operations on the same values.
constexpr int kTopLevels = 32;
struct TopLevels {
  int n;                     // levels
  int beg[kTopLevels + 1];   // level li: level_sn entries [beg[li], beg[li + 1])
};

// Threads per workgroup of the merged launches: 16 segments of 16 lanes or 4
// of 32 (32 KB of LDS either way).
constexpr int top_threads(int L) { return L == 16 ? 256 : 128; }

template <typename T, int L, bool FWD>
__global__ void __launch_bounds__(top_threads(L))
k_front_factor_top(SPlanDev p, TopLevels tl, const typename Num<T>::D* __restrict__ data,
                   T* __restrict__ front, int* __restrict__ piv,
                   typename Num<T>::A* __restrict__ x, typename Num<T>::A* __restrict__ uvec) {
  constexpr int SEGS = top_threads(L) / L;
  extern __shared__ double smem_d[];
  const int seg = threadIdx.x / L, c = threadIdx.x % L, b = blockIdx.y;
  T* F = reinterpret_cast<T*>(smem_d) + seg * L * L;
  for (int li = 0; li < tl.n; ++li) {
    const int beg = tl.beg[li], cnt = tl.beg[li + 1] - beg;
    for (int f0 = 0; f0 < cnt; f0 += SEGS) {
      const int fi = f0 + seg;
      const bool live = fi < cnt;
      factor_front_w<T, L, FWD, true>(p, p.level_sn[beg + (live ? fi : cnt - 1)], live, b, data,
                                      front, piv, x, uvec, F, c);
    }
    __syncthreads();
  }
}

template <typename T, int L>
__global__ void __launch_bounds__(top_threads(L))
k_front_fwd_top(SPlanDev p, TopLevels tl, const T* __restrict__ front,
                const int* __restrict__ piv, typename Num<T>::A* __restrict__ x,
                typename Num<T>::A* __restrict__ uvec) {
  constexpr int SEGS = top_threads(L) / L;
  extern __shared__ double smem_d[];
  const int seg = threadIdx.x / L, i = threadIdx.x % L, b = blockIdx.y;
  typename Num<T>::A* V = reinterpret_cast<typename Num<T>::A*>(smem_d) + seg * L;
  for (int li = 0; li < tl.n; ++li) {
    const int beg = tl.beg[li], cnt = tl.beg[li + 1] - beg;
    for (int f0 = 0; f0 < cnt; f0 += SEGS) {
      const int fi = f0 + seg;
      const bool live = fi < cnt;
      fwd_front_w<T, L, true>(p, p.level_sn[beg + (live ? fi : cnt - 1)], live, b, front, piv, x,
                              uvec, V, i);
    }
    __syncthreads();
  }
}

template <typename T, int L>
__global__ void __launch_bounds__(top_threads(L))
k_front_bwd_top(SPlanDev p, TopLevels tl, const T* __restrict__ front,
                typename Num<T>::A* __restrict__ x) {
  constexpr int SEGS = top_threads(L) / L;
  const int seg = threadIdx.x / L, i = threadIdx.x % L, b = blockIdx.y;
  for (int li = tl.n - 1; li >= 0; --li) {
    const int beg = tl.beg[li], cnt = tl.beg[li + 1] - beg;
    for (int f0 = 0; f0 < cnt; f0 += SEGS) {
      const int fi = f0 + seg;
      const bool live = fi < cnt;
      bwd_front_w<T, L>(p, p.level_sn[beg + (live ? fi : cnt - 1)], live, b, front, x, i);
    }
    __syncthreads();
  }
}

// Real factors (FP64, FP32) take the merged launches; complex ones keep the
// level kernels (their segments would need twice the LDS).
template <typename T>
constexpr bool kTopMerge = std::is_same<T, double>::value || std::is_same<T, float>::value;

// The levels merged into one launch: from the highest level down, the levels
// whose fronts are all of one register class L = 16 or 32 (each with a free
// lane for a right-hand side column when `free_lane`), at most two passes of
// the launch's segments each; level 0 never, and a single level is left to
// the level kernels.  Returns the first merged level (the level count if
// none); fills tl and L.
int top_levels(const SPlanHost& h, bool free_lane, TopLevels& tl, int& L) {
  const int nlev = (int)h.level_ptr.size() - 1;
  int l0 = nlev;
  L = 0;
  for (int l = nlev - 1; l >= 1 && nlev - l <= kTopLevels; --l) {
    const int* cls = &h.level_cls[5 * l];
    int which = -1, classes = 0;
    for (int c = 0; c < 4; ++c)
      if (cls[c + 1] > cls[c]) {
        which = c;
        ++classes;
      }
    if (classes != 1 || which < 2) break;
    const int Ll = which == 2 ? 32 : 16;
    const int cnt = cls[which + 1] - cls[which];
    if ((L && Ll != L) || cnt > 2 * (top_threads(Ll) / Ll)) break;
    bool fits = true;
    for (int i = cls[which]; fits && i < cls[which + 1]; ++i)
      fits = !free_lane || h.sn[h.level_sn[i]].k + h.sn[h.level_sn[i]].r < Ll;
    if (!fits) break;
    L = Ll;
    l0 = l;
  }
  if (nlev - l0 < 2) return nlev;
  tl.n = nlev - l0;
  for (int li = 0; li <= tl.n; ++li) tl.beg[li] = h.level_ptr[l0 + li];
  return l0;
}

template <typename T, bool FWD>
void launch_factor_top(const SPlanDev& d, const TopLevels& tl, int L, int B,
                       const typename Num<T>::D* data, T* front, int* piv,
                       typename Num<T>::A* x, typename Num<T>::A* uvec, hipStream_t st) {
  const dim3 grid(1, B), block(top_threads(L));
  const size_t lds = (size_t)top_threads(L) * L * sizeof(T);
  if (L == 16)
    hipLaunchKernelGGL((k_front_factor_top<T, 16, FWD>), grid, block, lds, st, d, tl, data,
                       front, piv, x, uvec);
  else
    hipLaunchKernelGGL((k_front_factor_top<T, 32, FWD>), grid, block, lds, st, d, tl, data,
                       front, piv, x, uvec);
}

template <typename T>
void launch_sweep_top(const SPlanDev& d, const TopLevels& tl, int L, int B, const T* front,
                      const int* piv, typename Num<T>::A* x, typename Num<T>::A* uvec, bool fwd,
                      hipStream_t st) {
  const dim3 grid(1, B), block(top_threads(L));
  const size_t lds = fwd ? (size_t)top_threads(L) * sizeof(typename Num<T>::A) : 0;
  if (fwd && L == 16)
    hipLaunchKernelGGL((k_front_fwd_top<T, 16>), grid, block, lds, st, d, tl, front, piv, x,
                       uvec);
  else if (fwd)
    hipLaunchKernelGGL((k_front_fwd_top<T, 32>), grid, block, lds, st, d, tl, front, piv, x,
                       uvec);
  else if (L == 16)
    hipLaunchKernelGGL((k_front_bwd_top<T, 16>), grid, block, 0, st, d, tl, front, x);
  else
    hipLaunchKernelGGL((k_front_bwd_top<T, 32>), grid, block, 0, st, d, tl, front, x);
}

// The same fronts factored with the forward sweep of x riding along (FP64).
template <int L>
void launch_factor_fwd_w(const SPlanDev& d, int beg, int cnt, int B, const double* data,
                         double* front, int* piv, double* x, double* uvec, hipStream_t st) {
  if (cnt <= 0 || B <= 0) return;
  constexpr int FPW = 64 / L;
  hipLaunchKernelGGL((k_front_factor_w<double, L, true>), dim3((cnt + FPW - 1) / FPW, B),
                     dim3(64), (size_t)64 * L * sizeof(double), st, d, d.level_sn + beg, cnt,
                     data, front, piv, x, uvec);
}

// Dynamic LDS the LDS variant of k_front_factor may use per workgroup: what
// the device grants a workgroup on request (up to kFrontLdsBytes, the whole
// 160 KB of a CU), else the default 64 KB.  Queried once per process and
// element type; a refused request leaves no error state behind.
template <typename T>
size_t front_lds_limit() {
  static const size_t limit = [] {
    size_t lim = 64 * 1024;
    int dev = 0, optin = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) !=
            hipSuccess) {
      (void)hipGetLastError();
      return lim;
    }
    const size_t want = std::min(kFrontLdsBytes, (size_t)optin);
    if (want > lim) {
      if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_front_factor<T, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(want - 64)) == hipSuccess)
        lim = want - 64;
      else
        (void)hipGetLastError();
    }
    return lim;
  }();
  return limit;
}

template <typename T>
int factor_impl(const SPlanHost& h, const SPlanDev& d, int B, const typename Num<T>::D* data,
                T* front, int* piv, hipStream_t st) {
  const int nlev = (int)h.level_ptr.size() - 1;
  const size_t lds_max = front_lds_limit<T>();
  TopLevels tl{};
  int Ltop = 0;
  const int l0 = kTopMerge<T> ? top_levels(h, false, tl, Ltop) : nlev;
  for (int l = 0; l < l0; ++l) {
    const int* cls = &h.level_cls[5 * l];
    // fronts larger than 64 (and, if so tuned, 33..64): one 256-thread workgroup each
    const int mid = h.mid_blocked ? 2 : 1;
    const int beg = cls[0], cnt = cls[mid] - cls[0];
    if (cnt > 0) {
      const SNode& s0 = h.sn[h.level_sn[beg]];  // largest first
      const int maxf = s0.k + s0.r;
      const size_t bytes = (size_t)maxf * maxf * sizeof(T);
      dim3 grid(cnt, B);
      // a front in HBM stages its panels in LDS while they fit the default
      // 64 KB of a workgroup (fronts of up to ~470 FP64 unknowns)
      const size_t panel = (size_t)maxf * 17 * sizeof(T);
      const bool staged = panel <= 64 * 1024 - 256;
      if (bytes <= lds_max)
        hipLaunchKernelGGL((k_front_factor<T, true>), grid, dim3(256), bytes, st, d, beg, data,
                           front, piv, 0);
      else
        hipLaunchKernelGGL((k_front_factor<T, false>), grid, dim3(256), staged ? panel : 0, st,
                           d, beg, data, front, piv, staged ? 1 : 0);
    }
    if (!h.mid_blocked)
      launch_factor_w<T, 64>(d, cls[1], cls[2] - cls[1], B, data, front, piv, st);
    launch_factor_w<T, 32>(d, cls[2], cls[3] - cls[2], B, data, front, piv, st);
    launch_factor_w<T, 16>(d, cls[3], cls[4] - cls[3], B, data, front, piv, st);
  }
  if constexpr (kTopMerge<T>) {
    if (l0 < nlev)
      launch_factor_top<T, false>(d, tl, Ltop, B, data, front, piv, nullptr, nullptr, st);
  }
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess && std::getenv("VXM_DEBUG"))
    std::fprintf(stderr, "vxm: factor launch failed: %s\n", hipGetErrorString(err));
  return err == hipSuccess ? 0 : -2;
}

template <typename T, int L>
void launch_solve_w(const SPlanDev& d, int beg, int cnt, int B, const T* front, const int* piv,
                    typename Num<T>::A* x, typename Num<T>::A* uvec, bool fwd, bool transpose,
                    hipStream_t st) {
  if (cnt <= 0 || B <= 0) return;
  const dim3 grid((cnt + 64 / L - 1) / (64 / L), B), block(64);
  const size_t lds = 64 * sizeof(typename Num<T>::A);
  const int* list = d.level_sn + beg;
  if (fwd && !transpose)
    hipLaunchKernelGGL((k_front_fwd_w<T, L>), grid, block, lds, st, d, list, cnt, front, piv, x,
                       uvec);
  else if (fwd)
    hipLaunchKernelGGL((k_front_fwd_t_w<T, L>), grid, block, lds, st, d, list, cnt, front, x,
                       uvec);
  else if (!transpose)
    hipLaunchKernelGGL((k_front_bwd_w<T, L>), grid, block, lds, st, d, list, cnt, front, x);
  else
    hipLaunchKernelGGL((k_front_bwd_t_w<T, L>), grid, block, lds, st, d, list, cnt, front, piv,
                       x);
}

// One level of a triangular sweep.
template <typename T>
void sweep_level(const SPlanHost& h, const SPlanDev& d, int l, int B, const T* front,
                 const int* piv, typename Num<T>::A* x, typename Num<T>::A* uvec, bool fwd,
                 bool transpose, hipStream_t st) {
    const int* cls = &h.level_cls[5 * l];
    // fronts larger than 64: one 128-thread workgroup each
    const int beg = cls[0], cnt = cls[1] - cls[0];
    if (cnt > 0) {
      const SNode& s0 = h.sn[h.level_sn[beg]];  // largest first
      const size_t lds = (size_t)(s0.k + s0.r) * sizeof(typename Num<T>::A);
      const dim3 grid(cnt, B), block(128);
      if (fwd && transpose)
        hipLaunchKernelGGL((k_front_fwd_t<T>), grid, block, lds, st, d, beg, front, x, uvec);
      else if (fwd)
        hipLaunchKernelGGL((k_front_fwd<T>), grid, block, lds, st, d, beg, front, piv, x, uvec);
      else if (transpose)
        hipLaunchKernelGGL((k_front_bwd_t<T>), grid, block, lds, st, d, beg, front, piv, x);
      else
        hipLaunchKernelGGL((k_front_bwd<T>), grid, block, lds, st, d, beg, front, x);
    }
    launch_solve_w<T, 64>(d, cls[1], cls[2] - cls[1], B, front, piv, x, uvec, fwd, transpose, st);
    launch_solve_w<T, 32>(d, cls[2], cls[3] - cls[2], B, front, piv, x, uvec, fwd, transpose, st);
    launch_solve_w<T, 16>(d, cls[3], cls[4] - cls[3], B, front, piv, x, uvec, fwd, transpose, st);
}

template <typename T>
int solve_impl(const SPlanHost& h, const SPlanDev& d, int B, const T* front, const int* piv,
               typename Num<T>::A* x, typename Num<T>::A* uvec, bool transpose, hipStream_t st) {
  const int nlev = (int)h.level_ptr.size() - 1;
  // the top levels of small fronts in one launch per sweep (not transposed)
  TopLevels tl{};
  int Ltop = 0;
  const int l0 = kTopMerge<T> && !transpose ? top_levels(h, false, tl, Ltop) : nlev;
  for (int l = 0; l < l0; ++l)
    sweep_level<T>(h, d, l, B, front, piv, x, uvec, true, transpose, st);
  if constexpr (kTopMerge<T>) {
    if (l0 < nlev) {
      launch_sweep_top<T>(d, tl, Ltop, B, front, piv, x, uvec, true, st);
      launch_sweep_top<T>(d, tl, Ltop, B, front, piv, x, uvec, false, st);
    }
  }
  for (int l = l0 - 1; l >= 0; --l)
    sweep_level<T>(h, d, l, B, front, piv, x, uvec, false, transpose, st);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// FP64 factorization and solve of A x = b (x holds b on entry) in one pass of
// the levels: register fronts that leave a lane free (k + r < L) factor with
// the forward sweep riding along (k_front_factor_w<double, L, true>); the other
// fronts of the level (fronts of exactly L unknowns -- first in their class,
// the level lists being sorted largest first -- and the blocked ones) get
// their forward sweep right after the level's factorization; then the
// backward sweep.  Same results as splu_factor followed by splu_solve.
int factor_solve_f64(const SPlanHost& h, const SPlanDev& d, int B, const double* data,
                     double* front, int* piv, double* x, double* uvec, hipStream_t st) {
  const int nlev = (int)h.level_ptr.size() - 1;
  const size_t lds_max = front_lds_limit<double>();
  // first entry of [beg, end) with fewer than L unknowns (sizes non-increasing)
  auto first_free = [&](int beg, int end, int L) {
    int i = beg;
    while (i < end && h.sn[h.level_sn[i]].k + h.sn[h.level_sn[i]].r >= L) ++i;
    return i;
  };
  // the top levels of small fronts with a free lane: one launch (k_front_factor_top)
  TopLevels tl{};
  int Ltop = 0;
  const int l0 = top_levels(h, true, tl, Ltop);
  for (int l = 0; l < l0; ++l) {
    const int* cls = &h.level_cls[5 * l];
    const int mid = h.mid_blocked ? 2 : 1;
    const int beg = cls[0], cnt = cls[mid] - cls[0];
    if (cnt > 0) {
      const SNode& s0 = h.sn[h.level_sn[beg]];  // largest first
      const int maxf = s0.k + s0.r;
      const size_t bytes = (size_t)maxf * maxf * sizeof(double);
      const size_t panel = (size_t)maxf * 17 * sizeof(double);
      const bool staged = panel <= 64 * 1024 - 256;
      dim3 grid(cnt, B);
      if (bytes <= lds_max)
        hipLaunchKernelGGL((k_front_factor<double, true>), grid, dim3(256), bytes, st, d, beg,
                           data, front, piv, 0);
      else
        hipLaunchKernelGGL((k_front_factor<double, false>), grid, dim3(256),
                           staged ? panel : 0, st, d, beg, data, front, piv, staged ? 1 : 0);
    }
    const int lanes[4] = {0, 64, 32, 16};
    int split[4] = {0, 0, 0, 0};  // class c: [cls[c], split[c]) unfused, the rest fused
    for (int c = mid; c <= 3; ++c) {
      const int b0 = cls[c], e = cls[c + 1];
      split[c] = first_free(b0, e, lanes[c]);
      const int n0 = split[c] - b0, n1 = e - split[c];
      if (c == 1) {
        launch_factor_w<double, 64>(d, b0, n0, B, data, front, piv, st);
        launch_factor_fwd_w<64>(d, split[c], n1, B, data, front, piv, x, uvec, st);
      } else if (c == 2) {
        launch_factor_w<double, 32>(d, b0, n0, B, data, front, piv, st);
        launch_factor_fwd_w<32>(d, split[c], n1, B, data, front, piv, x, uvec, st);
      } else {
        launch_factor_w<double, 16>(d, b0, n0, B, data, front, piv, st);
        launch_factor_fwd_w<16>(d, split[c], n1, B, data, front, piv, x, uvec, st);
      }
    }
    // forward sweep of the fronts not fused: blocked ones (> 64, and 33..64
    // when so configured), then the register classes' full fronts
    {
      const int beg = cls[0], cnt = cls[1] - cls[0];
      if (cnt > 0) {
        const SNode& s0 = h.sn[h.level_sn[beg]];
        const size_t lds = (size_t)(s0.k + s0.r) * sizeof(double);
        hipLaunchKernelGGL((k_front_fwd<double>), dim3(cnt, B), dim3(128), lds, st, d, beg,
                           front, piv, x, uvec);
      }
    }
    if (mid == 2)
      launch_solve_w<double, 64>(d, cls[1], cls[2] - cls[1], B, front, piv, x, uvec, true, false,
                                 st);
    for (int c = mid; c <= 3; ++c) {
      const int n0 = split[c] - cls[c];
      if (c == 1) launch_solve_w<double, 64>(d, cls[c], n0, B, front, piv, x, uvec, true, false, st);
      if (c == 2) launch_solve_w<double, 32>(d, cls[c], n0, B, front, piv, x, uvec, true, false, st);
      if (c == 3) launch_solve_w<double, 16>(d, cls[c], n0, B, front, piv, x, uvec, true, false, st);
    }
  }
  if (l0 < nlev) {
    launch_factor_top<double, true>(d, tl, Ltop, B, data, front, piv, x, uvec, st);
    launch_sweep_top<double>(d, tl, Ltop, B, front, piv, x, uvec, false, st);
  }
  for (int l = l0 - 1; l >= 0; --l)
    sweep_level<double>(h, d, l, B, front, piv, x, uvec, false, false, st);
  const hipError_t err = hipGetLastError();
  return err == hipSuccess ? 0 : -2;
}

int splu_factor(const SPlanHost& h, const SPlanDev& d, int B, const void* data, void* front,
                int* piv, int dtype, hipStream_t st) {
  if (dtype == kSpluF32)
    return factor_impl<float>(h, d, B, (const double*)data, (float*)front, piv, st);
  if (dtype == kSpluC128)
    return factor_impl<cplx>(h, d, B, (const cplx*)data, (cplx*)front, piv, st);
  return factor_impl<double>(h, d, B, (const double*)data, (double*)front, piv, st);
}

int splu_factor_solve(const SPlanHost& h, const SPlanDev& d, int B, const double* data,
                      double* front, int* piv, double* x, double* uvec, hipStream_t st) {
  return factor_solve_f64(h, d, B, data, front, piv, x, uvec, st);
}

int splu_solve(const SPlanHost& h, const SPlanDev& d, int B, const void* front, const int* piv,
               void* x, void* uvec, bool transpose, int dtype, hipStream_t st) {
  if (dtype == kSpluF32)
    return solve_impl<float>(h, d, B, (const float*)front, piv, (double*)x, (double*)uvec,
                             transpose, st);
  if (dtype == kSpluC128)
    return solve_impl<cplx>(h, d, B, (const cplx*)front, piv, (cplx*)x, (cplx*)uvec,
                            transpose, st);
  return solve_impl<double>(h, d, B, (const double*)front, piv, (double*)x, (double*)uvec,
                            transpose, st);
}

}  // namespace vxm
