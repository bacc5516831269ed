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

#include "splu.hpp"

namespace vxm {

template <typename T>
__device__ __forceinline__ void block_argmax(T& best, int& bi, T* sv, int* si) {
  // lanes: reduce within the wave, then across waves through LDS
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T ov = __shfl_xor(best, o);
    int oi = __shfl_xor(bi, o);
    if (ov > best || (ov == best && oi < bi)) {
      best = ov;
      bi = oi;
    }
  }
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) {
    sv[w] = best;
    si[w] = bi;
  }
  __syncthreads();
  best = sv[0];
  bi = si[0];
  for (int i = 1; i < nw; ++i)
    if (sv[i] > best || (sv[i] == best && si[i] < bi)) {
      best = sv[i];
      bi = si[i];
    }
  __syncthreads();
}

template <typename T, bool LDS>
__global__ void __launch_bounds__(256)
k_front_factor(SPlanDev p, int lbeg, const double* __restrict__ data, T* __restrict__ front,
               int* __restrict__ piv) {
  const int s = p.level_sn[lbeg + blockIdx.x];
  const int b = blockIdx.y;
  const SNode sn = p.sn[s];
  const int k = sn.k, nf = sn.k + sn.r;
  T* Fg = front + (int64_t)b * p.front_total + sn.front_off;
  extern __shared__ double smem_d[];
  T* F = LDS ? reinterpret_cast<T*>(smem_d) : Fg;
  __shared__ T sv[4];
  __shared__ int si[4];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int nf2 = nf * nf;
  for (int e = tid; e < nf2; e += nt) F[e] = T(0);
  __syncthreads();
  const double* db = data + (int64_t)b * p.nnz;
  for (int e = tid; e < sn.n_aent; e += nt)
    F[p.apos[sn.aent_off + e]] += (T)db[p.aslot[sn.aent_off + e]];
  __syncthreads();
  for (int ci = 0; ci < sn.n_child; ++ci) {
    const SNode cs = p.sn[p.children[sn.child_off + ci]];
    const int kc = cs.k, rc = cs.r, nfc = kc + rc;
    const T* Fc = front + (int64_t)b * p.front_total + cs.front_off;
    const int* cm = p.cmap + cs.cmap_off;
    for (int e = tid; e < rc * rc; e += nt) {
      const int i = e / rc, j = e - i * rc;
      F[cm[i] * nf + cm[j]] += Fc[(kc + i) * nfc + kc + j];
    }
    __syncthreads();
  }
  int* pv = piv + (int64_t)b * p.piv_total + sn.piv_off;
  for (int j = 0; j < k; ++j) {
    T best = T(-1);
    int bi = j;
    for (int i = j + tid; i < k; i += nt) {
      T v = F[i * nf + j];
      v = v < T(0) ? -v : v;
      if (v != v) v = T(INFINITY);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    block_argmax(best, bi, sv, si);
    const int pr = bi;
    if (tid == 0) pv[j] = pr;
    if (pr != j) {
      for (int c = tid; c < nf; c += nt) {
        T a = F[j * nf + c];
        F[j * nf + c] = F[pr * nf + c];
        F[pr * nf + c] = a;
      }
      __syncthreads();
    }
    const T d = F[j * nf + j];
    for (int i = j + 1 + tid; i < nf; i += nt) F[i * nf + j] /= d;
    __syncthreads();
    const int m = nf - j - 1;
    for (int e = tid; e < m * m; e += nt) {
      const int i = j + 1 + e / m, c = j + 1 + e % m;
      F[i * nf + c] -= F[i * nf + j] * F[j * nf + c];
    }
    __syncthreads();
  }
  if (LDS)
    for (int e = tid; e < nf2; e += nt) Fg[e] = F[e];
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
template <typename T, int L>
__global__ void __launch_bounds__(64)
k_front_factor_w(SPlanDev p, const int* __restrict__ list, int count,
                 const double* __restrict__ data, T* __restrict__ front, int* __restrict__ piv) {
  constexpr int FPW = 64 / L;
  extern __shared__ double smem_d[];
  const int lane = threadIdx.x, seg = lane / L, c = lane % L;
  const int fi = blockIdx.x * FPW + seg;
  const int b = blockIdx.y;
  const bool live = fi < count;
  const SNode sn = p.sn[list[live ? fi : count - 1]];
  const int k = live ? sn.k : 0, nf = live ? sn.k + sn.r : 0;
  T* F = reinterpret_cast<T*>(smem_d) + seg * L * L;  // pitch L
  for (int e = c; e < L * L; e += L) F[e] = T(0);
  __syncthreads();
  const double* db = data + (int64_t)b * p.nnz;
  for (int e = c; e < (live ? sn.n_aent : 0); e += L) {
    const int pos = p.apos[sn.aent_off + e];
    const int i = pos / nf, j = pos - i * nf;
    F[i * L + j] += (T)db[p.aslot[sn.aent_off + e]];
  }
  __syncthreads();
  for (int ci = 0; ci < (live ? sn.n_child : 0); ++ci) {
    const SNode cs = p.sn[p.children[sn.child_off + ci]];
    const int kc = cs.k, rc = cs.r, nfc = kc + rc;
    const T* Fc = front + (int64_t)b * p.front_total + cs.front_off;
    const int* cm = p.cmap + cs.cmap_off;
    for (int e = c; e < rc * rc; e += L) {
      const int i = e / rc, j = e - i * rc;
      F[cm[i] * L + cm[j]] += Fc[(kc + i) * nfc + kc + j];
    }
    __syncthreads();  // one wave per block: orders the LDS sums child by child
  }
  __syncthreads();
  T a[L];
#pragma unroll
  for (int i = 0; i < L; ++i) a[i] = F[i * L + c];
  int* pv = piv + (int64_t)b * p.piv_total + sn.piv_off;
#pragma unroll
  for (int j = 0; j < L; ++j) {
    if (!__any(j < k)) break;
    const bool act = j < k;
    // pivot search in column j (lane j's registers) among rows j..k-1
    T best = T(-1);
    int bi = j;
#pragma unroll
    for (int i = j; i < L; ++i) {
      T v = a[i] < T(0) ? -a[i] : a[i];
      if (v != v) v = T(INFINITY);
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
    const T rinv = T(1) / __shfl(a[j], j, L);
    const T uj = a[j];
#pragma unroll
    for (int i = j + 1; i < L; ++i) {
      const T l = __shfl(a[i], j, L) * rinv;
      if (act && c == j) a[i] = l;
      if (act && c > j) a[i] -= l * uj;
    }
  }
  if (!live || c >= nf) return;
  T* Fg = front + (int64_t)b * p.front_total + sn.front_off;
#pragma unroll
  for (int i = 0; i < L; ++i)
    if (i < nf) Fg[i * nf + c] = a[i];
}

// ---------------------------------------------------------------- solves
// Forward sweep of A x = b: y = L^-1 P b, border contributions to parents.
template <typename T>
__global__ void __launch_bounds__(128)
k_front_fwd(SPlanDev p, int lbeg, const T* __restrict__ front, const int* __restrict__ piv,
            double* __restrict__ x, double* __restrict__ uvec) {
  const int s = p.level_sn[lbeg + blockIdx.x];
  const int b = blockIdx.y;
  const SNode sn = p.sn[s];
  const int k = sn.k, nf = sn.k + sn.r;
  const T* F = front + (int64_t)b * p.front_total + sn.front_off;
  extern __shared__ double v[];
  const int* id = p.idx + sn.idx_off;
  double* xb = x + (int64_t)b * p.S;
  double* ub = uvec + (int64_t)b * p.r_total;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int j = tid; j < nf; j += nt) v[j] = j < k ? xb[id[j]] : 0.0;
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
      const double a = v[j];
      v[j] = v[q];
      v[q] = a;
    }
  }
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    const double vj = v[j];
    for (int i = j + 1 + tid; i < nf; i += nt) v[i] -= (double)F[i * nf + j] * vj;
    __syncthreads();
  }
  for (int j = tid; j < k; j += nt) xb[id[j]] = v[j];
  for (int i = tid; i < sn.r; i += nt) ub[sn.uvec_off + i] = v[k + i];
}

// Backward sweep of A x = b: x_K = U11^-1 (y_K - U12 x_R).
template <typename T>
__global__ void __launch_bounds__(128)
k_front_bwd(SPlanDev p, int lbeg, const T* __restrict__ front, double* __restrict__ x) {
  const int s = p.level_sn[lbeg + blockIdx.x];
  const int b = blockIdx.y;
  const SNode sn = p.sn[s];
  const int k = sn.k, nf = sn.k + sn.r;
  const T* F = front + (int64_t)b * p.front_total + sn.front_off;
  extern __shared__ double v[];
  const int* id = p.idx + sn.idx_off;
  double* xb = x + (int64_t)b * p.S;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int j = tid; j < nf; j += nt) v[j] = xb[id[j]];
  __syncthreads();
  for (int j = tid; j < k; j += nt) {
    double acc = 0.0;
    for (int i = k; i < nf; ++i) acc += (double)F[j * nf + i] * v[i];
    v[j] -= acc;
  }
  __syncthreads();
  for (int j = k - 1; j >= 0; --j) {
    if (tid == 0) v[j] /= (double)F[j * nf + j];
    __syncthreads();
    const double vj = v[j];
    for (int i = tid; i < j; i += nt) v[i] -= (double)F[i * nf + j] * vj;
    __syncthreads();
  }
  for (int j = tid; j < k; j += nt) xb[id[j]] = v[j];
}

// Forward sweep of A^T x = b: U^T w = b (U11^T lower), border via U12^T.
template <typename T>
__global__ void __launch_bounds__(128)
k_front_fwd_t(SPlanDev p, int lbeg, const T* __restrict__ front, double* __restrict__ x,
              double* __restrict__ uvec) {
  const int s = p.level_sn[lbeg + blockIdx.x];
  const int b = blockIdx.y;
  const SNode sn = p.sn[s];
  const int k = sn.k, nf = sn.k + sn.r;
  const T* F = front + (int64_t)b * p.front_total + sn.front_off;
  extern __shared__ double v[];
  const int* id = p.idx + sn.idx_off;
  double* xb = x + (int64_t)b * p.S;
  double* ub = uvec + (int64_t)b * p.r_total;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int j = tid; j < nf; j += nt) v[j] = j < k ? xb[id[j]] : 0.0;
  __syncthreads();
  for (int ci = 0; ci < sn.n_child; ++ci) {
    const SNode cs = p.sn[p.children[sn.child_off + ci]];
    const int* cm = p.cmap + cs.cmap_off;
    for (int i = tid; i < cs.r; i += nt) v[cm[i]] += ub[cs.uvec_off + i];
    __syncthreads();
  }
  for (int j = 0; j < k; ++j) {
    if (tid == 0) v[j] /= (double)F[j * nf + j];
    __syncthreads();
    const double vj = v[j];
    for (int i = j + 1 + tid; i < nf; i += nt) v[i] -= (double)F[j * nf + i] * vj;
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
              double* __restrict__ x) {
  const int s = p.level_sn[lbeg + blockIdx.x];
  const int b = blockIdx.y;
  const SNode sn = p.sn[s];
  const int k = sn.k, nf = sn.k + sn.r;
  const T* F = front + (int64_t)b * p.front_total + sn.front_off;
  extern __shared__ double v[];
  const int* id = p.idx + sn.idx_off;
  double* xb = x + (int64_t)b * p.S;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int j = tid; j < nf; j += nt) v[j] = xb[id[j]];
  __syncthreads();
  for (int j = tid; j < k; j += nt) {
    double acc = 0.0;
    for (int i = k; i < nf; ++i) acc += (double)F[i * nf + j] * v[i];
    v[j] -= acc;
  }
  __syncthreads();
  for (int j = k - 1; j >= 0; --j) {
    const double vj = v[j];
    for (int i = tid; i < j; i += nt) v[i] -= (double)F[j * nf + i] * vj;
    __syncthreads();
  }
  if (tid == 0) {
    const int* pv = piv + (int64_t)b * p.piv_total + sn.piv_off;
    for (int j = k - 1; j >= 0; --j) {
      const int q = pv[j];
      const double a = v[j];
      v[j] = v[q];
      v[q] = a;
    }
  }
  __syncthreads();
  for (int j = tid; j < k; j += nt) xb[id[j]] = v[j];
}

// ------------------------------------------------------------ launchers
template <typename T, int L>
void launch_factor_w(const SPlanDev& d, int beg, int cnt, int B, const double* data, T* front,
                     int* piv, hipStream_t st) {
  if (cnt <= 0 || B <= 0) return;
  constexpr int FPW = 64 / L;
  hipLaunchKernelGGL((k_front_factor_w<T, L>), dim3((cnt + FPW - 1) / FPW, B), dim3(64),
                     (size_t)64 * L * sizeof(T), st, d, d.level_sn + beg, cnt, data, front, piv);
}

// Lets the LDS variant of k_front_factor take up to kFrontLdsBytes (the whole
// 160 KB of a CU) per workgroup; once per process and element type.
template <typename T>
bool allow_large_lds() {
  static const bool ok =
      hipFuncSetAttribute(reinterpret_cast<const void*>(&k_front_factor<T, true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)kFrontLdsBytes) == hipSuccess;
  return ok;
}

template <typename T>
int factor_impl(const SPlanHost& h, const SPlanDev& d, int B, const double* data, T* front,
                int* piv, hipStream_t st) {
  const int nlev = (int)h.level_ptr.size() - 1;
  const size_t lds_max = allow_large_lds<T>() ? kFrontLdsBytes : 64 * 1024;
  for (int l = 0; l < nlev; ++l) {
    const int* cls = &h.level_cls[5 * l];
    // fronts larger than 64: one 256-thread workgroup each
    const int beg = cls[0], cnt = cls[1] - cls[0];
    if (cnt > 0) {
      const SNode& s0 = h.sn[h.level_sn[beg]];  // largest first
      const int maxf = s0.k + s0.r;
      const size_t bytes = (size_t)maxf * maxf * sizeof(T);
      dim3 grid(cnt, B);
      if (bytes <= lds_max)
        hipLaunchKernelGGL((k_front_factor<T, true>), grid, dim3(256), bytes, st, d, beg, data,
                           front, piv);
      else
        hipLaunchKernelGGL((k_front_factor<T, false>), grid, dim3(256), 0, st, d, beg, data,
                           front, piv);
    }
    launch_factor_w<T, 64>(d, cls[1], cls[2] - cls[1], B, data, front, piv, st);
    launch_factor_w<T, 32>(d, cls[2], cls[3] - cls[2], B, data, front, piv, st);
    launch_factor_w<T, 16>(d, cls[3], cls[4] - cls[3], B, data, front, piv, st);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <typename T>
int solve_impl(const SPlanHost& h, const SPlanDev& d, int B, const T* front, const int* piv,
               double* x, double* uvec, double* z, bool transpose, hipStream_t st) {
  const int nlev = (int)h.level_ptr.size() - 1;
  auto lds = [&](int l) {
    int maxf = 1;
    for (int i = h.level_ptr[l]; i < h.level_ptr[l + 1]; ++i) {
      const SNode& s = h.sn[h.level_sn[i]];
      maxf = std::max(maxf, s.k + s.r);
    }
    return (size_t)maxf * sizeof(double);
  };
  for (int l = 0; l < nlev; ++l) {
    const int beg = h.level_ptr[l], cnt = h.level_ptr[l + 1] - beg;
    if (transpose)
      hipLaunchKernelGGL((k_front_fwd_t<T>), dim3(cnt, B), dim3(128), lds(l), st, d, beg, front,
                         x, uvec);
    else
      hipLaunchKernelGGL((k_front_fwd<T>), dim3(cnt, B), dim3(128), lds(l), st, d, beg, front,
                         piv, x, uvec);
  }
  for (int l = nlev - 1; l >= 0; --l) {
    const int beg = h.level_ptr[l], cnt = h.level_ptr[l + 1] - beg;
    if (transpose)
      hipLaunchKernelGGL((k_front_bwd_t<T>), dim3(cnt, B), dim3(128), lds(l), st, d, beg, front,
                         piv, x);
    else
      hipLaunchKernelGGL((k_front_bwd<T>), dim3(cnt, B), dim3(128), lds(l), st, d, beg, front,
                         x);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int splu_factor(const SPlanHost& h, const SPlanDev& d, int B, const double* data, void* front,
                int* piv, int fp32, hipStream_t st) {
  if (fp32) return factor_impl<float>(h, d, B, data, (float*)front, piv, st);
  return factor_impl<double>(h, d, B, data, (double*)front, piv, st);
}

int splu_solve(const SPlanHost& h, const SPlanDev& d, int B, const void* front, const int* piv,
               double* x, double* uvec, double* z, bool transpose, int fp32, hipStream_t st) {
  if (fp32)
    return solve_impl<float>(h, d, B, (const float*)front, piv, x, uvec, z, transpose, st);
  return solve_impl<double>(h, d, B, (const double*)front, piv, x, uvec, z, transpose, st);
}

}  // namespace vxm
