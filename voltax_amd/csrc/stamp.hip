// Batched device evaluation + deterministic stamping (generic sizes).
//
// Reference: voltax/sparse.py:220-249 (jacobian: per-group JVPs, then
// data.at[pos].add), voltax/circuit.py:209-226 (_assemble: residual
// scatter-add), voltax/analysis.py:359-360 (gmin on node diagonals).
//
// Two phases, no atomics:
//   k_eval      one thread per (instance, device): FP64 physics, writes the
//               device's residual contributions and Jacobian block into a
//               per-instance contribution buffer;
//   k_gather_*  one thread per (instance, row) / (instance, slot): sums the
//               contributions of that row / slot in the fixed device order
//               recorded in the topology maps.
// Instance-major layouts: params [B][P], z [B][S], data [B][nnz], res [B][S],
// work [B][Cres + Cjac].
#include <hip/hip_runtime.h>

#include "devices.hpp"
#include "kernels.hpp"
#include "topology.hpp"

namespace vxm {

__global__ void __launch_bounds__(256)
k_eval(TopoDev t, int B, const double* __restrict__ params, const double* __restrict__ z,
       double tt, double cf, double cq, int want_jac, double* __restrict__ work) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)B * t.D) return;
  const int b = (int)(gid / t.D), d = (int)(gid % t.D);
  const int kind = t.kind[d];
  const int m = kind_m(kind);
  const double* zb = z + (long)b * t.S;
  double u[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int idx = (i < m) ? t.unk[4 * d + i] : -1;
    u[i] = idx >= 0 ? zb[idx] : 0.0;
  }
  const double* p = params + (long)b * t.P + t.poff[d];
  double res[4], jac[16];
  if (want_jac)
    device_eval<true>(kind, t.wave[d], p, u, tt, cf, cq, res, jac);
  else
    device_eval<false>(kind, t.wave[d], p, u, tt, cf, cq, res, jac);
  double* wb = work + (long)b * (t.Cres + t.Cjac);
  double* wr = wb + t.roff[d];
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (i < m) wr[i] = res[i];
  if (want_jac) {
    double* wj = wb + t.Cres + t.joff[d];
#pragma unroll
    for (int i = 0; i < 16; ++i)
      if (i < m * m) wj[i] = jac[i];
  }
}

// res[b][s] = sum of contributions + hscale*hist[b][s] + gmin*z (node rows)
__global__ void __launch_bounds__(256)
k_gather_res(TopoDev t, int B, const double* __restrict__ work, const double* __restrict__ z,
             const double* __restrict__ hist, double gmin, double* __restrict__ res) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)B * t.S) return;
  const int b = (int)(gid / t.S), s = (int)(gid % t.S);
  const double* wr = work + (long)b * (t.Cres + t.Cjac);
  double acc = 0.0;
  for (int c = t.rptr[s]; c < t.rptr[s + 1]; ++c) acc += wr[t.ridx[c]];
  if (hist) acc += hist[gid];
  if (s < t.nv) acc += gmin * z[gid];
  res[gid] = acc;
}

__global__ void __launch_bounds__(256)
k_gather_jac(TopoDev t, int B, const double* __restrict__ work, double gmin,
             double* __restrict__ data) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)B * t.nnz) return;
  const int b = (int)(gid / t.nnz), p = (int)(gid % t.nnz);
  const double* wj = work + (long)b * (t.Cres + t.Cjac) + t.Cres;
  double acc = 0.0;
  for (int c = t.jptr[p]; c < t.jptr[p + 1]; ++c) acc += wj[t.jidx[c]];
  // the diagonal slot of a node row also carries gmin (only one such slot)
  if (gmin != 0.0) {
    // row of slot p: binary search in indptr
    int lo = 0, hi = t.S;
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (t.indptr[mid] <= p) lo = mid; else hi = mid;
    }
    if (lo < t.nv && t.diag[lo] == p) acc += gmin;
  }
  data[gid] = acc;
}

static inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int launch_stamp(const TopoDev& t, int B, const double* params, const double* z, double tt,
                 double cf, double cq, double gmin, const double* hist, double* data,
                 double* res, double* work, hipStream_t stream) {
  if (B <= 0) return 0;
  const bool jac = data != nullptr;
  if (t.D > 0)
    hipLaunchKernelGGL(k_eval, dim3(nblocks((long)B * t.D, 256)), dim3(256), 0, stream, t, B,
                       params, z, tt, cf, cq, jac ? 1 : 0, work);
  if (res)
    hipLaunchKernelGGL(k_gather_res, dim3(nblocks((long)B * t.S, 256)), dim3(256), 0, stream,
                       t, B, work, z, hist, gmin, res);
  if (data)
    hipLaunchKernelGGL(k_gather_jac, dim3(nblocks((long)B * t.nnz, 256)), dim3(256), 0,
                       stream, t, B, work, gmin, data);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ------------------------------------------------------------------ SpMV
// y[b] = A[b] x[b] for CSR values data [B][nnz] (voltax/sparse.py:130-133).
__global__ void __launch_bounds__(256)
k_spmv(int S, int nnz, int B, const int* __restrict__ indptr, const int* __restrict__ indices,
       const double* __restrict__ data, const double* __restrict__ x, double* __restrict__ y,
       int transpose_unused) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)B * S) return;
  const int b = (int)(gid / S), s = (int)(gid % S);
  const double* db = data + (long)b * nnz;
  const double* xb = x + (long)b * S;
  double acc = 0.0;
  for (int p = indptr[s]; p < indptr[s + 1]; ++p) acc += db[p] * xb[indices[p]];
  y[gid] = acc;
}

int launch_spmv(const TopoDev& t, int B, const double* data, const double* x, double* y,
                hipStream_t stream) {
  if (B <= 0) return 0;
  hipLaunchKernelGGL(k_spmv, dim3(nblocks((long)B * t.S, 256)), dim3(256), 0, stream, t.S,
                     t.nnz, B, t.indptr, t.indices, data, x, y, 0);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace vxm
