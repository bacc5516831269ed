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
#include "devices_ad.hpp"
#include "kernels.hpp"
#include "splu.hpp"
#include "topology.hpp"
#include "va_codegen.hpp"
#include "va_eval.hpp"

namespace vxm {

// Scalars may come by value or, for graph replays, from device memory:
// scal = [t, cf, cq, gmin, source_scale] overrides the by-value arguments
// (source scale 1 when given by value).
__global__ void __launch_bounds__(256)
k_eval(TopoDev t, int B, const double* __restrict__ params, const double* __restrict__ z,
       double tt, double cf, double cq, int want_jac, double* __restrict__ work,
       const double* __restrict__ scal) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)B * t.D) return;
  double ss = 1.0;
  if (scal) {
    tt = scal[0];
    cf = scal[1];
    cq = scal[2];
    ss = scal[4];
  }
  const int b = (int)(gid / t.D), d = (int)(gid % t.D);
  const int kind = t.kind[d];
  if (kind_is_va(kind)) return;  // k_eval_va
  const int m = kind_m(kind);
  const double* zb = z + (long)b * t.S;
  double u[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int idx = (i < m) ? t.unk[t.uw * d + i] : -1;
    u[i] = idx >= 0 ? zb[idx] : 0.0;
  }
  const double* p = params + (long)b * t.P + t.poff[d];
  double res[4], jac[16];
  if (want_jac)
    device_eval<true>(kind, t.wave[d], p, u, tt, cf, cq, res, jac, ss);
  else
    device_eval<false>(kind, t.wave[d], p, u, tt, cf, cq, res, jac, ss);
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

// Verilog-A devices of a topology (kinds KIND_VA2..4): the compiled program of
// each device interpreted per (instance, device), contributions written to the
// same slots a built-in device uses.  A kernel of its own, so that the
// interpreter's register file (private memory) never burdens k_eval.
__global__ void __launch_bounds__(256)
k_eval_va(TopoDev t, int B, const double* __restrict__ params, const double* __restrict__ z,
          double tt, double cf, double cq, int want_jac, double* __restrict__ work,
          const double* __restrict__ scal) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)B * t.D) return;
  if (scal) {
    cf = scal[1];
    cq = scal[2];
  }
  const int b = (int)(gid / t.D), d = (int)(gid % t.D);
  const int kind = t.kind[d];
  if (!kind_is_interpreted(kind)) return;
  const int m = kind_m(kind);
  const double* zb = z + (long)b * t.S;
  double u[4];
  for (int i = 0; i < 4; ++i) {
    const int idx = (i < m) ? t.unk[t.uw * d + i] : -1;
    u[i] = idx >= 0 ? zb[idx] : 0.0;
  }
  const double* p = params + (long)b * t.P + t.poff[d];
  double res[4], jac[16];
  double* wb = work + (long)b * (t.Cres + t.Cjac);
  if (want_jac) {
    va_eval<true>(t.code + t.wave[d], p, u, cf, cq, res, jac);
    double* wj = wb + t.Cres + t.joff[d];
    for (int i = 0; i < m * m; ++i) wj[i] = jac[i];
  } else {
    va_eval<false>(t.code + t.wave[d], p, u, cf, cq, res, jac);
  }
  double* wr = wb + t.roff[d];
  for (int i = 0; i < m; ++i) wr[i] = res[i];
}

static inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Sum of a wavefront's 64 values in a fixed butterfly order (every lane gets it).
__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Gathers of the topology's contribution lists.  A thread sums one list of
// at most kHeavyGather terms in list order; the few longer lists (a supply
// node's row and diagonal collect a term from every device on it) are
// summed by one wavefront each, lane-strided then in a fixed butterfly, in
// the workgroups after the per-thread ones: the launch needs `n_light`
// blocks of 256 threads plus one block per four heavy (instance, list) pairs.
struct GatherGrid {
  unsigned light, heavy;
};
static inline GatherGrid gather_grid(long items, int B, int n_heavy) {
  return {nblocks(items, 256), nblocks((long)B * n_heavy, 4)};
}

// One wavefront's heavy item: (instance b, list h) of B * n_heavy, or -1.
__device__ __forceinline__ long heavy_item(unsigned light_blocks, int B, int n_heavy) {
  const long w = (long)(blockIdx.x - light_blocks) * 4 + threadIdx.x / 64;
  return w < (long)B * n_heavy ? w : -1;
}

// res[b][s] = sum of contributions + hscale*hist[b][s] + gmin*z (node rows)
__global__ void __launch_bounds__(256)
k_gather_res(TopoDev t, int B, const double* __restrict__ work, const double* __restrict__ z,
             const double* __restrict__ hist, double gmin, double* __restrict__ res,
             const double* __restrict__ scal, unsigned light_blocks) {
  if (scal) gmin = scal[3];
  long gid;
  int s;
  double acc = 0.0;
  if (blockIdx.x < light_blocks) {
    gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long)B * t.S) return;
    const int b = (int)(gid / t.S);
    s = (int)(gid % t.S);
    const int c0 = t.rptr[s], c1 = t.rptr[s + 1];
    if (c1 - c0 > kHeavyGather) return;
    const double* wr = work + (long)b * (t.Cres + t.Cjac);
    for (int c = c0; c < c1; ++c) acc += wr[t.ridx[c]];
  } else {
    const long w = heavy_item(light_blocks, B, t.n_heavy_r);
    if (w < 0) return;
    const int b = (int)(w / t.n_heavy_r);
    s = t.heavy_r[w % t.n_heavy_r];
    gid = (long)b * t.S + s;
    const double* wr = work + (long)b * (t.Cres + t.Cjac);
    for (int c = t.rptr[s] + (int)(threadIdx.x % 64); c < t.rptr[s + 1]; c += 64)
      acc += wr[t.ridx[c]];
    acc = wave_sum64(acc);
    if (threadIdx.x % 64) return;
  }
  if (hist) acc += hist[gid];
  if (s < t.nv) acc += gmin * z[gid];
  res[gid] = acc;
}

__global__ void __launch_bounds__(256)
k_gather_jac(TopoDev t, int B, const double* __restrict__ work, double gmin,
             double* __restrict__ data, const double* __restrict__ scal, unsigned light_blocks) {
  if (scal) gmin = scal[3];
  long gid;
  int p;
  double acc = 0.0;
  if (blockIdx.x < light_blocks) {
    gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long)B * t.nnz) return;
    const int b = (int)(gid / t.nnz);
    p = (int)(gid % t.nnz);
    const int c0 = t.jptr[p], c1 = t.jptr[p + 1];
    if (c1 - c0 > kHeavyGather) return;
    const double* wj = work + (long)b * (t.Cres + t.Cjac) + t.Cres;
    for (int c = c0; c < c1; ++c) acc += wj[t.jidx[c]];
  } else {
    const long w = heavy_item(light_blocks, B, t.n_heavy_j);
    if (w < 0) return;
    const int b = (int)(w / t.n_heavy_j);
    p = t.heavy_j[w % t.n_heavy_j];
    gid = (long)b * t.nnz + p;
    const double* wj = work + (long)b * (t.Cres + t.Cjac) + t.Cres;
    for (int c = t.jptr[p] + (int)(threadIdx.x % 64); c < t.jptr[p + 1]; c += 64)
      acc += wj[t.jidx[c]];
    acc = wave_sum64(acc);
    if (threadIdx.x % 64) return;
  }
  // the diagonal slot of a node row also carries gmin (only one such slot)
  const int r = t.slot_row[p];
  if (r < t.nv && t.diag[r] == p) acc += gmin;
  data[gid] = acc;
}

static void launch_gather_res(const TopoDev& t, int B, const double* work, const double* z,
                              const double* hist, double gmin, double* res, const double* scal,
                              hipStream_t stream) {
  const GatherGrid g = gather_grid((long)B * t.S, B, t.n_heavy_r);
  hipLaunchKernelGGL(k_gather_res, dim3(g.light + g.heavy), dim3(256), 0, stream, t, B, work, z,
                     hist, gmin, res, scal, g.light);
}

int launch_stamp(const TopoDev& t, int B, const double* params, const double* z, double tt,
                 double cf, double cq, double gmin, const double* hist, double* data,
                 double* res, double* work, hipStream_t stream, const double* scal) {
  if (B <= 0) return 0;
  const bool jac = data != nullptr;
  if (t.D > 0)
    hipLaunchKernelGGL(k_eval, dim3(nblocks((long)B * t.D, 256)), dim3(256), 0, stream, t, B,
                       params, z, tt, cf, cq, jac ? 1 : 0, work, scal);
  if (t.D > 0 && t.has_va)
    hipLaunchKernelGGL(k_eval_va, dim3(nblocks((long)B * t.D, 256)), dim3(256), 0, stream, t,
                       B, params, z, tt, cf, cq, jac ? 1 : 0, work, scal);
  if (t.n_gen > 0 && launch_va_generated(t, jac ? 1 : 0, B, params, z, nullptr, cf, cq, scal,
                                         0.0, work, nullptr, stream))
    return -2;
  if (res) launch_gather_res(t, B, work, z, hist, gmin, res, scal, stream);
  if (data) {
    const GatherGrid g = gather_grid((long)B * t.nnz, B, t.n_heavy_j);
    hipLaunchKernelGGL(k_gather_jac, dim3(g.light + g.heavy), dim3(256), 0, stream, t, B, work,
                       gmin, data, scal, g.light);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ------------------------------------------------------------------ SpMV
// Row s of instance b of A x (rows of more than kHeavyGather entries by a
// wavefront, as the gathers above); false for a thread without a result.
__device__ __forceinline__ bool row_product(const TopoDev& t, int B, const double* data,
                                            const double* x, unsigned light_blocks, long& gid,
                                            double& acc) {
  acc = 0.0;
  if (blockIdx.x < light_blocks) {
    gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long)B * t.S) return false;
    const int b = (int)(gid / t.S), s = (int)(gid % t.S);
    const int p0 = t.indptr[s], p1 = t.indptr[s + 1];
    if (p1 - p0 > kHeavyGather) return false;
    const double* db = data + (long)b * t.nnz;
    const double* xb = x + (long)b * t.S;
    for (int p = p0; p < p1; ++p) acc += db[p] * xb[t.indices[p]];
    return true;
  }
  const long w = heavy_item(light_blocks, B, t.n_heavy_a);
  if (w < 0) return false;
  const int b = (int)(w / t.n_heavy_a), s = t.heavy_a[w % t.n_heavy_a];
  gid = (long)b * t.S + s;
  const double* db = data + (long)b * t.nnz;
  const double* xb = x + (long)b * t.S;
  for (int p = t.indptr[s] + (int)(threadIdx.x % 64); p < t.indptr[s + 1]; p += 64)
    acc += db[p] * xb[t.indices[p]];
  acc = wave_sum64(acc);
  return threadIdx.x % 64 == 0;
}

// y[b] = A[b] x[b] for CSR values data [B][nnz] (voltax/sparse.py:130-133).
__global__ void __launch_bounds__(256)
k_spmv(TopoDev t, int B, const double* __restrict__ data, const double* __restrict__ x,
       double* __restrict__ y, unsigned light_blocks) {
  long gid;
  double acc;
  if (row_product(t, B, data, x, light_blocks, gid, acc)) y[gid] = acc;
}

int launch_spmv(const TopoDev& t, int B, const double* data, const double* x, double* y,
                hipStream_t stream) {
  if (B <= 0) return 0;
  const GatherGrid g = gather_grid((long)B * t.S, B, t.n_heavy_a);
  hipLaunchKernelGGL(k_spmv, dim3(g.light + g.heavy), dim3(256), 0, stream, t, B, data, x, y,
                     g.light);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ------------------------------------------------------------- adjoint
// grad[b][poff+k] += scale * lam^T d(cf*f + cq*q)/dp_k  (one thread per device;
// each device owns its parameter slots, so no two threads share an address)
__global__ void __launch_bounds__(256)
k_param_vjp(TopoDev t, int B, const double* __restrict__ params, const double* __restrict__ z,
            double tt, double cf, double cq, const double* __restrict__ lam, double scale,
            double* __restrict__ grad) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)B * t.D) return;
  const int b = (int)(gid / t.D), d = (int)(gid % t.D);
  const int kind = t.kind[d];
  if (kind_is_va(kind)) return;  // k_param_vjp_va
  const int m = kind_m(kind);
  const double* zb = z + (long)b * t.S;
  const double* lb = lam + (long)b * t.S;
  double u[4], l[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int idx = (i < m) ? t.unk[t.uw * d + i] : -1;
    u[i] = idx >= 0 ? zb[idx] : 0.0;
    l[i] = idx >= 0 ? lb[idx] : 0.0;
  }
  const double* p = params + (long)b * t.P + t.poff[d];
  double g[kMaxRowAd];
  const int np = device_row_vjp(kind, t.wave[d], p, u, tt, cf, cq, l, g);
  double* gb = grad + (long)b * t.P + t.poff[d];
  for (int k = 0; k < np; ++k) gb[k] += scale * g[k];
}

// out[b][d] += scale * lam^T d(cf*f_d)/dt for the independent sources (time
// enters the residuals through their waveforms only); other devices add 0.
__global__ void __launch_bounds__(256)
k_time_vjp(TopoDev t, int B, const double* __restrict__ params, double tt, double cf,
           const double* __restrict__ lam, double scale, double* __restrict__ out) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)B * t.D) return;
  const int b = (int)(gid / t.D), d = (int)(gid % t.D);
  const int kind = t.kind[d];
  if (kind != KIND_V && kind != KIND_I) return;
  const double* lb = lam + (long)b * t.S;
  double l[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int idx = (i < kind_m(kind)) ? t.unk[t.uw * d + i] : -1;
    l[i] = idx >= 0 ? lb[idx] : 0.0;
  }
  const double* p = params + (long)b * t.P + t.poff[d];
  out[gid] += scale * device_time_vjp(kind, t.wave[d], p, tt, cf, l);
}

int launch_time_vjp(const TopoDev& t, int B, const double* params, double tt, double cf,
                    const double* lam, double scale, double* out, hipStream_t stream) {
  if (B <= 0 || t.D == 0) return 0;
  hipLaunchKernelGGL(k_time_vjp, dim3(nblocks((long)B * t.D, 256)), dim3(256), 0, stream, t, B,
                     params, tt, cf, lam, scale, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// The same for Verilog-A devices: gradients with respect to every entry of
// the device's parameter row (real instance parameters and the multiplicity,
// physical units), four seeds per interpreter pass.
__global__ void __launch_bounds__(256)
k_param_vjp_va(TopoDev t, int B, const double* __restrict__ params,
               const double* __restrict__ z, double cf, double cq,
               const double* __restrict__ lam, double scale, double* __restrict__ grad) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)B * t.D) return;
  const int b = (int)(gid / t.D), d = (int)(gid % t.D);
  const int kind = t.kind[d];
  if (!kind_is_interpreted(kind)) return;
  const int m = kind_m(kind);
  const double* zb = z + (long)b * t.S;
  const double* lb = lam + (long)b * t.S;
  double u[4], l[4];
  for (int i = 0; i < 4; ++i) {
    const int idx = (i < m) ? t.unk[t.uw * d + i] : -1;
    u[i] = idx >= 0 ? zb[idx] : 0.0;
    l[i] = idx >= 0 ? lb[idx] : 0.0;
  }
  const int* prog = t.code + t.wave[d];
  const int npar = prog[4];
  const double* p = params + (long)b * t.P + t.poff[d];
  double* gb = grad + (long)b * t.P + t.poff[d];
  double res[4], jac[16];
  for (int base = 0; base < npar; base += 4) {
    va_eval<true, true>(prog, p, u, cf, cq, res, jac, base);
    for (int j = 0; j < 4 && base + j < npar; ++j) {
      double acc = 0.0;
      for (int i = 0; i < m; ++i) acc += l[i] * jac[i * 4 + j];
      gb[base + j] += scale * acc;
    }
  }
}

int launch_param_vjp(const TopoDev& t, int B, const double* params, const double* z, double tt,
                     double cf, double cq, const double* lam, double scale, double* grad,
                     hipStream_t stream) {
  if (B <= 0 || t.D == 0) return 0;
  hipLaunchKernelGGL(k_param_vjp, dim3(nblocks((long)B * t.D, 256)), dim3(256), 0, stream, t,
                     B, params, z, tt, cf, cq, lam, scale, grad);
  if (t.has_va)
    hipLaunchKernelGGL(k_param_vjp_va, dim3(nblocks((long)B * t.D, 256)), dim3(256), 0, stream,
                       t, B, params, z, cf, cq, lam, scale, grad);
  if (t.n_gen > 0 && launch_va_generated(t, 3, B, params, z, lam, cf, cq, nullptr, scale,
                                         nullptr, grad, stream))
    return -2;
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// Per device: c_j = sum_i J_loc[i][j] * lam[u_i], written to the residual
// contribution slots, so the row gather forms J^T lam deterministically.
__global__ void __launch_bounds__(256)
k_jt_contrib(TopoDev t, int B, const double* __restrict__ params, const double* __restrict__ z,
             double tt, double cf, double cq, const double* __restrict__ lam,
             double* __restrict__ work) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)B * t.D) return;
  const int b = (int)(gid / t.D), d = (int)(gid % t.D);
  const int kind = t.kind[d];
  if (kind == KIND_VA_GEN) return;  // generated code, mode 2
  const int m = kind_m(kind);
  const double* zb = z + (long)b * t.S;
  const double* lb = lam + (long)b * t.S;
  double u[4], l[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int idx = (i < m) ? t.unk[t.uw * d + i] : -1;
    u[i] = idx >= 0 ? zb[idx] : 0.0;
    l[i] = idx >= 0 ? lb[idx] : 0.0;
  }
  const double* p = params + (long)b * t.P + t.poff[d];
  double res[4], jac[16];
  if (kind_is_va(kind))
    va_eval<true>(t.code + t.wave[d], p, u, cf, cq, res, jac);
  else
    device_eval<true>(kind, t.wave[d], p, u, tt, cf, cq, res, jac);
  double* wr = work + (long)b * (t.Cres + t.Cjac) + t.roff[d];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (j >= m) continue;
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i < m) acc += jac[i * m + j] * l[i];
    wr[j] = acc;
  }
}

int launch_jt_vec(const TopoDev& t, int B, const double* params, const double* z, double tt,
                  double cf, double cq, double gmin, const double* lam, double* out,
                  double* work, hipStream_t stream) {
  if (B <= 0) return 0;
  if (t.D > 0)
    hipLaunchKernelGGL(k_jt_contrib, dim3(nblocks((long)B * t.D, 256)), dim3(256), 0, stream,
                       t, B, params, z, tt, cf, cq, lam, work);
  if (t.n_gen > 0 && launch_va_generated(t, 2, B, params, z, lam, cf, cq, nullptr, 0.0, work,
                                         nullptr, stream))
    return -2;
  // out = gathered contributions + gmin*lam on node rows (gmin*I is symmetric)
  launch_gather_res(t, B, work, lam, nullptr, gmin, out, nullptr, stream);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ---------------------------------------------- refinement / Newton helpers
__global__ void __launch_bounds__(256)
k_residual(TopoDev t, int B, const double* __restrict__ data, const double* __restrict__ x,
           const double* __restrict__ b, double* __restrict__ r, unsigned light_blocks) {
  long gid;
  double acc;
  if (row_product(t, B, data, x, light_blocks, gid, acc)) r[gid] = b[gid] - acc;
}

int launch_residual(const TopoDev& t, int B, const double* data, const double* x,
                    const double* b, double* r, hipStream_t st) {
  if (B <= 0) return 0;
  const GatherGrid g = gather_grid((long)B * t.S, B, t.n_heavy_a);
  hipLaunchKernelGGL(k_residual, dim3(g.light + g.heavy), dim3(256), 0, st, t, B, data, x, b, r,
                     g.light);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

__global__ void __launch_bounds__(256)
k_axpy(long n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

int launch_axpy(int64_t n, double a, const double* x, double* y, hipStream_t st) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_axpy, dim3(nblocks(n, 256)), dim3(256), 0, st, (long)n, a, x, y);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// One workgroup per instance: max |dz| over node rows -> damping scale,
// z += scale*dz (if the instance is active), convergence flag.
__global__ void __launch_bounds__(256)
k_newton_update(int S, int nv, double* __restrict__ z, const double* __restrict__ dz,
                double sgn, NewtonOpts o, const int8_t* __restrict__ active,
                int8_t* __restrict__ done, double* __restrict__ step_norm) {
  const int b = blockIdx.x;
  double* zb = z + (long)b * S;
  const double* db = dz + (long)b * S;
  __shared__ double sm[4];
  __shared__ int si[4];
  const int tid = threadIdx.x;
  double vmax = 0.0;
  for (int i = tid; i < nv; i += blockDim.x) vmax = fmax(vmax, fabs(db[i]));
  for (int off = 32; off > 0; off >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, off));
  if ((tid & 63) == 0) sm[tid >> 6] = vmax;
  __syncthreads();
  vmax = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
  const double scale = fmin(1.0, o.max_dv / (vmax + 1e-300));
  if (step_norm && tid == 0 && (active == nullptr || active[b])) step_norm[b] = vmax;
  int ok = 1;
  const bool act = active == nullptr || active[b];
  for (int i = tid; i < S; i += blockDim.x) {
    const double d = sgn * db[i];
    const double zn = zb[i] + scale * d;
    ok &= (fabs(d) <= o.atol + o.rtol * fabs(zn)) && isfinite(zn);
    if (act) zb[i] = zn;
  }
  for (int off = 32; off > 0; off >>= 1) ok &= __shfl_xor(ok, off);
  if ((tid & 63) == 0) si[tid >> 6] = ok;
  __syncthreads();
  if (tid == 0) done[b] = (si[0] & si[1] & si[2] & si[3]) && scale == 1.0;
}

// The update of an iteration whose linear solve may have used the factors of
// an earlier iteration of the same solve (`reused`).  w = max_i |d_i| /
// (atol + rtol |z_i|) is the tolerance-weighted size of the update; a reused-
// factor update converges linearly with rate theta = w / wnorm (the previous
// applied update's w) and ends the iteration only when it passes the plain
// test and theta w <= 1e-4 (1 - theta); theta > 1/2 asks for fresh factors
// (refactor = 1) and theta >= 1 also drops the update.  A fresh update is
// tested as in k_newton_update.  wnorm keeps the last applied update's w.
__global__ void __launch_bounds__(256)
k_newton_update_reuse(int S, int nv, double* __restrict__ z, const double* __restrict__ dz,
                      double sgn, NewtonOpts o, const int8_t* __restrict__ active,
                      int8_t* __restrict__ done, double* __restrict__ wnorm,
                      int8_t* __restrict__ refactor, int reused) {
  const int b = blockIdx.x;
  double* zb = z + (long)b * S;
  const double* db = dz + (long)b * S;
  __shared__ double sm[4], sw[4];
  __shared__ int si[4];
  const int tid = threadIdx.x;
  double vmax = 0.0, w = 0.0;
  for (int i = tid; i < S; i += blockDim.x) {
    const double d = fabs(db[i]);
    if (i < nv) vmax = fmax(vmax, d);
    const double wi = d / (o.atol + o.rtol * fabs(zb[i]));
    w = (wi > w || wi != wi) ? wi : w;  // keeps a NaN
  }
  for (int off = 32; off > 0; off >>= 1) {
    vmax = fmax(vmax, __shfl_xor(vmax, off));
    const double wo = __shfl_xor(w, off);
    w = (wo > w || wo != wo) ? wo : w;
  }
  if ((tid & 63) == 0) {
    sm[tid >> 6] = vmax;
    sw[tid >> 6] = w;
  }
  __syncthreads();
  vmax = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
  w = sw[0];
#pragma unroll
  for (int k = 1; k < 4; ++k) w = (sw[k] > w || sw[k] != sw[k]) ? sw[k] : w;
  const double scale = fmin(1.0, o.max_dv / (vmax + 1e-300));
  const bool act = active == nullptr || active[b];
  bool apply = true, extra = true, slow = false;
  if (reused) {
    const double theta = w / wnorm[b];
    slow = !(theta <= 0.5);
    extra = !slow && theta * w <= 1e-4 * (1.0 - theta);
    apply = theta < 1.0;
  }
  int ok = 1;
  for (int i = tid; i < S; i += blockDim.x) {
    const double d = sgn * db[i];
    const double zn = zb[i] + scale * d;
    ok &= (fabs(d) <= o.atol + o.rtol * fabs(zn)) && isfinite(zn);
    if (act && apply) zb[i] = zn;
  }
  for (int off = 32; off > 0; off >>= 1) ok &= __shfl_xor(ok, off);
  if ((tid & 63) == 0) si[tid >> 6] = ok;
  __syncthreads();
  if (tid == 0) {
    done[b] = (si[0] & si[1] & si[2] & si[3]) && scale == 1.0 && extra;
    refactor[b] = (act && slow) ? 1 : 0;
    if (act && apply) wnorm[b] = w;
  }
}

int launch_newton_update_reuse(int B, int S, int nv, double* z, const double* dz,
                               double dz_sign, const NewtonOpts& o, const int8_t* active,
                               int8_t* done, double* wnorm, int8_t* refactor, int reused,
                               hipStream_t st) {
  if (B <= 0) return 0;
  hipLaunchKernelGGL(k_newton_update_reuse, dim3(B), dim3(256), 0, st, S, nv, z, dz, dz_sign,
                     o, active, done, wnorm, refactor, reused);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_newton_update(int B, int S, int nv, double* z, const double* dz, double dz_sign,
                         const NewtonOpts& o, const int8_t* active, int8_t* done,
                         double* step_norm, hipStream_t st) {
  if (B <= 0) return 0;
  hipLaunchKernelGGL(k_newton_update, dim3(B), dim3(256), 0, st, S, nv, z, dz, dz_sign, o,
                     active, done, step_norm);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace vxm

// --------------------------------------------------------------- small signal
namespace vxm {
namespace {
// Complex systems of the AC analysis for F frequencies (voltax/analysis.py:
// 510-520): data[f][e] = g[e] + j 2 pi freqs[f] c[e] on the pattern,
// rhs[f] = -b; complex values as (re, im) pairs.
__global__ void k_ac_system(int nnz, int S, int F, const double* __restrict__ g,
                            const double* __restrict__ c, const double* __restrict__ freqs,
                            const double* __restrict__ b, double* __restrict__ data,
                            double* __restrict__ rhs) {
  const int64_t n1 = (int64_t)F * nnz, n2 = (int64_t)F * S;
  const double two_pi = 6.283185307179586;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n1 + n2;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i < n1) {
      const int f = static_cast<int>(i / nnz), e = static_cast<int>(i - (int64_t)f * nnz);
      data[2 * i] = g[e];
      data[2 * i + 1] = two_pi * freqs[f] * c[e];
    } else {
      const int64_t k = i - n1;
      const int s = static_cast<int>(k % S);
      rhs[2 * k] = -b[2 * s];
      rhs[2 * k + 1] = -b[2 * s + 1];
    }
  }
}
}  // namespace

int launch_ac_system(int nnz, int S, int F, const double* g, const double* c,
                     const double* freqs, const double* b, double* data, double* rhs,
                     hipStream_t st) {
  const int64_t n = (int64_t)F * (nnz + S);
  if (n == 0) return 0;
  const int blocks = static_cast<int>(std::min<int64_t>((n + 255) / 256, 65536));
  hipLaunchKernelGGL(k_ac_system, dim3(blocks), dim3(256), 0, st, nnz, S, F, g, c, freqs, b,
                     data, rhs);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
}  // namespace vxm
