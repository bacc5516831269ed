// Host dispatch of the fused dense kernel over its tile classes, and the
// per-run derived parameter rows its device evaluations read.
#include <algorithm>

#include "devices.hpp"
#include "kernels.hpp"

namespace vxm {

__global__ void k_copy_params(long n, const double* __restrict__ src, double* __restrict__ dst) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// One block per wave item: the rows of every device of its instances.
__global__ void __launch_bounds__(64)
k_derive_params(const PatternDev* __restrict__ pats, const WaveItem* __restrict__ items,
                const int64_t* __restrict__ ipoff, const double* __restrict__ params,
                double* __restrict__ dparams) {
  const WaveItem& it = items[blockIdx.x];
  const TopoDev& t = pats[it.pattern].t;
  for (int k = 0; k < it.ninst; ++k) {
    const int64_t o = ipoff[it.inst[k]];
    for (int d = threadIdx.x; d < t.D; d += 64)
      derive_device(t.kind[d], params + o + t.poff[d], dparams + o + t.poff[d]);
  }
}

int launch_derive_params(const PatternDev* pats, const WaveItem* items, int n_items,
                         const int64_t* ipoff, int64_t n_par, const double* params,
                         double* dparams, hipStream_t stream) {
  if (n_par <= 0 || n_items <= 0) return 0;
  const long blocks = std::min<long>((n_par + 255) / 256, 4096);
  hipLaunchKernelGGL(k_copy_params, dim3(blocks), dim3(256), 0, stream, (long)n_par, params,
                     dparams);
  hipLaunchKernelGGL(k_derive_params, dim3(n_items), dim3(64), 0, stream, pats, items, ipoff,
                     params, dparams);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int dense_class(int S) {
  for (int c = 0; c < kDenseClasses; ++c)
    if (S <= kDenseSmax[c]) return c;
  return -1;
}

int launch_dense_run(const PatternDev* pats, const WaveItem* items, const int* begin,
                     const int* count, const int* count_band, const int* lds_bytes,
                     const int64_t* ipoff,
                     const int64_t* izoff, const double* params, const double* z0,
                     const double* ts, int T, int method, int do_dc, double t_dc,
                     const NewtonOpts& o, int precision, double* out_z, int8_t* out_conv,
                     int32_t* out_iters, hipStream_t stream) {
  typedef int (*Fn)(const PatternDev*, const WaveItem*, int, int, int, const int64_t*,
                    const int64_t*, const double*, const double*, const double*, int, int, int,
                    double, const NewtonOpts&, int, double*, int8_t*, int32_t*, hipStream_t);
  static const Fn fns[kDenseClasses] = {
      launch_dense_class<8, 8>,   launch_dense_class<12, 16>, launch_dense_class<16, 16>,
      launch_dense_class<20, 32>, launch_dense_class<24, 32>, launch_dense_class<28, 32>,
      launch_dense_class<32, 32>, launch_dense_class<40, 64>, launch_dense_class<48, 64>,
      launch_dense_class<56, 64>, launch_dense_class<64, 64>};
  // largest tiles first: they run longest
  for (int c = kDenseClasses - 1; c >= 0; --c) {
    if (count[c] == 0) continue;
    int st = fns[c](pats, items + begin[c], count[c], count_band[c], lds_bytes[c], ipoff,
                    izoff, params, z0, ts, T, method, do_dc, t_dc, o, precision, out_z,
                    out_conv, out_iters, stream);
    if (st != 0) return st;
  }
  return 0;
}

}  // namespace vxm
