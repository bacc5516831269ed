// Host dispatch of the fused dense kernel over its tile classes.
#include "kernels.hpp"

namespace vxm {

int dense_class(int S) {
  for (int c = 0; c < kDenseClasses; ++c)
    if (S <= kDenseSmax[c]) return c;
  return -1;
}

int launch_dense_run(const PatternDev* pats, const WaveItem* items, const int* begin,
                     const int* count, const int* lds_bytes, const int64_t* ipoff,
                     const int64_t* izoff, const double* params, const double* z0,
                     const double* ts, int T, int method, int do_dc, double t_dc,
                     const NewtonOpts& o, int precision, double* out_z, int8_t* out_conv,
                     int32_t* out_iters, hipStream_t stream) {
  typedef int (*Fn)(const PatternDev*, const WaveItem*, int, int, const int64_t*,
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
    int st = fns[c](pats, items + begin[c], count[c], lds_bytes[c], ipoff, izoff, params, z0,
                    ts, T, method, do_dc, t_dc, o, precision, out_z, out_conv, out_iters,
                    stream);
    if (st != 0) return st;
  }
  return 0;
}

}  // namespace vxm
