// Batched small dense solves x = A^-1 b (or A^-T b), n <= 64, FP64: the dense
// linear solve of the reference's non-sparse path (`jnp.linalg.solve`, LU with
// partial pivoting; voltax/analysis.py:201-205) for callers with a matrix of
// their own (solve_root's Newton steps, linearized systems).
//
// One segment of L = SMAX lanes per system, lane r holding row r of A (column
// r for the transpose) in registers; the factorization and substitutions are
// the row-per-lane ones of the fused Newton kernel (dense_newton.hpp).  Rows
// n..SMAX-1 and systems past the batch are identity rows, so every lane of the
// wavefront takes part in every step and reaches every barrier.
#include <hip/hip_runtime.h>

#include "dense_newton.hpp"
#include "status.hpp"

namespace vxm {

template <int SMAX>
__global__ void __launch_bounds__(64)
k_dense_solve(int B, int n, const double* __restrict__ A, const double* b,
              double* x, int transpose) {
  constexpr int L = SMAX, PER = 64 / L;
  __shared__ int pivl[PER][SMAX];
  const int lane = threadIdx.x & 63, seg = lane / L, r = lane % L;
  const long sys = (long)blockIdx.x * PER + seg;
  const bool live = sys < B;
  const double* As = A + sys * (long)n * n;
  double a[SMAX];
#pragma unroll
  for (int j = 0; j < SMAX; ++j) {
    double v = (j == r) ? 1.0 : 0.0;
    if (live && r < n) v = (j < n) ? (transpose ? As[(long)j * n + r] : As[(long)r * n + j]) : 0.0;
    a[j] = v;
  }
  double rhs = (live && r < n) ? b[sys * n + r] : 0.0;
  int myk;
  double dinv;
  lu_factor<SMAX, L, double, true>(a, r, pivl[seg], myk, rhs, dinv);
  const double xv = lu_backsolve<SMAX, L, double>(a, rhs, r, pivl[seg], myk, dinv);
  if (live && r < n) x[sys * n + r] = xv;
}

template <int SMAX>
static int launch(int B, int n, const double* A, const double* b, double* x, int tr,
                  hipStream_t s) {
  constexpr int PER = 64 / SMAX;
  const int blocks = (int)((B + PER - 1) / PER);
  hipLaunchKernelGGL(k_dense_solve<SMAX>, dim3(blocks), dim3(64), 0, s, B, n, A, b, x, tr);
  return hipGetLastError() == hipSuccess ? VXM_OK : VXM_ERR_HIP;
}

}  // namespace vxm

extern "C" int vxm_dense_solve(int32_t batch, int32_t n, const double* A, const double* b,
                               double* x, int32_t transpose, vxm_stream stream) {
  VXM_TRY
  if (batch < 0 || n < 1 || n > 64 || !A || !b || !x) return VXM_ERR_INVALID;
  if (batch == 0) return VXM_OK;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int tr = transpose ? 1 : 0;
  if (n <= 8) return vxm::launch<8>(batch, n, A, b, x, tr, s);
  if (n <= 16) return vxm::launch<16>(batch, n, A, b, x, tr, s);
  if (n <= 32) return vxm::launch<32>(batch, n, A, b, x, tr, s);
  return vxm::launch<64>(batch, n, A, b, x, tr, s);
  VXM_CATCH
}
