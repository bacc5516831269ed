// BF16 MFMA kernels of the learned Newton warm start: a multilayer perceptron
// trained on the GPU (forward, backward, Adam) and evaluated in batches to
// predict Newton initial guesses.
//
// Storage: activations, weights and their transposes in bf16 (round to nearest
// even), accumulation and master weights in fp32.  Every product is the one
// "NT" GEMM below, C[M,N] = A[M,K] . B[N,K]^T with both operands K-contiguous;
// the backward pass gets its operands in that form from the transpose kernel.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "status.hpp"

namespace vxm {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) {
  return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

struct Epilogue {
  float alpha;            // C = alpha * A.B^T (+ bias[n]) (+ beta * C_in, fp32 C only)
  float beta;
  const float* bias;      // [N] or null
  const uint16_t* mask;   // [M][ldm] bf16: zero C where mask <= 0 (ReLU derivative), or null
  int ldm;
  int relu;               // C = max(C, 0)
  float* c;               // fp32 output [M][ldc] or null
  int ldc;
  uint16_t* cb;           // bf16 output [M][ldcb] or null
  int ldcb;
};

// Block tile BM x BN, four waves in a 2 x 2 grid, K in steps of 32 through two
// LDS buffers (global loads of step k+1 are in flight while step k multiplies).
// v_mfma_f32_16x16x32_bf16: lane l supplies row (l & 15) of each 16-row operand
// with its K elements 8*(l >> 4) .. +7; the product's lane l holds column
// (l & 15) of rows 4*(l >> 4) .. +3.  The weights operand goes first so that a
// lane ends up with four consecutive output columns of one row (vector stores).
template <int BM, int BN>
__global__ __launch_bounds__(256) void k_gemm_nt(const uint16_t* __restrict__ A, int lda,
                                                 const uint16_t* __restrict__ B, int ldb, int K,
                                                 Epilogue e) {
  constexpr int BK = 32, LD = BK + 8;  // LDS row pitch (bf16): 80 bytes, 16-byte aligned
  constexpr int WM = BM / 2, WN = BN / 2, TI = WM / 16, TJ = WN / 16;
  constexpr int CA = BM * 4 / 256, CB = BN * 4 / 256;  // 16-byte chunks per thread
  __shared__ __attribute__((aligned(16))) uint16_t sA[2][BM * LD];
  __shared__ __attribute__((aligned(16))) uint16_t sB[2][BN * LD];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  static_assert(CA >= 1 && CA <= 2 && CB >= 1 && CB <= 2, "one or two chunks per thread");
  // Thread tid moves 16-byte chunks tid and tid + 256 of each operand tile
  // (chunk c: row c / 4, K offset 8 (c % 4)); held in named registers so that
  // nothing is indexed at run time.
  const int c0 = tid, c1 = tid + 256;
  const uint16_t* pa0 = A + (size_t)(m0 + (c0 >> 2)) * lda + (c0 & 3) * 8;
  const uint16_t* pa1 = A + (size_t)(m0 + (c1 >> 2) % BM) * lda + (c1 & 3) * 8;
  const uint16_t* pb0 = B + (size_t)(n0 + (c0 >> 2)) * ldb + (c0 & 3) * 8;
  const uint16_t* pb1 = B + (size_t)(n0 + (c1 >> 2) % BN) * ldb + (c1 & 3) * 8;
  const int sa0 = (c0 >> 2) * LD + (c0 & 3) * 8, sa1 = ((c1 >> 2) % BM) * LD + (c1 & 3) * 8;
  const int sb0 = (c0 >> 2) * LD + (c0 & 3) * 8, sb1 = ((c1 >> 2) % BN) * LD + (c1 & 3) * 8;
  uint4 ra0, ra1, rb0, rb1;
#define VXM_GLOAD(k0)                                                   \
  ra0 = *reinterpret_cast<const uint4*>(pa0 + (k0));                    \
  if constexpr (CA == 2) ra1 = *reinterpret_cast<const uint4*>(pa1 + (k0)); \
  rb0 = *reinterpret_cast<const uint4*>(pb0 + (k0));                    \
  if constexpr (CB == 2) rb1 = *reinterpret_cast<const uint4*>(pb1 + (k0));
#define VXM_SSTORE(buf)                                                 \
  *reinterpret_cast<uint4*>(&sA[buf][sa0]) = ra0;                       \
  if constexpr (CA == 2) *reinterpret_cast<uint4*>(&sA[buf][sa1]) = ra1; \
  *reinterpret_cast<uint4*>(&sB[buf][sb0]) = rb0;                       \
  if constexpr (CB == 2) *reinterpret_cast<uint4*>(&sB[buf][sb1]) = rb1;

  f32x4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int nk = K / BK;
  VXM_GLOAD(0)
  VXM_SSTORE(0)
  __syncthreads();
  const int r = lane & 15, kg = (lane >> 4) * 8;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) {
      VXM_GLOAD((kt + 1) * BK)
    }
    bf16x8 af[TI], bw[TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
      af[i] = *reinterpret_cast<const bf16x8*>(&sA[cur][(wm * WM + i * 16 + r) * LD + kg]);
#pragma unroll
    for (int j = 0; j < TJ; ++j)
      bw[j] = *reinterpret_cast<const bf16x8*>(&sB[cur][(wn * WN + j * 16 + r) * LD + kg]);
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[j], af[i], acc[i][j], 0, 0, 0);
    if (kt + 1 < nk) {
      VXM_SSTORE(cur ^ 1)
    }
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < TI; ++i) {
    const int m = m0 + wm * WM + i * 16 + r;
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      const int n = n0 + wn * WN + j * 16 + (lane >> 4) * 4;
      float v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = e.alpha * acc[i][j][q] + (e.bias ? e.bias[n + q] : 0.f);
      if (e.beta != 0.f) {
        const float4 c0 = *reinterpret_cast<const float4*>(e.c + (size_t)m * e.ldc + n);
        v[0] += e.beta * c0.x;
        v[1] += e.beta * c0.y;
        v[2] += e.beta * c0.z;
        v[3] += e.beta * c0.w;
      }
      if (e.mask) {
        const uint2 mk = *reinterpret_cast<const uint2*>(e.mask + (size_t)m * e.ldm + n);
        const uint16_t mh[4] = {static_cast<uint16_t>(mk.x & 0xFFFF),
                                static_cast<uint16_t>(mk.x >> 16),
                                static_cast<uint16_t>(mk.y & 0xFFFF),
                                static_cast<uint16_t>(mk.y >> 16)};
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (!(bf2f(mh[q]) > 0.f)) v[q] = 0.f;
      }
      if (e.relu)
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = fmaxf(v[q], 0.f);
      if (e.c)
        *reinterpret_cast<float4*>(e.c + (size_t)m * e.ldc + n) = make_float4(v[0], v[1], v[2], v[3]);
      if (e.cb) {
        uint2 o;
        o.x = f2bf(v[0]) | (static_cast<uint32_t>(f2bf(v[1])) << 16);
        o.y = f2bf(v[2]) | (static_cast<uint32_t>(f2bf(v[3])) << 16);
        *reinterpret_cast<uint2*>(e.cb + (size_t)m * e.ldcb + n) = o;
      }
    }
  }
#undef VXM_GLOAD
#undef VXM_SSTORE
}

// dst[c][r] = src[r][c] through a 64 x 64 LDS tile.
__global__ __launch_bounds__(256) void k_transpose_bf16(int R, int Cc, const uint16_t* src,
                                                        int lds, uint16_t* dst, int ldd) {
  __shared__ uint16_t tile[64][65];
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int i = ty; i < 64; i += 4) {
    const int rr = r0 + i, cc = c0 + tx;
    tile[i][tx] = (rr < R && cc < Cc) ? src[(size_t)rr * lds + cc] : 0;
  }
  __syncthreads();
  for (int i = ty; i < 64; i += 4) {
    const int cc = c0 + i, rr = r0 + tx;
    if (cc < Cc && rr < R) dst[(size_t)cc * ldd + rr] = tile[tx][i];
  }
}

// Deterministic tree sum of one value per thread of a 256-thread block.
__device__ float block_sum(float v, float* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const float out = sh[0];
  __syncthreads();
  return out;
}

// out[r] = sum_c src[r][c] (bias gradients from the transposed output gradient).
__global__ __launch_bounds__(256) void k_rowsum_bf16(int Cc, const uint16_t* src, int ld,
                                                     float* out) {
  __shared__ float sh[256];
  const uint16_t* row = src + (size_t)blockIdx.x * ld;
  float s = 0.f;
  for (int c = threadIdx.x; c < Cc; c += 256) s += bf2f(row[c]);
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// Mean-squared-error gradient: g = scale (y - t) in bf16 for the first m_valid
// rows and n_valid columns (zero elsewhere), loss_rows[m] = sum_c (y - t)^2.
__global__ __launch_bounds__(256) void k_mse_grad(int m_valid, int N, int n_valid,
                                                  const float* y, int ldy, const float* t,
                                                  int ldt, float scale, uint16_t* g, int ldg,
                                                  float* loss_rows) {
  __shared__ float sh[256];
  const int m = blockIdx.x;
  float s = 0.f;
  for (int c = threadIdx.x; c < N; c += 256) {
    float d = 0.f;
    if (m < m_valid && c < n_valid) d = y[(size_t)m * ldy + c] - t[(size_t)m * ldt + c];
    s += d * d;
    g[(size_t)m * ldg + c] = f2bf(scale * d);
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0 && loss_rows) loss_rows[m] = s;
}

// Adam (Kingma & Ba 2015) with decoupled weight decay wd (Loshchilov & Hutter
// 2019) on fp32 master weights, then the bf16 copy used by the GEMMs.
// bc1 = 1 - b1^step, bc2 = 1 - b2^step.
__global__ void k_adam(int64_t n, float* w, const float* g, float* m1, float* m2, float lr,
                       float b1, float b2, float eps, float wd, float bc1, float bc2,
                       uint16_t* wb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float wi = w[i] - lr * ((a / bc1) / (sqrtf(b / bc2) + eps) + wd * w[i]);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

// out[m][c] = (x[m][c] - mu[c]) * inv_sd[c] for c < C, 0 for C <= c < Cpad.
__global__ void k_normalize(int M, int Cc, int Cpad, const double* x, int ldx, const double* mu,
                            const double* inv_sd, uint16_t* ob, float* of, int ldo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)M * Cpad) return;
  const int m = static_cast<int>(i / Cpad), c = static_cast<int>(i % Cpad);
  float v = 0.f;
  if (c < Cc) v = static_cast<float>((x[(size_t)m * ldx + c] - mu[c]) * inv_sd[c]);
  if (ob) ob[(size_t)m * ldo + c] = f2bf(v);
  if (of) of[(size_t)m * ldo + c] = v;
}

// out[m][c] = y[m][c] * sd[c] + mu[c] (FP64 initial guesses from the network).
__global__ void k_denormalize(int M, int Cc, const float* y, int ldy, const double* mu,
                              const double* sd, double* out, int ldo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)M * Cc) return;
  const int m = static_cast<int>(i / Cc), c = static_cast<int>(i % Cc);
  out[(size_t)m * ldo + c] = static_cast<double>(y[(size_t)m * ldy + c]) * sd[c] + mu[c];
}

inline int launch_status() { return hipGetLastError() == hipSuccess ? VXM_OK : VXM_ERR_HIP; }

}  // namespace
}  // namespace vxm

using namespace vxm;

extern "C" {

int vxm_gemm_bf16(int32_t m, int32_t n, int32_t k, const uint16_t* a, int32_t lda,
                  const uint16_t* b, int32_t ldb, float alpha, float beta, const float* bias,
                  const uint16_t* mask, int32_t ldm, int32_t relu, float* c, int32_t ldc,
                  uint16_t* cb, int32_t ldcb, vxm_stream stream) {
  VXM_TRY
  if (m <= 0 || n <= 0 || k <= 0 || m % 64 || n % 64 || k % 32 || !a || !b || (!c && !cb) ||
      (beta != 0.f && !c))
    return VXM_ERR_INVALID;
  if (lda < k || ldb < k || lda % 8 || ldb % 8) return VXM_ERR_INVALID;
  if ((c && (ldc < n || ldc % 4)) || (cb && (ldcb < n || ldcb % 4)) ||
      (mask && (ldm < n || ldm % 4)))
    return VXM_ERR_INVALID;
  Epilogue e{alpha, beta, bias, mask, ldm, relu, c, ldc, cb, ldcb};
  hipStream_t s = static_cast<hipStream_t>(stream);
  // 128 x 128 tiles when they still give every CU a workgroup, else 64 x 64
  if (m % 128 == 0 && n % 128 == 0 && (int64_t)(m / 128) * (n / 128) >= 256)
    hipLaunchKernelGGL((k_gemm_nt<128, 128>), dim3(n / 128, m / 128), dim3(256), 0, s, a, lda, b,
                       ldb, k, e);
  else
    hipLaunchKernelGGL((k_gemm_nt<64, 64>), dim3(n / 64, m / 64), dim3(256), 0, s, a, lda, b,
                       ldb, k, e);
  return launch_status();
  VXM_CATCH
}

int vxm_transpose_bf16(int32_t rows, int32_t cols, const uint16_t* src, int32_t lds,
                       uint16_t* dst, int32_t ldd, vxm_stream stream) {
  VXM_TRY
  if (rows <= 0 || cols <= 0 || !src || !dst || lds < cols || ldd < rows) return VXM_ERR_INVALID;
  hipLaunchKernelGGL(k_transpose_bf16, dim3((cols + 63) / 64, (rows + 63) / 64), dim3(256), 0,
                     static_cast<hipStream_t>(stream), rows, cols, src, lds, dst, ldd);
  return launch_status();
  VXM_CATCH
}

int vxm_rowsum_bf16(int32_t rows, int32_t cols, const uint16_t* src, int32_t ld, float* out,
                    vxm_stream stream) {
  VXM_TRY
  if (rows <= 0 || cols <= 0 || !src || !out || ld < cols) return VXM_ERR_INVALID;
  hipLaunchKernelGGL(k_rowsum_bf16, dim3(rows), dim3(256), 0, static_cast<hipStream_t>(stream),
                     cols, src, ld, out);
  return launch_status();
  VXM_CATCH
}

int vxm_mse_grad(int32_t m, int32_t m_valid, int32_t n, int32_t n_valid, const float* y,
                 int32_t ldy, const float* t, int32_t ldt, float scale, uint16_t* g,
                 int32_t ldg, float* loss_rows, vxm_stream stream) {
  VXM_TRY
  if (m <= 0 || n <= 0 || m_valid < 0 || m_valid > m || n_valid < 0 || n_valid > n || !y ||
      !t || !g || ldy < n_valid || ldt < n_valid || ldg < n)
    return VXM_ERR_INVALID;
  hipLaunchKernelGGL(k_mse_grad, dim3(m), dim3(256), 0, static_cast<hipStream_t>(stream),
                     m_valid, n, n_valid, y, ldy, t, ldt, scale, g, ldg, loss_rows);
  return launch_status();
  VXM_CATCH
}

int vxm_adam(int64_t n, float* w, const float* g, float* m1, float* m2, float lr, float b1,
             float b2, float eps, float wd, int32_t step, uint16_t* wb, vxm_stream stream) {
  VXM_TRY
  if (n <= 0 || !w || !g || !m1 || !m2 || step < 1 || !(b1 >= 0.f && b1 < 1.f) ||
      !(b2 >= 0.f && b2 < 1.f))
    return VXM_ERR_INVALID;
  const float bc1 = 1.f - powf(b1, static_cast<float>(step));
  const float bc2 = 1.f - powf(b2, static_cast<float>(step));
  hipLaunchKernelGGL(k_adam, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), n, w, g, m1, m2, lr, b1, b2, eps, wd,
                     bc1, bc2, wb);
  return launch_status();
  VXM_CATCH
}

int vxm_normalize(int32_t rows, int32_t cols, int32_t cols_pad, const double* x, int32_t ldx,
                  const double* mu, const double* inv_sd, uint16_t* out_bf16, float* out_f32,
                  int32_t ldo, vxm_stream stream) {
  VXM_TRY
  if (rows <= 0 || cols <= 0 || cols_pad < cols || !x || !mu || !inv_sd || ldx < cols ||
      ldo < cols_pad || (!out_bf16 && !out_f32))
    return VXM_ERR_INVALID;
  const int64_t total = (int64_t)rows * cols_pad;
  hipLaunchKernelGGL(k_normalize, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), rows, cols, cols_pad, x, ldx, mu, inv_sd,
                     out_bf16, out_f32, ldo);
  return launch_status();
  VXM_CATCH
}

int vxm_denormalize(int32_t rows, int32_t cols, const float* y, int32_t ldy, const double* mu,
                    const double* sd, double* out, int32_t ldo, vxm_stream stream) {
  VXM_TRY
  if (rows <= 0 || cols <= 0 || !y || !mu || !sd || !out || ldy < cols || ldo < cols)
    return VXM_ERR_INVALID;
  const int64_t total = (int64_t)rows * cols;
  hipLaunchKernelGGL(k_denormalize, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256),
                     0, static_cast<hipStream_t>(stream), rows, cols, y, ldy, mu, sd, out, ldo);
  return launch_status();
  VXM_CATCH
}

}  // extern "C"
