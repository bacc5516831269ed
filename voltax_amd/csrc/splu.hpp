// Device view of a supernodal plan and the numeric entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "kernels.hpp"
#include "splu_plan.hpp"

namespace vxm {

struct SPlanDev {
  int S, nnz, ns, piv_total, r_total;
  int64_t front_total;
  const SNode* sn;
  const int *idx, *cmap, *children, *aslot, *apos, *level_sn;
};

// Fronts larger than 64 unknowns and up to this many bytes are eliminated in
// LDS by one workgroup (all of a CU's 160 KB); larger ones in place in HBM.
constexpr size_t kFrontLdsBytes = 160 * 1024;

// Element types of factors: FP64, FP32 (values read as FP64, solves in FP64),
// complex FP64 (values, factors and solves complex, interleaved re/im).
constexpr int kSpluF64 = 0, kSpluF32 = 1, kSpluC128 = 2;

int splu_factor(const SPlanHost& h, const SPlanDev& d, int B, const void* data, void* front,
                int* piv, int dtype, hipStream_t st);
// FP64 factorization and solve of A x = b in place (x holds b on entry): the
// forward sweep of register fronts fused into their factorization.  uvec:
// B * r_total doubles.
int splu_factor_solve(const SPlanHost& h, const SPlanDev& d, int B, const double* data,
                      double* front, int* piv, double* x, double* uvec, hipStream_t st);
// uvec: B * r_total entries of the solve's element type (double or complex)
int splu_solve(const SPlanHost& h, const SPlanDev& d, int B, const void* front, const int* piv,
               void* x, void* uvec, bool transpose, int dtype, hipStream_t st);

// r = b - A x (FP64) for B instances of one pattern
int launch_residual(const TopoDev& t, int B, const double* data, const double* x,
                    const double* b, double* r, hipStream_t st);
// y += a * x over n entries
int launch_axpy(int64_t n, double a, const double* x, double* y, hipStream_t st);
// Damped Newton update + convergence test per instance (voltax/analysis.py:224-230)
int launch_newton_update_reuse(int B, int S, int nv, double* z, const double* dz,
                               double dz_sign, const NewtonOpts& o, const int8_t* active,
                               int8_t* done, double* wnorm, int8_t* refactor, int reused,
                               hipStream_t st);
int launch_newton_update(int B, int S, int nv, double* z, const double* dz, double dz_sign,
                         const NewtonOpts& o, const int8_t* active, int8_t* done,
                         double* step_norm, hipStream_t stream);

}  // namespace vxm
