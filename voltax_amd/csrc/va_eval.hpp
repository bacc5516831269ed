// Interpreter of compiled Verilog-A device programs (voltax_amd/va.py).
//
// A program is a register bytecode over dual numbers: each register holds a
// value and its derivatives with respect to the device's m <= 4 unknowns
// (terminal voltages, internal nodes, branch currents).  Layout (int32):
//   [magic, m, n_regs, n_ins, n_params, n_terminals] [n_ins x (op, dst, a, b, c)]
//   [2m outputs]; the first n_terminals unknowns are terminal voltages
// Outputs are the registers of the m residual rows of f (currents) and of q
// (charges).  Constants carry their double bits in (a, b).  Data-dependent
// branches were compiled into selects, which take value and derivatives
// from the chosen side only (a NaN or inf on the discarded side never
// leaks, as with the reference's guarded where).  Every program is checked
// on the host (vxm_topology_set_code) before it can reach this code, so the
// register, parameter and unknown indices below are in range.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "vxm_types.h"

namespace vxm {

enum VAOp : int {
  VA_CONST, VA_PARAM, VA_UNK, VA_ADD, VA_SUB, VA_MUL, VA_DIV, VA_NEG, VA_EXP, VA_LOG,
  VA_SQRT, VA_ABS, VA_POW, VA_MIN, VA_MAX, VA_LIMEXP, VA_TANH, VA_SIN, VA_COS, VA_LT,
  VA_LE, VA_GT, VA_GE, VA_EQ, VA_NE, VA_AND, VA_OR, VA_NOT, VA_SELECT, VA_TRUNC, VA_ROUND,
  VA_LIMIT, VA_NOPS
};
constexpr int kVAMagic = 0x5642, kVAMaxRegs = 64, kVAHead = 6, kVAInsWords = 5;
constexpr double kVALimexpMax = 80.0;  // voltax/va/runtime.py:49-56

// JAC: also the partials.  The seeds are the m unknowns (jac[i*m + j] =
// d res_i / d u_j), or with PSEED the parameters pbase..pbase+3 (jac[i*4 + j]
// = d res_i / d p_{pbase+j}, for parameter gradients).
template <bool JAC, bool PSEED = false>
__device__ void va_eval(const int* __restrict__ prog, const double* __restrict__ p,
                        const double* __restrict__ u, double cf, double cq, double* res,
                        double* jac, int pbase = 0) {
  constexpr int W = JAC ? 5 : 1;  // value + up to four partials
  const int m = prog[1], nins = prog[3];
  double R[kVAMaxRegs][W];
  const int* ins = prog + kVAHead;
  for (int k = 0; k < nins; ++k, ins += kVAInsWords) {
    const int op = ins[0], dst = ins[1];
    double v = 0.0, g[4] = {0.0, 0.0, 0.0, 0.0};
    if (op == VA_CONST) {
      v = __hiloint2double(ins[3], ins[2]);
    } else if (op == VA_PARAM) {
      v = p[ins[2]];
      const int k = ins[2] - pbase;
      if (JAC && PSEED && k >= 0 && k < 4) g[k] = 1.0;
    } else if (op == VA_UNK) {
      v = u[ins[2]];
      if (JAC && !PSEED) g[ins[2]] = 1.0;
    } else {
      const double* A = R[ins[2]];
      const double* B = R[ins[3]];
      const double a = A[0], b = B[0];
      switch (op) {
        case VA_ADD:
          v = a + b;
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = A[1 + j] + B[1 + j];
          break;
        case VA_SUB:
          v = a - b;
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = A[1 + j] - B[1 + j];
          break;
        case VA_MUL:
          v = a * b;
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = A[1 + j] * b + a * B[1 + j];
          break;
        case VA_DIV:
          v = a / b;
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = (A[1 + j] - v * B[1 + j]) / b;
          break;
        case VA_NEG:
          v = -a;
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = -A[1 + j];
          break;
        case VA_EXP:
          v = exp(a);
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = v * A[1 + j];
          break;
        case VA_LOG:
          v = log(a);
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = A[1 + j] / a;
          break;
        case VA_SQRT:
          v = sqrt(a);
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = A[1 + j] / (2.0 * v);
          break;
        case VA_ABS:
          v = fabs(a);
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = a < 0.0 ? -A[1 + j] : A[1 + j];
          break;
        case VA_POW: {
          v = pow(a, b);
          const double da = b * pow(a, b - 1.0);
          if (JAC)
            for (int j = 0; j < 4; ++j)
              g[j] = da * A[1 + j] + (B[1 + j] != 0.0 ? v * log(a) * B[1 + j] : 0.0);
          break;
        }
        case VA_MIN:
        case VA_MAX: {
          const bool pick_a = (op == VA_MIN) ? (a <= b) : (a >= b);
          v = pick_a ? a : b;
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = pick_a ? A[1 + j] : B[1 + j];
          break;
        }
        case VA_LIMEXP: {
          const double e = exp(fmin(a, kVALimexpMax));
          v = a > kVALimexpMax ? e * (1.0 + a - kVALimexpMax) : e;
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = e * A[1 + j];
          break;
        }
        case VA_TANH:
          v = tanh(a);
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = (1.0 - v * v) * A[1 + j];
          break;
        case VA_SIN:
          v = sin(a);
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = cos(a) * A[1 + j];
          break;
        case VA_COS:
          v = cos(a);
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = -sin(a) * A[1 + j];
          break;
        case VA_LT: v = a < b; break;
        case VA_LE: v = a <= b; break;
        case VA_GT: v = a > b; break;
        case VA_GE: v = a >= b; break;
        case VA_EQ: v = a == b; break;
        case VA_NE: v = a != b; break;
        case VA_AND: v = (a != 0.0) && (b != 0.0); break;
        case VA_OR: v = (a != 0.0) || (b != 0.0); break;
        case VA_NOT: v = a == 0.0; break;
        case VA_TRUNC: v = trunc(a); break;  // integer division (derivative 0)
        case VA_ROUND:  // real -> integer: half away from zero
          v = isfinite(a) ? copysign(floor(fabs(a) + 0.5), a) : a;
          break;
        case VA_SELECT: {  // c ? b : R[ins[4]]
          const double* C = R[ins[4]];
          const bool pick = a != 0.0;
          v = pick ? b : C[0];
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = pick ? B[1 + j] : C[1 + j];
          break;
        }
        case VA_LIMIT: {  // value of b; derivative of R[ins[4]] where a != 0 and it is finite
          const double* C = R[ins[4]];
          const bool general = a != 0.0 && isfinite(C[0]);
          v = b;
          if (JAC) for (int j = 0; j < 4; ++j) g[j] = general ? C[1 + j] : B[1 + j];
          break;
        }
        default: break;
      }
    }
    R[dst][0] = v;
    if (JAC)
      for (int j = 0; j < 4; ++j) R[dst][(1 + j) % W] = g[j];
  }
  const int* out = prog + kVAHead + nins * kVAInsWords;
  for (int i = 0; i < m; ++i) {
    const double* F = R[out[i]];
    const double* Q = R[out[m + i]];
    res[i] = cf * F[0] + cq * Q[0];
    if (JAC && PSEED)
      for (int j = 0; j < 4; ++j) jac[i * 4 + j] = cf * F[(1 + j) % W] + cq * Q[(1 + j) % W];
    else if (JAC)
      for (int j = 0; j < m; ++j) jac[i * m + j] = cf * F[(1 + j) % W] + cq * Q[(1 + j) % W];
  }
}

}  // namespace vxm
