// Generated code for large Verilog-A programs.
//
// The interpreter (va_eval.hpp) dispatches on every instruction over a
// register file of 64 duals with 4 partials in private memory: fine for
// models of tens of instructions, wrong for BSIM4-class models (about a
// thousand statements, internal nodes and branch currents, up to 16
// unknowns).  Programs with more than 4 unknowns or 64 registers therefore
// become straight-line HIP C++, one kernel per program, compiled at run time
// by hiprtc for the device's architecture: each SSA register is a local
// value the compiler places in VGPRs, each dual carries exactly m partials,
// and the kernel computes one (instance, device) per thread, in one of four
// modes that mirror the interpreter's kernels:
//   0  residual contributions               (k_eval_va, no Jacobian)
//   1  residual and Jacobian contributions  (k_eval_va)
//   2  J^T lam contributions                (k_jt_contrib)
//   3  parameter VJP                        (k_param_vjp_va)
// Compiled code objects are cached per program text for the process.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "topology.hpp"

namespace vxm {

constexpr int kInterpretedRegs = 64;      // the interpreter's register file
constexpr int kGeneratedRegs = 1 << 16;   // SSA registers of a generated program

// VXM_VA_CODEGEN=all in the environment: every Verilog-A program runs as
// generated code (the tests run both paths on the same models).
bool va_codegen_for_all();

// Programs launched together: one kernel per group of up to kVAGroup
// programs of a topology, block range g running program g (the per-program
// launches of a batch each fill a fraction of the chip).
constexpr int kVAGroup = 8;
struct VAGroupPart {
  int nd, blk0;       // devices of the program, its first block
  const int* devs;    // its device list
};
struct VAGroupArgs {  // the generated kernel's first argument (same layout)
  int n;
  VAGroupPart g[kVAGroup];
};
static_assert(sizeof(VAGroupPart) == 16 && sizeof(VAGroupArgs) == 8 + 16 * kVAGroup,
              "the layout the generated source declares");

// HIP C++ source of the kernel of checked programs progs[0..n) (layout
// va_eval.hpp), n <= kVAGroup.
std::string va_source(const int* const* progs, int n);

// Compiles (once per source text per process) and loads on the current
// device the kernel of programs progs[0..n).
int load_va_kernel(const int* const* progs, int n, VAGenKernel** out);

// Unloads the kernels a topology loaded.
void release_va_kernels(Topology* t);

// One launch of every generated program of topology `t` in `mode` (above);
// lam is read in modes 2 and 3, grad written in mode 3, work in 0..2.
int launch_va_generated(const TopoDev& t, int mode, int B, const double* params,
                        const double* z, const double* lam, double cf, double cq,
                        const double* scal, double scale, double* work, double* grad,
                        hipStream_t stream);

}  // namespace vxm
