// Host-side topology compilation and the device-resident tables the kernels
// read.  A topology is one circuit pattern: its devices (kind, waveform,
// local unknowns, parameter offsets), the CSR pattern of its Jacobian
// (union of the per-device blocks plus the full diagonal, column indices
// sorted; voltax/sparse.py:161-193) and two gather maps used for
// deterministic, atomic-free stamping:
//   residual row s  <- contributions ridx[rptr[s] .. rptr[s+1])
//   Jacobian slot p <- contributions jidx[jptr[p] .. jptr[p+1])
// Contributions are listed in device order, then local order: the order in
// which the reference's sequential scatter-adds visit them.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "vxm_types.h"

namespace vxm {

// Plain-old-data view handed to kernels (all pointers are device pointers).
struct TopoDev {
  int S, nv, D, P, nnz, Cres, Cjac;
  int uw;              // unknowns per device-table row (4 .. 16)
  const int* kind;     // [D] (KIND_VA_GEN for devices run by generated code)
  const int* wave;     // [D]
  const int* unk;      // [D*uw], -1 = ground / unused
  const int* poff;     // [D] offset of the device's parameters in a row
  const int* roff;     // [D] offset of its residual contributions
  const int* joff;     // [D] offset of its Jacobian contributions
  const int* indptr;   // [S+1]
  const int* indices;  // [nnz]
  const int* rptr;     // [S+1]
  const int* ridx;     // [Cres]
  const int* jptr;     // [nnz+1]
  const int* jidx;     // [Cjac]
  const int* diag;     // [S] slot of the diagonal entry
  const int* slot_row; // [nnz] row of each slot
  // rows / slots whose sums run over more than kHeavyGather terms (a supply
  // node's row, its diagonal): summed by a whole wavefront each
  const int* heavy_r;  // [n_heavy_r] residual rows (rptr)
  const int* heavy_j;  // [n_heavy_j] Jacobian slots (jptr)
  const int* heavy_a;  // [n_heavy_a] pattern rows (indptr), for products
  int n_heavy_r, n_heavy_j, n_heavy_a;
  const int* code;     // compiled Verilog-A programs (devices of kinds KIND_VA2..4)
  const int* va_devs;  // devices of each generated program, concatenated
  int n_gen;           // generated programs (host-side list, for launches)
  const struct VAGenProgram* gen;
  int has_va;          // any Verilog-A device in the table
};

// Longest sum a thread of the gather and product kernels takes on alone.
constexpr int kHeavyGather = 32;

struct VAGenKernel;  // va_codegen.cpp

// One Verilog-A program of a topology that runs as generated code: its
// devices are va_devs[dev_off .. dev_off + n_devs).
struct VAGenProgram {
  int code_off = 0, m = 0, npar = 0, dev_off = 0, n_devs = 0;
  // the first program of each group of up to kVAGroup holds the group's
  // kernel (loaded on upload, owned by the topology) and the group's size
  VAGenKernel* kernel = nullptr;
  int group_n = 0;
};

struct Topology {
  // host copies
  int S = 0, nv = 0, D = 0, P = 0, nnz = 0, Cres = 0, Cjac = 0;
  int uw = 4;  // unknowns per device-table row
  std::vector<int> kind, wave, unk, poff, roff, joff;
  // terminals of each device (its first nterm unknowns; the rest are internal
  // nodes and branch currents): fixed per built-in kind, from the program
  // header for Verilog-A devices
  std::vector<int> nterm;
  std::vector<int> code;  // Verilog-A programs, set before the first upload
  bool code_checked = false;  // set_va_code accepted the programs
  std::vector<VAGenProgram> gen;  // programs run as generated code
  std::vector<int> va_devs;       // their device lists
  std::vector<int> indptr, indices, rptr, ridx, jptr, jidx, diag;
  std::vector<int> slot_row, heavy_r, heavy_j, heavy_a;
  // device copies (one allocation), made on first GPU use
  void* dev_block = nullptr;
  int dev_ordinal = -1;  // the device holding dev_block
  TopoDev dev{};
  std::mutex mu;
  ~Topology();
};

// Builds the pattern and maps from the device list (host only, no GPU needed).
// Returns 0 on success, a negative vxm status code otherwise.
// unknowns: [D * width] rows (width 4..16; entries past a device's m ignored).
int build_topology(int S, int nv, int D, int P, int width, const int32_t* kind,
                   const int32_t* wave, const int32_t* unknowns, const int32_t* param_off,
                   Topology** out);

// Uploads the tables to the current device once; later calls are no-ops.
int upload_topology(Topology* t);
// Verilog-A programs of the VA devices, checked; before the first upload only.
int set_va_code(Topology* t, const int* words, int n_words);

}  // namespace vxm
