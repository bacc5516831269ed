// Host-side launch wrappers shared between the C-ABI and the kernel files.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "topology.hpp"

namespace vxm {

// Newton settings (voltax/analysis.py:50-84).
struct NewtonOpts {
  int max_steps;
  double rtol, atol, max_dv, gmin;
  int gmin_steps;
  double gmin_start;
  int direct_steps;
  double band_max_mult;  // static-pivot band LU: multiplier bound
};

// Static-pivot band LU of the fused dense kernel: a pattern qualifies when its
// unknowns can be numbered so that, apart from the last kBandBorder positions
// (voltage-source nodes and branch currents, hubs), every row has its
// nonzeros right of the diagonal within kBandWidth columns.
constexpr int kBandWidth = 2;
constexpr int kBandBorder = 2;

// One pattern of a dense batch: its tables plus the register tile it runs in.
// The kernel works on positions 0..sk-1: position p holds unknown q[p] of the
// circuit (-1: padding, an identity row).  Unbanded patterns use p = unknown
// (sk = S); banded ones a band order with the border at smax-kBandBorder..
// (sk = smax).  The kernel copies `blob` (int16) into LDS once; it holds
//   dev   [dev_slots*10]  kind wave u0 u1 u2 u3 poff roff joff -   (kind -1 =
//                         pad; u = positions)
//   jpos  [Cjac]          position col*sk + row of a device-major Jacobian
//                         entry in the column-major system matrix, -1 if its
//                         row or column is ground
//   rpos  [Cres]          row of a device-major residual entry, -1 for ground
//   q     [smax]          unknown at each position, -1 for padding
struct PatternDev {
  TopoDev t;
  int smax;          // padded system size (rows held one per lane)
  int lanes;         // lanes per instance (power of two >= smax)
  int seg_doubles;   // LDS doubles per instance
  int cls;           // tile class
  const short* blob;
  int blob_words;    // blob length in int32 words (the blob is zero padded)
  int o_jpos, o_rpos, o_q;  // section offsets in int16 units
  int dev_slots;     // device-table rows: kinds padded to whole passes of `lanes`
  int sk;            // positions of the kernel's system (S, or smax if banded)
  int band;          // 1: numbered for the static-pivot band LU
  unsigned long long nodemask;  // positions holding node voltages
  unsigned long long realmask;  // positions holding unknowns
};

// One workgroup (one wavefront) of a dense batch: up to 8 instances, all of
// pattern `pattern` (segment k of the wave runs instance inst[k]).
struct WaveItem {
  int pattern, ninst;
  int inst[8];
};

int launch_stamp(const TopoDev& t, int B, const double* params, const double* z, double tt,
                 double cf, double cq, double gmin, const double* hist, double* data,
                 double* res, double* work, hipStream_t stream, const double* scal = nullptr);
int launch_spmv(const TopoDev& t, int B, const double* data, const double* x, double* y,
                hipStream_t stream);
int launch_param_vjp(const TopoDev& t, int B, const double* params, const double* z, double tt,
                     double cf, double cq, const double* lam, double scale, double* grad,
                     hipStream_t stream);
int launch_time_vjp(const TopoDev& t, int B, const double* params, double tt, double cf,
                    const double* lam, double scale, double* out, hipStream_t stream);
int launch_jt_vec(const TopoDev& t, int B, const double* params, const double* z, double tt,
                  double cf, double cq, double gmin, const double* lam, double* out,
                  double* work, hipStream_t stream);

// Complex AC systems for F frequencies: data[f] = g + j 2 pi freqs[f] c, rhs[f] = -b
// (complex values as re/im pairs).
int launch_ac_system(int nnz, int S, int F, const double* g, const double* c,
                     const double* freqs, const double* b, double* data, double* rhs,
                     hipStream_t stream);

// Tile classes of the fused dense kernel: padded sizes and lanes per instance.
// Steps of 4 (8) unknowns keep the padding of the unrolled LU and of the
// row registers small; lanes are the next power of two.
constexpr int kDenseClasses = 11;
constexpr int kDenseSmax[kDenseClasses] = {8, 12, 16, 20, 24, 28, 32, 40, 48, 56, 64};
constexpr int kDenseLanes[kDenseClasses] = {8, 16, 16, 32, 32, 32, 32, 64, 64, 64, 64};
// Class index for S unknowns (-1 if S is too large for the dense kernel).
int dense_class(int S);

// Items [begin[c], begin[c]+count[c]) all belong to class c.
// precision: 0 = FP64 LU, 1 = FP32 LU + FP64 iterative refinement, 2 = FP64 LU whose
// factors later iterations of a time step reuse (simplified Newton).
int launch_dense_run(const PatternDev* pats, const WaveItem* items, const int* begin,
                     const int* count, const int* count_band, const int* lds_bytes,
                     const int64_t* inst_poff, const int64_t* inst_zoff, const double* params, const double* z0,
                     const double* ts, int T, int method, int do_dc, double t_dc,
                     const NewtonOpts& o, int precision, double* out_z, int8_t* out_conv,
                     int32_t* out_iters, hipStream_t stream);

// Derived parameter rows for the fused kernel (device_eval<.., true>):
// dparams[0, n_par) = params[0, n_par), then the EKV, resistor and capacitor
// rows of every instance of the items replaced by their derived rows.
int launch_derive_params(const PatternDev* pats, const WaveItem* items, int n_items,
                         const int64_t* ipoff, int64_t n_par, const double* params,
                         double* dparams, hipStream_t stream);

template <int SMAX, int L>
int launch_dense_class(const PatternDev* pats, const WaveItem* items, int n_items, int n_band,
                       int lds_bytes, const int64_t* ipoff, const int64_t* izoff,
                       const double* params, const double* z0, const double* ts, int T,
                       int method, int do_dc, double t_dc, const NewtonOpts& o, int precision,
                       double* out_z, int8_t* out_conv, int32_t* out_iters,
                       hipStream_t stream);

}  // namespace vxm
