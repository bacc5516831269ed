/* voltax_mi355x: C ABI of the MI355X-native Newton inner loop.
 *
 * Plain pointers and sizes only.  Pointers documented "device" are HIP device
 * pointers; "host" pointers are read during the call only.  Every function
 * returns an int status (VXM_OK = 0, negative on error), never throws and
 * never aborts.  Work is enqueued on the given HIP stream (NULL = default
 * stream); nothing synchronises except the *_create calls, which upload
 * their tables with a blocking copy.
 *
 * Batched layouts are instance-major: params [B][P], states [B][S],
 * Jacobian values [B][nnz].  A batch of 0 instances is valid: the call
 * checks its arguments, enqueues nothing and returns VXM_OK.
 */
#ifndef VOLTAX_MI355X_H
#define VOLTAX_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VXM_OK 0
#define VXM_ERR_INVALID (-1)
#define VXM_ERR_HIP (-2)
#define VXM_ERR_MISMATCH (-3)
#define VXM_ERR_UNSUPPORTED (-4)
#define VXM_ERR_TOO_LARGE (-5) /* beyond the fixed scratch of vxm_stamp_noalloc */
#define VXM_ERR_HOST_MEMORY (-6) /* host memory exhausted while building tables */

/* Device kinds and waveform kinds of the device table. */
#define VXM_KIND_R 0   /* Resistor:  unknowns (p, n),       params [log_r]           */
#define VXM_KIND_C 1   /* Capacitor: unknowns (p, n),       params [log_c]           */
#define VXM_KIND_V 2   /* VoltageSource: unknowns (p, n, i), params = waveform       */
#define VXM_KIND_I 3   /* CurrentSource: unknowns (p, n),   params = waveform        */
#define VXM_KIND_EKV 4 /* EKVMOSFET: unknowns (d, g, s, b), params
                          [log_w log_l dvth sign kp vth n lam vt cox cov]            */
#define VXM_KIND_VA2 5 /* compiled Verilog-A device with m = 2..16 unknowns       */
#define VXM_KIND_VA3 6 /* (terminals, internal nodes, branch currents) has kind    */
#define VXM_KIND_VA4 7 /* 3 + m; dev_wave = offset of its program in the words of */
#define VXM_KIND_VA(m) (3 + (m)) /* vxm_topology_set_code                          */
#define VXM_MAX_DEVICE_UNKNOWNS 16
#define VXM_PREDICT 2 /* flag of vxm_batch_run's method (see there) */
#define VXM_REUSE 4   /* flag of vxm_batch_run's method (see there) */
#define VXM_BAND_LU 8 /* flag of vxm_batch_run's method (see there) */
#define VXM_WAVE_NONE 0
#define VXM_WAVE_CONST 1 /* [value]                                        */
#define VXM_WAVE_STEP 2  /* [v0 v1 delay rise]                             */
#define VXM_WAVE_PULSE 3 /* [v1 v2 delay rise fall width period]           */
#define VXM_WAVE_SINE 4  /* [offset amplitude freq delay damping phase]    */
#define VXM_WAVE_PWL 5   /* [npts delay repeat t0..t(n-1) v0..v(n-1)]     */

typedef struct vxm_topology vxm_topology;
typedef struct vxm_batch vxm_batch;
typedef struct vxm_splu vxm_splu;
typedef void* vxm_stream; /* a hipStream_t */

typedef struct {
  int32_t max_steps;    /* Options.max_steps      (voltax/analysis.py:74)    */
  double rtol, atol;    /* Options.rtol / atol    (voltax/analysis.py:75-76) */
  double max_dv;        /* Options.max_dv         (voltax/analysis.py:77)    */
  double gmin;          /* Options.gmin           (voltax/analysis.py:78)    */
  int32_t gmin_steps;   /* Options.gmin_steps     (voltax/analysis.py:79)    */
  double gmin_start;    /* Options.gmin_start     (voltax/analysis.py:80)    */
  int32_t direct_steps; /* Options.direct_steps   (voltax/analysis.py:81)    */
  double band_max_mult; /* VXM_BAND_LU: largest |multiplier| a static pivot may
                         * produce before the factorization falls back to
                         * partial pivoting (0: 100); no reference field     */
} vxm_newton_options;

const char* vxm_status_string(int status);

/* ---------------------------------------------------------------- topology
 * Replaces voltax/sparse.py:150-193 (`structure`: CSR pattern of df/dz and
 * dq/dz = diagonal + every device block, per-group stamp positions).
 * dev_unknowns: [n_devices*4] host int32 state indices per device (terminals
 * then internals; -1 = ground or unused).  dev_param_off: offset of each
 * device's parameters inside one instance's parameter row of n_params. */
int vxm_topology_create(int32_t size, int32_t n_nodes, int32_t n_devices, int32_t n_params,
                        const int32_t* dev_kind, const int32_t* dev_wave,
                        const int32_t* dev_unknowns, const int32_t* dev_param_off,
                        vxm_topology** out);
/* The same with device rows of `width` unknowns (4 <= width <= 16):
 * dev_unknowns [n_devices*width], for Verilog-A devices of more than 4
 * unknowns (BSIM4-class models with internal nodes and branch currents). */
int vxm_topology_create_wide(int32_t size, int32_t n_nodes, int32_t n_devices,
                             int32_t n_params, int32_t width, const int32_t* dev_kind,
                             const int32_t* dev_wave, const int32_t* dev_unknowns,
                             const int32_t* dev_param_off, vxm_topology** out);
/* Programs of the Verilog-A devices (kinds VXM_KIND_VA*) of a topology, compiled
 * by voltax_amd/va.py from the reference's Verilog-A subset (voltax/va/
 * compiler.py:342-1525): int32 words [magic, m, n_regs, n_ins, n_params,
 * n_terminals] [n_ins x (op, dst, a, b, c)] [2m output registers].  Programs
 * of at most 4 unknowns and 64 registers run on the GPU interpreter; larger
 * ones (and all of them with VXM_VA_CODEGEN=all in the environment) become
 * straight-line HIP C++ compiled by hiprtc when the topology is first used.  A device's parameter
 * row is its real instance parameters then its multiplicity m (physical
 * units); vxm_param_vjp differentiates every entry of it.  Checked here (every
 * index in range) and uploaded with the topology; required before the first
 * GPU use of a topology with Verilog-A devices (which returns
 * VXM_ERR_INVALID without it).  Such topologies run on the stamped (sparse)
 * path only (vxm_batch_create returns VXM_ERR_UNSUPPORTED). */
int vxm_topology_set_code(vxm_topology* topo, int32_t n_words, const int32_t* words);
int vxm_topology_destroy(vxm_topology* topo);
/* sizes[0..6] = S, n_nodes, n_devices, n_params, nnz, work doubles per instance */
int vxm_topology_sizes(const vxm_topology* topo, int64_t* sizes);
/* host copies of the CSR pattern: indptr [S+1], indices [nnz] */
int vxm_topology_pattern(const vxm_topology* topo, int32_t* indptr, int32_t* indices);

/* ---------------------------------------------------------------- stamping
 * Replaces voltax/sparse.py:220-249 (`jacobian`: cf*df/dz + cq*dq/dz on the
 * pattern), voltax/circuit.py:196-226 (`f`, `q`) and the residuals /
 * Jacobians of voltax/analysis.py:325-382.  For a batch of B instances:
 *   res[b]  = cf*f(z[b],t) + cq*q(z[b]) + hist[b] + gmin*v(z[b])   (if res)
 *   data[b] = cf*df/dz + cq*dq/dz + gmin on node diagonals          (if data)
 * hist may be NULL.  work: device scratch of B * sizes[5] doubles. */
int vxm_stamp(const vxm_topology* topo, int32_t batch, const double* params, const double* z,
              double t, double cf, double cq, double gmin, const double* hist, double* data,
              double* res, double* work, vxm_stream stream);

/* ------------------------------------------------------------------ adjoint
 * Replace the VJPs of the implicit-function backward pass
 * (voltax/analysis.py:282-288, `_root_bwd`):
 * vxm_param_vjp:  grad[b] += scale * lam[b]^T d(cf*f + cq*q)/dp at (z[b], t);
 *   grad has the layout of params ([B][P]) and every entry of every device's
 *   row is differentiated: log_r (R), log_c (C), each waveform entry of a
 *   source (a PWL's point count gets 0), the EKV row log_w, log_l, dvth, kp,
 *   vth, n, lam, vt, cox, cov (physical units; the polarity entry gets 0) --
 *   the per-device and MOSProcess leaves of voltax/elements/mosfet.py:26-87
 *   and the signal leaves of voltax/signals.py:56-189 -- and every slot of a
 *   Verilog-A device.  Built-in rows longer than 67 entries (PWL sources of
 *   more than 32 points) are left untouched.
 * vxm_jt_vec:     out[b] = (cf*df/dz + cq*dq/dz + gmin on node diagonals)^T lam[b]
 *   work: B * sizes[5] doubles. */
int vxm_param_vjp(const vxm_topology* topo, int32_t batch, const double* params,
                  const double* z, double t, double cf, double cq, const double* lam,
                  double scale, double* grad, vxm_stream stream);
int vxm_jt_vec(const vxm_topology* topo, int32_t batch, const double* params, const double* z,
               double t, double cf, double cq, double gmin, const double* lam, double* out,
               double* work, vxm_stream stream);
/* The time-point cotangents of the same backward pass (the reference's step
 * residuals take t_prev and t as differentiable parameters,
 * voltax/analysis.py:338-350): out[b][d] += scale * lam[b]^T d(cf*f_d)/dt
 * for each device d at time t ([B][n_devices]; only the independent sources
 * depend on t, through their waveforms; others add 0).  The charge part,
 * -/+ lam^T (q(z_n) - q(z_n-1)) / h^2, comes from two charge stamps. */
int vxm_time_vjp(const vxm_topology* topo, int32_t batch, const double* params, double t,
                 double cf, const double* lam, double scale, double* out, vxm_stream stream);

/* vxm_stamp with t, cf, cq, gmin, source_scale read from device memory
 * scal[0..5) at run time, so a captured HIP graph can be replayed across time
 * steps.  source_scale multiplies every independent source (the source-stepping
 * homotopy of the DC fallback; 1 = plain). */
int vxm_stamp_dev(const vxm_topology* topo, int32_t batch, const double* params,
                  const double* z, const double* scal, const double* hist, double* data,
                  double* res, double* work, vxm_stream stream);

/* vxm_stamp without the work argument and without any allocation: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The device contributions go to a fixed scratch array of
 * VXM_NOALLOC_SCRATCH_DOUBLES doubles (32 MiB) that the library holds on every
 * device it runs on until the process ends; batch * sizes[5] beyond it returns
 * VXM_ERR_TOO_LARGE (split the batch).  The topology's tables must already be
 * resident on the current device (any earlier vxm_stamp, vxm_spmv, ... call on
 * it uploads them; VXM_ERR_INVALID otherwise, VXM_ERR_MISMATCH if they live on
 * another device).  All calls on one device share its scratch: calls that
 * overlap on one device (other streams, other threads) race; serialize them.
 * Results equal vxm_stamp's bit for bit (the same kernels). */
#define VXM_NOALLOC_SCRATCH_DOUBLES (4 * 1024 * 1024)
int vxm_stamp_noalloc(const vxm_topology* topo, int32_t batch, const double* params,
                      const double* z, double t, double cf, double cq, double gmin,
                      const double* hist, double* data, double* res, vxm_stream stream);

/* Batched small dense solves: x[b] = A[b]^-1 rhs[b] (A[b]^-T rhs[b] if
 * transpose), A row-major [B][n][n], rhs and x [B][n] (x may alias rhs),
 * 1 <= n <= 64, FP64 LU with partial pivoting.  Replaces the dense
 * `jnp.linalg.solve` of voltax/analysis.py:201-205 for matrices the caller
 * builds itself (solve_root, voltax/analysis.py:291-306). */
int vxm_dense_solve(int32_t batch, int32_t n, const double* A, const double* rhs, double* x,
                    int32_t transpose, vxm_stream stream);

/* y[b] = A[b] x[b] in FP64 (voltax/sparse.py:130-133, Structure.matvec). */
int vxm_spmv(const vxm_topology* topo, int32_t batch, const double* data, const double* x,
             double* y, vxm_stream stream);

/* ------------------------------------------------- fused small-circuit batch
 * A batch of instances of one or more patterns with S <= 64 unknowns, run by
 * one kernel launch (one wavefront per 64/L instances of one pattern).
 * inst_topo[i] picks the pattern of instance i; inst_param_off[i] is the
 * offset of its parameter row in `params`.  State arrays are concatenations
 * over instances in order (offset of instance i = sum of S of instances < i).
 * Replaces the dense path of voltax/analysis.py:201-205 and 385-402 under
 * `dc` (:410-431, Newton + gmin stepping) and `transient` (:434-480). */
int vxm_batch_create(int32_t n_topos, vxm_topology* const* topos, int32_t n_inst,
                     const int32_t* inst_topo, const int64_t* inst_param_off,
                     vxm_batch** out);
int vxm_batch_destroy(vxm_batch* batch);
/* Runs (optionally) the DC operating point at t_dc from z0, then T-1 time
 * steps over ts[0..T) (device) with method 0 = backward Euler, 1 = trap,
 * optionally OR VXM_PREDICT (backward Euler only; ignored for trap): each
 * step's Newton iteration then starts from the value at the new time of the
 * cubic through the last four accepted points (fewer on the first steps)
 * instead of the last point (the reference's guess; an initial guess only: the step's
 * equations and convergence test are unchanged, and a step whose Newton
 * iteration fails from the extrapolated guess is redone from the last point).
 * Optionally OR VXM_REUSE (precision 0 only): within each time step, Newton
 * iterations after the first reuse the step's last LU factors (residual-only
 * device evaluation, forward/backward substitution) while each update shrinks
 * the previous one by half or more in the tolerance-weighted norm; otherwise
 * the next iteration factors afresh.  The equations and the convergence test
 * are unchanged (DC solves always factor afresh); out_iters then holds the
 * iterations in bits 0..15 and the LU factorizations in bits 16..30.  No
 * counterpart in the reference, which factors every iteration.
 * Optionally OR VXM_BAND_LU (precision 0 only; ignored for precision 1): the
 * instances whose pattern has a band numbering (vxm_batch_band) factor with
 * static pivots in that order -- bandwidth 2 outside a border of 2 unknowns
 * (voltage-source nodes and branch currents), the border block with partial
 * pivoting -- as long as every multiplier stays within opts->band_max_mult
 * in magnitude; a factorization with a larger one (a pivot small against
 * its column) is redone with partial pivoting.  Reusing runs (VXM_REUSE)
 * take the band LU in the tile classes of 64 lanes only.  The same roots as the reference's LU with
 * partial pivoting (jnp.linalg.solve, voltax/analysis.py:201-205) up to
 * round-off; no counterpart in the reference.
 * out_z: concatenated [T][S_i] per instance; out_conv [n_inst][T] int8;
 * out_iters [n_inst][T] int32 (Newton iterations; the DC entry counts all
 * gmin stages).  precision 0 = FP64 LU, 1 = FP32 LU + FP64 refinement.
 * params is read once, at the start of the run on `stream`, into the batch's
 * own rows of the parameter-only factors of the device models (after the
 * previous run on the handle is done with them); it may change afterwards. */
int vxm_batch_run(const vxm_batch* batch, const double* params, const double* z0,
                  const double* ts, int32_t T, int32_t method, int32_t do_dc, double t_dc,
                  const vxm_newton_options* opts, int32_t precision, double* out_z,
                  int8_t* out_conv, int32_t* out_iters, vxm_stream stream);
/* The tile classes of one run are launched concurrently on streams owned by
 * the batch and joined back to `stream`.  ms[0..n_classes): duration of each
 * class launch of the last run, from HIP events recorded on the stream it ran
 * on (0 for empty classes); waits for them. */
int vxm_batch_timings(const vxm_batch* batch, float* ms);
/* band[p] = 1 if topology p of the batch (vxm_batch_create order) has a band
 * numbering for VXM_BAND_LU, else 0. */
int vxm_batch_band(const vxm_batch* batch, int32_t* band);
/* The band numbering VXM_BAND_LU uses for a topology padded to smax
 * unknowns (host only): pos[u] = position of unknown u (positions 0..smax-1;
 * the border at the last 2).  VXM_ERR_UNSUPPORTED if the pattern has none.
 * No counterpart in the reference. */
int vxm_topology_band(const vxm_topology* topo, int32_t smax, int32_t* pos);
/* Tile classes of the fused dense kernel: *n_classes, and (if not NULL) the
 * padded system size smax[c] and lanes per instance lanes[c] of class c; an
 * instance of S unknowns runs in the first class with smax >= S. */
int vxm_dense_classes(int32_t* n_classes, int32_t* smax, int32_t* lanes);

/* ------------------------------------------------ supernodal sparse LU
 * Symbolic analysis once per topology (replaces voltax/sparse.py:257-535:
 * pivot groups, nested dissection, supernodes, level schedule) and the
 * numeric multifrontal factorization / solves (replaces :543-595).
 * leaf: largest nested-dissection leaf kept as one supernode. */
int vxm_splu_create(const vxm_topology* topo, int32_t leaf, vxm_splu** out);
/* vxm_splu_create with flags: VXM_SPLU_COMPACT groups only what pivoting
 * needs -- a two-terminal device's single internal unknown matched to a
 * terminal as above, the other devices' internal unknowns together (their
 * terminals free), and every unknown without a diagonal contribution (and
 * no gmin) with the unknowns it couples to -- instead of gluing each such
 * device to all its terminals (voltax/sparse.py:274-300), which makes one
 * front of every circuit whose devices with internal nodes share a supply.
 * Same solutions up to round-off; no counterpart in the reference. */
#define VXM_SPLU_COMPACT 1
int vxm_splu_create_ex(const vxm_topology* topo, int32_t leaf, int32_t flags, vxm_splu** out);
int vxm_splu_destroy(vxm_splu* plan);
/* sizes[0..8] = front values per instance, pivots per instance, border
 * entries per instance, supernodes, levels, largest front, largest pivot
 * block, solve work doubles per instance (= border entries + S) */
int vxm_splu_sizes(const vxm_splu* plan, int64_t* sizes);
/* stats[0..1] = LU flops per factorization, factor entries */
int vxm_splu_stats(const vxm_splu* plan, double* stats);
/* host copies: supernode of every unknown [S], parent of every supernode
 * [n_supernodes], level of every supernode [n_supernodes] */
int vxm_splu_structure(const vxm_splu* plan, int32_t* sn_of, int32_t* parent, int32_t* level);
/* Host copies of the whole symbolic layout (for inspection and CPU checks):
 * counts[0..5] = supernodes, idx entries, cmap entries, children entries,
 * A entries, level-list entries (call with all arrays NULL to get counts).
 * sn [supernodes*10] int64: k, r, parent, level, front_off, piv_off,
 * idx_off, cmap_off, child_off, n_child; then aent_off, n_aent, uvec_off
 * packed as sn2 [supernodes*3] int32. */
int vxm_splu_export(const vxm_splu* plan, int64_t* counts, int64_t* sn, int32_t* sn2,
                    int32_t* idx, int32_t* cmap, int32_t* children, int32_t* aslot,
                    int32_t* apos, int32_t* level_ptr, int32_t* level_sn);
/* Schedule knob of the numeric phase (no reference counterpart; set by the
 * autotuner): mid_blocked = 1 factors fronts of 33..64 unknowns with the
 * blocked 256-thread LDS kernel instead of the register-resident wavefront
 * kernel.  Results agree to rounding; only speed differs. */
int vxm_splu_configure(vxm_splu* plan, int32_t mid_blocked);
/* Element types of vxm_splu_factor / vxm_splu_solve. */
#define VXM_SPLU_F64 0  /* FP64 values, factors and solves                           */
#define VXM_SPLU_F32 1  /* FP64 values, FP32 factors, FP64 solves                    */
#define VXM_SPLU_C128 2 /* complex FP64 throughout (re, im interleaved): AC systems   */

/* Factor B matrices data [B][nnz] (values on the topology's pattern; complex
 * pairs for VXM_SPLU_C128) into front [B][sizes[0]] elements of the factor
 * type and piv [B][sizes[1]]. */
int vxm_splu_factor(const vxm_splu* plan, int32_t batch, const double* data, void* front,
                    int32_t* piv, int32_t dtype, vxm_stream stream);
/* In place x[B][S] (complex pairs for VXM_SPLU_C128): b -> A^-1 b (or A^-T b
 * if transpose; the plain transpose, not the conjugate one).  work: B * sizes[7]
 * doubles of device scratch (twice that for VXM_SPLU_C128). */
int vxm_splu_solve(const vxm_splu* plan, int32_t batch, const void* front, const int32_t* piv,
                   double* x, double* work, int32_t transpose, int32_t dtype,
                   vxm_stream stream);
/* vxm_splu_factor (FP64) then vxm_splu_solve (transpose 0) of x in one pass
 * over the levels: where every front of a register-front size class leaves
 * a lane free, the forward sweep rides along in the factorization (the
 * fronts are not read back for it).  Same factors and the same x, bit for
 * bit; no counterpart in the reference (factor then solve,
 * voltax/sparse.py:549-595).  work as for vxm_splu_solve. */
int vxm_splu_factor_solve(const vxm_splu* plan, int32_t batch, const double* data, double* front,
                          int32_t* piv, double* x, double* work, vxm_stream stream);
/* x = A^-1 b with `iters` steps of FP64 iterative refinement against the FP64
 * values `data` (r = b - A x; x += A^-1 r).  work: B * (sizes[7] + 2 S). */
int vxm_splu_solve_refined(const vxm_splu* plan, int32_t batch, const double* data,
                           const void* front, const int32_t* piv, const double* b, double* x,
                           double* work, int32_t iters, int32_t transpose, int32_t fp32,
                           vxm_stream stream);

/* Small-signal systems of `ac` (voltax/analysis.py:510-520) for n_freq
 * frequencies (Hz, device): data[f][e] = g[e] + j 2 pi freqs[f] c[e] on the
 * pattern (g = df/dz, c = dq/dz values from vxm_stamp with (cf, cq) = (1, 0)
 * and (0, 1), gmin 0), rhs[f] = -b with b the complex stimulus [size].
 * Complex arrays are (re, im) pairs; factor/solve them with VXM_SPLU_C128. */
int vxm_ac_system(int32_t nnz, int32_t size, int32_t n_freq, const double* g, const double* c,
                  const double* freqs, const double* b, double* data, double* rhs,
                  vxm_stream stream);

/* Damped Newton update of B instances (voltax/analysis.py:224-230): with
 * d = dz_sign*dz and scale = min(1, max_dv / max|d[:n_nodes]|), z += scale*d
 * where active[b] (NULL = all), and
 * done[b] = all(|d| <= atol + rtol|z_new|) & finite & scale == 1.
 * step_norm (NULL = not written): max|d[:n_nodes]| of each active instance, the
 * contraction measure of the warm-start fallback test. */
int vxm_newton_update(int32_t batch, int32_t size, int32_t n_nodes, double* z,
                      const double* dz, double dz_sign, const vxm_newton_options* opts,
                      const int8_t* active, int8_t* done, double* step_norm,
                      vxm_stream stream);
/* The same update for an iteration whose linear solve used the LU factors of
 * an earlier iteration of the same solve (reused = 1) or fresh ones (0); no
 * counterpart in the reference, which factors every iteration
 * (voltax/analysis.py:221-230).  wnorm[b]: tolerance-weighted size
 * max_i |d_i| / (atol + rtol |z_i|) of the last applied update (written by
 * every applied update, read by reused ones).  A reused-factor update
 * converges linearly with rate theta = its size over wnorm[b]: it sets done
 * only if it passes the plain test and theta * size <= 1e-4 (1 - theta);
 * theta > 1/2 sets refactor[b] = 1 (the caller should factor afresh), and
 * theta >= 1 also leaves z unchanged.  refactor[b] = 0 otherwise. */
int vxm_newton_update_reuse(int32_t batch, int32_t size, int32_t n_nodes, double* z,
                            const double* dz, double dz_sign, const vxm_newton_options* opts,
                            const int8_t* active, int8_t* done, double* wnorm,
                            int8_t* refactor, int32_t reused, vxm_stream stream);

/* ---- Learned warm start: BF16 MFMA multilayer perceptron ------------------
 * No reference counterpart (the reference starts Newton from the caller's
 * guess, voltax/analysis.py:291-317).  bf16 buffers are uint16 bit patterns
 * (round to nearest even); all pointers are device pointers.
 *
 * vxm_gemm_bf16: C[m][n] = epi(alpha * sum_k a[m][k] b[n][k] + bias[n]
 * + beta * C[m][n]) with fp32 accumulation (beta != 0 reads the fp32 c);
 * epi zeroes entries where mask[m][n] <= 0 (mask may be NULL), then applies
 * max(., 0) if relu.  Writes fp32 c and/or bf16 cb.  m, n multiples of 64,
 * k a multiple of 32, lda/ldb multiples of 8, ldc/ldcb/ldm multiples of 4. */
int vxm_gemm_bf16(int32_t m, int32_t n, int32_t k, const uint16_t* a, int32_t lda,
                  const uint16_t* b, int32_t ldb, float alpha, float beta, const float* bias,
                  const uint16_t* mask, int32_t ldm, int32_t relu, float* c, int32_t ldc,
                  uint16_t* cb, int32_t ldcb, vxm_stream stream);
/* dst[c][r] = src[r][c] for r < rows, c < cols. */
int vxm_transpose_bf16(int32_t rows, int32_t cols, const uint16_t* src, int32_t lds,
                       uint16_t* dst, int32_t ldd, vxm_stream stream);
/* out[r] = sum_{c < cols} src[r][c] (fp32, fixed summation order). */
int vxm_rowsum_bf16(int32_t rows, int32_t cols, const uint16_t* src, int32_t ld, float* out,
                    vxm_stream stream);
/* Mean-squared-error gradient over rows r < m: g[r][c] = bf16(scale * (y - t))
 * for r < m_valid and c < n_valid, 0 elsewhere (c < n);
 * loss_rows[r] = sum_c of the same (y - t)^2 (loss_rows may be NULL). */
int vxm_mse_grad(int32_t m, int32_t m_valid, int32_t n, int32_t n_valid, const float* y,
                 int32_t ldy, const float* t, int32_t ldt, float scale, uint16_t* g,
                 int32_t ldg, float* loss_rows, vxm_stream stream);
/* One Adam step (bias-corrected, step >= 1) with decoupled weight decay wd
 * (w -= lr * (m^ / (sqrt(v^) + eps) + wd * w)) on n fp32 weights; also writes
 * their bf16 copy wb when wb is not NULL. */
int vxm_adam(int64_t n, float* w, const float* g, float* m1, float* m2, float lr, float b1,
             float b2, float eps, float wd, int32_t step, uint16_t* wb, vxm_stream stream);
/* out[r][c] = (x[r][c] - mu[c]) * inv_sd[c] for c < cols, 0 for cols <= c < cols_pad,
 * as bf16 (out_bf16) and/or fp32 (out_f32), both with row pitch ldo. */
int vxm_normalize(int32_t rows, int32_t cols, int32_t cols_pad, const double* x, int32_t ldx,
                  const double* mu, const double* inv_sd, uint16_t* out_bf16, float* out_f32,
                  int32_t ldo, vxm_stream stream);
/* out[r][c] = y[r][c] * sd[c] + mu[c] in FP64 (network output -> Newton guess). */
int vxm_denormalize(int32_t rows, int32_t cols, const float* y, int32_t ldy, const double* mu,
                    const double* sd, double* out, int32_t ldo, vxm_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* VOLTAX_MI355X_H */
