"""Benchmark: batched Newton solves/sec on one MI355X.

Workload (the batched configuration the metric is quoted on): 4096
Monte-Carlo EKV ring oscillators of uneven length (odd stage count uniform
in 3..31), VTH0 offset ~ N(0, 30 mV) and W lognormal (sigma 5 %) per device,
20 fF loads, VDD = 1.2 V, 2000 backward-Euler steps of 2 ps, started from the
alternating 0.9/0.1 VDD state.  One "step" of this benchmark is one full
pass: all 4096 transients (8.19 M Newton solves), one fused launch per tile
class, inputs resident in HBM.

Paths: plain FP64 LU ("fp64", the reference's algorithm), FP32 LU + FP64
refinement ("mixed"), Newton started from the cubic extrapolation of the last
four points ("pred"), and that start with the step's LU factors reused by its
later iterations while they contract ("pred_reuse").  The headline `value` is
the fastest path whose node voltages stay within 1e-9 of the FP64 path
relative to VDD (else the FP64 path).  A compiled C++ restatement of the
reference's dense path (oracle/cpu_dense.cpp, pinned to the numpy oracle by
the CPU tests) is timed on a bounded sample of the same workload on one host
core as `cpu_baseline`, and on the box's CPU share as `cpu_baseline_all_cores`.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VDD = 1.2
PEAK_FP64_TFLOPS = 78.6
EKV_FLOPS = 200.0  # one EKV evaluation with charges and its 4x4 Jacobian block
EKV_RES_FLOPS = 90.0  # the same evaluation without the Jacobian block
CAP_FLOPS = 12.0


def flops_per_iter(n_stages: int, mixed: bool) -> float:
    """Algorithmic FP64 work of one Newton iteration of an n-stage ring."""
    S = n_stages + 2
    lu = 2.0 / 3.0 * S ** 3 + 2.0 * S ** 2
    if mixed:
        lu += 2 * (4.0 * S ** 2)  # two refinement sweeps: residual + solve
    return EKV_FLOPS * 2 * n_stages + CAP_FLOPS * n_stages + 10.0 + lu + 26.0 * n_stages


_BAND_LU = {}


def band_lu_flops(n_stages: int) -> tuple[float, float]:
    """(factorization, forward + back substitution) FP64 operations of the
    static-pivot band LU for an n-stage ring: a structural replay of the
    kernel's elimination in the band numbering (voltax_amd's
    vxm_topology_band), counting one division-free multiply per multiplier
    and a multiply-add (2) per update of a structurally nonzero entry, the
    2 x 2 border block dense."""
    if n_stages in _BAND_LU:
        return _BAND_LU[n_stages]
    import ctypes

    import voltax_amd as vx
    from voltax_amd import engine, workloads as wl
    from voltax_amd._lib import lib

    c, _ = wl.ring(vx.CircuitBuilder(), n_stages)
    S = c.size
    smax = int(class_tables()[0][tile_class(S)])
    topo, _ = engine.topology_of(c)
    pos = np.zeros(S, np.int32)
    if lib.vxm_topology_band(topo.handle, smax, pos.ctypes.data_as(ctypes.c_void_p)) != 0:
        raise RuntimeError(f"no band numbering for the {n_stages}-stage ring")
    nz = np.eye(smax, dtype=bool)
    for u in c.device_table()[2]:
        u = [pos[x] for x in u if x >= 0]
        nz[np.ix_(u, u)] = True
    fact = 0.0
    for k in range(smax):
        rows = np.flatnonzero(nz[k + 1:, k]) + k + 1
        cols = np.flatnonzero(nz[k, k + 1:]) + k + 1
        fact += len(rows) * (1 + 2 * len(cols))
        nz[np.ix_(rows, cols)] = True
    solve = 2.0 * (nz.sum() - smax) + smax  # both sweeps over L and U, one scaling each
    _BAND_LU[n_stages] = (fact, solve)
    return fact, solve


def flops_per_band_iter(n_stages: int) -> float:
    """flops_per_iter with the band LU instead of the dense one."""
    fact, solve = band_lu_flops(n_stages)
    return EKV_FLOPS * 2 * n_stages + CAP_FLOPS * n_stages + 10.0 + fact + solve + 26.0 * n_stages


def flops_per_band_reused_iter(n_stages: int) -> float:
    """flops_per_reused_iter with the band LU's substitutions."""
    S = n_stages + 2
    return flops_per_reused_iter(n_stages) - 2.0 * S ** 2 + band_lu_flops(n_stages)[1]


def flops_per_reused_iter(n_stages: int) -> float:
    """The same for an iteration that reuses the step's LU factors: residual-
    only device evaluation, residual assembly and two triangular solves."""
    S = n_stages + 2
    return EKV_RES_FLOPS * 2 * n_stages + 4.0 * n_stages + 10.0 + 2.0 * S ** 2 + 8.0 * n_stages


def class_tables():
    from voltax_amd.engine import dense_classes

    return dense_classes()


def measured_traffic(kernel_class: int, prec: int = 0, band: bool = False):
    """HBM bytes per launch of a tile class from the committed rocprofv3 PMC
    summary (FETCH_SIZE x 2 + WRITE_SIZE, calibrated on a 1 GiB copy), for the
    default 4096-ring / 2000-step workload; None if not available."""
    path = os.path.join(ROOT, "profiles", "k_dense_run_rings4096_traffic.md")
    smax, lanes = class_tables()
    key = (f"k_dense_run<{smax[kernel_class]}, {lanes[kernel_class]}, {prec}, "
           f"{'true' if band else 'false'}>")
    try:
        for line in open(path):
            if key in line:
                return float(line.strip().strip("|").split("|")[-1])
    except OSError:
        pass
    return None


def kernel_prec(path: str) -> int:
    """Template argument PREC of the kernels a bench path launches."""
    return 1 if path == "mixed" else (2 if "reuse" in path else 0)


def tile_class(S: int) -> int:
    for c, smax in enumerate(class_tables()[0]):
        if S <= smax:
            return c
    raise ValueError(S)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_ring(spec, ts):
    """One ring of the workload through the compiled CPU restatement of the
    reference's dense path (oracle/cpu_dense.cpp): (solves, z [T, S])."""
    import voltax_amd as vx
    from oracle import cpu_dense
    from voltax_amd import workloads as wl

    n, dv, ws = spec
    c, nodes = wl.ring(vx.CircuitBuilder(), n, dv, ws)
    z, _, conv = cpu_dense.transient_be(c.size, c.n_nodes, c.device_table(),
                                        wl.ring_ic(c, nodes), ts)
    return len(ts) - 1, z, bool(conv.all())


def cpu_baseline(specs, ts, target_s: float):
    """The reference's dense path restated in compiled single-thread C++ (the
    same equations, forward-mode derivatives, dense LU with partial pivoting
    and damped Newton; pinned to the numpy oracle by the CPU tests), timed on
    one host core over the first rings of the workload, whole 2000-step
    transients, until about target_s seconds."""
    from oracle import cpu_dense

    cpu_dense.lib()  # load (or build) before the clock starts
    t0 = time.perf_counter()
    solves, done, sizes, zs, ok = 0, 0, [], [], True
    for spec in specs:
        k, z, conv = _cpu_ring(spec, ts)
        solves += k
        done += 1
        sizes.append(spec[0])
        zs.append(z)
        ok &= conv
        if time.perf_counter() - t0 > target_s:
            break
    el = time.perf_counter() - t0
    return {"value": solves / el, "unit": "Newton solves/s", "cores": 1, "kind": "port",
            "sample": f"the first {done} rings of the workload (stages {min(sizes)}.."
                      f"{max(sizes)}), {len(ts) - 1} BE steps each, compiled C++ restatement "
                      f"of the reference dense path (oracle/cpu_dense.cpp), 1 thread, "
                      f"{el:.1f} s",
            "cpu_model": _cpu_model(), "host_nproc": os.cpu_count(), "all_converged": ok,
            "rings": done, "_z": zs}


def cpu_baseline_all_cores(specs, ts, target_s: float, workers: int):
    """The same restatement on `workers` host threads (ctypes releases the
    interpreter lock during the call, so the rings run in parallel), over as
    many rings as fill about target_s seconds."""
    import concurrent.futures as cf

    from oracle import cpu_dense

    cpu_dense.lib()
    t0 = time.perf_counter()
    solves = rings = 0
    with cf.ThreadPoolExecutor(workers) as ex:
        it = iter(specs)
        pending = set()
        for _ in range(2 * workers):
            spec = next(it, None)
            if spec is not None:
                pending.add(ex.submit(_cpu_ring, spec, ts))
        while pending:
            done, pending = cf.wait(pending, return_when=cf.FIRST_COMPLETED)
            for f in done:
                solves += f.result()[0]
                rings += 1
            if time.perf_counter() - t0 < target_s:
                for _ in done:
                    spec = next(it, None)
                    if spec is not None:
                        pending.add(ex.submit(_cpu_ring, spec, ts))
    el = time.perf_counter() - t0
    return {"value": solves / el, "unit": "Newton solves/s", "cores": workers, "kind": "port",
            "sample": f"{rings} rings of the workload, {len(ts) - 1} BE steps each, the same "
                      f"compiled restatement on {workers} threads, {el:.1f} s",
            "cpu_model": _cpu_model(), "host_nproc": os.cpu_count()}


def grid_workload(args, emit=True):
    """Config 2 (100x100 R mesh): DC + 200-step transient + adjoint gradient of
    the worst IR drop with respect to every segment width (R = R0 / w)."""
    import torch

    import voltax_amd as vx
    from voltax_amd import adjoint, sparse, workloads as wl
    from voltax_amd.sparse_newton import transient_sparse

    torch.cuda.set_device(0)
    c = wl.power_grid(vx.CircuitBuilder(), args.grid, args.grid, 0, vx.Pulse, 100e-12)
    ts = np.linspace(0, 2e-9, args.tsteps_grid + 1)
    opts = vx.Options()
    # warm-up: plans, autotuning of leaf size and front kernels, allocations
    sol = transient_sparse(c, ts, "dc", "be", opts, leaf="auto")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sol = transient_sparse(c, ts, "dc", "be", opts, leaf="auto")
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - t0) / args.steps
    print(f"# grid forward {fwd * 1e3:.1f} ms, {len(ts)} solves, iters {int(sol.iters.sum())}",
          file=sys.stderr, flush=True)
    # the same transient with the LU factors of a step reused by its later iterations
    ropts = vx.Options(reuse=True)
    rsol = transient_sparse(c, ts, "dc", "be", ropts, leaf="auto")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rsol = transient_sparse(c, ts, "dc", "be", ropts, leaf="auto")
    torch.cuda.synchronize()
    fwd_reuse = (time.perf_counter() - t0) / args.steps
    print(f"# grid forward, reused factors {fwd_reuse * 1e3:.1f} ms, iters "
          f"{int(rsol.iters.sum())}, factorizations {rsol.factorizations}",
          file=sys.stderr, flush=True)
    drop = 1.0 - sol.z[:, : c.n_nodes]
    n, k = np.unravel_index(int(np.argmax(drop)), drop.shape)
    g = np.zeros_like(sol.z)
    g[n, k] = -1.0  # d(1 - v)/dv
    t0 = time.perf_counter()
    gr = adjoint.transient_vjp(c, sol, g, ic="dc")
    torch.cuda.synchronize()
    bwd = time.perf_counter() - t0
    dw = -gr["Resistor"]["log_r"]  # d(max drop)/d log w
    solves = len(ts)
    line = {
        "metric": "Newton solves/sec (1 GPU), plain FP64 GPU path, single 100x100 grid",
        "value": solves / fwd, "unit": "Newton solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": fwd * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.grid}x{args.grid} R mesh power grid, DC + "
                               f"{len(ts) - 1} BE steps of 10 ps, adjoint of max IR drop",
                   "unknowns": c.size, "nnz": int(sparse.structure(c).nnz)},
        "newton_iters": int(sol.iters.sum()), "all_converged": bool(sol.converged.all()),
        "reuse_solves_per_s": solves / fwd_reuse,
        "reuse_newton_iters": int(rsol.iters.sum()),
        "reuse_lu_factorizations_steps": int(rsol.factorizations),
        "reuse_all_converged": bool(rsol.converged.all()),
        "max_abs_v_diff_reuse_vs_plain": float(np.abs(rsol.z - sol.z).max()),
        "max_ir_drop_V": float(drop.max()), "adjoint_s": bwd,
        "grad_max_abs": float(np.abs(dw).max()), "grad_params": int(dw.size),
    }
    if emit:
        print(json.dumps(line))
    return line


def bsim_workload(args, emit=True):
    """Rings of the BSIM4-sized Verilog-A MOSFET (10 unknowns per device, the
    stand-in for BSIM4, whose source is not available): a batch of
    `bsim_count` 51-stage rings at VDD = 1.0 V, W_n 0.2 um, W_p 0.4 um,
    L 45 nm, 1 fF per stage, per-instance VTH0 offset ~ N(0, 20 mV), 1 ps
    backward-Euler steps on the sparse path with generated device code: plain
    FP64 Newton, FP32 LU + FP64 refinement, and the line's value, FP64 Newton
    started from the cubic extrapolated through the last four points; periods,
    and the adjoint of one instance's period with respect to every device's
    VTH0.  `bsim_grouping` picks the sparse plan's pivot groups (the
    reference's glue every PMOS to vdd: one front of the whole ring)."""
    import torch

    import voltax_amd as vx
    from voltax_amd import adjoint, measure, va, workloads as wl
    from voltax_amd.analysis import Solution
    from voltax_amd.sparse_newton import transient_sparse_batch

    torch.cuda.set_device(0)
    log = lambda m: print("# " + m, file=sys.stderr, flush=True)  # noqa: E731
    model = va.load(wl.BSIM_SIZED_VA)
    N = model.element(polarity="n", differentiable=["vth0"], name="bsim_n")
    P = model.element(polarity="p", differentiable=["vth0"], name="bsim_p")
    c, nodes = wl.bsim_ring(vx.CircuitBuilder(), args.bsim_stages, N, P)
    row = c.device_table()[4]
    cols = wl.vth0_columns(c)
    B = args.bsim_count
    dv = np.random.default_rng(4).normal(0.0, 0.02, B)
    params = np.tile(row, (B, 1))
    params[:, cols["n"]] += dv[:, None]
    params[:, cols["p"]] -= dv[:, None]  # PMOS VTH0 is negative: the same signed shift
    z0 = np.tile(wl.bsim_ring_ic(c, nodes), (B, 1))
    ts = np.arange(args.bsim_steps + 1) * 1e-12
    res = {}
    # reused factors do not pay here: the stiff device-internal nodes make a
    # reused LU contract slowly, and a batch steps at its slowest instance
    variants = {"fp64": dict(), "mixed": dict(precision="mixed"), "pred": dict(predict=True)}
    for name, kw in variants.items():
        opts = vx.Options(solver="sparse", grouping=args.bsim_grouping, **kw)
        transient_sparse_batch(c, params, z0, ts[:5], options=opts)  # plan, code, graphs
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out, conv, its = transient_sparse_batch(c, params, z0, ts, options=opts)
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0, out, conv, its)
        log(f"bsim rings {name}: {res[name][0]:.2f} s, iters/solve "
            f"{float(its[1:].float().mean()):.2f}, converged {bool(conv.all())}")
    zf = res["fp64"][1].cpu().numpy()
    zm = res["mixed"][1].cpu().numpy()
    zp = res["pred"][1].cpu().numpy()
    wall, _, conv, its = res["pred"]
    nv, n0 = c.n_nodes, c.node(nodes[0])
    periods = np.array([measure.period(ts, zf[:, b, n0], level=0.5) for b in range(B)])
    pm = np.array([measure.period(ts, zm[:, b, n0], level=0.5) for b in range(B)])
    # adjoint of instance 0's period (level fixed at VDD/2) with respect to VTH0
    for name, el in c.elements.items():
        if getattr(el, "va", None) is not None:
            el.set(slice(None), vth0=0.4 - dv[0] if name.endswith("_p") else 0.4 + dv[0])
    sol = Solution(ts, zf[:, 0, :], np.ones(len(ts), bool), c.layout)
    g = np.zeros_like(sol.z)
    g[:, n0] = measure.period_vjp(ts, zf[:, 0, n0], 0.5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gr = adjoint.transient_vjp(c, sol, g, ic="zero",
                               options=vx.Options(solver="sparse", grouping=args.bsim_grouping))
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    dvth = float(gr["bsim_n_n"]["vth0"].sum() - gr["bsim_p_p"]["vth0"].sum())
    solves = B * (len(ts) - 1)
    pp = np.array([measure.period(ts, zp[:, b, n0], level=0.5) for b in range(B)])
    line = {
        "metric": "Newton solves/sec (batched, 1 GPU), rings of a BSIM4-sized Verilog-A model",
        "value": solves / wall, "unit": "Newton solves/s", "n_gpus": 1, "steps": 1,
        "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{B} rings of {args.bsim_stages} stages of a 10-unknown "
                               f"Verilog-A MOSFET (BSIM4 stand-in), {args.bsim_steps} BE steps "
                               f"of 1 ps, VTH0 N(0,20mV) per instance",
                   "unknowns": c.size, "pivot_groups": args.bsim_grouping,
                   "path": "pred (Newton started from the cubic through the last four "
                           "points)"},
        "newton_iters_per_solve": float(its[1:].float().mean()),
        "all_converged": bool(conv.all()),
        "plain_fp64_solves_per_s": solves / res["fp64"][0],
        "plain_fp64_iters_per_solve": float(res["fp64"][3][1:].float().mean()),
        "plain_fp64_all_converged": bool(res["fp64"][2].all()),
        "speedup_vs_plain_fp64": res["fp64"][0] / wall,
        "max_abs_v_diff_vs_plain_fp64": float(np.abs(zp[:, :, :nv] - zf[:, :, :nv]).max()),
        "max_rel_period_diff_vs_plain_fp64": float(np.nanmax(np.abs(pp - periods) / periods)),
        "mixed_solves_per_s": solves / res["mixed"][0],
        "mixed_all_converged": bool(res["mixed"][2].all()),
        "max_rel_v_err_mixed_vs_fp64": float(np.abs(zm[:, :, :nv] - zf[:, :, :nv]).max()),
        "period_ps_mean": float(np.nanmean(periods) * 1e12),
        "period_ps_std": float(np.nanstd(periods) * 1e12),
        "max_rel_period_err_mixed_vs_fp64": float(np.nanmax(np.abs(pm - periods) / periods)),
        "dperiod_dvth0_ps_per_V": dvth * 1e12, "adjoint_s": t_adj,
    }
    if emit:
        print(json.dumps(line))
    return line


def warm_workload(args, emit=True):
    """DC operating points of variants of a 64x64 power grid loaded by 512 EKV
    inverters: the plain method (source stepping from the all-zero guess) and
    the learned warm start (4 x 512 MLP trained on the GPU on plain-method
    solutions; fallback to source stepping when the Newton step fails to
    contract), both converged to rtol 1e-9."""
    import torch

    import voltax_amd as vx
    from voltax_amd.dcbatch import DCBatch
    from voltax_amd.warmstart import MLP
    from voltax_amd.workloads import GridInverters

    torch.cuda.set_device(0)
    log = lambda m: print("# " + m, file=sys.stderr, flush=True)  # noqa: E731
    fam = GridInverters(vx.CircuitBuilder, args.grid_warm, args.grid_warm, args.inverters)
    c = fam.circuit
    opts = vx.Options(rtol=1e-9, atol=1e-12)
    dcb = DCBatch(c, chunk=args.chunk, options=opts)
    Ptr, Ftr = fam.variants(args.train, seed=1)
    Pte, Fte = fam.variants(args.warm_count, seed=2)
    Ptr, Ftr, Pte, Fte = (torch.from_numpy(a).cuda() for a in (Ptr, Ftr, Pte, Fte))
    log(f"S={c.size}, train {args.train}, test {args.warm_count}, chunk {args.chunk}")
    dcb.plain(Pte[: min(256, args.warm_count)])  # warm-up: plan, graph capture, allocations
    train = dcb.plain(Ptr)
    log(f"training data: {train.seconds:.1f} s, iters/solve {train.iters.mean():.2f}, "
        f"converged {train.converged.mean():.4f}")
    ok = torch.from_numpy(np.flatnonzero(train.converged)).cuda()
    net = MLP(Ftr.shape[1], c.size, width=512, layers=4, seed=0, lr=args.lr,
              weight_decay=args.wd, skip=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = net.fit(Ftr.index_select(0, ok), train.z.index_select(0, ok), epochs=args.epochs,
                     batch=args.batch, seed=0, lr_decay="cosine", log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    steps = args.epochs * (len(ok) // args.batch)
    dims = net.dims
    mlp_flops = 2 * sum(dims[i] * dims[i + 1] for i in range(len(dims) - 1))
    mlp_flops += 2 * dims[0] * dims[-1]  # linear skip
    guess = lambda f, out: net.predict(f, out=out)  # noqa: E731
    # the warm start may run with FP32 LU + FP64 refinement; the plain method is FP64 LU
    wopts = vx.Options(rtol=1e-9, atol=1e-12, precision=args.warm_precision)
    dcw = dcb if args.warm_precision == "fp64" else DCBatch(c, chunk=args.chunk, options=wopts)
    dcw.warm(Pte[: min(256, args.warm_count)], guess, Fte[: min(256, args.warm_count)])  # warm-up
    dcb.direct(Pte[: min(256, args.warm_count)])  # warm-up of the reference's algorithm
    direct = dcb.direct(Pte)
    log(f"direct (reference DC: Newton from zero, gmin stepping on failure) "
        f"{direct.seconds:.2f} s iters {direct.iters.mean():.2f} "
        f"stepped {direct.stats['gmin_stepped']}")
    plain = dcb.plain(Pte)
    plain_tuned = dict(dcb.tuned)
    if dcw is not dcb:
        dcb._sn.clear()  # free the plain method's buffers before the mixed-precision run
    warm = dcw.warm(Pte, guess, Fte)
    log(f"plain {plain.seconds:.2f} s iters {plain.iters.mean():.2f}; warm {warm.seconds:.2f} s "
        f"iters {warm.iters.mean():.2f} fallback {warm.fallback.mean():.4f}")
    nv = c.n_nodes
    both = plain.converged & warm.converged
    zp, zw = plain.z[:, :nv], warm.z[:, :nv]
    err = float(((zw - zp).abs().max(1).values / zp.abs().max(1).values).max())
    pred = net.predict(Fte)
    guess_err = float((pred[:, :nv] - zp).abs().mean())
    line = {
        "metric": "Newton solves/sec (batched, 1 GPU), learned warm start vs the reference's "
                  "DC algorithm (direct Newton, gmin stepping on failure) at 1e-9 rel",
        "value": args.warm_count / warm.seconds, "unit": "DC solves/s", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": warm.seconds * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 solve, bf16 MLP",
        "data": "synthetic",
        "config": {"workload": f"{args.warm_count} variants of a {args.grid_warm}x{args.grid_warm} "
                               f"power grid + {args.inverters} EKV inverters, DC",
                   "unknowns": c.size, "train_solutions": int(len(ok)),
                   "mlp": f"{len(dims) - 1} layers, width 512, + linear skip, bf16 MFMA"},
        "direct_solves_per_s": args.warm_count / direct.seconds,
        "warm_vs_direct_speedup": direct.seconds / warm.seconds,
        "newton_iters_direct": float(direct.iters.mean()),
        "direct_gmin_stepped": direct.stats["gmin_stepped"],
        "direct_all_converged": bool(direct.converged.all()),
        "plain_solves_per_s": args.warm_count / plain.seconds,
        "warm_vs_plain_speedup": plain.seconds / warm.seconds,
        "newton_iters_plain": float(plain.iters.mean()),
        "newton_iters_warm": float(warm.iters.mean()),
        "fallback_rate": float(warm.fallback.mean()),
        "warm_precision": args.warm_precision,
        "all_converged": bool(plain.converged.all() and warm.converged.all()),
        "converged_both": int(both.sum()),
        "max_rel_v_err_warm_vs_plain": err,
        "mean_abs_guess_err_V": guess_err,
        "guess_seconds": warm.stats["guess_seconds"],
        "train_seconds": t_train, "train_steps": steps, "train_loss_first_last":
            [losses[0], losses[-1]],
        "train_tflops": 3 * mlp_flops * args.batch * steps / t_train / 1e12,
        "training_data_seconds": train.seconds,
        "autotuned_leaf": {str(k): v for k, v in plain_tuned.items()},
    }
    if emit:
        print(json.dumps(line))
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="rings", choices=["rings", "grid", "warm", "bsim"])
    ap.add_argument("--bsim-count", type=int, default=256)
    ap.add_argument("--bsim-stages", type=int, default=51)
    ap.add_argument("--bsim-grouping", default="compact", choices=["reference", "compact"])
    ap.add_argument("--bsim-steps", type=int, default=5000,
                    help="time steps of the BSIM-sized rings (0: skip them in the default run)")
    ap.add_argument("--grid-warm", type=int, default=64)
    ap.add_argument("--inverters", type=int, default=512)
    ap.add_argument("--train", type=int, default=50000)
    ap.add_argument("--warm-count", type=int, default=10000,
                    help="DC variants of the warm-start workload (0: skip it in the default run)")
    ap.add_argument("--chunk", type=int, default=10000)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--wd", type=float, default=0.01)
    ap.add_argument("--warm-precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--grid", type=int, default=100,
                    help="side of the power-grid mesh (0: skip it in the default run)")
    ap.add_argument("--tsteps-grid", type=int, default=200)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--tsteps", type=int, default=2000)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--paths", default="fp64,mixed,band,pred,pred_band,pred_reuse,pred_reuse_band",
                    help="fp64: FP64 LU with partial pivoting (the reference's algorithm); "
                         "mixed: FP32 LU + FP64 refinement; band: FP64 static-pivot band LU "
                         "(checked multipliers, partial pivoting as fallback); pred: each step's "
                         "Newton started from the extrapolated cubic through the last four "
                         "points; reuse: FP64 LU factors reused by later iterations of a step "
                         "while the updates contract; combinations joined by _")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device copy of this size (counter calibration)")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES for this process (the tile classes run on their "
                         "own streams; 0 keeps the environment's value)")
    args = ap.parse_args()
    if args.hw_queues > 0:  # read by the HIP runtime when it initialises (not yet)
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(args.hw_queues, 32))
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part "
                                   "of it (instances are independent and would shard by "
                                   "pattern group)", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "grid":
        return grid_workload(args)
    if args.workload == "warm":
        return warm_workload(args)
    if args.workload == "bsim":
        return bsim_workload(args)

    import torch

    import voltax_amd as vx
    from voltax_amd import workloads as wl
    from voltax_amd.engine import DenseBatch

    torch.cuda.set_device(0)
    specs = wl.ring_batch_specs(args.count, seed=0)
    ts = np.arange(args.tsteps + 1) * 2e-12
    circuits, ics, _ = wl.build_rings(vx.CircuitBuilder, specs)
    ns = np.array([s[0] for s in specs])
    # one batch: the kernel groups instances by pattern and launches the tile
    # classes concurrently
    batch = DenseBatch(circuits)
    z0 = torch.as_tensor(np.concatenate(ics)).cuda()
    ts_dev = torch.as_tensor(ts).cuda()
    stream = torch.cuda.current_stream()
    total_solves = len(specs) * args.tsteps
    cls_of = np.array([tile_class(int(n) + 2) for n in ns])

    if args.calib_copy_mib:
        src = torch.empty(args.calib_copy_mib << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()

    results = {}
    opts = vx.Options()
    for path in args.paths.split(","):
        prec = 1 if path == "mixed" else 0
        pred = "pred" in path
        reuse = "reuse" in path
        lu = "band" if "band" in path else "partial"
        for _ in range(args.warmup):
            batch.run(z0, ts_dev, False, 0.0, opts, "be", precision=prec, stream=stream,
                      predict=pred, reuse=reuse, lu=lu)
        torch.cuda.synchronize()
        kern_runs = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = batch.run(z0, ts_dev, False, 0.0, opts, "be", precision=prec, stream=stream,
                            predict=pred, reuse=reuse, lu=lu)
            kern_runs.append(batch.timings())
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        kern = np.mean(np.stack(kern_runs), axis=0)
        iters = out[2].cpu().numpy()
        facts = batch.last_factorizations.cpu().numpy()
        conv = bool(out[1].cpu().numpy().all())
        zs = batch.split(out[0], len(ts))
        results[path] = dict(wall=wall, kern=kern, iters=iters, facts=facts, conv=conv, zs=zs)
        print(f"# {path}: {wall * 1e3:.1f} ms/pass, {total_solves / wall / 1e6:.2f} M solves/s, "
              f"converged={conv}, per-class kernel ms={np.round(kern, 2).tolist()}",
              file=sys.stderr, flush=True)

    base = results.get("fp64")
    errs = {}  # path -> (max rel voltage error vs FP64, max rel period error)
    for name, res in results.items():
        if base is None or name == "fp64":
            continue
        ev = max(float(np.max(np.abs(a - b_))) for a, b_ in zip(base["zs"], res["zs"])) / VDD
        pe = []
        for i in range(0, len(specs), max(1, len(specs) // 256)):
            node = 1  # n0 (index 0 is the supply node)
            p0 = vx.measure.period(ts, base["zs"][i][:, node])
            p1 = vx.measure.period(ts, res["zs"][i][:, node])
            if np.isfinite(p0) and np.isfinite(p1):
                pe.append(abs(p1 - p0) / p0)
        errs[name] = (ev, float(max(pe)) if pe else None)
    # the headline: the fastest path whose waveforms stay within 1e-9 of FP64
    head = "fp64" if "fp64" in results else list(results)[0]
    for name, (ev, _) in errs.items():
        if ev <= 1e-9 and results[name]["conv"] and results[name]["wall"] < results[head]["wall"]:
            head = name
    err_v, err_period = errs.get("mixed", (None, None))
    r = results[head]
    value = total_solves / r["wall"]
    # roofline of the dominant kernel (the tile-class launch with the longest time)
    dom = int(np.argmax(r["kern"]))
    idx = np.flatnonzero(cls_of == dom)
    # the band LU runs in plain-Newton kernels of every class and in the
    # reusing kernels of the 64-lane classes (voltax_amd launch_dense_class)
    lanes_dom = int(class_tables()[1][dom])
    band_ran = "band" in head and ("reuse" not in head or lanes_dom == 64)
    per_iter = (flops_per_band_iter if band_ran
                else (lambda n: flops_per_iter(n, head == "mixed")))
    per_reused = flops_per_band_reused_iter if band_ran else flops_per_reused_iter
    flops = sum(float(r["facts"][i, 1:].sum()) * per_iter(int(ns[i]))
                + float((r["iters"][i, 1:] - r["facts"][i, 1:]).sum())
                * per_reused(int(ns[i])) for i in idx)
    achieved = flops / (float(r["kern"][dom]) * 1e-3) / 1e12
    line = {
        "metric": "Newton solves/sec (batched, 1 GPU) vs plain FP64 GPU path at 1e-9 rel error",
        "value": value, "unit": "Newton solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["wall"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (f32 LU + f64 refinement)" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(specs)} EKV ring oscillators, 3..31 stages, "
                               f"{args.tsteps} BE steps of 2 ps, VTH0 N(0,30mV), W lognormal 5%",
                   "instances": len(specs), "time_steps": args.tsteps, "path": head},
        "plain_fp64_solves_per_s": (total_solves / results["fp64"]["wall"]
                                    if "fp64" in results else None),
        "mixed_solves_per_s": (total_solves / results["mixed"]["wall"]
                               if "mixed" in results else None),
        "mixed_vs_fp64_speedup": (results["fp64"]["wall"] / results["mixed"]["wall"]
                                  if "fp64" in results and "mixed" in results else None),
        "max_rel_v_err_mixed_vs_fp64": err_v, "max_rel_period_err_mixed_vs_fp64": err_period,
        "predicted_start_solves_per_s": (total_solves / results["pred"]["wall"]
                                         if "pred" in results else None),
        "max_rel_v_err_predicted_vs_fp64": errs.get("pred", (None, None))[0],
        "max_rel_period_err_predicted_vs_fp64": errs.get("pred", (None, None))[1],
        "solves_per_s_by_path": {k: total_solves / v["wall"] for k, v in results.items()},
        "max_rel_v_err_vs_fp64_by_path": {k: v[0] for k, v in errs.items()},
        "max_rel_period_err_vs_fp64_by_path": {k: v[1] for k, v in errs.items()},
        "lu_factorizations_per_solve_by_path": {k: float(np.mean(v["facts"][:, 1:]))
                                                for k, v in results.items()},
        "newton_iters_per_solve_by_path": {k: float(np.mean(v["iters"][:, 1:]))
                                           for k, v in results.items()},
        "newton_iters_per_solve": float(np.mean(r["iters"][:, 1:])),
        "all_converged": r["conv"],
        "roofline": {"bound": "valu_fp64", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": (measured_traffic(dom, kernel_prec(head), band_ran)
                                 if (len(specs) == 4096 and args.tsteps == 2000) else None),
                     "kernel": f"k_dense_run tile class {class_tables()[0][dom]}",
                     "kernel_ms": float(r["kern"][dom]),
                     "class_ms": [round(float(x), 3) for x in r["kern"]]},
    }
    if args.cpu_seconds > 0:
        cb = cpu_baseline(specs, ts, args.cpu_seconds)
        zs_cpu = cb.pop("_z")
        cb.pop("rings")
        if "fp64" in results:  # the plain GPU path against the CPU restatement, same rings
            cb["max_abs_v_diff_gpu_fp64_vs_cpu"] = float(max(
                np.abs(results["fp64"]["zs"][i] - zc).max() for i, zc in enumerate(zs_cpu)))
        line["cpu_baseline"] = cb
        if line["plain_fp64_solves_per_s"]:
            line["plain_fp64_gpu_vs_cpu_1core"] = line["plain_fp64_solves_per_s"] / cb["value"]
        # the box's CPU share for one GPU is 16 cores (whatever nproc shows)
        workers = min(16, os.cpu_count() or 1)
        if workers > 1:
            line["cpu_baseline_all_cores"] = cpu_baseline_all_cores(specs, ts,
                                                                    args.cpu_seconds, workers)
    if args.grid > 0:
        # the single large matrix on the sparse path, same run
        g = grid_workload(args, emit=False)
        line["power_grid"] = {k: g[k] for k in (
            "metric", "value", "unit", "ms_per_step", "config", "newton_iters", "all_converged",
            "reuse_solves_per_s", "reuse_newton_iters", "reuse_lu_factorizations_steps",
            "max_abs_v_diff_reuse_vs_plain", "max_ir_drop_V", "adjoint_s", "grad_params")}
    if args.warm_count > 0:
        # the learned warm start on the DC family it is built for, same run
        del batch, results, base, out, z0
        torch.cuda.empty_cache()
        w = warm_workload(args, emit=False)
        line["warm_start"] = {k: w[k] for k in (
            "metric", "value", "unit", "config", "direct_solves_per_s", "warm_vs_direct_speedup",
            "newton_iters_direct", "direct_gmin_stepped", "plain_solves_per_s",
            "warm_vs_plain_speedup", "newton_iters_plain", "newton_iters_warm", "fallback_rate", "all_converged",
            "max_rel_v_err_warm_vs_plain", "train_seconds", "train_tflops",
            "training_data_seconds", "warm_precision")}
    if args.bsim_steps > 0:
        # rings of the BSIM-sized Verilog-A model on the sparse path, same run
        torch.cuda.empty_cache()
        b = bsim_workload(args, emit=False)
        line["bsim_rings"] = {k: b[k] for k in (
            "metric", "value", "unit", "ms_per_step", "config", "newton_iters_per_solve",
            "all_converged", "plain_fp64_solves_per_s", "plain_fp64_iters_per_solve",
            "speedup_vs_plain_fp64", "max_abs_v_diff_vs_plain_fp64",
            "max_rel_period_diff_vs_plain_fp64", "mixed_solves_per_s",
            "max_rel_v_err_mixed_vs_fp64", "period_ps_mean", "period_ps_std",
            "dperiod_dvth0_ps_per_V", "adjoint_s")}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
