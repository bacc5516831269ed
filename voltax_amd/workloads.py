"""Synthetic workloads of the benchmark configurations, written against the
builder API (so the oracle's builder can build the same circuits).

* `ring_batch`   4096 Monte-Carlo EKV ring oscillators of uneven length
                 (odd stage count uniform in 3..31), VTH0 offset ~ N(0, 30 mV)
                 and W lognormal (sigma 5 %) per device, 20 fF loads,
                 VDD = 1.2 V (benchmarks/benchmark_ring_oscillator.py:30-49
                 gives the topology and the alternating initial condition).
* `power_grid`   R mesh with pads, DC sinks, decaps and pulsed sinks.
* `inverter_chain`  the 3-stage EKV chain of the reference README.
* `bsim_ring`    ring oscillators of a BSIM4-sized Verilog-A MOSFET (10
                 unknowns per device: terminal resistances to internal nodes,
                 a gate inductance with its branch current, a binning-style
                 preamble of a few hundred instructions and an EKV-style core),
                 the stand-in for BSIM4 (whose source is not available here)
                 in the configuration of 51-stage rings at VDD = 1.0 V.
"""

from __future__ import annotations

import numpy as np

WN, WP, L = 1e-6, 2e-6, 0.13e-6  # voltax/library/cmos.py:23, builder default L


def ring(b, n, dvth=None, wscale=None, cl=20e-15, vdd=1.2):
    b.vsource("vdd", "0", vdd, name="Vdd")
    nodes = [f"n{k}" for k in range(n)]
    for i, a in enumerate(nodes):
        y = nodes[(i + 1) % n]
        b.pmos(y, a, "vdd", "vdd", w=WP * (1.0 if wscale is None else wscale[2 * i]), l=L,
               dvth=0.0 if dvth is None else dvth[2 * i])
        b.nmos(y, a, "0", "0", w=WN * (1.0 if wscale is None else wscale[2 * i + 1]), l=L,
               dvth=0.0 if dvth is None else dvth[2 * i + 1])
    for node in nodes:
        b.capacitor(node, "0", cl)
    return b.build(), nodes


def ring_ic(c, nodes, vdd=1.2):
    """Alternating 0.9/0.1 VDD start (benchmarks/benchmark_ring_oscillator.py:44-49)."""
    z = c.zeros()
    z[c.node("vdd")] = vdd
    for k, nd in enumerate(nodes):
        z[c.node(nd)] = (0.9 if k % 2 == 0 else 0.1) * vdd
    return z


def ring_batch_specs(count=4096, seed=0, sigma_vth=0.03, sigma_w=0.05):
    """(n, dvth, wscale) per instance."""
    rng = np.random.default_rng(seed)
    ns = 2 * rng.integers(1, 16, size=count) + 1  # 3, 5, ..., 31
    out = []
    for n in ns:
        out.append((int(n), rng.normal(0.0, sigma_vth, 2 * n),
                    np.exp(rng.normal(0.0, sigma_w, 2 * n))))
    return out


def build_rings(builder_cls, specs):
    circuits, ics, nodes = [], [], []
    for n, dv, ws in specs:
        c, nd = ring(builder_cls(), n, dv, ws)
        circuits.append(c)
        nodes.append(nd)
        ics.append(ring_ic(c, nd))
    return circuits, ics, nodes


def power_grid(b, nx=100, ny=100, seed=0, pulse_cls=None, pulse_width=100e-12):
    """R mesh nx*ny: segments logU[0.05, 0.5] ohm, 1.0 V pads at every 10th
    boundary node, per-node DC sink U[0, 1 mA], decap logU[10, 100] fF, and
    (if pulse_cls) a trapezoidal sink amp U[0, 2 mA], start U[0, 1 ns], 20 ps
    edges, `pulse_width` flat top."""
    rng = np.random.default_rng(seed)
    name = lambda i, j: f"g{i}_{j}"  # noqa: E731
    for i in range(nx):
        for j in range(ny):
            if j + 1 < ny:
                b.resistor(name(i, j), name(i, j + 1), float(np.exp(rng.uniform(np.log(0.05),
                                                                               np.log(0.5)))))
            if i + 1 < nx:
                b.resistor(name(i, j), name(i + 1, j), float(np.exp(rng.uniform(np.log(0.05),
                                                                               np.log(0.5)))))
    boundary = ([(0, j) for j in range(ny)] + [(i, ny - 1) for i in range(1, nx)]
                + [(nx - 1, j) for j in range(ny - 2, -1, -1)]
                + [(i, 0) for i in range(nx - 2, 0, -1)])
    for k, (i, j) in enumerate(boundary):
        if k % 10 == 0:
            b.vsource(name(i, j), "0", 1.0)
    for i in range(nx):
        for j in range(ny):
            b.isource(name(i, j), "0", float(rng.uniform(0, 1e-3)))
            b.capacitor(name(i, j), "0", float(np.exp(rng.uniform(np.log(10e-15),
                                                                  np.log(100e-15)))))
            if pulse_cls is not None:
                amp, start = float(rng.uniform(0, 2e-3)), float(rng.uniform(0, 1e-9))
                b.isource(name(i, j), "0", pulse_cls(0.0, amp, start, 20e-12, 20e-12,
                                                     pulse_width, 1.0))
    return b.build()


class GridInverters:
    """A power grid of nx*ny nodes (the distributions of `power_grid`, DC only)
    loaded by n_inv EKV inverters whose PMOS source and body sit on grid nodes
    chosen once (layout_seed), so every variant shares one pattern.  Inverter k
    has input node in{k} driven by a V source at bit*vdd and output out{k} with
    a load capacitor.

    A variant redraws every segment resistance (logU[0.05, 0.5] ohm), every
    per-node DC sink (U[0, 1 mA]) and every input bit (Bernoulli(0.5));
    `variants` writes them straight into parameter rows of the one pattern."""

    def __init__(self, builder_cls, nx=64, ny=64, n_inv=512, layout_seed=1, vdd=1.0,
                 cl=5e-15):
        self.nx, self.ny, self.n_inv, self.vdd = nx, ny, n_inv, vdd
        b = builder_cls()
        name = lambda i, j: f"g{i}_{j}"  # noqa: E731
        for i in range(nx):
            for j in range(ny):
                if j + 1 < ny:
                    b.resistor(name(i, j), name(i, j + 1), 0.1)
                if i + 1 < nx:
                    b.resistor(name(i, j), name(i + 1, j), 0.1)
        boundary = ([(0, j) for j in range(ny)] + [(i, ny - 1) for i in range(1, nx)]
                    + [(nx - 1, j) for j in range(ny - 2, -1, -1)]
                    + [(i, 0) for i in range(nx - 2, 0, -1)])
        for k, (i, j) in enumerate(boundary):
            if k % 10 == 0:
                b.vsource(name(i, j), "0", vdd)
        for i in range(nx):
            for j in range(ny):
                b.isource(name(i, j), "0", 0.5e-3)
        lrng = np.random.default_rng(layout_seed)
        self.sites = lrng.choice(nx * ny, n_inv, replace=False)
        for k, site in enumerate(self.sites):
            g = name(int(site) // ny, int(site) % ny)
            b.pmos(f"out{k}", f"in{k}", g, g, w=WP, l=L)
            b.nmos(f"out{k}", f"in{k}", "0", "0", w=WN, l=L)
            b.capacitor(f"out{k}", "0", cl)
            b.vsource(f"in{k}", "0", 0.0)
        self.circuit = c = b.build()
        kind, wave, unk, poff, row = c.device_table()
        self.row = row
        from .circuit import KIND_I, KIND_R, KIND_V

        self.r_slots = poff[kind == KIND_R]
        self.i_slots = poff[kind == KIND_I]
        vdev = np.flatnonzero(kind == KIND_V)
        in_nodes = {c.node(f"in{k}"): k for k in range(n_inv)}
        self.in_slots = np.zeros(n_inv, np.int64)
        for d in vdev:
            k = in_nodes.get(int(unk[d, 0]))
            if k is not None:
                self.in_slots[k] = poff[d]
        assert len(self.r_slots) == nx * (ny - 1) + (nx - 1) * ny and len(self.i_slots) == nx * ny
        self.n_features = len(self.r_slots) + len(self.i_slots) + n_inv

    def variants(self, count, seed):
        """(params [count, P], features [count, n_features]) as float64 numpy:
        parameter rows of the pattern and the network inputs (log r, sink in mA,
        input bit) of `count` variants drawn from `seed`."""
        rng = np.random.default_rng(seed)
        log_r = rng.uniform(np.log(0.05), np.log(0.5), (count, len(self.r_slots)))
        sink = rng.uniform(0.0, 1e-3, (count, len(self.i_slots)))
        bits = rng.integers(0, 2, (count, self.n_inv)).astype(float)
        params = np.tile(self.row, (count, 1))
        params[:, self.r_slots] = log_r
        params[:, self.i_slots] = sink
        params[:, self.in_slots] = bits * self.vdd
        return params, np.concatenate([log_r, sink * 1e3, bits], axis=1)


# A BSIM4-sized synthetic MOSFET in the reference's Verilog-A subset.
BSIM_SIZED_VA = """
`include "disciplines.vams"
module bsimish(d, g, s, b);
  inout d, g, s, b;
  electrical d, g, s, b, di, si, gi, bi, gx;
  parameter real rd = 20 from (0:inf);
  parameter real rs = 20 from (0:inf);
  parameter real rg = 50 from (0:inf);
  parameter real rb = 200 from (0:inf);
  parameter real vth0 = 0.4;
  parameter real kp = 4e-4;
  parameter real nfac = 1.3;
  parameter real lam = 0.05;
  parameter real vt = 0.025852;
  parameter real cgs0 = 0.3f;
  parameter real cgd0 = 0.2f;
  parameter real lg = 1p;
  parameter integer type = 1 from [-1:1] exclude 0;
  parameter integer nbin = 40;
  (* type="instance" *) parameter real w = 1u from (0:inf);
  (* type="instance" *) parameter real l = 0.1u from (0:inf);
  real vgs, vds, vbs, vth, vp, ff, fr, ids, acc, t, sabs;
  integer k;
  analog begin
    I(d, di) <+ V(d, di) / rd;
    I(s, si) <+ V(s, si) / rs;
    I(g, gx) <+ V(g, gx) / rg;
    V(gx, gi) <+ lg * ddt(I(gx, gi));
    I(b, bi) <+ V(b, bi) / rb;
    vgs = type * V(gi, si);
    vds = type * V(di, si);
    vbs = type * V(bi, si);
    acc = 0.0;
    t = 1.0;
    for (k = 0; k < nbin; k = k + 1) begin
      t = t * (1.0 + 0.01 * vbs) / (1.0 + 0.002 * k);
      acc = acc + t * exp(-0.05 * k) * (1.0 + 0.001 * vgs * vgs);
    end
    vth = vth0 - 0.05 * vbs + 1e-3 * acc / nbin;
    vp = (vgs - vth) / nfac;
    ff = ln(1.0 + limexp(vp / (2.0 * vt)));
    fr = ln(1.0 + limexp((vp - vds) / (2.0 * vt)));
    sabs = sqrt(vds * vds + 1e-6) - 1e-3;
    ids = 2.0 * nfac * kp * (w / l) * vt * vt * (ff * ff - fr * fr) * (1.0 + lam * sabs);
    I(di, si) <+ type * ids;
    I(gi, si) <+ ddt(cgs0 * (w / 1u) * V(gi, si));
    I(gi, di) <+ ddt(cgd0 * (w / 1u) * V(gi, di));
    I(bi, si) <+ 1e-9 * V(bi, si);
    I(bi, di) <+ 1e-9 * V(bi, di);
  end
endmodule
"""


def bsim_ring(b, n, nmos, pmos, vdd=1.0, wn=0.2e-6, wp=0.4e-6, l=45e-9, cl=1e-15):
    """n-stage ring of the BSIM4-sized model: `nmos` / `pmos` are the product's
    element classes (`va.load(BSIM_SIZED_VA).element(polarity=...)`) or the
    oracle's models (placed with `b.va`); a load `cl` on every stage."""
    b.vsource("vdd", "0", vdd, name="Vdd")
    nodes = [f"n{k}" for k in range(n)]

    def place(model, terminals, **inst):
        if hasattr(b, "va"):  # the oracle's builder
            return b.va(model, terminals, **inst)
        return b.add(model(terminals, **inst))

    for i, a in enumerate(nodes):
        y = nodes[(i + 1) % n]
        place(pmos, (y, a, "vdd", "vdd"), w=wp, l=l)
        place(nmos, (y, a, "0", "0"), w=wn, l=l)
        b.capacitor(y, "0", cl)
    return b.build(), nodes


def bsim_ring_ic(c, nodes, vdd=1.0):
    """Alternating 0.9/0.1 VDD start on the ring nodes; internal nodes of each
    device start at 0 (the first steps settle them)."""
    z = c.zeros()
    z[c.node("vdd")] = vdd
    for k, nd in enumerate(nodes):
        z[c.node(nd)] = (0.9 if k % 2 == 0 else 0.1) * vdd
    return z


def vth0_columns(c):
    """Indices of every device's vth0 entry in the circuit's parameter row
    (models made with ``differentiable=["vth0"]``), NMOS and PMOS apart."""
    cols = {"n": [], "p": []}
    off = 0
    for name, el in c.elements.items():
        rows = el.param_rows()
        w = rows.shape[1]
        if getattr(el, "va", None) is not None:
            j = el._col("vth0")
            pol = "p" if name.endswith("_p") else "n"
            cols[pol].extend(off + i * w + j for i in range(el.size))
        off += w * el.size
    return {k: np.asarray(v, np.int64) for k, v in cols.items()}
