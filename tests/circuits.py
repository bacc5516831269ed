"""Test circuits, written against the builder API that the product
(voltax_amd.CircuitBuilder) and the oracle (oracle.circuit.CircuitBuilder)
share, so each test builds the same topology for both.  The circuits follow
the reference's own test circuits (tests/test_analysis.py:13-20,
tests/test_sparse.py:128-151) and the workloads' descriptions."""

import numpy as np


def rc_ladder(b, R1=1e3, R2=2e3, C1=1e-9, C2=0.5e-9, Rload=1e4):
    b.vsource("in", "0", 1.0, name="Vin")
    b.resistor("in", "a", R1, name="R1")
    b.resistor("a", "out", R2, name="R2")
    b.resistor("out", "0", Rload, name="Rload")
    b.capacitor("a", "0", C1, name="C1")
    b.capacitor("out", "0", C2, name="C2")
    return b.build()


def inverter_chain(b, pulse, stages=3, wn=1e-6, wp=2e-6, l=0.13e-6, cl=5e-15, vdd=1.2, ac=0.0):
    """The 3-stage EKV inverter chain of the reference README (README.md:60-71)."""
    b.vsource("vdd", "0", vdd, name="Vdd")
    b.vsource("in", "0", pulse, ac=ac, name="Vin")
    names = ["in"] + [chr(ord("a") + i) for i in range(stages - 1)] + ["out"]
    for i, (a, y) in enumerate(zip(names[:-1], names[1:])):
        b.pmos(y, a, "vdd", "vdd", w=wp, l=l, name=f"X{i + 1}.M1")
        b.nmos(y, a, "0", "0", w=wn, l=l, name=f"X{i + 1}.M2")
        b.capacitor(y, "0", cl, name=f"X{i + 1}.Cl")
    return b.build()


def ring(b, n, cl=20e-15, vdd=1.2, dvth=None, wscale=None, wn=1e-6, wp=2e-6, l=0.13e-6):
    """n-stage EKV ring oscillator with per-device mismatch."""
    b.vsource("vdd", "0", vdd, name="Vdd")
    nodes = [f"n{k}" for k in range(n)]
    for i, a in enumerate(nodes):
        y = nodes[(i + 1) % n]
        dp = 0.0 if dvth is None else dvth[2 * i]
        dn = 0.0 if dvth is None else dvth[2 * i + 1]
        sp = 1.0 if wscale is None else wscale[2 * i]
        sn = 1.0 if wscale is None else wscale[2 * i + 1]
        b.pmos(y, a, "vdd", "vdd", w=wp * sp, l=l, dvth=dp)
        b.nmos(y, a, "0", "0", w=wn * sn, l=l, dvth=dn)
    for node in nodes:
        b.capacitor(node, "0", cl)
    return b.build(), nodes


def ring_ic(c, nodes, vdd=1.2):
    return c.state(v={"vdd": vdd, **{nd: (0.9 if k % 2 == 0 else 0.1) * vdd
                                     for k, nd in enumerate(nodes)}})


def grid(b, n=10, seed=0):
    """Random resistive grid (tests/test_sparse.py:128-141)."""
    rng = np.random.default_rng(seed)
    for i in range(n):
        for j in range(n):
            if j + 1 < n:
                b.resistor(f"n{i}_{j}", f"n{i}_{j + 1}", 0.1 * (1 + rng.random()))
            if i + 1 < n:
                b.resistor(f"n{i}_{j}", f"n{i + 1}_{j}", 0.1 * (1 + rng.random()))
            if rng.random() < 0.2:
                b.isource(f"n{i}_{j}", "0", 1e-3 * (1 + rng.random()))
    for i, j in [(0, 0), (0, n - 1), (n - 1, 0), (n - 1, n - 1)]:
        b.vsource(f"n{i}_{j}", "0", 1.0)
    return b.build()


def star(b, n=80, seed=0):
    """A hub node tied to n leaves (resistor, load capacitor and current source
    each) and driven through an inverter: the hub's residual row, its
    diagonal slot and its pattern row each sum over more than 32 terms."""
    rng = np.random.default_rng(seed)
    b.vsource("vdd", "0", 1.2)
    b.vsource("in", "0", 0.3)
    b.nmos("hub", "in", "0", "0", w=1e-6, l=0.13e-6)
    b.pmos("hub", "in", "vdd", "vdd", w=2e-6, l=0.13e-6)
    for k in range(n):
        b.resistor("hub", f"l{k}", 100.0 * (1 + rng.random()))
        b.capacitor(f"l{k}", "0", 1e-15 * (1 + rng.random()))
        b.isource(f"l{k}", "0", 1e-5 * rng.random())
        b.capacitor("hub", f"l{k}", 0.2e-15)
    return b.build()


def ladder(b, n=160, seed=0):
    """RC ladder of n nodes driven by a 1 V source: with a leaf size above n
    the whole matrix is one supernode (one large dense front)."""
    rng = np.random.default_rng(seed)
    b.vsource("n0", "0", 1.0)
    for i in range(n - 1):
        b.resistor(f"n{i}", f"n{i + 1}", 10.0 * (1 + rng.random()))
        b.capacitor(f"n{i + 1}", "0", 1e-15 * (1 + rng.random()))
        b.resistor(f"n{i + 1}", "0", 1e3 * (1 + rng.random()))
    return b.build()


def grid_inverters(b, log_r, sink, bits, nx, ny, sites, vdd=1.0, cl=5e-15, wn=1e-6, wp=2e-6,
                   l=0.13e-6):
    """The grid-with-inverters family written out device by device with given
    values (same construction order as voltax_amd.workloads.GridInverters)."""
    name = lambda i, j: f"g{i}_{j}"  # noqa: E731
    k = 0
    for i in range(nx):
        for j in range(ny):
            if j + 1 < ny:
                b.resistor(name(i, j), name(i, j + 1), float(np.exp(log_r[k])))
                k += 1
            if i + 1 < nx:
                b.resistor(name(i, j), name(i + 1, j), float(np.exp(log_r[k])))
                k += 1
    boundary = ([(0, j) for j in range(ny)] + [(i, ny - 1) for i in range(1, nx)]
                + [(nx - 1, j) for j in range(ny - 2, -1, -1)]
                + [(i, 0) for i in range(nx - 2, 0, -1)])
    for m, (i, j) in enumerate(boundary):
        if m % 10 == 0:
            b.vsource(name(i, j), "0", vdd)
    for i in range(nx):
        for j in range(ny):
            b.isource(name(i, j), "0", float(sink[i * ny + j]))
    for q, site in enumerate(sites):
        g = name(int(site) // ny, int(site) % ny)
        b.pmos(f"out{q}", f"in{q}", g, g, w=wp, l=l)
        b.nmos(f"out{q}", f"in{q}", "0", "0", w=wn, l=l)
        b.capacitor(f"out{q}", "0", cl)
        b.vsource(f"in{q}", "0", float(bits[q] * vdd))
    return b.build()


def va_grid(b, n, place, diode, rl, seed=0):
    """n x n resistive grid with Verilog-A devices that own internal unknowns:
    a diode with series resistance (one internal node: a two-terminal device
    matched to one terminal) at every 7th node, an RL branch (internal node
    plus branch current: glued whole) across every 11th horizontal segment,
    and a row of floating 10 mV sources in series (two candidate partners
    each, so the matching has choices).  `place(b, pair, terminals)` adds a
    Verilog-A device to either builder."""
    rng = np.random.default_rng(seed)
    nm = lambda i, j: f"n{i}_{j}"  # noqa: E731
    k = 0
    for i in range(n):
        for j in range(n):
            if j + 1 < n:
                if k % 11 == 5:
                    place(b, rl, (nm(i, j), nm(i, j + 1)))
                else:
                    b.resistor(nm(i, j), nm(i, j + 1), 0.1 * (1 + rng.random()))
                k += 1
            if i + 1 < n:
                b.resistor(nm(i, j), nm(i + 1, j), 0.1 * (1 + rng.random()))
            if (i * n + j) % 7 == 3:
                place(b, diode, (nm(i, j), "0"))
            b.isource(nm(i, j), "0", 1e-3 * rng.random())
    for j in range(1, n // 2):
        b.vsource(f"f{j - 1}" if j > 1 else nm(n - 1, 0), f"f{j}", 0.01)
    b.resistor(f"f{n // 2 - 1}", nm(n - 1, n // 2), 1.0)
    for i, j in [(0, 0), (0, n - 1), (n - 1, n - 1)]:
        b.vsource(nm(i, j), "0", 1.0)
    return b.build()


# --- static CMOS gates (voltax/library/cmos.py:27-94): NMOS bulks on gnd,
# PMOS bulks on vdd, a series stack of k devices widened k-fold
WN, WP = 1e-6, 2e-6


def cmos_inverter(b, a, y, vdd, gnd="0", wn=WN, wp=WP):
    b.pmos(y, a, vdd, vdd, w=wp)
    b.nmos(y, a, gnd, gnd, w=wn)


def cmos_nand(b, inputs, y, vdd, gnd="0", wn=WN, wp=WP):
    k = len(inputs)
    for a in inputs:
        b.pmos(y, a, vdd, vdd, w=wp)
    lower = gnd
    for i, a in enumerate(inputs):
        upper = y if i == k - 1 else b.node()
        b.nmos(upper, a, lower, gnd, w=k * wn)
        lower = upper


def cmos_nor(b, inputs, y, vdd, gnd="0", wn=WN, wp=WP):
    k = len(inputs)
    for a in inputs:
        b.nmos(y, a, gnd, gnd, w=wn)
    upper = vdd
    for i, a in enumerate(inputs):
        lower = y if i == k - 1 else b.node()
        b.pmos(lower, a, upper, vdd, w=k * wp)
        upper = lower


def cmos_and(b, inputs, y, vdd, gnd="0"):
    mid = b.node()
    cmos_nand(b, inputs, mid, vdd, gnd)
    cmos_inverter(b, mid, y, vdd, gnd)


def cmos_or(b, inputs, y, vdd, gnd="0"):
    mid = b.node()
    cmos_nor(b, inputs, mid, vdd, gnd)
    cmos_inverter(b, mid, y, vdd, gnd)


def cmos_xor(b, a, c, y, vdd, gnd="0"):
    n1, n2, n3 = b.node(), b.node(), b.node()
    cmos_nand(b, [a, c], n1, vdd, gnd)
    cmos_nand(b, [a, n1], n2, vdd, gnd)
    cmos_nand(b, [c, n1], n3, vdd, gnd)
    cmos_nand(b, [n2, n3], y, vdd, gnd)


def cmos_full_adder(b, a, c, cin, s, cout, vdd, gnd="0"):
    p, g, t = b.node(), b.node(), b.node()
    cmos_xor(b, a, c, p, vdd, gnd)
    cmos_xor(b, p, cin, s, vdd, gnd)
    cmos_and(b, [a, c], g, vdd, gnd)
    cmos_and(b, [cin, p], t, vdd, gnd)
    cmos_or(b, [g, t], cout, vdd, gnd)


def full_adder_dc(b, a=0.0, c=0.0):
    """The reference's gmin-stepping fixture (tests/test_analysis.py:138-149):
    a CMOS full adder with inputs a, c at 0 V (or the given values), carry-in
    grounded, VDD 1.2 V."""
    b.vsource("vdd", "0", 1.2)
    b.vsource("a", "0", a)
    b.vsource("c", "0", c)
    cmos_full_adder(b, "a", "c", "0", "s", "co", "vdd")
    return b.build()
