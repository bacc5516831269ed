"""Verilog-A test models and a builder-neutral way to place them.

Each model exercises one part of the subset: instance parameters (RES,
SQNMOS), internal-node collapse and limexp (DIODE), a voltage branch with an
internal node and ddt of a probed branch current (RL), loops with a constant
bound (SUMK) and with a data-dependent exit (HERON), data-dependent if/else
chains and the ternary (SQNMOS), single-node voltage branches (VCVS), and
the math functions (MIXFN)."""

RES = """
`include "disciplines.vams"
module res(p, n);
  inout p, n;
  electrical p, n;
  (* type="instance" *) parameter real r = 1k from (0:inf);
  analog I(p, n) <+ V(p, n) / r;
endmodule
"""

DIODE = """
module diode(a, c);
  inout a, c;
  electrical a, c, ai;
  parameter real is = 1e-14;
  parameter real rs = 0 from [0:inf);
  parameter real cj = 1p;
  parameter real vt = 0.025852;
  (* type="instance" *) parameter real area = 1;
  analog begin
    if (rs > 0)
      I(a, ai) <+ V(a, ai) / rs;
    else
      V(a, ai) <+ 0;
    I(ai, c) <+ area * is * (limexp(V(ai, c) / vt) - 1.0);
    I(ai, c) <+ ddt(area * cj * V(ai, c));
  end
endmodule
"""

RL = """
module rl(p, n);
  inout p, n;
  electrical p, n, mid;
  parameter real r = 10;
  parameter real l = 1m;
  analog begin
    I(p, mid) <+ V(p, mid) / r;
    V(mid, n) <+ l * ddt(I(mid, n));
  end
endmodule
"""

SUMK = """
module sumk(p, n);
  inout p, n;
  electrical p, n;
  parameter real g = 1;
  real s;
  integer k;
  analog begin
    s = 0;
    k = 1;
    while (k <= 6) begin
      s = s + pow(2, -k);
      k = k + 1;
    end
    I(p, n) <+ g * s * V(p, n);
  end
endmodule
"""

HERON = """
module heron(p, n);
  inout p, n;
  electrical p, n;
  parameter real g = 1m;
  parameter integer iters = 8;
  real x, y;
  integer k;
  analog begin
    x = 1 + V(p, n) * V(p, n);
    y = x;
    k = 0;
    while (k < iters && abs(y * y - x) > 1e-14 * x) begin
      y = 0.5 * (y + x / y);
      k = k + 1;
    end
    I(p, n) <+ g * (y - 1);
  end
endmodule
"""

SQNMOS = """
module sqnmos(d, g, s);
  inout d, g, s;
  electrical d, g, s;
  parameter real kp = 2m;
  parameter real vt0 = 0.5;
  parameter real lambda = 0.02;
  (* type="instance" *) parameter real w = 1;
  real vgs, vds, vov, id;
  analog begin
    vgs = V(g, s);
    vds = V(d, s);
    vov = vgs - vt0;
    if (vov <= 0)
      id = 0;
    else if (vds < vov)
      id = w * kp * (vov - 0.5 * vds) * vds;
    else
      id = 0.5 * w * kp * vov * vov;
    id = id * (1 + lambda * (vds > 0 ? vds : -vds));
    I(d, s) <+ id;
  end
endmodule
"""

VCVS = """
module vcvs(i, o);
  inout i, o;
  electrical i, o;
  parameter real a = 4;
  analog V(o) <+ a * V(i);
endmodule
"""

MIXFN = """
module mixfn(p, n);
  inout p, n;
  electrical p, n;
  parameter real g = 1m;
  real v;
  analog begin
    v = V(p, n);
    I(p, n) <+ g * (tanh(v) + 0.1 * sin(v) + 0.01 * ln(1 + v * v) + 0.05 * log(2 + cos(v))
                    + 0.001 * pow(1 + v * v, 1.5) + 0.01 * min(v, 0.3) + 0.01 * max(v, -0.2)
                    + 0.01 * sqrt(1 + v * v) + 0.001 * exp(v / 3) + 0.002 * $abs(v)
                    + (v > 0.1 && v < 0.7 ? 0.003 * v : -0.001 * v) + 0.004 * v / (2 - v * 0.1));
    I(p, n) <+ ddt(1p * v + 2p * V(p) - 0.5p * V(n)) / 2;
  end
endmodule
"""

# a data-dependent branch whose untaken side has a NaN derivative
# (sqrt at negative v): the select must keep it out (tests/test_va.py:209-247
# of the reference)
GUARDED = """
module guarded(p, n);
  inout p, n;
  electrical p, n;
  real v;
  analog begin
    v = V(p, n);
    if (v > 0)
      I(p, n) <+ 1m * sqrt(v);
    else
      I(p, n) <+ 1m * v;
  end
endmodule
"""

# fixed-point iteration with a constant counter and a data-dependent
# tolerance test: the first test is constant (x=0, xo=1), five are traced
# (tests/test_va.py:395-412 of the reference)
ITER = """
module iter(p, n);
  inout p, n;
  electrical p, n;
  real x, xo;
  integer k;
  analog begin
    x = 0.0;
    xo = 1.0;
    k = 0;
    while ((k < 6) && (abs(x - xo) > 1e-3)) begin
      xo = x;
      x = 0.5 * (x + V(p, n));
      k = k + 1;
    end
    I(p, n) <+ x;
  end
endmodule
"""

# four-terminal square-law MOSFET with a polarity parameter (the reference's
# netlist-hook fixture, tests/test_va.py:430-467)
SQUARE_LAW = """
module sqmos(d, g, s, b);
  inout d, g, s, b;
  electrical d, g, s, b;
  parameter integer type = 1 from [-1:1] exclude 0;
  parameter real kp = 200u;
  parameter real vth = 0.4;
  (* type="instance" *) parameter real w = 1u from (0:inf);
  (* type="instance" *) parameter real l = 100n from (0:inf);
  real vgs, vds, vov, id;
  analog begin
    vgs = type * V(g, s);
    vds = type * V(d, s);
    vov = vgs - vth;
    if (vov < 0) vov = 0;
    if (vds < vov) id = kp * w / l * (vov - 0.5 * vds) * vds;
    else id = 0.5 * kp * w / l * vov * vov;
    I(d, s) <+ type * id;
  end
endmodule
"""

# the language features around the analog block: macros with arguments,
# named branches, an aliased parameter, an analog function with an integer
# argument and a for loop, case, integer division and remainder, integer
# rounding, $vt / $temperature / $param_given / $simparam, a system task
# (tests/test_va.py:32-146, 368-390 of the reference)
LANG = """
`include "constants.vams"
`define CLAMP(x, lo, hi) (((x) < (lo)) ? (lo) : (((x) > (hi)) ? (hi) : (x)))
`define NTERMS 5
module lang(p, n);
  inout p, n;
  electrical p, n, mid;
  branch (p, mid) bp;
  branch (mid, n) bn;
  parameter real g0 = 1m from (0:inf);
  parameter integer order = 3 from [1:8];
  parameter integer mode = 1 from [0:2];
  parameter real tc = 0.0;
  parameter real c0 = 1p;
  aliasparam gzero = g0;
  (* type="instance" *) parameter real scale = 1 from (0:inf);
  real v, s, acc;
  integer k, half, rem, nround;

  analog function real poly;  /* sum_j x^j / (j + 1), j < n */
    input x, n;
    real x;
    integer n;
    integer j;
    real t;
    begin
      poly = 0;
      t = 1;
      for (j = 0; j < n; j = j + 1) begin
        poly = poly + t / (j + 1);
        t = t * x;
      end
    end
  endfunction

  analog begin : main
    v = `CLAMP(V(bp) / $vt, -10, 10);
    half = order / 2;
    rem = order % 2;
    nround = 2.5;
    acc = 0;
    for (k = 0; k < `NTERMS; k = k + 1)
      acc = acc + poly(v * 0.1, order + k % 2);
    case (mode)
      0: s = 1;
      1, 2: s = 1 + 0.5 * tanh(v);
      default: s = 0;
    endcase
    if ($param_given(tc))
      s = s * (1 + tc * ($temperature - 300.15));
    $strobe("acc = %g", acc);
    I(bp) <+ scale * g0 * (half + rem + nround / 4.0) * acc * s * V(bp)
             + $simparam("gmin", 0) * V(bp);
    I(bn) <+ 2m * V(bn) + ddt(c0 * V(bn));
  end
endmodule
"""

# integer semantics on device-dependent values: assignment to an integer
# rounds half away from zero, / and % of integers truncate toward zero
INTOPS = """
module intops(p, n);
  inout p, n;
  electrical p, n;
  integer q;
  real x;
  analog begin
    x = 7 * V(p, n);
    q = x;
    I(p, n) <+ 1m * (q / 2 + q % 3 + 0.25 * (q / 2.0)) + 1m * x + 1m * (x % 2.5);
  end
endmodule
"""

# A BSIM4-sized synthetic MOSFET: drain, source, gate and body resistances
# to internal nodes, a series gate inductance (a branch current), an EKV-style
# core and a binning-style preamble whose loop unrolls into a few hundred
# instructions: 10 unknowns per device, the size class of BSIM4 with
# rdsmod/rgatemod/rbodymod (more than the interpreter's 4, so it runs as
# generated code).
BSIMISH = """
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

ALL = {"res": RES, "diode": DIODE, "rl": RL, "sumk": SUMK, "heron": HERON,
       "sqnmos": SQNMOS, "vcvs": VCVS, "mixfn": MIXFN, "guarded": GUARDED,
       "iter": ITER, "lang": LANG}


class Pair:
    """The product's element class and the oracle's model for one source."""

    def __init__(self, source, card=None, **kw):
        from oracle import va as ova
        from voltax_amd import va

        self.product = va.load(source).element(card, **kw)
        self.oracle = ova.Model(source, card, **kw)


def place(b, pair, terminals, name=None, **inst):
    """Add one device to a product or oracle builder."""
    from oracle import circuit as oc

    if isinstance(b, oc.CircuitBuilder):
        return b.va(pair.oracle, terminals, name=name, **inst)
    return b.add(pair.product(terminals, **inst), name=name)
