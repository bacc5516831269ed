"""voltax_amd: MI355X-native Newton inner loop of the voltax circuit simulator.

The public names mirror the reference package (``import voltax as vx``):
circuits are built with `CircuitBuilder`, analysed with `dc` / `transient`,
measured with `measure`; `sparse` mirrors voltax/sparse.py.  Every
numerical kernel runs in HIP on the GPU (libvoltax_mi355x.so).
"""

from . import measure, signals, va
from .analysis import (ACSolution, Options, Solution, ac, dc, dc_batch, dc_sweep, dense_solve,
                       linearize, solve_root, solve_root_vjp, transient, transient_batch)
from .circuit import (VT_300K, Capacitor, Circuit, CircuitBuilder, CurrentSource, EKVMOSFET,
                      Layout, MOSProcess, Resistor, VoltageSource)
from .netlist import NetlistError, parse_netlist, parse_value
from .signals import PWL, Constant, Pulse, Sine, Step

__all__ = [
    "ACSolution", "Options", "Solution", "ac", "dc", "dc_batch", "dc_sweep", "transient",
    "transient_batch", "linearize", "solve_root", "solve_root_vjp", "dense_solve",
    "Capacitor",
    "Circuit", "CircuitBuilder", "CurrentSource", "EKVMOSFET", "Layout", "MOSProcess",
    "Resistor", "VoltageSource", "VT_300K", "PWL", "Constant", "Pulse", "Sine", "Step",
    "measure", "signals", "va", "parse_netlist", "parse_value", "NetlistError",
]
