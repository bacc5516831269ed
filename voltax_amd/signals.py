"""Source waveforms (the reference's voltax/signals.py:56-189).

A waveform is data: its parameters are packed into the device parameter row
of the source that uses it and evaluated on the GPU every Newton iteration
(`devices.hpp: wave_value`).  The layouts are listed in
include/voltax_mi355x.h (VXM_WAVE_*).
"""

from __future__ import annotations

import numpy as np

WAVE_CONST, WAVE_STEP, WAVE_PULSE, WAVE_SINE, WAVE_PWL = 1, 2, 3, 4, 5


class Signal:
    """Base class; ``fields`` name the per-device parameters in row order."""

    code = 0
    fields: tuple[str, ...] = ()

    def __init__(self, **values):
        self.values = {k: np.asarray(v, dtype=float) for k, v in values.items()}

    def row(self) -> np.ndarray:
        """Parameter row of one device."""
        return np.array([float(self.values[f]) for f in self.fields])

    def key(self) -> tuple:
        """Devices with equal keys can share a group (same waveform type and
        parameter shapes, like the reference's group_key)."""
        return (type(self).__name__,)

    def __getattr__(self, name):
        values = self.__dict__.get("values", {})
        if name in values:
            return values[name]
        raise AttributeError(name)


class Constant(Signal):
    code, fields = WAVE_CONST, ("value",)

    def __init__(self, value=0.0):
        super().__init__(value=value)


class Step(Signal):
    """Linear ramp from v0 to v1 starting at delay, lasting rise."""

    code, fields = WAVE_STEP, ("v0", "v1", "delay", "rise")

    def __init__(self, v0=0.0, v1=1.0, delay=0.0, rise=1e-12):
        super().__init__(v0=v0, v1=v1, delay=delay, rise=rise)


class Pulse(Signal):
    """SPICE ``PULSE(v1 v2 delay rise fall width period)``."""

    code = WAVE_PULSE
    fields = ("v1", "v2", "delay", "rise", "fall", "width", "period")

    def __init__(self, v1=0.0, v2=1.0, delay=0.0, rise=1e-9, fall=1e-9, width=1e-6,
                 period=2e-6):
        super().__init__(v1=v1, v2=v2, delay=delay, rise=rise, fall=fall, width=width,
                         period=period)


class Sine(Signal):
    """SPICE ``SIN(offset amplitude freq delay damping phase_deg)``."""

    code = WAVE_SINE
    fields = ("offset", "amplitude", "freq", "delay", "damping", "phase")

    def __init__(self, offset=0.0, amplitude=1.0, freq=1.0, delay=0.0, damping=0.0,
                 phase=0.0):
        super().__init__(offset=offset, amplitude=amplitude, freq=freq, delay=delay,
                         damping=damping, phase=phase)


class PWL(Signal):
    """Piecewise-linear waveform through (times[k], values[k]), with SPICE
    ``td=`` (delay) and ``r=`` (repeat) semantics."""

    code, fields = WAVE_PWL, ("times", "values", "delay", "repeat")

    def __init__(self, times, values, delay=0.0, repeat=-1.0):
        times, values = np.asarray(times, float), np.asarray(values, float)
        if times.shape != values.shape or times.ndim != 1 or len(times) < 2:
            raise ValueError("PWL needs matching 1-D times/values with >= 2 points")
        super().__init__(times=times, values=values, delay=delay, repeat=repeat)

    def row(self):
        t, v = self.values["times"], self.values["values"]
        return np.concatenate([[len(t), float(self.values["delay"]),
                                float(self.values["repeat"])], t, v])

    def key(self):
        return ("PWL", len(self.values["times"]))


def as_signal(value) -> Signal:
    return value if isinstance(value, Signal) else Constant(value)
