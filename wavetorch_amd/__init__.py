"""wavetorch_amd -- wavetorch's WaveRNN time loop on AMD Instinct MI355X (gfx950).

Drop-in for the reference's module API (fancompute/wavetorch v0.2.1,
``wavetorch/__init__.py:2-10``): the same class names, constructor
signatures, buffers and ``state_dict`` keys.  The time loop, its adjoint and
the geometry parameterisation run in hand-written HIP kernels loaded from
``wavetorch_amd/_lib/libwavetorch_hip.so`` through the C ABI declared in
``include/wavetorch_hip.h``.
"""
from . import cell, data, geom, io, loss, operators, probe, rnn, source, utils
from .cell import TimeStep, WaveCell
from .geom import WaveGeometryFreeForm, WaveGeometryHoley
from .probe import WaveIntensityProbe, WaveProbe
from .rnn import WaveRNN
from .source import WaveLineSource, WaveSource
from .train import GraphedTrainStep, train

__all__ = ["WaveCell", "WaveGeometryHoley", "WaveGeometryFreeForm", "WaveProbe",
           "WaveIntensityProbe", "WaveRNN", "WaveSource", "WaveLineSource"]

__version__ = "0.2.1"
