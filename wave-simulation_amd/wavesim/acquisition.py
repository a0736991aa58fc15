"""Source wavelets and acquisition helpers (host side, once per shot).

ricker() restates src/Acquisition/SourceSignal/Ricker.cpp:29-53 in float32:
    tau = pi*fc*(t - 1.5/fc - tshift),  s = amp*(1 - 2 tau^2) exp(-tau^2),  t = k*DT.
source_signal() returns any of the nine synthetic shapes of
Sources::generateSyntheticSignal (src/Acquisition/Sources.cpp:157-220) from the library
(wv_source_signal, host code).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi

SHAPES = {"ricker": 1, "sinw": 2, "sinthree": 3, "fgaussian": 4, "spike": 5, "intgsinthree": 6,
          "ricker_gprmax": 7, "berlage": 8, "sin": 9}


def source_signal(shape, NT, DT, fc, amp=1.0, tshift=0.0):
    """Synthetic signal of wavelet shape 1..9 (or its name, SHAPES) as float32 [NT]."""
    shape = SHAPES[shape.lower()] if isinstance(shape, str) else int(shape)
    out = np.zeros(max(int(NT), 1), np.float32)
    _abi.check(_abi.load().wv_source_signal(shape, int(NT), float(DT), float(fc), float(amp), float(tshift),
                                            out.ctypes.data_as(C.POINTER(C.c_float))))
    return out


def ricker(NT, DT, fc, amp=1.0, tshift=0.0):
    f32 = np.float32
    t = np.arange(NT, dtype=f32) * f32(DT)
    help_ = f32(1.5 / float(f32(fc)) + float(f32(tshift)))
    tau = (t - help_) * f32(np.pi * float(f32(fc)))
    h2 = tau * tau
    e = np.exp(-h2.astype(np.float64)).astype(f32)
    w = f32(1) - f32(2) * h2
    return (f32(amp) * w * e).astype(f32)


def nt_from(T, DT):
    """NT = (IndexType)(T/DT + 0.5) in the value type (Sources.cpp:26, Common.hpp:240-243)."""
    return int(np.float32(T) / np.float32(DT) + np.float32(0.5))
