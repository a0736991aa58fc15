"""Synthetic source signals, shapes 1..9 of Sources::generateSyntheticSignal
(src/Acquisition/Sources.cpp:157-220, src/Acquisition/SourceSignal/*.cpp), through
wv_source_signal (host code of libwavehip.so; no GPU needed).

Each shape is checked by what its mathematics fixes -- support, zeros, symmetry,
extreme values, integrals, derivative relations -- not by re-evaluating the formula:
  1 Ricker        equals the closed form already pinned by test_oracle_math (oracle Ricker)
  2 SinW          sin t - sin(2t)/2 on one period after Tshift: zero elsewhere, zero mean,
                  maximum 3 sqrt(3)/4 at t = 2 pi / 3
  3 SinThree      sin^3 on one half period: non-negative, peak 1 at the middle, integral 4/(3 pi fc)
  4 FGaussian     odd about 1.2/fc + Tshift, extrema +-sqrt(2/e) at +-1/(sqrt(2) pi fc)
  5 Spike         one sample AMP at floor(Tshift/DT)
  6 IntgSinThree  on its window a running integral of SinThree (difference quotient = SinThree),
                  from 0 to 1/(0.75 pi fc); the reference adds its constant 0.5/(0.75 pi fc) outside too
  7 Ricker_GprMax odd about 1/fc + Tshift, extrema at tau = +-1/sqrt(2 zeta)
  8 Berlage       0 before 1/fc + Tshift, largest magnitude AMP
  9 Sin           AMP sin(2 pi fc (t - Tshift)): zero crossings every 1/(2 fc)
"""
import numpy as np
import pytest

import wavesim
from oracle import pyoracle as po


def sig(shape, NT=4000, DT=1e-4, fc=25.0, amp=2.0, ts=0.01):
    return wavesim.source_signal(shape, NT, DT, fc, amp, ts)


def test_ricker_equals_oracle_closed_form():
    s = sig(1)
    ref = po.ricker(4000, 1e-4, 25.0, 2.0, 0.01)
    assert np.allclose(s, ref, rtol=0, atol=1e-6)


def test_sinw_one_period():
    DT, fc, amp, ts = 1e-4, 25.0, 2.0, 0.01
    s = sig(2, DT=DT, fc=fc, amp=amp, ts=ts)
    i1, n = int(np.floor(ts / DT)), int(np.floor(1 / fc / DT))
    assert np.all(s[:i1] == 0) and np.all(s[i1 + n + 1:] == 0)
    w = s[i1:i1 + n + 1]
    assert abs(w.sum()) < 1e-3 * np.abs(w).sum()
    k = int(np.argmax(w))
    assert abs(w[k] - amp * 3 * np.sqrt(3) / 4) < 1e-3 * amp
    assert abs(k - n / 3) <= 1


def test_sinthree_half_period():
    DT, fc, amp, ts = 1e-4, 25.0, 2.0, 0.01
    s = sig(3, DT=DT, fc=fc, amp=amp, ts=ts)
    i1, n = int(np.floor(ts / DT)), int(np.floor(1 / fc / DT))
    w = s[i1:i1 + n + 1]
    assert np.all(w >= 0) and np.all(s[:i1] == 0) and np.all(s[i1 + n + 1:] == 0)
    assert abs(w.max() - amp) < 1e-6 * amp and abs(int(np.argmax(w)) - n / 2) <= 1
    assert abs(w.sum() * DT - amp * 4 / (3 * np.pi * fc)) < 1e-3 * amp / fc


def test_fgaussian_odd_with_known_extrema():
    DT, fc, amp, ts = 1e-4, 25.0, 2.0, 0.0
    s = sig(4, DT=DT, fc=fc, amp=amp, ts=ts)
    c = int(round((1.2 / fc + ts) / DT))
    assert abs(s[c]) < 1e-4 * amp
    m = 300
    assert np.allclose(s[c - m:c], -s[c + m:c:-1], atol=1e-5 * amp)
    d = int(round(1 / (np.sqrt(2) * np.pi * fc) / DT))
    assert abs(s[c - d] - amp * np.sqrt(2 / np.e)) < 1e-3 * amp
    assert abs(s[c + d] + amp * np.sqrt(2 / np.e)) < 1e-3 * amp


def test_spike():
    s = sig(5, ts=0.0123)
    assert np.count_nonzero(s) == 1 and s[123] == np.float32(2.0)


def test_intgsinthree_is_the_running_integral():
    DT, fc, amp, ts = 1e-4, 25.0, 2.0, 0.01
    s = sig(6, DT=DT, fc=fc, amp=amp, ts=ts)
    s3 = sig(3, DT=DT, fc=fc, amp=amp, ts=ts)
    i1, n = int(np.floor(ts / DT)), int(np.floor(1 / fc / DT))
    w = s[i1:i1 + n + 1]
    unit = amp / (0.75 * np.pi * fc)
    assert abs(w[0]) < 1e-6 * unit and abs(w[-1] - unit) < 1e-3 * unit
    assert np.all(np.diff(w) >= -1e-7 * unit)
    # d/dt (0.5 - 0.75 cos x + 0.25 cos^3 x) / (0.75 pi fc) = sin^3 x for x = pi fc t: the centred
    # difference quotient of the integral is the SinThree signal itself
    dq = (w[2:] - w[:-2]) / (2 * DT)
    assert np.allclose(dq, s3[i1 + 1:i1 + n], atol=2e-3 * amp)
    assert np.allclose(s[:i1], 0.5 * unit) and np.allclose(s[i1 + n + 1:], 0.5 * unit)


def test_ricker_gprmax_odd_with_known_extrema():
    DT, fc, amp, ts = 1e-5, 25.0, 2.0, 0.0
    s = sig(7, NT=20000, DT=DT, fc=fc, amp=amp, ts=ts)
    c = int(round((1 / fc + ts) / DT))
    assert abs(s[c]) < 1e-3 * np.abs(s).max()
    assert np.allclose(s[c - 2000:c], -s[c + 2000:c:-1], atol=1e-4 * np.abs(s).max())
    zeta = 2 * np.pi ** 2 * fc ** 2
    d = 1 / np.sqrt(2 * zeta)
    assert abs(int(np.argmax(s)) * DT - (c * DT - d)) < 2 * DT
    assert abs(int(np.argmin(s)) * DT - (c * DT + d)) < 2 * DT


def test_berlage_causal_and_normalised():
    DT, fc, amp, ts = 1e-4, 25.0, 2.0, 0.005
    s = sig(8, DT=DT, fc=fc, amp=amp, ts=ts)
    c = int(np.ceil((1 / fc + ts) / DT - 1e-3))
    assert np.all(s[:c] == 0)
    assert abs(np.abs(s).max() - amp) < 1e-6 * amp


def test_sin_zero_crossings():
    DT, fc, amp, ts = 1e-4, 25.0, 2.0, 0.004
    s = sig(9, DT=DT, fc=fc, amp=amp, ts=ts)
    half = 1 / (2 * fc)
    for j in range(5):
        k = int(round((ts + j * half) / DT))
        assert abs(s[k]) < 1e-3 * amp
    assert abs(s.max() - amp) < 1e-4 * amp


def test_unknown_shape_and_bad_arguments():
    with pytest.raises(wavesim.WaveError, match="Unknown wavelet shape"):
        sig(10)
    with pytest.raises(wavesim.WaveError, match="Tshift > Tsimulation"):
        sig(2, NT=100, ts=0.5)
