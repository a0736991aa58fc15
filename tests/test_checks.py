"""Host checks of the shot loop: CheckParameter::checkNumericalArtefactsAndInstabilities
(src/CheckParameter/CheckParameter.hpp:60-246).  Host code only: no GPU needed.

The reference computes h of its stability criterion as 7 / 6, 149 / 120, 2161 / 1680, ...
in integer arithmetic (CheckParameter.hpp:75-92), so h = 1 for every order: the bound
is DH / (sqrt(D) vmax) whatever the order.  These tests pin that, and that a DT which
only the true h would reject passes, as in the reference.
"""
import math

import numpy as np
import pytest

from wavesim import WaveError
from wavesim import cases
from wavesim.solver import check_parameters


def _settings(**kw):
    s = dict(NX=40, NY=30, NZ=1, DH=1.0, DT=1e-4, spatial_fd_order=8)
    s.update(kw)
    return s


def _homog(n, vp, vs, rho=2000.0):
    return np.full(n, vp, np.float32), np.full(n, vs, np.float32), np.full(n, rho, np.float32)


@pytest.mark.parametrize("order", [2, 4, 6, 8, 10, 12])
def test_cfl_bound_is_order_independent(order):
    vp = 3000.0
    bound = 1.0 / (math.sqrt(2.0) * vp)  # DH / (h sqrt(D) vmax), h = 1
    m = _homog(40 * 30, vp, vp / 2)
    ok, _ = check_parameters("2D", "elastic", _settings(DT=bound * 0.999, spatial_fd_order=order), *m, fc_max=10.0)
    assert ok
    with pytest.raises(WaveError, match="Courant-Friedrichs-Lewy"):
        check_parameters("2D", "elastic", _settings(DT=bound * 1.001, spatial_fd_order=order), *m, fc_max=10.0)


def test_cfl_integer_h_quirk_accepts_what_true_h_rejects():
    vp = 3000.0
    true_h8 = 2161.0 / 1680.0
    dt = 1.0 / (true_h8 * math.sqrt(2.0) * vp) * 1.1  # above the true-h bound, below the h = 1 bound
    assert dt < 1.0 / (math.sqrt(2.0) * vp)
    check_parameters("2D", "acoustic", _settings(DT=dt), np.full(1200, vp, np.float32), None,
                     np.full(1200, 2000.0, np.float32), fc_max=10.0)


def test_cfl_3d_uses_sqrt3_and_max_velocity():
    n = 10 * 12 * 14
    vp = np.full(n, 2000.0, np.float32)
    vp[77] = 4000.0  # vmax is the largest cell
    vs, rho = vp / 1.8, np.full(n, 2000.0, np.float32)
    s = _settings(NX=10, NY=12, NZ=14, DH=2.0)
    bound = 2.0 / (math.sqrt(3.0) * 4000.0)
    check_parameters("3D", "elastic", dict(s, DT=bound * 0.999), vp, vs, rho, fc_max=5.0)
    with pytest.raises(WaveError):
        check_parameters("3D", "elastic", dict(s, DT=bound * 1.001), vp, vs, rho, fc_max=5.0)


def test_vp_vs_ratio_assertion():
    n = 1200
    vp, vs, rho = _homog(n, 3000.0, 1500.0)
    vs[5 + 3 * 40] = 2200.0  # vp/vs = 1.36 < sqrt(2)
    with pytest.raises(WaveError, match=r"sqrt\(2.0\) at X,Y,Z =5,3,0"):
        check_parameters("2D", "elastic", _settings(), vp, vs, rho, fc_max=10.0)


def test_dispersion_is_a_warning():
    m = _homog(1200, 3000.0, 1500.0)
    # vmin = 1500 (vs), order 8: N = 6 -> DH must be below 1500 / (12 fc)
    ok, msg = check_parameters("2D", "elastic", _settings(), *m, fc_max=1500.0 / 12.0 * 0.99)
    assert ok and msg == ""
    ok, msg = check_parameters("2D", "elastic", _settings(), *m, fc_max=1500.0 / 12.0 * 1.01)
    assert not ok and "numerical dispersion" in msg
    # acoustic uses vp as vmin
    ok, _ = check_parameters("2D", "acoustic", _settings(), m[0], None, m[2], fc_max=3000.0 / 12.0 * 0.99)
    assert ok


def test_dispersion_ignores_zero_velocities():
    vp, vs, rho = _homog(1200, 3000.0, 1500.0)
    vs[:40] = 0.0  # a fluid row: vs = 0 is skipped by the vmin search
    ok, _ = check_parameters("2D", "elastic", _settings(), vp, vs, rho, fc_max=1500.0 / 12.0 * 0.99)
    assert ok


def test_em_velocity_bound():
    n = 1200
    epsr = np.full(n, 4.0, np.float32)
    mur = np.ones(n, np.float32)
    sig = np.zeros(n, np.float32)
    c = 1.0 / math.sqrt(8.8541878176e-12 * 1.2566370614e-6)
    epsr[100] = 1.0  # fastest cell: vacuum
    bound = 0.02 / (math.sqrt(2.0) * c)
    s = _settings(DH=0.02)
    check_parameters("2D", "tmem", dict(s, DT=bound * 0.999), epsr, sig, mur, fc_max=1e8)
    with pytest.raises(WaveError):
        check_parameters("2D", "tmem", dict(s, DT=bound * 1.001), epsr, sig, mur, fc_max=1e8)


@pytest.mark.parametrize("make", [lambda: cases.acoustic_c1(N=60, NT=10), lambda: cases.elastic_c2(N=60, NT=10, W=10),
                                  lambda: cases.tmem_c3(NX=60, NY=40, NT=10, nshots=1, W=8),
                                  lambda: cases.elastic_c4(N=60, NT=10, nshots=1, W=8)])
def test_benchmark_cases_pass(make):
    c = make()
    fc = {"acoustic": 50.0, "elastic": 30.0, "tmem": 2e8}[c.type]
    ok, msg = check_parameters(c.dimension, c.type, c.settings, c.m0, c.m1, c.m2, fc_max=fc)
    assert ok, msg
