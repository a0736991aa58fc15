"""The oracle's FP64 adjoint-state gradient against central finite differences.

The reference has no adjoint, so the gradient is pinned by the mathematics:
the discrete adjoint of the oracle's time loop must equal the derivative of
the oracle's own FP64 misfit.  Central differences in FP64 with a relative
step of 1e-4 are accurate to ~1e-8 here; a dropped CPML term, a wrong
transpose or a wrong chain-rule factor shows up as an O(1) error.
"""
import os
import sys

import numpy as np
import pytest

from oracle import pyoracle as po

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "wave-simulation_amd"))
from wavesim import cases  # noqa: E402  (case generator only; no GPU)


def _setup(damping=2, stencil=0, N=40, NT=220):
    c = cases.elastic_c4(N=N, NT=NT, nshots=1, W=8)
    c0 = cases.elastic_c4(N=N, NT=NT, nshots=1, W=8, anomaly=False)
    p = po.make_params(dim=2, physics=1, NX=N, NY=N, NT=NT, order=8, stencil=stencil, damping=damping, width=8,
                       DH=1.0, DT=1e-4, npower=4.0, fc_cpml=30.0, vmax_cpml=4000.0, nthreads=min(8, os.cpu_count()))
    sh = c.shots[0]
    sh.rec_type = sh.rec_type.copy()
    sh.rec_type[::4] = 1  # some P receivers
    obs, _ = po.run(p, c0.m0, c0.m1, c0.m2, sh.src_index, sh.src_type, sh.src_signal, sh.rec_index, sh.rec_type,
                    "f64")
    return c, p, sh, obs


def _misfit(p, sh, obs, m):
    s, _ = po.run(p, m[0], m[1], m[2], sh.src_index, sh.src_type, sh.src_signal, sh.rec_index, sh.rec_type, "f64")
    return 0.5 * np.sum((s - obs) ** 2)


@pytest.mark.parametrize("damping,stencil", [(2, 0), (0, 1)])
def test_adjoint_gradient_equals_central_differences(damping, stencil):
    c, p, sh, obs = _setup(damping, stencil)
    seis, J, gcoef, gmod = po.adjoint_elastic2d(p, c.m0, c.m1, c.m2, sh.src_index, sh.src_type, sh.src_signal,
                                                sh.rec_index, sh.rec_type, obs)
    assert J > 0 and np.isclose(J, _misfit(p, sh, obs, [c.m0.astype(np.float64), c.m1.astype(np.float64),
                                                         c.m2.astype(np.float64)]), rtol=1e-12)
    N = p.NX
    rng = np.random.default_rng(7)
    cells = [int(y * N + x) for x, y in rng.integers(10, N - 10, (4, 2))]
    for pi in range(3):
        scale = np.abs(gmod[pi]).max()
        for i in cells:
            m = [c.m0.astype(np.float64).copy(), c.m1.astype(np.float64).copy(), c.m2.astype(np.float64).copy()]
            h = 1e-4 * m[pi][i]
            m[pi][i] += h
            jp = _misfit(p, sh, obs, m)
            m[pi][i] -= 2 * h
            jm = _misfit(p, sh, obs, m)
            fd = (jp - jm) / (2 * h)
            assert abs(fd - gmod[pi][i]) <= 1e-6 * scale + 1e-5 * abs(fd), (pi, i, fd, gmod[pi][i])


def test_adjoint_rejects_free_surface():
    c, p, sh, obs = _setup()
    p.free_surface = 1
    with pytest.raises(RuntimeError):
        po.adjoint_elastic2d(p, c.m0, c.m1, c.m2, sh.src_index, sh.src_type, sh.src_signal, sh.rec_index,
                             sh.rec_type, obs)


@pytest.mark.parametrize("K", [7, 50, 219])
def test_checkpointed_backward_pass_is_bit_identical(K):
    """Large runs recompute each segment from an FP64 checkpoint instead of keeping every
    step's record; the recomputation repeats the same arithmetic, so the gradient is the
    same to the last bit (segments of 7 and 50 steps, and one ragged final segment)."""
    c, p, sh, obs = _setup()
    args = (p, c.m0, c.m1, c.m2, sh.src_index, sh.src_type, sh.src_signal, sh.rec_index, sh.rec_type, obs)
    lib = po.lib("f64")
    try:
        lib.oracle_adjoint_segment(0)
        s0, J0, gc0, gm0 = po.adjoint_elastic2d(*args)
        lib.oracle_adjoint_segment(K)
        s1, J1, gc1, gm1 = po.adjoint_elastic2d(*args)
    finally:
        lib.oracle_adjoint_segment(0)
    assert np.array_equal(s0, s1) and J0 == J1
    assert np.array_equal(gc0, gc1) and np.array_equal(gm0, gm1)
