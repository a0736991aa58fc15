"""The fused-loop CPU mode (the CPU baseline's best-effort loop) computes what the
operator-sequence oracle computes: the same per-cell arithmetic in the same order,
so the seismograms are equal to the last bit (CPML and no boundary, orders 4 and 8,
a rectangular grid, VX/VY/P receivers)."""
import numpy as np
import pytest

from oracle import pyoracle as po

from _cases import oracle_params
from wavesim import cases


@pytest.mark.parametrize("damping,order,precision", [(2, 8, "f32"), (0, 8, "f32"), (2, 4, "f64"), (2, 12, "f32")])
def test_fused_loop_equals_operator_sequence(damping, order, precision):
    c = cases.elastic_c2(N=90, NT=220, W=12)
    c.settings.update(NX=90, NY=70, damping_boundary=damping, spatial_fd_order=order)
    vp, vs, rho = (m.reshape(90, 90)[:70].ravel().copy() for m in (c.m0, c.m1, c.m2))
    sh = c.shots[0]
    x = np.arange(20, 70)
    rec = np.concatenate([x + 15 * 90, x + 15 * 90, x + 60 * 90])
    rty = np.concatenate([np.full(50, 2), np.full(50, 3), np.full(50, 1)]).astype(np.int32)
    src = np.array([45 + 35 * 90])
    p = oracle_params(c)
    args = (vp, vs, rho, src, sh.src_type, sh.src_signal, rec, rty)
    ref, _ = po.run(p, *args, precision=precision)
    fused = po.run_fused_elastic2d(p, *args, precision=precision)
    assert np.abs(ref).max() > 0
    assert np.array_equal(fused, ref)


def test_fused_loop_rejects_free_surface():
    c = cases.elastic_c2(N=40, NT=10, W=8)
    p = oracle_params(c)
    p.free_surface = 1
    sh = c.shots[0]
    with pytest.raises(RuntimeError):
        po.run_fused_elastic2d(p, c.m0, c.m1, c.m2, sh.src_index, sh.src_type, sh.src_signal, sh.rec_index,
                               sh.rec_type)
