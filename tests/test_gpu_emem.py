"""2D TE-EM (ForwardSolver2Demem.cpp:108-171, the "emem" type) on the GPU, through the C ABI.

No reference fixture covers EM; the oracle's TE loop is pinned by the line-source twin with the
TM loop (test_oracle_invariants.py).  Here the GPU against the oracle's float build (relative L2
<= 1e-5) and FP64 run on a blocky epsilon / sigma / mu model (170 x 97, ragged against the
64-column tiles), three shots with EX, EY and HZ sources and receivers of all three components:
CPML at orders 2, 4, 8, 12 (final wavefields at order 8), ABS at every tile height,
batched == single-shot, graph segments, and the reference's rejections (no free surface, no EZ).
"""
import numpy as np
import pytest

import wavesim
from wavesim import cases

from _cases import oracle_shot, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _check(case, fields=False, **kw):
    sol = cases.make_solver(case, **kw)
    sol.run(0, case.settings["NT"])
    sol.check_finite()
    for k, sh in enumerate(case.shots):
        o32, f32 = oracle_shot(case, sh, "f32", want_fields=fields)
        o64, _ = oracle_shot(case, sh, "f64")
        g = sol.seismograms(k)
        e32, e64 = rel_l2(g, o32), rel_l2(g, o64)
        print(f"{case.name} order {case.settings['spatial_fd_order']} damping {case.settings['damping_boundary']} "
              f"{kw} shot {k}: relL2 vs f32 oracle {e32:.2e}, vs f64 {e64:.2e}", flush=True)
        assert np.abs(o32).max() > 0
        assert e32 <= TOL, (k, e32)
        assert e64 <= 1e-4, (k, e64)
        if fields and k == len(case.shots) - 1:
            for f, name in enumerate(("hz", "ex", "ey")):
                ef = rel_l2(sol.wavefield(name, shot=k), f32[f])
                assert ef <= TOL, (name, ef)


@pytest.mark.parametrize("order", [2, 4, 8, 12])
def test_emem_cpml_orders(order):
    _check(cases.emem_2d(order=order), fields=order == 8)


@pytest.mark.parametrize("tile_y", [8, 16, 32])
def test_emem_abs_tiles(tile_y):
    _check(cases.emem_2d(damping=1, W=12), tile_y=tile_y)


def test_emem_batched_equals_single_and_graphs():
    c = cases.emem_2d(NT=400)
    batched = cases.make_solver(c)
    batched.run(0, 400)
    graphed = cases.make_solver(c, steps_per_graph=50)
    graphed.run(0, 400)
    for k, sh in enumerate(c.shots):
        one = cases.make_solver(c, shots=[sh])
        one.run(0, 400)
        assert np.array_equal(one.seismograms(0), batched.seismograms(k))
        assert np.array_equal(graphed.seismograms(k), batched.seismograms(k))


def test_emem_rejections():
    c = cases.emem_2d(nshots=1, NT=50)
    fs = dict(c.settings, free_surface=1)
    sol = wavesim.create("2D", "emem")
    with pytest.raises(wavesim.WaveError, match="FreeSurface"):
        sol.init_forward_solver(nshots=1, **fs)
    pl = wavesim.create("2D", "emem")
    with pytest.raises(wavesim.WaveError, match="plain path"):
        pl.init_forward_solver(nshots=1, path="plain", **c.settings)
    ok = cases.make_solver(c)
    sh = c.shots[0]
    with pytest.raises(wavesim.WaveError, match="unsupported source type"):
        ok.set_acquisition(0, sh.src_index, np.ones(len(sh.src_index), np.int32), sh.src_signal, sh.rec_index,
                           sh.rec_type)
    with pytest.raises(wavesim.WaveError, match="not provided"):
        wavesim.create("3D", "tmem")  # the reference factory has no 3D TM solver
