"""2D visco TM-EM (ForwardSolver2Dviscotmem.cpp:77-200) and visco TE-EM
(ForwardSolver2Dviscoemem.cpp:62-220) on the GPU, through the C ABI.

No reference fixture exists for either; the oracle is pinned by the Debye model's limits
(test_oracle_invariants.py).  The GPU runs blocky static epsilon_r / sigma / mu_r models (170 x 97,
ragged against the 64-column tiles) with the two tau per block, three batched shots with uneven
receiver counts, against the oracle's float build (relative L2 <= 1e-5) and FP64 run: CPML at
orders 2, 4, 8, 12; one to three mechanisms; ABS at every tile height; batched == single-shot;
graph segments; the final wavefields.
"""
import numpy as np
import pytest

import wavesim
from wavesim import cases

from _cases import oracle_shot, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5
FIELDS = {"viscotmem": ("hx", "hy", "ez"), "viscoemem": ("hz", "ex", "ey")}


def _check(case, fields=False, **kw):
    sol = cases.make_solver(case, **kw)
    sol.run(0, case.settings["NT"])
    sol.check_finite()
    out = []
    for k, sh in enumerate(case.shots):
        o32, f32 = oracle_shot(case, sh, "f32", want_fields=fields)
        o64, _ = oracle_shot(case, sh, "f64")
        g = sol.seismograms(k)
        e32, e64 = rel_l2(g, o32), rel_l2(g, o64)
        print(f"{case.name} order {case.settings['spatial_fd_order']} mechanisms {len(case.visco['freqs'])} {kw} "
              f"shot {k}: relL2 vs f32 oracle {e32:.2e}, vs f64 {e64:.2e}", flush=True)
        assert np.abs(o32).max() > 0
        assert e32 <= TOL, (k, e32)
        assert e64 <= 1e-4, (k, e64)
        if fields and k == len(case.shots) - 1:
            for f, name in enumerate(FIELDS[case.type]):
                ef = rel_l2(sol.wavefield(name, shot=k), f32[f])
                assert ef <= TOL, (name, ef)
        out.append(g)
    return out


@pytest.mark.parametrize("kind", ["viscotmem", "viscoemem"])
@pytest.mark.parametrize("order", [2, 4, 8, 12])
def test_viscoem_cpml_orders(kind, order):
    _check(getattr(cases, kind + "_2d")(order=order), fields=order == 8)


@pytest.mark.parametrize("kind", ["viscotmem", "viscoemem"])
@pytest.mark.parametrize("freqs", [(1e8, 6e8), (5e7, 3e8, 1.2e9)])
def test_viscoem_several_mechanisms(kind, freqs):
    _check(getattr(cases, kind + "_2d")(freqs=freqs))


@pytest.mark.parametrize("kind", ["viscotmem", "viscoemem"])
@pytest.mark.parametrize("tile_y", [8, 16, 32])
def test_viscoem_abs_tiles(kind, tile_y):
    _check(getattr(cases, kind + "_2d")(damping=1, W=10, freqs=(2e8, 8e8)), tile_y=tile_y)


@pytest.mark.parametrize("kind", ["viscotmem", "viscoemem"])
def test_viscoem_batched_equals_single_and_graphs(kind):
    c = getattr(cases, kind + "_2d")(NT=400, freqs=(2e8, 7e8))
    batched = cases.make_solver(c)
    batched.run(0, 400)
    graphed = cases.make_solver(c, steps_per_graph=50)
    graphed.run(0, 400)
    for k, sh in enumerate(c.shots):
        one = cases.make_solver(c, shots=[sh])
        one.run(0, 400)
        assert np.array_equal(one.seismograms(0), batched.seismograms(k)), k
        assert np.array_equal(graphed.seismograms(k), batched.seismograms(k)), k


def test_viscoem_rejections():
    c = cases.viscotmem_2d(nshots=1, NT=50)
    sol = wavesim.create("2D", "viscotmem")
    sol.init_forward_solver(nshots=1, **c.settings)
    with pytest.raises(wavesim.WaveError, match="wv_set_relaxation first"):
        sol.prepare_for_modelling(c.m0, c.m1, c.m2)
    with pytest.raises(wavesim.WaveError, match="relaxation mechanisms"):
        sol.set_relaxation([1e8] * 5, 0.1, 1e-10)
    with pytest.raises(wavesim.WaveError):
        wavesim.create("3D", "viscotmem")
    pl = wavesim.create("2D", "viscoemem")
    with pytest.raises(wavesim.WaveError, match="plain path"):
        pl.init_forward_solver(nshots=1, path="plain", **c.settings)
