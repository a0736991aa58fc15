"""2D viscoelastic P-SV (ForwardSolver2Dviscoelastic.cpp:120-300) and visco-SH
(ForwardSolver2Dviscosh.cpp:110-190) on the GPU, through the C ABI.

* The reference's CI fixture par/ci/seismogram.2D.visco.ref.vy.mtx from its own configuration
  (order 8, one mechanism at 5 Hz, free surface, ABS, the two-layer model with tauP = tauS = 0.1)
  with the zero-padded operators the fixture matches (see test_oracle_golden.py), at the
  reference's misfit bound, and against the oracle.
* A layered, perturbed model (170 x 97: ragged against the 64-column tiles) with tauP / tauS
  varying by layer, three batched shots of P, VX and VY sources and receivers of all three
  types, against the oracle's float build (relative L2 <= 1e-5) and FP64 run: CPML at orders
  2, 4, 8, 12; one, two and three mechanisms at distinct frequencies; ABS with the free surface
  at every tile height; batched == single-shot; graph segments; the final wavefields.
"""
import os

import numpy as np
import pytest

import wavesim
from wavesim import cases
from oracle import pyoracle as po

from _cases import oracle_shot, rel_l2

pytestmark = pytest.mark.gpu

CI = os.path.join(os.path.dirname(__file__), "golden", "ci")
TOL = 1e-5


def test_visco_golden_seismogram():
    case = po.ci_case(os.path.join(CI, "configuration_ci.2D.visco.txt"))
    p = case["params"]
    sol = wavesim.create("2D", "viscoelastic")
    sol.init_forward_solver(NX=p.NX, NY=p.NY, NT=p.NT, DH=p.DH, DT=p.DT, spatial_fd_order=p.order,
                            ops_mode="stencil", free_surface=p.free_surface, damping_boundary=p.damping,
                            boundary_width=p.width, damping_coeff=p.damping_coeff, npower=p.npower,
                            fc_cpml=p.fc_cpml, vmax_cpml=p.vmax_cpml)
    v = case["visco"]
    sol.set_relaxation(v["freqs"], v["tauP"], v["tauS"])
    sol.prepare_for_modelling(case["vp"], case["vs"], case["rho"])
    sig = np.stack([wavesim.ricker(p.NT, p.DT, s["fc"], s["amp"], s["tshift"]) for s in case["sources"]])
    sol.set_acquisition(0, case["src_idx"], case["src_type"], sig, case["rec_idx"], case["rec_type"])
    sol.reset()
    sol.run(0, p.NT)
    seis = sol.seismograms(0)
    ref = po.read_mtx_array(case["ref_base"] + ".vy.mtx")
    m = po.ci_misfit(seis, ref)
    p.stencil = 1
    p.nthreads = 8
    o32, _ = po.run(p, case["vp"], case["vs"], case["rho"], case["src_idx"], case["src_type"], sig, case["rec_idx"],
                    case["rec_type"], "f32", visco=v)
    e = rel_l2(seis, o32)
    print(f"2D visco CI fixture: misfit {m:.3e} (bound 5e-7), relL2 vs fixture {rel_l2(seis, ref):.2e}, "
          f"vs f32 oracle {e:.2e}", flush=True)
    assert m <= 5e-7
    assert e <= TOL


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
            for f, name in enumerate(("vx", "vy", "sxx", "syy", "sxy")):
                ef = rel_l2(sol.wavefield(name, shot=k), f32[f])
                assert ef <= TOL, (name, ef)
        out.append(g)
    return out


@pytest.mark.parametrize("order", [2, 4, 8, 12])
def test_visco_cpml_orders(order):
    _check(cases.visco_2d(order=order), fields=order == 8)


@pytest.mark.parametrize("freqs", [(12.0, 45.0), (8.0, 30.0, 90.0)])
def test_visco_several_mechanisms(freqs):
    _check(cases.visco_2d(freqs=freqs, tau=0.1))


@pytest.mark.parametrize("tile_y", [8, 16, 32])
def test_visco_abs_free_surface_tiles(tile_y):
    _check(cases.visco_2d(damping=1, free_surface=1, W=12, freqs=(20.0, 60.0)), tile_y=tile_y)


def test_visco_batched_equals_single_and_graphs():
    c = cases.visco_2d(NT=400, freqs=(25.0, 70.0))
    batched = cases.make_solver(c)
    batched.run(0, 400)
    graphed = cases.make_solver(c, steps_per_graph=50)
    graphed.run(0, 400)
    for k, sh in enumerate(c.shots):
        one = cases.make_solver(c, shots=[sh])
        one.run(0, 400)
        assert np.array_equal(one.seismograms(0), batched.seismograms(k))
        assert np.array_equal(graphed.seismograms(k), batched.seismograms(k))


def test_visco_rejections():
    c = cases.visco_2d(nshots=1, NT=50)
    sol = wavesim.create("2D", "viscoelastic")
    sol.init_forward_solver(nshots=1, **c.settings)
    with pytest.raises(wavesim.WaveError, match="wv_set_relaxation first"):
        sol.prepare_for_modelling(c.m0, c.m1, c.m2)
    with pytest.raises(wavesim.WaveError, match="relaxation mechanisms"):
        sol.set_relaxation([10.0] * 5, 0.1, 0.1)
    with pytest.raises(wavesim.WaveError, match="relaxation frequency"):
        sol.set_relaxation([0.0], 0.1, 0.1)
    sol2 = cases.make_solver(c)
    with pytest.raises(wavesim.WaveError, match="adjoint"):
        sol2.set_checkpointing(10)
    el = wavesim.create("2D", "elastic")
    el.init_forward_solver(nshots=1, **c.settings)
    with pytest.raises(wavesim.WaveError, match="viscoelastic solver"):
        el.set_relaxation([10.0], 0.1, 0.1)
    st = wavesim.create("2D", "viscoelastic")
    with pytest.raises(wavesim.WaveError, match="row-strip"):
        st.init_forward_solver(nshots=1, path="strip", **c.settings)


# ---- 2D visco-SH (ForwardSolver2Dviscosh.cpp:110-190): the sh_2d acquisition and model with
# tauS by layer and one to three mechanisms, against the oracle (no reference fixture exists)


def _check_sh(case, **kw):
    sol = cases.make_solver(case, **kw)
    sol.run(0, case.settings["NT"])
    sol.check_finite()
    out = []
    for k, sh in enumerate(case.shots):
        o32, f32 = oracle_shot(case, sh, "f32", want_fields=True)
        o64, _ = oracle_shot(case, sh, "f64")
        g = sol.seismograms(k)
        e32, e64 = rel_l2(g, o32), rel_l2(g, o64)
        print(f"{case.name} order {case.settings['spatial_fd_order']} mechanisms {len(case.visco['freqs'])} {kw} "
              f"shot {k}: relL2 vs f32 oracle {e32:.2e}, vs f64 {e64:.2e}", flush=True)
        assert np.abs(o32).max() > 0
        assert e32 <= TOL, (k, e32)
        assert e64 <= 1e-4, (k, e64)
        if k == len(case.shots) - 1:
            for f, name in enumerate(("vz", "sxz", "syz")):
                assert rel_l2(sol.wavefield(name, shot=k), f32[f]) <= TOL, name
        out.append(g)
    return out


@pytest.mark.parametrize("order,freqs", [(2, (30.0,)), (8, (12.0, 45.0)), (12, (8.0, 30.0, 90.0))])
def test_viscosh_cpml_orders_and_mechanisms(order, freqs):
    _check_sh(cases.viscosh_2d(order=order, freqs=freqs))


@pytest.mark.parametrize("tile_y", [8, 16, 32])
def test_viscosh_abs_free_surface_tiles(tile_y):
    _check_sh(cases.viscosh_2d(damping=1, free_surface=1, W=12, freqs=(20.0, 60.0)), tile_y=tile_y)


def test_viscosh_batched_equals_single():
    c = cases.viscosh_2d(NT=400, freqs=(25.0, 70.0))
    batched = cases.make_solver(c, steps_per_graph=50)
    batched.run(0, 400)
    for k, sh in enumerate(c.shots):
        one = cases.make_solver(c, shots=[sh])
        one.run(0, 400)
        assert np.array_equal(one.seismograms(0), batched.seismograms(k))
