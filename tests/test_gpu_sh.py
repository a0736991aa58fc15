"""2D SH (Love waves, ForwardSolver2Dsh.cpp:101-190) on the GPU, through the C ABI.

* The reference's CI fixture par/ci/seismogram.2D.sh.ref.vz.mtx (order 2, no
  boundary, VZ source and receiver) from its own configuration -- sparse operators,
  as configured -- at the reference's misfit bound, and the GPU against the oracle.
* A layered, perturbed model with three batched shots of uneven receiver counts
  (180 x 120: ragged against the 64-column tiles), against the oracle's float build
  (relative L2 <= 1e-5) and FP64 run: CPML at orders 2, 4, 8, 12; ABS with the free
  surface (the image operator for syz); every tile height; final wavefields.
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


def test_sh_golden_seismogram_from_its_configuration():
    case = po.ci_case(os.path.join(CI, "configuration_ci.2D.sh.txt"))
    p = case["params"]
    sol = wavesim.create("2D", "sh")
    sol.init_forward_solver(NX=p.NX, NY=p.NY, NT=p.NT, DH=p.DH, DT=p.DT, spatial_fd_order=p.order,
                            ops_mode="sparse", free_surface=p.free_surface, damping_boundary=p.damping,
                            boundary_width=p.width, damping_coeff=p.damping_coeff, npower=p.npower,
                            fc_cpml=p.fc_cpml, vmax_cpml=p.vmax_cpml)
    sol.prepare_for_modelling(None, case["vs"], case["rho"])
    sig = np.stack([wavesim.ricker(p.NT, p.DT, s["fc"], s["amp"], s["tshift"]) for s in case["sources"]])
    sol.set_acquisition(0, case["src_idx"], case["src_type"], sig, case["rec_idx"], case["rec_type"])
    sol.reset()
    sol.run(0, p.NT)
    seis = sol.seismograms(0)
    ref = po.read_mtx_array(case["ref_base"] + ".vz.mtx")
    m = po.ci_misfit(seis, ref)
    p.nthreads = 8
    o32, _ = po.run(p, case["vp"], case["vs"], case["rho"], case["src_idx"], case["src_type"], sig, case["rec_idx"],
                    case["rec_type"], "f32")
    e = rel_l2(seis, o32)
    print(f"2D SH CI fixture: misfit {m:.3e} (bound 5e-7), relL2 vs fixture {rel_l2(seis, ref):.2e}, "
          f"vs f32 oracle {e:.2e}")
    assert m <= 5e-7
    assert e <= TOL


def _check(case, **kw):
    sol = cases.make_solver(case, **kw)
    sol.run(0, case.settings["NT"])
    sol.check_finite()
    for k, sh in enumerate(case.shots):
        o32, f32 = oracle_shot(case, sh, "f32", want_fields=True)
        o64, _ = oracle_shot(case, sh, "f64")
        g = sol.seismograms(k)
        e32, e64 = rel_l2(g, o32), rel_l2(g, o64)
        print(f"{case.name} {case.settings['spatial_fd_order']} {kw} shot {k} ({len(sh.rec_index)} receivers): "
              f"relL2 vs f32 oracle {e32:.2e}, vs f64 {e64:.2e}", flush=True)
        assert np.abs(o32).max() > 0
        assert e32 <= TOL, (k, e32)
        assert e64 <= 1e-4, (k, e64)
        if k == len(case.shots) - 1:
            for f, name in enumerate(("vz", "sxz", "syz")):
                ef = rel_l2(sol.wavefield(name, shot=k), f32[f])
                assert ef <= TOL, (name, ef)


@pytest.mark.parametrize("order", [2, 4, 8, 12])
def test_sh_cpml_orders(order):
    _check(cases.sh_2d(order=order))


@pytest.mark.parametrize("tile_y", [8, 16, 32])
def test_sh_abs_free_surface_tiles(tile_y):
    _check(cases.sh_2d(damping=1, free_surface=1, W=12), tile_y=tile_y)


def test_sh_stencil_mode_and_graph_segments():
    c = cases.sh_2d(order=8, NT=500)
    sol = cases.make_solver(c, ops_mode="stencil", steps_per_graph=50)
    sol.run(0, 500)
    for k, sh in enumerate(c.shots):
        o32, _ = oracle_shot(c, sh, "f32", ops_mode="stencil")
        assert rel_l2(sol.seismograms(k), o32) <= TOL


def test_sh_rejects_non_vz_acquisition():
    c = cases.sh_2d(nshots=1)
    sol = cases.make_solver(c)
    with pytest.raises(wavesim.WaveError):
        sol.set_acquisition(0, [10 + 30 * 180], [1], np.zeros((1, c.settings["NT"]), np.float32), [5], [4])
