"""3D acoustic (ForwardSolver3Dacoustic.cpp:95-229) on the GPU, through the C ABI.

* The reference's CI fixture par/ci/seismogram.3D.acoustic.ref.p.mtx (order 4, ABS,
  two-layer model, P source / P receiver) with zero-padded operators, which the
  fixture was made with (test_oracle_golden.py: its configuration's sparse mode
  misses it), at the reference's misfit bound.
* A layered, perturbed model (72 x 60 x 52: no two axes equal, ragged against the
  64-column and z tiles), two batched shots (P and VY sources) with P/VX/VY/VZ
  receivers, against the oracle's float build (relative L2 <= 1e-5) and FP64 run:
  CPML at orders 2, 4, 8, 12; ABS with the free surface; every z-tile height;
  batched equals single-shot; final wavefields.
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


def test_acoustic3d_golden_seismogram():
    case = po.ci_case(os.path.join(CI, "configuration_ci.3D.acoustic.txt"))
    p = case["params"]
    sol = wavesim.create("3D", "acoustic")
    sol.init_forward_solver(NX=p.NX, NY=p.NY, NZ=p.NZ, NT=p.NT, DH=p.DH, DT=p.DT, spatial_fd_order=p.order,
                            ops_mode="stencil", free_surface=p.free_surface, damping_boundary=p.damping,
                            boundary_width=p.width, damping_coeff=p.damping_coeff, npower=p.npower,
                            fc_cpml=p.fc_cpml, vmax_cpml=p.vmax_cpml, steps_per_graph=50)
    sol.prepare_for_modelling(case["vp"], None, case["rho"])
    sig = np.stack([wavesim.ricker(p.NT, p.DT, s["fc"], s["amp"], s["tshift"]) for s in case["sources"]])
    sol.set_acquisition(0, case["src_idx"], case["src_type"], sig, case["rec_idx"], case["rec_type"])
    sol.reset()
    sol.run(0, p.NT)
    seis = sol.seismograms(0)
    ref = po.read_mtx_array(case["ref_base"] + ".p.mtx")
    m = po.ci_misfit(seis, ref)
    print(f"3D acoustic CI fixture: misfit {m:.3e} (bound 5e-7), relL2 {rel_l2(seis, ref):.2e}")
    assert m <= 5e-7
    assert rel_l2(seis, ref) <= 1e-5


def _check(case, **kw):
    sol = cases.make_solver(case, **kw)
    sol.run(0, case.settings["NT"])
    sol.check_finite()
    out = []
    for k, sh in enumerate(case.shots):
        o32, f32 = oracle_shot(case, sh, "f32", want_fields=True)
        o64, _ = oracle_shot(case, sh, "f64")
        g = sol.seismograms(k)
        out.append(g)
        assert np.abs(o32).max() > 0
        # per receiver component: pressure and particle velocities differ by orders of magnitude
        for ty in np.unique(sh.rec_type):
            m = sh.rec_type == ty
            if not np.abs(o32[m]).max() > 0:
                assert np.abs(g[m]).max() == 0, ty  # pressure on the free surface stays 0
                continue
            e32, e64 = rel_l2(g[m], o32[m]), rel_l2(g[m], o64[m])
            print(f"{case.name} order {case.settings['spatial_fd_order']} {kw} shot {k} component {ty}: "
                  f"relL2 vs f32 oracle {e32:.2e}, vs f64 {e64:.2e}", flush=True)
            assert e32 <= TOL, (k, ty, e32)
            assert e64 <= 1e-4, (k, ty, e64)
        for f, name in enumerate(("vx", "vy", "vz", "p")):
            ef = rel_l2(sol.wavefield(name, shot=k), f32[f])
            assert ef <= TOL, (name, ef)
    return out


@pytest.mark.parametrize("order", [2, 4, 8, 12])
def test_acoustic3d_cpml_orders(order):
    _check(cases.acoustic_3d(order=order))


@pytest.mark.parametrize("tile_y", [4, 8, 16])
def test_acoustic3d_abs_free_surface_tiles(tile_y):
    _check(cases.acoustic_3d(damping=1, free_surface=1, W=8), tile_y=tile_y)


def test_acoustic3d_batched_equals_single_and_graphs():
    c = cases.acoustic_3d(nshots=3, NT=200)
    sol = cases.make_solver(c, steps_per_graph=50)
    sol.run(0, 200)
    for k in range(3):
        single = cases.make_solver(c, shots=[c.shots[k]])
        single.run(0, 200)
        assert np.array_equal(single.seismograms(0), sol.seismograms(k)), k
