"""GPU parity of the 3D elastic kernels (ForwardSolver3Delastic.cpp:120-416).

Against the reference's golden 3D seismogram (StencilMatrix order 2, free
surface, ABS) and against the oracle on reduced C5-shaped cases (order 8,
sparse operators with edge order reduction, CPML), with the same tolerance as
the 2D tests: relative L2 <= 1e-5 against the oracle's float build.
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


def _small_c5(**kw):
    args = dict(N=40, NT=160, nshots=1, W=8, nrec_side=8, spacing=2)
    args.update(kw)
    return cases.elastic_c5(**args)


def _check(case, tol=TOL, **kw):
    sol = cases.make_solver(case, **kw)
    sol.run(0, case.settings["NT"])
    out = []
    for k, sh in enumerate(case.shots):
        g = sol.seismograms(k)
        o32, _ = oracle_shot(case, sh, "f32", kw.get("ops_mode", "sparse"))
        o64, _ = oracle_shot(case, sh, "f64", kw.get("ops_mode", "sparse"))
        e32, e64, ref = rel_l2(g, o32), rel_l2(g, o64), rel_l2(o32, o64)
        print(f"{case.name} shot {k}: relL2 vs f32 oracle {e32:.2e}, vs f64 {e64:.2e} (oracle f32 vs f64 {ref:.2e})")
        assert np.abs(o32).max() > 0
        assert e32 <= tol, (k, e32)
        assert e64 <= max(10 * ref, tol), (k, e64)
        out.append(g)
    return sol, out


def test_golden_3d_elastic():
    case = po.ci_case(os.path.join(CI, "configuration_ci.3D.elastic.txt"))
    p = case["params"]
    sol = wavesim.create("3D", "elastic")
    sol.init_forward_solver(NX=p.NX, NY=p.NY, NZ=p.NZ, NT=p.NT, DH=p.DH, DT=p.DT, spatial_fd_order=p.order,
                            ops_mode="stencil", free_surface=p.free_surface, damping_boundary=p.damping,
                            boundary_width=p.width, damping_coeff=p.damping_coeff, npower=p.npower,
                            fc_cpml=p.fc_cpml, vmax_cpml=p.vmax_cpml)
    sol.prepare_for_modelling(case["vp"], case["vs"], case["rho"])
    sig = np.stack([wavesim.ricker(p.NT, p.DT, s["fc"], s["amp"], s["tshift"]) for s in case["sources"]])
    sol.set_acquisition(0, case["src_idx"], case["src_type"], sig, case["rec_idx"], case["rec_type"])
    sol.reset()
    sol.run(0, p.NT)
    seis = sol.seismograms(0)
    ref = po.read_mtx_array(case["ref_base"] + ".vy.mtx")
    m = po.ci_misfit(seis, ref)
    print(f"3D.elastic: CI misfit {m:.3e}, relL2 {rel_l2(seis, ref):.2e}")
    assert m <= 5e-7
    assert rel_l2(seis, ref) <= 1e-5


def test_3d_cpml_order8():
    _check(_small_c5())


@pytest.mark.parametrize("tile_y", [4, 8, 16])
def test_3d_tiles_batched_p_sources(tile_y):
    c = _small_c5(N=36, NT=140, nshots=2, W=6)
    c.settings["spatial_fd_order"] = 4
    sh = c.shots[1]
    sh.src_type = np.array([1])  # P source into sxx, syy, szz
    sh.rec_type = sh.rec_type.copy()
    sh.rec_type[::3] = 1        # P receivers (sxx+syy+szz)/3
    _check(c, tile_y=tile_y)


@pytest.mark.parametrize("order", [2, 6, 12])
def test_3d_orders_free_surface(order):
    c = _small_c5(N=36, NT=140, W=6)
    c.settings.update(spatial_fd_order=order, free_surface=1, DT=1.2e-4)
    c.shots[0].src_index = np.array([18 + 18 * 36 + 2 * 36 * 36])  # two cells below the surface
    c.shots[0].src_type = np.array([4])
    c.shots[0].src_signal = wavesim.ricker(140, 1.2e-4, 15.0, 1.0)[None]
    _check(c)


def _noncubic_case(NX=44, NY=36, NZ=30, NT=160, W=6, DT=2e-4):
    """Three different axis lengths, none a multiple of the x or z tile."""
    s = dict(NX=NX, NY=NY, NZ=NZ, NT=NT, DH=2.0, DT=DT, spatial_fd_order=8, damping_boundary=2, boundary_width=W,
             npower=4.0, fc_cpml=60.0, vmax_cpml=3600.0)
    g = cases.gaussian_field((NY, NZ, NX), 4.0, 21)
    gr = cases.gaussian_field((NY, NZ, NX), 4.0, 22)
    vp = (1800.0 + 40.0 * np.arange(NY))[:, None, None] * (1.0 + 0.03 * g)
    rho = (1900.0 + 15.0 * np.arange(NY))[:, None, None] * (1.0 + 0.03 * gr)
    c = cases.Case("elastic_noncubic", "3D", "elastic", s, vp.astype(np.float32).ravel(),
                   (vp / 1.8).astype(np.float32).ravel(), rho.astype(np.float32).ravel())
    idx = lambda x, y, z: x + z * NX + y * NX * NZ  # noqa: E731
    w = wavesim.ricker(NT, DT, 60.0, 1.0)
    rec = np.array([idx(x, W + 2, z) for x in range(W + 1, NX - W - 1, 5) for z in range(W + 1, NZ - W - 1, 5)])
    rec = np.concatenate([rec, rec, rec, rec])
    rty = np.repeat(np.array([1, 2, 3, 4], np.int32), len(rec) // 4)
    c.shots.append(cases.Shot(np.array([idx(13, 10, 17)]), np.array([3]), w[None, :], rec, rty))
    c.shots.append(cases.Shot(np.array([idx(30, 20, 9)]), np.array([1]), w[None, :], rec, rty))
    return c


@pytest.mark.parametrize("tile_y", [4, 16])
def test_3d_noncubic_ragged_grid(tile_y):
    """NX, NY, NZ all different, so a stride or axis mix-up cannot cancel."""
    _check(_noncubic_case(), tile_y=tile_y)


def test_3d_abs():
    c = _small_c5(N=36, NT=140, W=6)
    c.settings.update(damping_boundary=1, damping_coeff=8.0)
    _check(c)


def test_3d_graph_batched_single_bit_exact():
    c = _small_c5(N=32, NT=90, nshots=2, W=6)
    eager = cases.make_solver(c)
    eager.run(0, 90)
    graph = cases.make_solver(c, steps_per_graph=20)
    graph.run(0, 90)
    for k, sh in enumerate(c.shots):
        assert np.array_equal(eager.seismograms(k), graph.seismograms(k))
        single = cases.make_solver(c, shots=[sh])
        single.run(0, 90)
        assert np.array_equal(eager.seismograms(k), single.seismograms(0))


def test_3d_final_wavefields():
    c = _small_c5(N=32, NT=100, W=6)
    sol = cases.make_solver(c)
    sol.run(0, 100)
    _, f32 = oracle_shot(c, c.shots[0], "f32", want_fields=True)
    for k, name in enumerate(sol.FIELDS[("3D", "elastic")]):
        assert rel_l2(sol.wavefield(name), f32[k]) <= TOL, name


def test_3d_tuner():
    c = _small_c5(N=32, NT=100, W=6)
    sol = cases.make_solver(c)
    assert sol.tune(10) in ("3d-z4", "3d-z8", "3d-z16")
    sol.run(0, 100)
    o32, _ = oracle_shot(c, c.shots[0], "f32")
    assert rel_l2(sol.seismograms(0), o32) <= TOL
