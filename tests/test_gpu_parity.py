"""GPU parity: the HIP kernels (through the C ABI) against the oracle and the golden files.

Tolerances (north star): relative L2 seismogram error <= 1e-5 against the
oracle's float build (the reference float restatement); the error against the
FP64 oracle is asserted below a looser bound and printed.  Integer/index work
(which receiver gets which trace, batching) must be bit-exact.
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


def _gpu_all(case, **kw):
    sol = cases.make_solver(case, **kw)
    sol.run(0, case.settings["NT"])
    return [sol.seismograms(k) for k in range(len(case.shots if kw.get("shots") is None else kw["shots"]))]


def _check_case(case, ops_mode="sparse", shots=None, tol=TOL, **kw):
    shots = case.shots if shots is None else shots
    gpu = _gpu_all(case, ops_mode=ops_mode, shots=shots, **kw)
    for k, sh in enumerate(shots):
        o32, _ = oracle_shot(case, sh, "f32", ops_mode)
        o64, _ = oracle_shot(case, sh, "f64", ops_mode)
        e32, e64, ref64 = rel_l2(gpu[k], o32), rel_l2(gpu[k], o64), rel_l2(o32, o64)
        print(f"{case.name} shot {k}: relL2 vs f32 oracle {e32:.2e}, vs f64 {e64:.2e} (oracle f32 vs f64 {ref64:.2e})")
        assert np.abs(o32).max() > 0
        assert e32 <= tol, (k, e32)
        assert e64 <= max(10 * ref64, tol), (k, e64)
    return gpu


# --------------------------------------------------------------------------- golden files
def _ci(name):
    case = po.ci_case(os.path.join(CI, f"configuration_ci.{name}.txt"))
    p = case["params"]
    phys = {0: "acoustic", 1: "elastic", 2: "tmem"}[p.physics]
    sol = wavesim.create("2D", phys)
    sol.init_forward_solver(NX=p.NX, NY=p.NY, NT=p.NT, DH=p.DH, DT=p.DT, spatial_fd_order=p.order,
                            ops_mode="stencil", free_surface=p.free_surface, damping_boundary=p.damping,
                            boundary_width=p.width, damping_coeff=p.damping_coeff, npower=p.npower,
                            fc_cpml=p.fc_cpml, vmax_cpml=p.vmax_cpml)
    sol.prepare_for_modelling(case["vp"], case["vs"] if phys != "acoustic" else None, case["rho"])
    sig = np.stack([wavesim.ricker(p.NT, p.DT, s["fc"], s["amp"], s["tshift"]) for s in case["sources"]])
    sol.set_acquisition(0, case["src_idx"], case["src_type"], sig, case["rec_idx"], case["rec_type"])
    sol.reset()
    sol.run(0, p.NT)
    return sol.seismograms(0), case


@pytest.mark.parametrize("name,comp", [("2D.elastic", "vy"), ("2D.acoustic", "p")])
def test_golden_seismogram(name, comp):
    seis, case = _ci(name)
    ref = po.read_mtx_array(case["ref_base"] + f".{comp}.mtx")
    m = po.ci_misfit(seis, ref)
    print(f"{name}: CI misfit {m:.3e}, relL2 {rel_l2(seis, ref):.2e}")
    assert m <= 5e-7
    assert rel_l2(seis, ref) <= 1e-5


# --------------------------------------------------------------------------- configs (reduced size)
def test_acoustic_cpml():
    _check_case(cases.acoustic_c1(N=120, NT=700))


def test_elastic_cpml_order8():
    _check_case(cases.elastic_c2(N=160, NT=900, W=20))


def test_elastic_free_surface_cpml_sparse():
    c = cases.elastic_c2(N=128, NT=700, W=16)
    c.settings["free_surface"] = 1
    sh = c.shots[0]
    sh.src_index = np.array([64 + 3 * 128])  # source just below the surface
    _check_case(c)


def test_elastic_abs_p_source_receivers():
    c = cases.elastic_c2(N=100, NT=600, W=12)
    c.settings.update(damping_boundary=1, damping_coeff=8.0)
    sh = c.shots[0]
    sh.src_type = np.array([1])  # P source into sxx and syy
    sh.rec_type = sh.rec_type.copy()
    sh.rec_type[::7] = 1        # some pressure receivers: (sxx+syy)/2
    _check_case(c)


@pytest.mark.parametrize("order", [2, 4, 6, 12])
def test_elastic_orders(order):
    c = cases.elastic_c2(N=96, NT=500, W=12)
    c.settings["spatial_fd_order"] = order
    c.settings["DT"] = 5e-5
    c.shots[0].src_signal = wavesim.ricker(500, 5e-5, 30.0, 1.0)[None]
    _check_case(c)


def test_tmem_batched_uneven_receivers():
    c = cases.tmem_c3(NX=200, NY=120, NT=900, nshots=5, block=10, W=12)
    gpu = _check_case(c)
    assert [g.shape[0] for g in gpu] == [len(s.rec_index) for s in c.shots]


def test_batched_equals_single_shot():
    c = cases.elastic_c4(N=120, NT=300, nshots=3, W=15)
    batched = _gpu_all(c)
    for k, sh in enumerate(c.shots):
        single = _gpu_all(c, shots=[sh])[0]
        assert np.array_equal(batched[k], single)


def test_graph_segments_equal_eager():
    c = cases.elastic_c2(N=96, NT=230, W=12)
    eager = _gpu_all(c)[0]
    graph = _gpu_all(c, steps_per_graph=50)[0]
    assert np.array_equal(eager, graph)


def test_run_in_pieces_equals_one_run():
    """run(t) one step at a time, as Simulation.cpp:444-476 calls it, equals run(0, NT)."""
    c = cases.acoustic_c1(N=64, NT=120)
    whole = _gpu_all(c)[0]
    sol = cases.make_solver(c)
    for t in range(120):
        sol.run(t)
    assert np.array_equal(sol.seismograms(0), whole)


def test_plain_path_matches_oracle():
    _check_case(cases.elastic_c2(N=96, NT=400, W=12), path="plain")
    _check_case(cases.tmem_c3(NX=96, NY=64, NT=400, nshots=1, block=8, W=10), path="plain")


def test_final_wavefields_match_oracle():
    c = cases.elastic_c2(N=80, NT=300, W=10)
    sol = cases.make_solver(c)
    sol.run(0, 300)
    _, f32 = oracle_shot(c, c.shots[0], "f32", want_fields=True)
    for k, name in enumerate(("vx", "vy", "sxx", "syy", "sxy")):
        g = sol.wavefield(name)
        assert rel_l2(g, f32[k]) <= TOL, name


def test_invalid_arguments_raise():
    sol = wavesim.create("2D", "elastic")
    with pytest.raises(wavesim.WaveError):
        sol.run(0, 1)  # before init
    with pytest.raises(wavesim.WaveError):
        sol.init_forward_solver(NX=10, NY=10, NT=5, DH=1.0, DT=1e-4, spatial_fd_order=7)
    sol.init_forward_solver(NX=40, NY=40, NT=5, DH=1.0, DT=1e-4, spatial_fd_order=4)
    with pytest.raises(wavesim.WaveError):
        sol.set_acquisition(0, [40 * 40], [3], np.zeros((1, 5)), [1], [3])  # source outside the grid
    with pytest.raises(wavesim.WaveError):
        sol.set_acquisition(0, [10], [4], np.zeros((1, 5)), [1], [3])  # VZ in 2D


@pytest.mark.parametrize("path,tile_y", [("fused", 8), ("fused", 16), ("fused", 24), ("phased", 0), ("phased", 8),
                                         ("phased", 32)])
def test_elastic_kernel_variants(path, tile_y):
    """Single-pass kernel at every tile height and the two-launch kernels, CPML + VY and P sources."""
    c = cases.elastic_c2(N=150, NT=600, W=16)
    sh = c.shots[0]
    sh.src_index = np.array([75 + 75 * 150, 40 + 100 * 150])
    sh.src_type = np.array([3, 1])
    sh.src_signal = np.concatenate([sh.src_signal, 0.5 * sh.src_signal])
    sh.rec_type = sh.rec_type.copy()
    sh.rec_type[::5] = 1
    _check_case(c, path=path, tile_y=tile_y)


def _elastic_rect(NX=170, NY=97, NT=500, W=12):
    """NX != NY, neither a multiple of the 64-column strips or of any tile height."""
    s = dict(NX=NX, NY=NY, NT=NT, DH=1.0, DT=1e-4, spatial_fd_order=8, damping_boundary=2, boundary_width=W,
             npower=4.0, fc_cpml=30.0, vmax_cpml=4000.0)
    vp, vs, rho = cases.elastic_model(NX, NY, 7, 107, sigma=8.0)
    c = cases.Case("elastic_rect", "2D", "elastic", s, vp, vs, rho)
    x = np.arange(W + 2, NX - W - 2, 3)
    rec = np.concatenate([x + (W + 4) * NX, x + (W + 4) * NX, x + (NY - W - 5) * NX])
    rty = np.concatenate([np.full(len(x), 2), np.full(len(x), 3), np.full(len(x), 1)]).astype(np.int32)
    w = wavesim.ricker(NT, 1e-4, 30.0, 1.0)
    c.shots.append(cases.Shot(np.array([60 + 50 * NX, 130 + 30 * NX]), np.array([3, 1]),
                              np.stack([w, 0.5 * w]), rec, rty))
    return c


@pytest.mark.parametrize("path,tile_y", [("strip", 16), ("strip", 32), ("fused", 16), ("phased", 8), ("plain", 0)])
def test_elastic_rectangular_grid(path, tile_y):
    """A rectangular, ragged grid through every kernel family, so an x/y mix-up cannot cancel."""
    _check_case(_elastic_rect(), path=path, tile_y=tile_y)


@pytest.mark.parametrize("path,tile_y", [("fused", 8), ("fused", 32), ("plain", 0)])
def test_acoustic_rectangular_grid(path, tile_y):
    NX, NY, NT = 150, 83, 500
    s = dict(NX=NX, NY=NY, NT=NT, DH=1.0, DT=2e-4, spatial_fd_order=4, damping_boundary=2, boundary_width=12,
             npower=4.0, fc_cpml=50.0, vmax_cpml=2000.0)
    vp = (1800.0 + 4.0 * np.arange(NY))[:, None].repeat(NX, 1).astype(np.float32).ravel()
    rho = np.full(NX * NY, 2000.0, np.float32)
    c = cases.Case("acoustic_rect", "2D", "acoustic", s, vp, None, rho)
    x = np.arange(15, NX - 15, 2)
    rec = np.concatenate([x + 16 * NX, x + 16 * NX, x + 60 * NX])
    rty = np.concatenate([np.full(len(x), 1), np.full(len(x), 2), np.full(len(x), 3)]).astype(np.int32)
    c.shots.append(cases.Shot(np.array([100 + 40 * NX]), np.array([1]), wavesim.ricker(NT, 2e-4, 50.0, 1.0)[None],
                              rec, rty))
    _check_case(c, path=path, tile_y=tile_y)


@pytest.mark.parametrize("order", [2, 4, 12])
def test_single_pass_orders_free_surface(order):
    c = cases.elastic_c2(N=90, NT=400, W=12)
    c.settings.update(spatial_fd_order=order, DT=5e-5, free_surface=1)
    sh = c.shots[0]
    sh.src_index = np.array([45 + 2 * 90])
    sh.src_type = np.array([2])
    sh.src_signal = wavesim.ricker(400, 5e-5, 30.0, 1.0)[None]
    _check_case(c, tile_y=8)


@pytest.mark.parametrize("tile_y", [8, 32])
def test_two_launch_tile_heights_acoustic_tmem(tile_y):
    _check_case(cases.acoustic_c1(N=100, NT=500), tile_y=tile_y)
    _check_case(cases.tmem_c3(NX=150, NY=90, NT=500, nshots=2, block=10, W=10), tile_y=tile_y)


@pytest.mark.parametrize("fs,damping", [(1, 2), (1, 1), (0, 1)])
def test_acoustic_free_surface_and_abs(fs, damping):
    """Acoustic free surface (image operator, pressure zero at the surface) with CPML or ABS, and ABS alone."""
    c = cases.acoustic_c1(N=100, NT=500)
    c.settings.update(free_surface=fs, damping_boundary=damping, damping_coeff=8.0)
    if fs:
        c.shots[0].src_index = np.array([50 + 3 * 100])  # three cells below the surface
    _check_case(c)


@pytest.mark.parametrize("order", [2, 4, 12])
def test_tmem_orders(order):
    c = cases.tmem_c3(NX=150, NY=90, NT=500, nshots=1, block=10, W=10)
    c.settings.update(spatial_fd_order=order)
    _check_case(c)


@pytest.mark.parametrize("order", [8, 12])
def test_acoustic_orders(order):
    c = cases.acoustic_c1(N=100, NT=500)
    c.settings.update(spatial_fd_order=order)
    _check_case(c)


def test_tmem_abs():
    c = cases.tmem_c3(NX=150, NY=90, NT=500, nshots=2, block=10, W=10)
    c.settings.update(damping_boundary=1, damping_coeff=8.0)
    _check_case(c)


def test_single_pass_batched_shots_with_abs():
    c = cases.elastic_c4(N=110, NT=350, nshots=3, W=12)
    c.settings.update(damping_boundary=1, damping_coeff=8.0)
    _check_case(c)


def test_autotuner_keeps_parity():
    c = cases.elastic_c2(N=130, NT=500, W=14)
    sol = cases.make_solver(c)
    choice = sol.tune(20)
    assert choice in sol.TUNE_CHOICES
    sol.run(0, 500)
    o32, _ = oracle_shot(c, c.shots[0], "f32")
    assert rel_l2(sol.seismograms(0), o32) <= TOL


# ---- persistent TM-EM kernel (one cooperative launch per run call, every step of every shot)


@pytest.mark.parametrize("tile_y", [8, 16, 32])
def test_tmem_persistent_equals_two_launch_and_oracle(tile_y):
    """The persistent loop runs the same tile code as the two-launch kernels in another order, so
    its seismograms equal theirs to the bit; and the oracle's at the usual bar."""
    c = cases.tmem_c3(NX=200, NY=120, NT=600, nshots=5, block=10, W=12)
    pers = cases.make_solver(c, path="persistent", tile_y=tile_y)
    pers.run(0, 300)
    pers.run(300, 600)  # two cooperative launches: the step counter carries over
    two = cases.make_solver(c, path="phased", tile_y=tile_y, steps_per_graph=50)
    two.run(0, 600)
    for k, sh in enumerate(c.shots):
        assert np.array_equal(pers.seismograms(k), two.seismograms(k)), k
        o32, _ = oracle_shot(c, sh, "f32")
        assert rel_l2(pers.seismograms(k), o32) <= TOL, k


def test_tmem_persistent_rejections():
    c = cases.tmem_c3(NX=120, NY=80, NT=50, nshots=1, block=10, W=10)
    s = dict(c.settings, damping_boundary=1, damping_coeff=8.0)
    sol = wavesim.create("2D", "tmem")
    with pytest.raises(wavesim.WaveError, match="persistent"):
        sol.init_forward_solver(nshots=1, path="persistent", **s)
    el = wavesim.create("2D", "elastic")
    e = cases.elastic_c2(N=96, NT=50, W=12)
    with pytest.raises(wavesim.WaveError, match="persistent"):
        el.init_forward_solver(nshots=1, path="persistent", **e.settings)
