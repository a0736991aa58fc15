"""GPU parity of the 2D elastic row-strip kernels (path "strip") against the oracle.

Same bar as the other kernels: relative L2 seismogram error <= 1e-5 against
the oracle's float build.  The cases put strips on both sides of the interior
test (grid edges with reduced-order operators, CPML frames, free-surface image
operators, grids that are not a multiple of the block), so both the interior
body (half-weight operators, register columns) and the edge body (table
operators, one row at a time) run.
"""
import numpy as np
import pytest

import wavesim
from wavesim import cases

from _cases import oracle_shot, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _run(case, **kw):
    sol = cases.make_solver(case, path="strip", **kw)
    sol.run(0, case.settings["NT"])
    return sol, [sol.seismograms(k) for k in range(len(case.shots))]


def _check(case, ops_mode="sparse", **kw):
    _, gpu = _run(case, ops_mode=ops_mode, **kw)
    for k, sh in enumerate(case.shots):
        o32, _ = oracle_shot(case, sh, "f32", ops_mode)
        o64, _ = oracle_shot(case, sh, "f64", ops_mode)
        e32, e64, ref = rel_l2(gpu[k], o32), rel_l2(gpu[k], o64), rel_l2(o32, o64)
        print(f"{case.name} strip shot {k}: relL2 vs f32 oracle {e32:.2e}, vs f64 {e64:.2e} (f32 vs f64 {ref:.2e})",
              flush=True)
        assert np.abs(o32).max() > 0
        assert e32 <= TOL, (k, e32)
        assert e64 <= max(10 * ref, TOL), (k, e64)
    return gpu


@pytest.mark.parametrize("tile_y", [8, 16, 32])
def test_strip_cpml_mixed_sources(tile_y):
    c = cases.elastic_c2(N=150, NT=600, W=16)
    sh = c.shots[0]
    sh.src_index = np.array([75 + 75 * 150, 40 + 100 * 150])
    sh.src_type = np.array([3, 1])
    sh.src_signal = np.concatenate([sh.src_signal, 0.5 * sh.src_signal])
    sh.rec_type = sh.rec_type.copy()
    sh.rec_type[::5] = 1
    _check(c, tile_y=tile_y)


@pytest.mark.parametrize("order,ops_mode", [(2, "sparse"), (4, "stencil"), (12, "sparse"), (8, "stencil")])
def test_strip_orders_free_surface_odd_grid(order, ops_mode):
    c = cases.elastic_c2(N=90, NT=400, W=12)
    c.settings.update(spatial_fd_order=order, DT=5e-5, free_surface=1)
    sh = c.shots[0]
    sh.src_index = np.array([45 + 2 * 90, 30 + 50 * 90])
    sh.src_type = np.array([2, 3])
    w = wavesim.ricker(400, 5e-5, 30.0, 1.0)[None]
    sh.src_signal = np.concatenate([w, w])
    _check(c, ops_mode=ops_mode, tile_y=16)


def test_strip_batched_abs_equals_single():
    c = cases.elastic_c4(N=110, NT=350, nshots=3, W=12)
    c.settings.update(damping_boundary=1, damping_coeff=8.0)
    batched = _check(c, tile_y=32)
    for k, sh in enumerate(c.shots):
        one = cases.make_solver(c, path="strip", tile_y=32, shots=[sh])
        one.run(0, c.settings["NT"])
        assert np.array_equal(one.seismograms(0), batched[k])


def test_strip_final_wavefields_and_graphs():
    c = cases.elastic_c2(N=133, NT=240, W=10)
    sol, gpu = _run(c, tile_y=16)
    _, f32 = oracle_shot(c, c.shots[0], "f32", want_fields=True)
    for k, name in enumerate(("vx", "vy", "sxx", "syy", "sxy")):
        assert rel_l2(sol.wavefield(name), f32[k]) <= TOL, name
    _, graphed = _run(c, tile_y=16, steps_per_graph=50)
    assert np.array_equal(graphed[0], gpu[0])


def test_tuner_with_strip_candidates_keeps_parity():
    c = cases.elastic_c2(N=130, NT=400, W=14)
    sol = cases.make_solver(c)
    choice = sol.tune(20)
    print("tuner choice", choice, flush=True)
    sol.run(0, 400)
    o32, _ = oracle_shot(c, c.shots[0], "f32")
    assert rel_l2(sol.seismograms(0), o32) <= TOL
    strip = cases.make_solver(c, path="strip", tile_y=32)
    assert strip.tune(20) in ("strip-8", "strip-16", "strip-32")


# ---- TM-EM row-strip kernels (tm2d_h_strip, tm2d_e_strip): the ForwardSolver2Dtmem.cpp:108-171
# step with the elastic strips' split of the work


@pytest.mark.parametrize("tile_y", [8, 16, 32])
def test_tmem_strip_batched_uneven_receivers(tile_y):
    """Five batched shots of uneven receiver counts on a blocky lossy model (200 x 120: ragged
    against the 64-column blocks and every strip height), CPML, order 8."""
    c = cases.tmem_c3(NX=200, NY=120, NT=900, nshots=5, block=10, W=12)
    gpu = _check(c, tile_y=tile_y)
    assert [g.shape[0] for g in gpu] == [len(s.rec_index) for s in c.shots]


@pytest.mark.parametrize("order", [2, 4, 12])
def test_tmem_strip_orders_and_abs(order):
    c = cases.tmem_c3(NX=150, NY=90, NT=500, nshots=2, block=10, W=10)
    c.settings.update(spatial_fd_order=order, damping_boundary=1 if order == 4 else 2, damping_coeff=8.0)
    _check(c)


def test_tmem_strip_batched_equals_single_graphs_and_fields():
    c = cases.tmem_c3(NX=170, NY=97, NT=400, nshots=3, block=10, W=12)
    sol, batched = _run(c)
    graphed, gb = _run(c, steps_per_graph=50)
    for k, sh in enumerate(c.shots):
        _, one = _run(c, shots=[sh])
        assert np.array_equal(one[0], batched[k]), k
        assert np.array_equal(gb[k], batched[k]), k
    _, f32 = oracle_shot(c, c.shots[2], "f32", want_fields=True)
    for f, name in enumerate(("hx", "hy", "ez")):
        assert rel_l2(sol.wavefield(name, shot=2), f32[f]) <= TOL, name


def test_tuner_with_tmem_strip_candidates_keeps_parity():
    c = cases.tmem_c3(NX=200, NY=120, NT=500, nshots=3, block=10, W=12)
    sol = cases.make_solver(c)
    choice = sol.tune(20)
    assert choice in sol.TUNE_CHOICES
    sol.run(0, 500)
    for k, sh in enumerate(c.shots):
        o32, _ = oracle_shot(c, sh, "f32")
        assert rel_l2(sol.seismograms(k), o32) <= TOL, (choice, k)
