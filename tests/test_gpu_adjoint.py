"""GPU adjoint-state gradient (2D elastic) against the FP64 oracle adjoint and
central finite differences of the FP64 oracle forward model.

Tolerance (north star): gradient relative error <= 1e-3 against FP64, with
FP16 checkpoints (one scale per field and shot) and with FP32 checkpoints.
"""
import os

import numpy as np
import pytest

import wavesim
from wavesim import cases
from oracle import pyoracle as po

from _cases import oracle_params, rel_l2

pytestmark = pytest.mark.gpu

GTOL = 1e-3


def _c4(N=64, NT=300, nshots=2, W=10, **settings):
    c = cases.elastic_c4(N=N, NT=NT, nshots=nshots, W=W)
    c0 = cases.elastic_c4(N=N, NT=NT, nshots=nshots, W=W, anomaly=False)
    c.settings.update(settings)
    c0.settings.update(settings)
    for sh in c.shots:
        sh.rec_type = sh.rec_type.copy()
        sh.rec_type[::5] = 1  # some P receivers
    for sh0, sh in zip(c0.shots, c.shots):
        sh0.rec_type = sh.rec_type
    p = oracle_params(c)
    obs = [po.run(p, c0.m0, c0.m1, c0.m2, sh.src_index, sh.src_type, sh.src_signal, sh.rec_index, sh.rec_type,
                  "f64")[0] for sh in c0.shots]
    return c, obs


def _oracle_gradient(c, obs, resid=None):
    """FP64 adjoint gradient; with `resid` (per shot) the adjoint source is that residual instead of seis - obs."""
    p = oracle_params(c)
    g = np.zeros((3, c.cells))
    J = 0.0
    for k, (sh, o) in enumerate(zip(c.shots, obs)):
        if resid is not None:
            d64, _ = po.run(p, c.m0, c.m1, c.m2, sh.src_index, sh.src_type, sh.src_signal, sh.rec_index, sh.rec_type,
                            "f64")
            o = d64 - resid[k]
        _, j, _, gm = po.adjoint_elastic2d(p, c.m0, c.m1, c.m2, sh.src_index, sh.src_type, sh.src_signal,
                                           sh.rec_index, sh.rec_type, o)
        g += gm
        J += j
    return J, g


def _gpu_gradient(c, obs, every=50, fp16=True, store=0.0, **kw):
    """store: device-memory budget [GiB] for recording every step (0: checkpoints and recomputation)."""
    sol = cases.make_solver(c, **kw)
    sol.set_adjoint_store(store)
    sol.set_checkpointing(every, fp16)
    sol.run(0, c.settings["NT"])
    J = 0.0
    resid = []
    for k, o in enumerate(obs):
        r = sol.seismograms(k).astype(np.float64) - o
        J += 0.5 * np.sum(r * r)
        resid.append(r.astype(np.float32).astype(np.float64))
        sol.set_adjoint_source(k, r.astype(np.float32))
    sol.run_adjoint()
    return J, np.stack([sol.gradient(p) for p in ("vp", "vs", "rho")]), sol, resid


@pytest.mark.parametrize("fp16", [False, True])
def test_gradient_matches_fp64_oracle(fp16):
    """C4-shaped misfit against the unperturbed model.  The adjoint is compared
    with the FP64 adjoint driven by the same adjoint source (the GPU residual):
    the residual here is ~1e-3 of the direct wave, so FP32 forward rounding
    alone moves it by ~1e-3 -- that part is reported, not asserted."""
    c, obs = _c4()
    Jg, gg, _, resid = _gpu_gradient(c, obs, every=50, fp16=fp16)
    Jo, go = _oracle_gradient(c, obs, resid=resid)
    J64, g64 = _oracle_gradient(c, obs)
    print(f"misfit gpu {Jg:.6e}, FP64 forward {J64:.6e} (rel diff {abs(Jg - J64) / J64:.1e})")
    assert abs(Jg - J64) <= 2e-2 * J64
    for k, name in enumerate(("vp", "vs", "rho")):
        e = rel_l2(gg[k], go[k])
        print(f"fp16={fp16} {name}: relL2 vs FP64 adjoint (same adjoint source) {e:.2e}, "
              f"vs all-FP64 pipeline {rel_l2(gg[k], g64[k]):.2e}")
        assert e <= GTOL, (name, e)


@pytest.mark.parametrize("tile_y", [8, 16])
def test_gradient_with_row_strip_recomputation(tile_y):
    """The recomputation runs on the row-strip kernels (they record the derivative sums too)."""
    c, obs = _c4()
    _, gg, _, resid = _gpu_gradient(c, obs, every=40, fp16=True, path="strip", tile_y=tile_y)
    _, go = _oracle_gradient(c, obs, resid=resid)
    for k, name in enumerate(("vp", "vs", "rho")):
        e = rel_l2(gg[k], go[k])
        print(f"strip-{tile_y} {name}: relL2 vs FP64 adjoint {e:.2e}")
        assert e <= GTOL, (name, e)


@pytest.mark.parametrize("tile_y", [8, 16, 32])
def test_gradient_from_recorded_run(tile_y):
    """Every step's derivative sums recorded in FP16 during the forward run (no
    checkpoints, no recomputation), uneven shots per block (3 shots)."""
    c, obs = _c4(N=70, NT=260, nshots=3)
    _, gg, sol, resid = _gpu_gradient(c, obs, every=40, fp16=True, store=-1, path="strip", tile_y=tile_y)
    # the adjoint pass read the record: no recomputation launches
    import ctypes as C
    import wavesim

    L = wavesim.load()
    wavesim._abi.check(L.wv_set_kernel_timing(sol._h, 1))
    sol.run_adjoint()
    sol.synchronize()
    name, avg, n, b = C.c_char_p(), C.c_double(), C.c_longlong(), C.c_double()
    wavesim._abi.check(L.wv_kernel_stats(sol._h, 3, C.byref(name), C.byref(avg), C.byref(n), C.byref(b)))
    assert n.value == 0
    wavesim._abi.check(L.wv_kernel_stats(sol._h, 5, C.byref(name), C.byref(avg), C.byref(n), C.byref(b)))
    assert n.value == c.settings["NT"]
    _, go = _oracle_gradient(c, obs, resid=resid)
    for k, name in enumerate(("vp", "vs", "rho")):
        e = rel_l2(gg[k], go[k])
        print(f"strip-{tile_y} {name}: relL2 vs FP64 adjoint {e:.2e}")
        assert e <= GTOL, (name, e)


def _rect(NX=90, NY=61, NT=260, W=10):
    """NX != NY, ragged against the 64-column strips and the 16-row blocks; two shots."""
    def make(anomaly):
        s = dict(NX=NX, NY=NY, NT=NT, DH=1.0, DT=1e-4, spatial_fd_order=8, damping_boundary=2, boundary_width=W,
                 npower=4.0, fc_cpml=30.0, vmax_cpml=4000.0)
        vp = cases.layered(NY, NX, 10, 1500.0, 4000.0)
        rho = cases.layered(NY, NX, 10, 1800.0, 2600.0)
        if anomaly:
            yy, xx = np.mgrid[0:NY, 0:NX]
            vp = np.where((xx - 52) ** 2 + (yy - 36) ** 2 <= 100, vp * 1.10, vp)
        c = cases.Case("elastic_rect_adjoint", "2D", "elastic", s, vp.astype(np.float32).ravel(),
                       (vp / 1.8).astype(np.float32).ravel(), rho.astype(np.float32).ravel())
        y = W + 4
        xr = np.arange(W + 2, NX - W - 2, 2)
        rec = np.concatenate([xr + y * NX, xr + y * NX])
        rty = np.concatenate([np.full(len(xr), 2), np.full(len(xr), 3)]).astype(np.int32)
        rty[::5] = 1
        w = wavesim.ricker(NT, 1e-4, 30.0, 1.0)
        for sx in (30, 65):
            c.shots.append(cases.Shot(np.array([sx + y * NX]), np.array([3]), w[None, :], rec, rty))
        return c

    c, c0 = make(True), make(False)
    p = oracle_params(c)
    obs = [po.run(p, c0.m0, c0.m1, c0.m2, sh.src_index, sh.src_type, sh.src_signal, sh.rec_index, sh.rec_type,
                  "f64")[0] for sh in c0.shots]
    return c, obs


@pytest.mark.parametrize("store", [-1.0, 0.0])
def test_gradient_rectangular_grid(store):
    """A rectangular grid, from the per-step record (store -1) and from checkpoints + recomputation (store 0)."""
    c, obs = _rect()
    _, gg, _, resid = _gpu_gradient(c, obs, every=40, fp16=True, store=store, path="strip", tile_y=16)
    _, go = _oracle_gradient(c, obs, resid=resid)
    for k, name in enumerate(("vp", "vs", "rho")):
        e = rel_l2(gg[k], go[k])
        print(f"rectangular store={store} {name}: relL2 vs FP64 adjoint {e:.2e}")
        assert e <= GTOL, (name, e)


@pytest.mark.parametrize("order,store", [(2, -1.0), (4, 0.0), (12, -1.0), (12, 0.0)])
def test_gradient_orders(order, store):
    """Adjoint kernels at other stencil orders, from the per-step record (store -1) or checkpoints (store 0)."""
    c, obs = _c4(N=60, NT=240, nshots=2, spatial_fd_order=order)
    _, gg, _, resid = _gpu_gradient(c, obs, every=40, fp16=True, store=store, path="strip", tile_y=16)
    _, go = _oracle_gradient(c, obs, resid=resid)
    for k, name in enumerate(("vp", "vs", "rho")):
        e = rel_l2(gg[k], go[k])
        print(f"order {order} store={store} {name}: relL2 vs FP64 adjoint {e:.2e}")
        assert e <= GTOL, (name, e)


def test_tune_after_checkpointing_keeps_gradient():
    """wv_tune on a handle with checkpointing on (odd interval 25, the handle starting on the
    single-pass kernel): the tuner skips the single-pass candidates, lays the checkpoints
    out again for the variant it keeps, and the forward seismograms and the gradient
    still match the oracle."""
    c, obs = _c4(N=64, NT=250, nshots=2)
    sol = cases.make_solver(c, path="fused", tile_y=16)
    sol.set_adjoint_store(0.0)
    sol.set_checkpointing(25, True)
    choice = sol.tune(20)
    print("kept after tune:", choice)
    assert not choice.startswith("fused")
    sol.run(0, c.settings["NT"])
    resid = []
    for k, o in enumerate(obs):
        seis = sol.seismograms(k)
        o32, _ = po.run(oracle_params(c), c.m0, c.m1, c.m2, c.shots[k].src_index, c.shots[k].src_type,
                        c.shots[k].src_signal, c.shots[k].rec_index, c.shots[k].rec_type, "f32")
        assert rel_l2(seis, o32) <= 1e-5
        r = seis.astype(np.float64) - o
        resid.append(r.astype(np.float32).astype(np.float64))
        sol.set_adjoint_source(k, r.astype(np.float32))
    sol.run_adjoint()
    gg = np.stack([sol.gradient(p) for p in ("vp", "vs", "rho")])
    _, go = _oracle_gradient(c, obs, resid=resid)
    for k, name in enumerate(("vp", "vs", "rho")):
        e = rel_l2(gg[k], go[k])
        print(f"tune after checkpointing {name}: relL2 vs FP64 adjoint {e:.2e}")
        assert e <= GTOL, (name, e)


def test_gradient_uneven_segments_and_stencil_operators():
    """Checkpoint interval not dividing NT, no CPML, stencil-mode operators."""
    c, obs = _c4(N=56, NT=230, nshots=1)
    c.settings["damping_boundary"] = 0
    _, gg, _, resid = _gpu_gradient(c, obs, every=37, fp16=False, ops_mode="stencil")
    p = oracle_params(c, ops_mode="stencil")
    sh = c.shots[0]
    d64, _ = po.run(p, c.m0, c.m1, c.m2, sh.src_index, sh.src_type, sh.src_signal, sh.rec_index, sh.rec_type, "f64")
    _, _, _, go = po.adjoint_elastic2d(p, c.m0, c.m1, c.m2, sh.src_index, sh.src_type, sh.src_signal, sh.rec_index,
                                       sh.rec_type, d64 - resid[0])
    for k in range(3):
        assert rel_l2(gg[k], go[k]) <= GTOL


def test_gradient_central_differences_20_cells():
    """GPU gradient (FP16 checkpoints) of J = 1/2 ||d||^2 against FP64 central
    differences of the oracle at 20 seeded interior cells (seed 40, h = 1e-3
    relative), for vp, vs and rho."""
    c, _ = _c4(N=64, NT=300, nshots=2)
    zero = [np.zeros((len(sh.rec_index), c.settings["NT"])) for sh in c.shots]
    _, gg, _, _ = _gpu_gradient(c, zero, every=50, fp16=True)
    p = oracle_params(c)
    N = c.settings["NX"]
    rng = np.random.Generator(np.random.PCG64(40))
    cells = [int(y * N + x) for x, y in rng.integers(12, N - 12, (20, 2))]
    base = [c.m0.astype(np.float64), c.m1.astype(np.float64), c.m2.astype(np.float64)]

    def J(m):
        tot = 0.0
        for sh in c.shots:
            s, _ = po.run(p, m[0], m[1], m[2], sh.src_index, sh.src_type, sh.src_signal, sh.rec_index,
                          sh.rec_type, "f64")
            tot += 0.5 * np.sum(s * s)
        return tot

    worst = 0.0
    for pi in range(3):
        scale = np.abs(gg[pi]).max()
        for i in cells:
            m = [b.copy() for b in base]
            h = 1e-3 * m[pi][i]
            m[pi][i] += h
            jp = J(m)
            m[pi][i] -= 2 * h
            jm = J(m)
            fd = (jp - jm) / (2 * h)
            err = abs(gg[pi][i] - fd) / scale
            worst = max(worst, err)
            assert err <= 1e-3, (pi, i, gg[pi][i], fd)
    print(f"worst |gpu - fd| / max|gpu| over 20 cells x 3 parameters: {worst:.2e}")


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    from wavesim.autograd import elastic_seismograms

    c, obs = _c4(N=48, NT=200, nshots=2, W=8)
    sol = cases.make_solver(c)
    shape = (c.settings["NY"], c.settings["NX"])
    vp = torch.tensor(c.m0.reshape(shape), requires_grad=True)
    vs = torch.tensor(c.m1.reshape(shape), requires_grad=True)
    rho = torch.tensor(c.m2.reshape(shape), requires_grad=True)
    d = elastic_seismograms(vp, vs, rho, sol, every=40, fp16=True)
    target = torch.from_numpy(np.concatenate(obs).astype(np.float32))
    loss = 0.5 * ((d.double() - target.double()) ** 2).sum()
    loss.backward()
    r = (d.detach().double() - target.double()).float().double().numpy()
    off, resid = 0, []
    for sh in c.shots:
        resid.append(r[off:off + len(sh.rec_index)])
        off += len(sh.rec_index)
    _, go = _oracle_gradient(c, obs, resid=resid)
    for t, g in zip((vp, vs, rho), go):
        assert t.grad.shape == t.shape
        assert rel_l2(t.grad.numpy().ravel(), g) <= GTOL


@pytest.mark.parametrize("store", [0.0, -1.0])
def test_adjoint_cost_ratio(store):
    """Time of (checkpointed or recording forward run + adjoint) over the forward run, printed."""
    c, obs = _c4(N=200, NT=400, nshots=4, W=20)
    sol = cases.make_solver(c, steps_per_graph=50, path="strip", tile_y=16)
    sol.run(0, 400)
    sol.synchronize()
    sol.reset()
    sol.run(0, 400)
    t_fwd = sol.last_run_ms()
    _, _, sadj, _ = _gpu_gradient(c, obs, every=50, fp16=True, store=store, path="strip", tile_y=16)
    t_adj = sadj.last_run_ms()
    sadj.reset()
    sadj.run(0, 400)
    t_ck = sadj.last_run_ms()
    ratio = (t_ck + t_adj) / t_fwd
    print(f"store={store}: forward {t_fwd:.2f} ms, forward with checkpoints/record {t_ck:.2f} ms, "
          f"adjoint {t_adj:.2f} ms, ratio {ratio:.2f}")
    assert ratio < 20
