"""3D elastic adjoint-state gradient on the GPU (no reference counterpart).

Forward run with checkpoints, adjoint pass (recomputation with the recording
forward kernels, transposed stress and velocity steps), gradient w.r.t. vp, vs,
rho and the eight coefficient planes, on a 28 x 24 x 20 grid (no two axes equal,
ragged against the 64-column tiles and the 8-row z tiles) with CPML, two
batched shots with P/VX/VY/VZ receivers, compared with:
  * the oracle's FP64 adjoint (oracle_adjoint_elastic3d, itself pinned by central
    differences in test_oracle_adjoint3d.py) fed the same residual, <= 1e-3 relative;
  * the all-FP64 pipeline (FP64 forward, FP64 residual, FP64 adjoint), <= 1e-3;
  * FP64 central differences of the oracle's misfit on seeded cells.
Each memory plan is exercised: checkpoints on the device with one recomputation
per segment, sub-segments (a budget that fits 10 recorded steps of a 40-step
segment), checkpoints in pinned host memory, and device-sized checkpoints moved to the
host because they leave no room for a sub-segment.
"""
import numpy as np
import pytest

from wavesim import cases
from oracle import pyoracle as po

from _cases import oracle_params, rel_l2

pytestmark = pytest.mark.gpu

NX, NY, NZ, NT, W = 28, 24, 20, 120, 6
GTOL = 1e-3


def _idx(x, y, z):
    return x + z * NX + y * NX * NZ


def _case(order=8, anomaly=True):
    n = NX * NY * NZ
    s = dict(NX=NX, NY=NY, NZ=NZ, NT=NT, DH=1.0, DT=1e-4, spatial_fd_order=order, damping_boundary=2,
             boundary_width=W, npower=4.0, fc_cpml=400.0, vmax_cpml=3000.0)
    vp = np.full(n, 2500.0)
    vs = np.full(n, 2500.0 / 1.8)
    rho = np.full(n, 2100.0)
    if anomaly:
        rng = np.random.default_rng(21)
        vp *= 1 + 0.04 * rng.standard_normal(n)
        vs *= 1 + 0.03 * rng.standard_normal(n)
        rho *= 1 + 0.04 * rng.standard_normal(n)
    c = cases.Case("el3d_adjoint", "3D", "elastic", s, vp.astype(np.float32), vs.astype(np.float32),
                   rho.astype(np.float32))
    w = po.ricker(NT, 1e-4, 400.0, 1.0, precision="f32")
    cells = [_idx(19, 14, 12), _idx(8, 17, 6), _idx(16, 9, 14), _idx(11, 12, 9)]
    rec = np.repeat(np.array(cells), 4)
    rty = np.tile(np.array([1, 2, 3, 4], np.int32), len(cells))
    c.shots.append(cases.Shot(np.array([_idx(10, 9, 8)]), np.array([3]), w[None, :], rec, rty))
    c.shots.append(cases.Shot(np.array([_idx(17, 13, 11), _idx(12, 8, 13)]), np.array([1, 2]),
                              np.stack([w, 0.6 * w]), rec[4:], rty[4:]))
    return c


def _oracle_obs(c0, p):
    return [po.run(p, c0.m0, c0.m1, c0.m2, sh.src_index, sh.src_type, sh.src_signal, sh.rec_index, sh.rec_type,
                   "f64")[0] for sh in c0.shots]


def _sizes(c, nshots, fp16):
    """Bytes of the memory plan's pieces (the layout of elastic3d.hpp: x rows padded to 64 + 2 x 32
    floats, z rows to 16 + 2H, y planes + 2H)."""
    H = c.settings["spatial_fd_order"] // 2
    pitch = -(-NX // 64) * 64 + 64
    zrows = H + -(-NZ // 16) * 16 + H
    vol = (NY + 2 * H) * zrows * pitch
    fld = vol * nshots
    psi = 6 * nshots * 2 * W * (NY * NZ + NZ * NX + NY * NX)
    ck = 9 * fld * (2 if fp16 else 4) + psi * 4
    return dict(ck=ck, state=(9 * fld + psi) * 4, rec=9 * fld * 4)


def _gpu_gradient(c, obs, every, fp16, budget_gib=-1.0):
    sol = cases.make_solver(c)
    sol.set_adjoint_store(budget_gib)
    sol.set_checkpointing(every, fp16)
    plan = sol.adjoint_plan()
    sol.run(0, NT)
    sol.check_finite()
    res = []
    for k in range(len(c.shots)):
        r = sol.seismograms(k).astype(np.float64) - obs[k]
        res.append(r)
        sol.set_adjoint_source(k, r.astype(np.float32))
    sol.run_adjoint()
    g = {name: sol.gradient(name) for name in sol.GRADIENTS3D}
    return g, res, plan


def _oracle_gradient(c, p, residuals):
    """FP64 oracle adjoint of each shot fed the given residual (obs_eff = FP64 seis - residual), summed."""
    gc, gm = 0.0, 0.0
    for sh, r in zip(c.shots, residuals):
        seis64, _ = po.run(p, c.m0, c.m1, c.m2, sh.src_index, sh.src_type, sh.src_signal, sh.rec_index, sh.rec_type,
                           "f64")
        _, _, gcoef, gmodel = po.adjoint_elastic3d(p, c.m0, c.m1, c.m2, sh.src_index, sh.src_type, sh.src_signal,
                                                   sh.rec_index, sh.rec_type, seis64 - r)
        gc = gc + gcoef
        gm = gm + gmodel
    return gc, gm


def _compare(tag, g, gc, gm):
    names = ("vp", "vs", "rho", "inv_rho_x", "inv_rho_y", "inv_rho_z", "M", "mu", "mu_xy", "mu_xz", "mu_yz")
    ref = [gm[0], gm[1], gm[2]] + [gc[k] for k in range(8)]
    worst = 0.0
    for name, r in zip(names, ref):
        e = rel_l2(g[name], r)
        worst = max(worst, e)
        assert np.abs(r).max() > 0, name
        assert e <= GTOL, (tag, name, e)
    print(f"{tag}: worst relL2 over vp, vs, rho and the 8 coefficient planes {worst:.2e}", flush=True)
    return worst


@pytest.mark.parametrize("order,fp16", [(8, False), (8, True), (4, False)])
def test_adjoint3d_matches_fp64_oracle_adjoint(order, fp16):
    c = _case(order)
    p = oracle_params(c)
    obs = _oracle_obs(_case(order, anomaly=False), p)
    g, res, plan = _gpu_gradient(c, obs, 40, fp16)
    assert plan == (40, 1, False), plan
    gc, gm = _oracle_gradient(c, p, res)
    _compare(f"3D adjoint order {order}, {'FP16' if fp16 else 'FP32'} checkpoints every 40, same residual", g, gc,
             gm)
    # the whole pipeline against FP64 forward + FP64 residual + FP64 adjoint
    gc64, gm64 = 0.0, 0.0
    for sh, o in zip(c.shots, obs):
        _, _, gcoef, gmodel = po.adjoint_elastic3d(p, c.m0, c.m1, c.m2, sh.src_index, sh.src_type, sh.src_signal,
                                                   sh.rec_index, sh.rec_type, o)
        gc64, gm64 = gc64 + gcoef, gm64 + gmodel
    _compare(f"3D adjoint order {order}, full pipeline vs FP64", g, gc64, gm64)


def test_adjoint3d_memory_plans_agree():
    """Sub-segments (10 recorded steps of a 40-step segment) and host-memory checkpoints
    give the same gradient as the one-recomputation plan (FP32 checkpoints: to rounding)."""
    c = _case(8)
    p = oracle_params(c)
    obs = _oracle_obs(_case(8, anomaly=False), p)
    base, _, plan = _gpu_gradient(c, obs, 40, False)
    assert plan == (40, 1, False)
    sz = _sizes(c, 2, False)
    gib = 1 << 30
    nck = -(-NT // 40)
    budget = (nck * sz["ck"] + 10 * sz["rec"] + 4 * sz["state"] + sz["rec"] // 2) / gib
    sub, _, plan = _gpu_gradient(c, obs, 40, False, budget)
    assert plan == (10, 4, False), plan
    nck = -(-NT // 10)
    budget = (sz["ck"] + 10 * sz["rec"] + sz["rec"] // 2) / gib
    host, _, plan = _gpu_gradient(c, obs, 10, False, budget)
    assert plan == (10, 1, True), plan
    # checkpoints that fit half the budget but leave no room for any sub-segment next to
    # them (6 FP32 states against >= 8.6 states of record + sub-segment states) go to the host
    nck = -(-NT // 20)
    budget = 2 * nck * sz["ck"] / gib
    moved, _, plan = _gpu_gradient(c, obs, 20, False, budget)
    assert plan[2] and plan[0] < 20 and plan[1] > 1, plan
    for name in ("vp", "vs", "rho"):
        e1, e2, e3 = (rel_l2(g[name], base[name]) for g in (sub, host, moved))
        print(f"d/d{name}: sub-segments vs one recomputation {e1:.1e}, host checkpoints (every 10) {e2:.1e}, "
              f"moved to the host (every 20, plan {plan}) {e3:.1e}", flush=True)
        assert e1 <= 1e-6 and e2 <= 1e-6 and e3 <= 1e-6, (name, e1, e2, e3)


def test_adjoint3d_against_central_differences():
    """GPU gradient at seeded cells against FP64 central differences of the oracle's misfit
    (relative step 1e-4 of the cell's value)."""
    c = _case(8)
    p = oracle_params(c)
    obs = _oracle_obs(_case(8, anomaly=False), p)
    g, _, _ = _gpu_gradient(c, obs, 40, False)

    def misfit(m):
        J = 0.0
        for sh, o in zip(c.shots, obs):
            s, _ = po.run(p, m[0], m[1], m[2], sh.src_index, sh.src_type, sh.src_signal, sh.rec_index, sh.rec_type,
                          "f64")
            J += 0.5 * np.sum((s - o) ** 2)
        return J

    rng = np.random.default_rng(40)
    cells = [_idx(int(x), int(y), int(z)) for x, y, z in zip(rng.integers(W, NX - W, 4), rng.integers(W, NY - W, 4),
                                                               rng.integers(W, NZ - W, 4))]
    cells += [_idx(3, 12, 10), _idx(14, 2, 9)]  # inside the CPML frame
    m0 = [c.m0.astype(np.float64), c.m1.astype(np.float64), c.m2.astype(np.float64)]
    for k, name in enumerate(("vp", "vs", "rho")):
        scale = np.abs(g[name]).max()
        for i in cells:
            h = 1e-4 * m0[k][i]
            mp = [m.copy() for m in m0]
            mm = [m.copy() for m in m0]
            mp[k][i] += h
            mm[k][i] -= h
            fd = (misfit(mp) - misfit(mm)) / (2 * h)
            err = abs(fd - g[name][i]) / scale
            print(f"d/d{name} at cell {i}: GPU {g[name][i]:.6e}, central difference {fd:.6e}, |diff|/max {err:.1e}")
            assert err <= 1e-3, (name, i, err)
